// gpu_info: device properties of every visible GPU, the counterpart of the
// reference's oclInfo.cpp:27-69 (which lists OpenCL platforms and devices).
#include <hip/hip_runtime.h>

#include <iostream>

int main() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        std::cout << "no GPU visible to HIP" << std::endl;
        return 1;
    }
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, d) != hipSuccess) continue;
        std::cout << "device " << d << ": " << p.name << " (" << p.gcnArchName << ")\n"
                  << "  compute units        " << p.multiProcessorCount << "\n"
                  << "  wavefront size       " << p.warpSize << "\n"
                  << "  max work-group size  " << p.maxThreadsPerBlock << "\n"
                  << "  LDS per work-group   " << p.sharedMemPerBlock << " bytes\n"
                  << "  global memory        " << p.totalGlobalMem << " bytes\n"
                  << "  L2 cache             " << p.l2CacheSize << " bytes\n"
                  << "  clock                " << p.clockRate / 1000 << " MHz\n"
                  << "  memory clock         " << p.memoryClockRate / 1000 << " MHz, bus " << p.memoryBusWidth
                  << " bits\n";
    }
    return 0;
}
