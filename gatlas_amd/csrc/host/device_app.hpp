// DeviceApp: the HIP counterpart of the reference's OCLApp buffer/queue API
// (OCLApp.hpp:84-330, OCLAppUtil.hpp:63-390).  Buffers are integer handles
// owning a device allocation plus a host shadow; the app owns all of them
// and releaseBuffers() drops every one (OCLApp::releaseBuffers).
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../context.hpp"

namespace gat {

class DeviceApp {
public:
    // Takes an existing context (not owned) or creates one for `device`.
    explicit DeviceApp(int device);
    explicit DeviceApp(Context* shared);
    ~DeviceApp();

    bool ok() const { return ctx_ != nullptr; }
    Context& ctx() { return *ctx_; }
    hipStream_t stream() const { return ctx_->stream; }
    size_t maxWorkGroupSize() const { return (size_t)ctx_->max_workgroup; }

    // createBuffer<T>(n, name, value): host shadow filled with `value`,
    // copied to the device (OCLAppUtil.hpp:335-390 createBufferR/W/RW).
    template <typename T>
    int createBuffer(size_t n, const char* name, T value) {
        int h = alloc(n * sizeof(T), name);
        if (h < 0) return h;
        T* p = (T*)bufs_[h].host.data();
        for (size_t i = 0; i < n; ++i) p[i] = value;
        return syncToDevice(h) ? h : -1;
    }
    int createRaw(size_t bytes, const char* name);  // device + host shadow, contents undefined

    template <typename T>
    T* bufferPtr(int h) {
        return (T*)bufs_[h].host.data();
    }
    void* devicePtr(int h) { return bufs_[h].dev; }
    size_t bytes(int h) const { return bufs_[h].bytes; }

    bool syncToDevice(int h);
    bool syncFromDevice(int h);
    // clearBuffer (OCLAppUtil.hpp:63-72): zero the shadow and the device copy.
    bool clearBuffer(int h);
    void releaseBuffers();
    bool wait();

private:
    struct Buf {
        void* dev = nullptr;
        size_t bytes = 0;
        std::vector<unsigned char> host;
        std::string name;
    };
    int alloc(size_t bytes, const char* name);

    Context* ctx_ = nullptr;
    std::unique_ptr<Context> owned_;
    std::vector<Buf> bufs_;
};

}  // namespace gat
