// Loads and stores of device memory through a global-address-space pointer.
// A pointer read from memory (a problem table) carries no address space, so
// plain accesses through it compile to flat instructions, whose completion
// the compiler must then wait for together with LDS traffic (every wait
// covers both counters).  Casting at the access lets it emit global
// instructions; for kernel-argument pointers nothing changes.
#pragma once

namespace gat {

template <typename V, typename T>
__device__ __forceinline__ V gload(const T* p) {
    return *(const __attribute__((address_space(1))) V*)p;
}
template <typename V, typename T>
__device__ __forceinline__ V gload_nt(const T* p) {
    return __builtin_nontemporal_load((const __attribute__((address_space(1))) V*)p);
}
template <typename V, typename T>
__device__ __forceinline__ void gstore(T* p, V v) {
    *(__attribute__((address_space(1))) V*)p = v;
}

}  // namespace gat
