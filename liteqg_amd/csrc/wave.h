// wave.h -- wavefront-level helpers shared by the search kernels (gfx950,
// 64-wide wavefronts).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lqg_internal.h"

namespace lqg {

// Orders LDS accesses of the lanes of one wavefront (no workgroup barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Number of set bits of `mask` below the calling lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// 64-bit value of lane l (uniform result).
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), l);
    const uint32_t hi =
        __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v >> 32)), l);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint32_t readlane32(uint32_t v, int l) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), l));
}

// A value the caller knows to be equal on every lane (a query index taken by
// lane 0 and broadcast, say), moved to a scalar register.  The compiler then
// keeps what is derived from it (per-query ef and k, pool sizes, loop bounds)
// scalar too, instead of in vector registers under exec masks.
__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

// A global-memory pointer the caller knows to be equal on every lane, in
// scalar registers and typed as global memory, so that a load from it at a
// per-lane 32-bit index takes the scalar-base addressing form (no 64-bit
// address arithmetic per lane).
template <typename T>
using gptr = const __attribute__((address_space(1))) T*;

template <typename T>
__device__ __forceinline__ gptr<T> uniform_gptr(const void* p) {
    const uint64_t a = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = uniform(static_cast<uint32_t>(a));
    const uint32_t hi = uniform(static_cast<uint32_t>(a >> 32));
    return (gptr<T>)((static_cast<uint64_t>(hi) << 32) | lo);
}

// Sets flag to `value` once every lane's earlier writes are visible to the
// host (system-scope fence on every lane, then one store).
__device__ __forceinline__ void signal_host(uint32_t* flag, uint32_t lane,
                                            uint32_t value = kFlagDone) {
    __threadfence_system();
    if (lane == 0) __hip_atomic_store(flag, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace lqg
