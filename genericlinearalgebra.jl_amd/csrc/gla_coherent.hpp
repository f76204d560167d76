// Device-coherent memory helpers for kernels whose workgroups exchange data
// within one launch (grid barriers, dataflow flags).  gfx950 has one L2 per
// XCD; agent-scope (sc1) loads and stores go to the device-coherent level,
// so data written by a workgroup on one XCD is seen by a workgroup on
// another without a whole-L2 write-back/invalidate fence.
#pragma once

#include <hip/hip_runtime.h>

namespace gla {

template <typename T>
__device__ __forceinline__ void store_coherent(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T load_coherent(const T *p)
{
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Eight agent-coherent (sc1) loads in flight at once, then one wait.  The
// compiler issues relaxed atomic loads one round trip at a time; this block
// keeps the memory-level parallelism.  Outputs are early-clobber and the
// wait is inside the block, so no register is read before its load lands.
__device__ __forceinline__ void load8_coherent(const double *const p[8], double v[8])
{
    asm volatile(
        "global_load_dwordx2 %0, %8, off sc1\n\t"
        "global_load_dwordx2 %1, %9, off sc1\n\t"
        "global_load_dwordx2 %2, %10, off sc1\n\t"
        "global_load_dwordx2 %3, %11, off sc1\n\t"
        "global_load_dwordx2 %4, %12, off sc1\n\t"
        "global_load_dwordx2 %5, %13, off sc1\n\t"
        "global_load_dwordx2 %6, %14, off sc1\n\t"
        "global_load_dwordx2 %7, %15, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]),
          "=&v"(v[6]), "=&v"(v[7])
        : "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3]), "v"(p[4]), "v"(p[5]), "v"(p[6]), "v"(p[7])
        : "memory");
}
__device__ __forceinline__ void load8_coherent(const float *const p[8], float v[8])
{
    asm volatile(
        "global_load_dword %0, %8, off sc1\n\t"
        "global_load_dword %1, %9, off sc1\n\t"
        "global_load_dword %2, %10, off sc1\n\t"
        "global_load_dword %3, %11, off sc1\n\t"
        "global_load_dword %4, %12, off sc1\n\t"
        "global_load_dword %5, %13, off sc1\n\t"
        "global_load_dword %6, %14, off sc1\n\t"
        "global_load_dword %7, %15, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]),
          "=&v"(v[6]), "=&v"(v[7])
        : "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3]), "v"(p[4]), "v"(p[5]), "v"(p[6]), "v"(p[7])
        : "memory");
}

// load8_coherent plus a ninth agent-coherent load in the same round trip.
__device__ __forceinline__ void load9_coherent(const double *const p[8], double v[8],
                                               const double *p9, double &v9)
{
    asm volatile(
        "global_load_dwordx2 %0, %9, off sc1\n\t"
        "global_load_dwordx2 %1, %10, off sc1\n\t"
        "global_load_dwordx2 %2, %11, off sc1\n\t"
        "global_load_dwordx2 %3, %12, off sc1\n\t"
        "global_load_dwordx2 %4, %13, off sc1\n\t"
        "global_load_dwordx2 %5, %14, off sc1\n\t"
        "global_load_dwordx2 %6, %15, off sc1\n\t"
        "global_load_dwordx2 %7, %16, off sc1\n\t"
        "global_load_dwordx2 %8, %17, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]),
          "=&v"(v[6]), "=&v"(v[7]), "=&v"(v9)
        : "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3]), "v"(p[4]), "v"(p[5]), "v"(p[6]), "v"(p[7]),
          "v"(p9)
        : "memory");
}
__device__ __forceinline__ void load9_coherent(const float *const p[8], float v[8], const float *p9,
                                               float &v9)
{
    asm volatile(
        "global_load_dword %0, %9, off sc1\n\t"
        "global_load_dword %1, %10, off sc1\n\t"
        "global_load_dword %2, %11, off sc1\n\t"
        "global_load_dword %3, %12, off sc1\n\t"
        "global_load_dword %4, %13, off sc1\n\t"
        "global_load_dword %5, %14, off sc1\n\t"
        "global_load_dword %6, %15, off sc1\n\t"
        "global_load_dword %7, %16, off sc1\n\t"
        "global_load_dword %8, %17, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]),
          "=&v"(v[6]), "=&v"(v[7]), "=&v"(v9)
        : "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3]), "v"(p[4]), "v"(p[5]), "v"(p[6]), "v"(p[7]),
          "v"(p9)
        : "memory");
}

// Waits until this thread's outstanding vector memory operations, stores
// included, have completed (s_waitcnt vmcnt(0)).
__device__ __forceinline__ void stores_done() { __builtin_amdgcn_s_waitcnt(0x0F70); }

}  // namespace gla
