// bgf_common.h -- device helpers shared by the one-level kernels
// (bgf_kernels.hip) and the multi-level kernels (bgf_amr.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "bgf_kernels.h"

namespace bgf {

constexpr int BW = 16;      // cells per block side (src/Param.h:52)
constexpr int MW = 18;      // stored tile side: one halo cell each way
constexpr int BS = MW * MW; // values per block

// src/MemManagement.cu:37-40
__device__ __forceinline__ int memloc(int i, int j, int ib) { return (i + 1) + (j + 1) * MW + ib * BS; }

// Workgroup minimum over 256 threads (four wave64s); every thread gets it.
template <class T> __device__ __forceinline__ T block_min(T v, T* red, int t) {
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    T m = red[0];
#pragma unroll
    for (int k = 1; k < 4; k++) m = red[k] < m ? red[k] : m;
    return m;
}

// Positive floating-point values order like their bit patterns, so a block
// minimum merges into a dt slot with an integer atomic min.
__device__ __forceinline__ void slot_min(DevLoop* L, int k, float v) {
    atomicMin(&L->slot32[k & (NSLOT - 1)], __float_as_uint(v));
}
__device__ __forceinline__ void slot_min(DevLoop* L, int k, double v) {
    atomicMin(&L->slot64[k & (NSLOT - 1)], (unsigned long long)__double_as_longlong(v));
}

// calcres (src/Util_CPU.cu:194-198): cell size of a level.
template <class T> __device__ __forceinline__ T calcres(T dx, int level) {
    return level < 0 ? dx * T(1 << -level) : dx / T(1 << level);
}

}  // namespace bgf
