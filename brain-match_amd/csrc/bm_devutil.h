/*
 * Small device helpers shared by the kernel translation units.
 */
#ifndef BM_DEVUTIL_H
#define BM_DEVUTIL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define BM_ABSENT INT32_MIN
#define BM_WAVE 64

namespace bmd {

/* Multiplicative hash of a vertex id into 2^bits slots (bits >= 1). */
__device__ __forceinline__ uint32_t slot_hash(int32_t key, uint32_t bits) {
    return ((uint32_t)key * 0x9E3779B1u) >> (32u - bits);
}

/* Exact contribution of a male edge of weight w whose image is a female
 * lookup result fw: min(w, fw) when the female edge exists, else 0. */
__device__ __forceinline__ int32_t contrib(int32_t w, int32_t fw) {
    return fw == BM_ABSENT ? 0 : (w < fw ? w : fw);
}

__device__ __forceinline__ int32_t uni(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint32_t uniu(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ int64_t uni64(int64_t x) {
    uint32_t lo = uniu((uint32_t)x), hi = uniu((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

/* Compiler fence for wave-private LDS: a wave's LDS operations execute in
 * order, so only reordering by the compiler has to be prevented. */
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int64_t wave_sum64(int64_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        int lo = __shfl_xor((int)(uint32_t)x, o, BM_WAVE);
        int hi = __shfl_xor((int)(uint32_t)((uint64_t)x >> 32), o, BM_WAVE);
        x += (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    }
    return x;
}

/* New female index (0..2) of position i under permutation p (1..5), in the
 * numbering of calculateThreeNodeSwapKernel's table
 * (src/graph-re-alignment-gpu-3-swap.cu:167-174):
 * 1 (f1,f3,f2)  2 (f2,f1,f3)  3 (f2,f3,f1)  4 (f3,f1,f2)  5 (f3,f2,f1). */
__host__ __device__ __forceinline__ int perm_of(int p, int i) {
    switch (p) {
        case 1: return i == 0 ? 0 : (i == 1 ? 2 : 1);
        case 2: return i == 0 ? 1 : (i == 1 ? 0 : 2);
        case 3: return i == 0 ? 1 : (i == 1 ? 2 : 0);
        case 4: return i == 0 ? 2 : (i == 1 ? 0 : 1);
        case 5: return i == 0 ? 2 : (i == 1 ? 1 : 0);
        default: return i;
    }
}

/* Entry i (< off[n+1]) of a u-major pair stream with offsets off[1..n+1]:
 * the vertex u whose range [off[u], off[u+1]) holds i. */
__device__ __forceinline__ int32_t umajor_owner(const int64_t* __restrict__ off, int32_t n, int64_t i) {
    int32_t lo = 1, hi = n;   /* smallest u with off[u + 1] > i */
    while (lo < hi) {
        const int32_t m = (lo + hi) >> 1;
        if (off[m + 1] > i) hi = m; else lo = m + 1;
    }
    return lo;
}

template <typename T>
__device__ __forceinline__ T sel3(int j, const T& a, const T& b, const T& c) {
    return j == 0 ? a : (j == 1 ? b : c);
}

}  // namespace bmd

#endif
