/*
 * Greedy initial matching on the device (bm_greedy_match_device): the same
 * matching as bm_greedy_match (bm_greedy.c) from the context's resident
 * graphs, in the manner of greedy-match (src/greedy-match.c:213-530).
 *
 * 1. k_greedy_features: one block per vertex marks an (n+1)-bit LDS bitmap
 *    with the vertices two out-edges away, counts it, then does the same
 *    over two in-edges (compute_2hop_metrics, :213-252); with the in- and
 *    out-degree that is the vertex's four features, and its sort key is
 *    (2^32 - 1 - total) << 32 | id for vertices with an incident edge.
 * 2. A radix sort of the keys orders each graph by total descending, ties to
 *    the smaller id (:489-490 with the tie rule of bm_greedy.c).
 * 3. k_greedy_assign: one 1024-thread block walks the male vertices in that
 *    order; for each, every thread scans its share of the female list for the
 *    unmatched vertex of smallest L1 feature distance, and a block reduction
 *    of (distance << 32 | position) picks the first such vertex in female
 *    order (:495-522).  When the features fit in 16 bits and the female list
 *    in LDS (8 bytes per vertex, n <= 18,900), the list is staged there.
 */
#include "bm_device.h"
#include "bm_devutil.h"
#include "bm_host.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <stdio.h>
#include <string.h>

using namespace bmd;

namespace {

constexpr int kFeatThreads = 256;
constexpr int kAssignThreads = 1024;
constexpr size_t kLdsBudget = 160 * 1024;

__device__ __forceinline__ uint32_t block_sum(uint32_t x, uint32_t* red) {
    const int lane = threadIdx.x & (BM_WAVE - 1), wave = threadIdx.x / BM_WAVE;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, BM_WAVE);
    if (lane == 0) red[wave] = x;
    __syncthreads();
    uint32_t t = 0;
    for (int w = 0; w < (int)(blockDim.x / BM_WAVE); w++) t += red[w];
    __syncthreads();
    return t;
}

/* Features {in, out, reach2, coreach2} of every vertex and its sort key;
 * *present counts the vertices with an incident edge. */
__global__ void __launch_bounds__(kFeatThreads) k_greedy_features(int32_t n, const int64_t* __restrict__ optr,
                                                                  const int2* __restrict__ oadj,
                                                                  const int64_t* __restrict__ iptr,
                                                                  const int2* __restrict__ iadj, int4* __restrict__ feat,
                                                                  uint64_t* __restrict__ keys,
                                                                  uint32_t* __restrict__ present) {
    extern __shared__ uint32_t bits[];
    __shared__ uint32_t red[kFeatThreads / BM_WAVE];
    const int lane = threadIdx.x & (BM_WAVE - 1), wave = threadIdx.x / BM_WAVE;
    const int nwaves = blockDim.x / BM_WAVE;
    const int32_t words = (n + 1 + 31) / 32;
    for (int32_t v = blockIdx.x + 1; v <= n; v += gridDim.x) {
        uint32_t cnt[2];
        for (int pass = 0; pass < 2; pass++) {
            const int64_t* ptr = pass ? iptr : optr;
            const int2* adj = pass ? iadj : oadj;
            for (int32_t w = threadIdx.x; w < words; w += blockDim.x) bits[w] = 0;
            __syncthreads();
            /* waves take the first hop, lanes stride the second */
            for (int64_t a = ptr[v] + wave; a < ptr[v + 1]; a += nwaves) {
                const int32_t x = uni(adj[a].x);
                for (int64_t c = ptr[x] + lane; c < ptr[x + 1]; c += BM_WAVE) {
                    const int32_t y = adj[c].x;
                    atomicOr(&bits[y >> 5], 1u << (y & 31));
                }
            }
            __syncthreads();
            uint32_t local = 0;
            for (int32_t w = threadIdx.x; w < words; w += blockDim.x) local += (uint32_t)__popc(bits[w]);
            cnt[pass] = block_sum(local, red);
        }
        if (threadIdx.x == 0) {
            const int32_t od = (int32_t)(optr[v + 1] - optr[v]), id = (int32_t)(iptr[v + 1] - iptr[v]);
            feat[v] = make_int4(id, od, (int32_t)cnt[0], (int32_t)cnt[1]);
            const uint64_t total = (uint64_t)od + (uint64_t)id + cnt[0] + cnt[1];
            const bool has = od + id > 0;
            keys[v - 1] = has ? ((0xFFFFFFFFull - total) << 32) | (uint64_t)v : ~0ull;
            if (has) atomicAdd(present, 1u);
        }
    }
}

/* Sorted order -> ids and features. */
__global__ void k_greedy_gather(const uint64_t* __restrict__ sorted, int32_t k, const int4* __restrict__ feat,
                                int32_t* __restrict__ ids, int4* __restrict__ sfeat) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
        const int32_t v = (int32_t)(uint32_t)sorted[i];
        ids[i] = v;
        sfeat[i] = feat[v];
    }
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t x, int o) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, o, BM_WAVE);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), o, BM_WAVE);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int32_t absd(int32_t a, int32_t b) { return a > b ? a - b : b - a; }

/* The greedy walk over k male vertices (sorted order); pi[male id] = female
 * id.  STAGED: the female features as four uint16 in LDS. */
template <bool STAGED>
__global__ void __launch_bounds__(kAssignThreads) k_greedy_assign(int32_t k, const int4* __restrict__ mf,
                                                                  const int32_t* __restrict__ mid,
                                                                  const int4* __restrict__ ff,
                                                                  const int32_t* __restrict__ fid,
                                                                  int32_t* __restrict__ pi) {
    extern __shared__ uint4 smem[];
    __shared__ uint64_t red[kAssignThreads / BM_WAVE];
    const int32_t words = (k + 31) / 32;
    uint32_t* used = reinterpret_cast<uint32_t*>(smem);
    ushort4* F = reinterpret_cast<ushort4*>(used + ((words + 1) & ~1));
    const int lane = threadIdx.x & (BM_WAVE - 1), wave = threadIdx.x / BM_WAVE;
    for (int32_t w = threadIdx.x; w < words; w += blockDim.x) used[w] = 0;
    if (STAGED)
        for (int32_t j = threadIdx.x; j < k; j += blockDim.x) {
            const int4 f = ff[j];
            F[j] = make_ushort4((unsigned short)f.x, (unsigned short)f.y, (unsigned short)f.z, (unsigned short)f.w);
        }
    __syncthreads();
    for (int32_t i = 0; i < k; i++) {
        const int4 m = mf[i];
        uint64_t best = ~0ull;
        for (int32_t j = threadIdx.x; j < k; j += blockDim.x) {
            if (used[j >> 5] & (1u << (j & 31))) continue;
            int32_t d;
            if (STAGED) {
                const ushort4 f = F[j];
                d = absd(m.x, f.x) + absd(m.y, f.y) + absd(m.z, f.z) + absd(m.w, f.w);
            } else {
                const int4 f = ff[j];
                d = absd(m.x, f.x) + absd(m.y, f.y) + absd(m.z, f.z) + absd(m.w, f.w);
            }
            const uint64_t key = ((uint64_t)(uint32_t)d << 32) | (uint32_t)j;
            best = key < best ? key : best;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const uint64_t other = shfl_xor64(best, o);
            best = other < best ? other : best;
        }
        if (lane == 0) red[wave] = best;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t b = red[0];
            for (int w = 1; w < kAssignThreads / BM_WAVE; w++) b = red[w] < b ? red[w] : b;
            if (b != ~0ull) {
                const int32_t j = (int32_t)(uint32_t)b;
                used[j >> 5] |= 1u << (j & 31);
                pi[mid[i]] = fid[j];
            }
        }
        __syncthreads();
    }
}

template <typename T>
hipError_t dmalloc(T** p, size_t count) {
    return hipMalloc((void**)p, (count ? count : 1) * sizeof(T));
}

}  // namespace

extern "C" int bm_greedy_match_device(bm_ctx* ctx, int32_t* pi) {
    if (!ctx || !pi) return bm_ctx_set_error(ctx, BM_ERR_ARG, "bm_greedy_match_device: null argument");
    const BmDevGraph g = bm_ctx_devgraph(ctx);
    const int32_t n = g.n;
    const hipStream_t s = bm_ctx_stream(ctx);
    const int32_t words = (n + 1 + 31) / 32;
    const size_t feat_lds = (size_t)words * sizeof(uint32_t);
    if (feat_lds > kLdsBudget)
        return bm_ctx_set_error(ctx, BM_ERR_ARG, "bm_greedy_match_device: n too large for the LDS 2-hop bitmap");
    int4 *feat = nullptr, *mf = nullptr, *ff = nullptr;
    uint64_t *keys = nullptr, *sorted = nullptr;
    int32_t *mid = nullptr, *fid = nullptr, *dpi = nullptr;
    uint32_t* present = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
    hipError_t e = hipSetDevice(bm_ctx_device(ctx));
    int rc = BM_OK;
    char msg[256] = "";
#define GCHK(expr)                                                                                    \
    do {                                                                                              \
        if (e == hipSuccess) {                                                                        \
            e = (expr);                                                                               \
            if (e != hipSuccess) snprintf(msg, sizeof msg, "%s failed: %s", #expr, hipGetErrorString(e)); \
        }                                                                                             \
    } while (0)
    GCHK(dmalloc(&feat, 2 * ((size_t)n + 1)));
    GCHK(dmalloc(&keys, 2 * (size_t)n));
    GCHK(dmalloc(&sorted, 2 * (size_t)n));
    GCHK(dmalloc(&mf, (size_t)n));
    GCHK(dmalloc(&ff, (size_t)n));
    GCHK(dmalloc(&mid, (size_t)n));
    GCHK(dmalloc(&fid, (size_t)n));
    GCHK(dmalloc(&dpi, (size_t)n + 1));
    GCHK(dmalloc(&present, 2));
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortKeys(nullptr, temp_bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr, n);
    GCHK(hipMalloc(&temp, temp_bytes ? temp_bytes : 1));
    GCHK(hipMemsetAsync(present, 0, 2 * sizeof(uint32_t), s));
    GCHK(hipMemsetAsync(dpi, 0, ((size_t)n + 1) * sizeof(int32_t), s));
    const int32_t cus = bm_ctx_num_cus(ctx);
    const int blocks = (int)((int64_t)(cus > 0 ? cus : 256) * 8 < n ? (int64_t)(cus > 0 ? cus : 256) * 8 : (n > 0 ? n : 1));
    GCHK(hipFuncSetAttribute((const void*)k_greedy_features, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)feat_lds));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_greedy_features, dim3(blocks), dim3(kFeatThreads), feat_lds, s, n, g.m_out_ptr, g.m_out,
                           g.m_in_ptr, g.m_in, feat, keys, present);
        hipLaunchKernelGGL(k_greedy_features, dim3(blocks), dim3(kFeatThreads), feat_lds, s, n, g.f_out_ptr, g.f_out,
                           g.f_in_ptr, g.f_in, feat + n + 1, keys + n, present + 1);
        GCHK(hipGetLastError());
    }
    if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortKeys(temp, temp_bytes, keys, sorted, n, 0, 64, s);
    if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortKeys(temp, temp_bytes, keys + n, sorted + n, n, 0, 64, s);
    uint32_t counts[2] = {0, 0};
    GCHK(hipMemcpyAsync(counts, present, sizeof counts, hipMemcpyDeviceToHost, s));
    GCHK(hipStreamSynchronize(s));
    if (e == hipSuccess && counts[0] != counts[1]) {
        snprintf(msg, sizeof msg, "bm_greedy_match_device: graphs have different vertex counts (%u vs %u)", counts[0],
                 counts[1]);
        rc = BM_ERR_ARG;
    }
    const int32_t k = (int32_t)counts[0];
    if (e == hipSuccess && rc == BM_OK && k > 0) {
        hipLaunchKernelGGL(k_greedy_gather, dim3((k + 255) / 256), dim3(256), 0, s, sorted, k, feat, mid, mf);
        hipLaunchKernelGGL(k_greedy_gather, dim3((k + 255) / 256), dim3(256), 0, s, sorted + n, k, feat + n + 1, fid,
                           ff);
        GCHK(hipGetLastError());
        const size_t used_bytes = (size_t)((((k + 31) / 32) + 1) & ~1) * sizeof(uint32_t);
        const size_t staged = used_bytes + (size_t)k * sizeof(ushort4);
        /* features fit in 16 bits when every count is at most n <= 65535 */
        if (n <= 65535 && staged + 1024 <= kLdsBudget) {
            GCHK(hipFuncSetAttribute((const void*)k_greedy_assign<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)staged));
            if (e == hipSuccess)
                hipLaunchKernelGGL(k_greedy_assign<true>, dim3(1), dim3(kAssignThreads), staged, s, k, mf, mid, ff, fid,
                                   dpi);
        } else {
            GCHK(hipFuncSetAttribute((const void*)k_greedy_assign<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)used_bytes));
            if (e == hipSuccess)
                hipLaunchKernelGGL(k_greedy_assign<false>, dim3(1), dim3(kAssignThreads), used_bytes, s, k, mf, mid, ff,
                                   fid, dpi);
        }
        GCHK(hipGetLastError());
    }
    if (rc == BM_OK) {
        GCHK(hipMemcpyAsync(pi, dpi, ((size_t)n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GCHK(hipStreamSynchronize(s));
    }
#undef GCHK
    (void)hipFree(feat);
    (void)hipFree(keys);
    (void)hipFree(sorted);
    (void)hipFree(mf);
    (void)hipFree(ff);
    (void)hipFree(mid);
    (void)hipFree(fid);
    (void)hipFree(dpi);
    (void)hipFree(present);
    (void)hipFree(temp);
    if (rc != BM_OK) return bm_ctx_set_error(ctx, rc, msg);
    if (e != hipSuccess)
        return bm_ctx_set_error(ctx, e == hipErrorOutOfMemory ? BM_ERR_NOMEM : BM_ERR_HIP, msg[0] ? msg : "HIP error");
    return BM_OK;
}
