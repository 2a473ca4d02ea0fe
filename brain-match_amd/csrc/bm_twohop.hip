/*
 * 2-hop proposal generator (gfx950).
 *
 * The 2-hop set of a vertex v is the neighbourhood refine-twohop-random
 * collects (collect_neighborhood, src/refine-twohop-random.c:582-625):
 *   N+(v) u N+(N+(v)) u N-(v) u N-(N-(v)),  minus v itself,
 * in the male graph.  A proposal batch pairs each anchor v with up to K
 * targets drawn uniformly without replacement from its set: the K members x
 * with the smallest keys  hi32(rand(q_v, x))  (ties broken by the smaller
 * id), written in ascending id order; slots beyond the set size hold (v, v),
 * whose delta is 0.  The reference draws up to 1000 targets by a partial
 * Fisher-Yates shuffle with rand_r (:686-691); the key rule is the same
 * distribution (a uniform K-subset) with a counter-based stream, so a CPU
 * replay gives the same pairs.
 *
 * One 256-thread block per anchor, persistent over the anchors:
 *   1. an (n+1)-bit membership bitmap in LDS is cleared and marked with LDS
 *      atomics: each wave takes one first-hop neighbour at a time and its
 *      lanes stride that neighbour's list (second hop);
 *   2. every thread owns a contiguous run of bitmap words (so thread order is
 *      id order); a block sum gives the set size S;
 *   3. if S > K, a 4-pass 8-bit radix select over the members' 32-bit keys
 *      finds the K-th smallest key T and how many members with key == T are
 *      still needed;
 *   4. a block scan of the selected members' counts gives each thread its
 *      output offset; pairs are written in id order, then the padding.
 */
#include "bm_device.h"
#include "bm_devutil.h"
#include "bm_rng.h"

using namespace bmd;

namespace {

constexpr int kT = 256;
constexpr int kWaves = kT / BM_WAVE;

__device__ __forceinline__ uint32_t member_key(uint64_t q, int32_t x) {
    return (uint32_t)(bm_rand_at(q, (uint64_t)x) >> 32);
}

/* Marks N(v) and N(N(v)) along one list direction. */
__device__ __forceinline__ void mark_two_hops(uint32_t* bits, const int64_t* __restrict__ ptr,
                                              const int2* __restrict__ adj, int32_t v, int lane, int wave) {
    const int64_t b = ptr[v], e = ptr[v + 1];
    for (int64_t i = b + wave; i < e; i += kWaves) {
        const int32_t x = uni(adj[i].x);
        if (lane == 0) atomicOr(&bits[x >> 5], 1u << (x & 31));
        const int64_t xb = uni64(ptr[x]), xe = uni64(ptr[x + 1]);
        for (int64_t j = xb + lane; j < xe; j += BM_WAVE) {
            const int32_t y = adj[j].x;
            atomicOr(&bits[y >> 5], 1u << (y & 31));
        }
    }
}

/* Exclusive block scan of one int per thread; returns the prefix and sets
 * *total.  `tmp` holds kWaves ints; ends with a barrier. */
__device__ __forceinline__ int32_t block_scan(int32_t x, int32_t* tmp, int32_t* total) {
    const int lane = threadIdx.x & (BM_WAVE - 1), wave = threadIdx.x / BM_WAVE;
    int32_t incl = x;
#pragma unroll
    for (int o = 1; o < BM_WAVE; o <<= 1) {
        const int32_t y = __shfl_up(incl, o, BM_WAVE);
        if (lane >= o) incl += y;
    }
    if (lane == BM_WAVE - 1) tmp[wave] = incl;
    __syncthreads();
    int32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
        const int32_t s = tmp[w];
        if (w < wave) base += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + incl - x;
}

/* K = per-anchor slot count (sampling) or 0 (the whole set into out_list). */
__global__ void __launch_bounds__(kT) k_twohop(BmDevGraph g, const int32_t* __restrict__ anchors, uint64_t seed,
                                              int64_t n_anchors, int32_t K, int2* __restrict__ pairs,
                                              int32_t* __restrict__ out_list, int32_t* __restrict__ out_count) {
    extern __shared__ uint32_t bits[];
    __shared__ int32_t hist[256];
    __shared__ int32_t tmp[kWaves];
    __shared__ int32_t sel[2];
    const int tid = threadIdx.x, lane = tid & (BM_WAVE - 1), wave = tid / BM_WAVE;
    const int32_t words = (g.n + 1 + 31) >> 5;
    const int32_t per = (words + kT - 1) / kT;
    const int32_t w0 = tid * per < words ? tid * per : words;
    const int32_t w1 = w0 + per < words ? w0 + per : words;
    for (int64_t a = blockIdx.x; a < n_anchors; a += gridDim.x) {
        const int32_t v = anchors ? anchors[a] : 1 + (int32_t)bm_range32((uint32_t)bm_rand_at(seed, 2 * (uint64_t)a),
                                                                          (uint32_t)g.n);
        const uint64_t q = bm_rand_at(seed, 2 * (uint64_t)a + 1);
        for (int32_t w = tid; w < words; w += kT) bits[w] = 0;
        __syncthreads();
        mark_two_hops(bits, g.m_out_ptr, g.m_out, v, lane, wave);
        mark_two_hops(bits, g.m_in_ptr, g.m_in, v, lane, wave);
        __syncthreads();
        if (tid == 0) bits[v >> 5] &= ~(1u << (v & 31));
        __syncthreads();

        int32_t mine = 0;
        for (int32_t w = w0; w < w1; w++) mine += __popc(bits[w]);
        int32_t S;
        (void)block_scan(mine, tmp, &S);

        /* threshold: members with key < T, plus the first `need_eq` with key == T */
        uint32_t T = 0xffffffffu;
        int32_t need_eq = 0x7fffffff;
        const bool sample = K > 0 && S > K;
        if (sample) {
            uint32_t prefix = 0, pmask = 0;
            int32_t remaining = K;
            for (int shift = 24; shift >= 0; shift -= 8) {
                for (int i = tid; i < 256; i += kT) hist[i] = 0;
                __syncthreads();
                for (int32_t w = w0; w < w1; w++) {
                    uint32_t m = bits[w];
                    while (m) {
                        const int b = __ffs(m) - 1;
                        m &= m - 1;
                        const uint32_t k = member_key(q, (w << 5) + b);
                        if ((k & pmask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1);
                    }
                }
                __syncthreads();
                if (tid == 0) {
                    int32_t cum = 0, d = 0;
                    for (; d < 256; d++) {
                        if (cum + hist[d] >= remaining) break;
                        cum += hist[d];
                    }
                    sel[0] = d;
                    sel[1] = remaining - cum;
                }
                __syncthreads();
                prefix |= (uint32_t)sel[0] << shift;
                pmask |= 255u << shift;
                remaining = sel[1];
                __syncthreads();
            }
            T = prefix;
            need_eq = remaining;
        }

        /* rank of this thread's key == T members among all of them (id order) */
        int32_t eq = 0;
        if (sample)
            for (int32_t w = w0; w < w1; w++) {
                uint32_t m = bits[w];
                while (m) {
                    const int b = __ffs(m) - 1;
                    m &= m - 1;
                    eq += member_key(q, (w << 5) + b) == T;
                }
            }
        int32_t eq_tot;
        int32_t eq_rank = block_scan(eq, tmp, &eq_tot);

        /* count selected, scan, write in id order */
        int32_t cnt = 0;
        {
            int32_t r = eq_rank;
            for (int32_t w = w0; w < w1; w++) {
                uint32_t m = bits[w];
                while (m) {
                    const int b = __ffs(m) - 1;
                    m &= m - 1;
                    if (!sample) { cnt++; continue; }
                    const uint32_t k = member_key(q, (w << 5) + b);
                    if (k < T) cnt++;
                    else if (k == T) { cnt += r < need_eq; r++; }
                }
            }
        }
        int32_t total;
        int32_t off = block_scan(cnt, tmp, &total);
        {
            int32_t r = eq_rank;
            for (int32_t w = w0; w < w1; w++) {
                uint32_t m = bits[w];
                while (m) {
                    const int b = __ffs(m) - 1;
                    m &= m - 1;
                    const int32_t x = (w << 5) + b;
                    bool take = true;
                    if (sample) {
                        const uint32_t k = member_key(q, x);
                        take = k < T || (k == T && r++ < need_eq);
                    }
                    if (!take) continue;
                    if (K > 0) pairs[a * K + off] = make_int2(v, x);
                    else out_list[off] = x;
                    off++;
                }
            }
        }
        if (K > 0)
            for (int32_t s = total + tid; s < K; s += kT) pairs[a * K + s] = make_int2(v, v);
        else if (tid == 0)
            out_count[a] = total;
        __syncthreads();
    }
}

}  // namespace

size_t bm_twohop_lds_bytes(int32_t n) { return (size_t)((n + 1 + 31) >> 5) * 4; }

hipError_t bm_launch_twohop(const BmDevGraph& g, const int32_t* anchors, uint64_t seed, int64_t n_anchors,
                            int32_t K, int2* pairs, int32_t* out_list, int32_t* out_count, int32_t num_cus,
                            hipStream_t s) {
    if (n_anchors <= 0) return hipSuccess;
    const size_t lds = bm_twohop_lds_bytes(g.n);
    hipError_t e = hipFuncSetAttribute((const void*)k_twohop, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_twohop, kT, lds);
    if (e != hipSuccess) return e;
    if (occ < 1) occ = 1;
    int64_t blocks = (int64_t)occ * (num_cus > 0 ? num_cus : 256);
    if (blocks > n_anchors) blocks = n_anchors;
    hipLaunchKernelGGL(k_twohop, dim3((unsigned)blocks), dim3(kT), lds, s, g, anchors, seed, n_anchors, K, pairs,
                       out_list, out_count);
    return hipGetLastError();
}
