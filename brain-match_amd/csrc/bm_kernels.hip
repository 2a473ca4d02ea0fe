/*
 * gfx950 support kernels around the delta kernels (bm_delta.hip):
 * candidate generation (uniform streams and the reference's changed-first
 * enumerations), candidate validation, reductions over deltas (stats, rank
 * picks), deterministic disjoint acceptance, and single-move application.
 */
#include "bm_device.h"
#include "bm_devutil.h"
#include "bm_rng.h"

#include <hipcub/hipcub.hpp>

using namespace bmd;

namespace {

/* ------------------------------------------------------------------ */
/* Candidate generation and validation                                 */
/* ------------------------------------------------------------------ */

__global__ void k_gen_pairs(int32_t n, uint64_t seed, int64_t k, int2* __restrict__ cand) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < k; c += stride) {
        int32_t u, v;
        bm_uniform_pair(seed, (uint64_t)c, n, &u, &v);
        cand[c] = make_int2(u, v);
    }
}

__global__ void k_gen_triples(int32_t n, uint64_t seed, int64_t k, int4* __restrict__ cand) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < k; c += stride) {
        int32_t a, b, d;
        bm_uniform_triple(seed, (uint64_t)c, n, &a, &b, &d);
        cand[c] = make_int4(a, b, d, 0);
    }
}

/* Lexicographic pair t of C(c,2) -> (i, j), i < j < c. */
__device__ __forceinline__ void pair_of(int64_t t, int64_t c, int64_t* i, int64_t* j) {
    /* row i starts at s(i) = i*(2c-i-1)/2 */
    const double b = 2.0 * (double)c - 1.0;
    int64_t r = (int64_t)((b - sqrt(b * b - 8.0 * (double)t)) * 0.5);
    if (r < 0) r = 0;
    if (r > c - 2) r = c - 2;
    while (r > 0 && r * (2 * c - r - 1) / 2 > t) r--;
    while (r + 1 <= c - 2 && (r + 1) * (2 * c - r - 2) / 2 <= t) r++;
    *i = r;
    *j = r + 1 + (t - r * (2 * c - r - 1) / 2);
}

__global__ void k_gen_changed_pairs(int32_t n, int32_t c, const int32_t* __restrict__ reorder, int64_t P,
                                    int2* __restrict__ cand) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t cc = (int64_t)c * (c - 1) / 2;
    const int64_t nu = (int64_t)n - c;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < P; idx += stride) {
        int64_t i, j;
        if (idx < cc) {
            pair_of(idx, c, &i, &j);
            cand[idx] = make_int2(reorder[i], reorder[j]);
        } else {
            const int64_t t = idx - cc;
            i = t / nu;
            j = t - i * nu;
            cand[idx] = make_int2(reorder[i], reorder[c + j]);
        }
    }
}

/* The 3-section enumeration of calculateThreeNodeSwapKernel
 * (src/graph-re-alignment-gpu-3-swap.cu:67-150), with the same order. */
__global__ void k_gen_changed_triples(int32_t n, int32_t c, const int32_t* __restrict__ reorder, int64_t T,
                                      int4* __restrict__ cand) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t s1 = (int64_t)c * (c - 1) * (c - 2) / 6;
    const int64_t cc = (int64_t)c * (c - 1) / 2;
    const int64_t s2 = s1 + cc * ((int64_t)n - c);
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < T; idx += stride) {
        int32_t a, b, d;
        if (idx < s1) {
            /* lexicographic triple of the c changed positions */
            int64_t t = idx, i = 0, rem = (int64_t)(c - 1) * (c - 2) / 2;
            while (t >= rem) { t -= rem; i++; rem = (int64_t)(c - 1 - i) * (c - 2 - i) / 2; }
            int64_t j = i + 1, r2 = c - i - 2;
            while (t >= r2) { t -= r2; j++; r2--; }
            a = reorder[i];
            b = reorder[j];
            d = reorder[j + 1 + t];
        } else if (idx < s2) {
            const int64_t t = idx - s1;
            const int64_t m3 = t / cc;
            int64_t i, j;
            pair_of(t - m3 * cc, c, &i, &j);
            a = reorder[i];
            b = reorder[j];
            d = reorder[c + m3];
        } else {
            const int64_t t = idx - s2;
            const int64_t rest = (int64_t)n - c - 1;
            const int64_t i = t / rest, j = t - i * rest;
            a = reorder[i];
            b = reorder[c + j];
            d = reorder[c + j + 1];
        }
        cand[idx] = make_int4(a, b, d, 0);
    }
}

__global__ void k_check_pairs(int32_t n, const int2* __restrict__ cand, int64_t k, int32_t* bad) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int badl = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < k; c += stride) {
        const int2 p = cand[c];
        badl |= (p.x < 1) | (p.x > n) | (p.y < 1) | (p.y > n);
    }
    if (badl) atomicOr(bad, 1);
}

__global__ void k_check_triples(int32_t n, const int4* __restrict__ cand, int64_t k, int32_t* bad) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int badl = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < k; c += stride) {
        const int4 p = cand[c];
        badl |= (p.x < 1) | (p.x > n) | (p.y < 1) | (p.y > n) | (p.z < 1) | (p.z > n);
        badl |= (p.x == p.y) | (p.x == p.z) | (p.y == p.z);
    }
    if (badl) atomicOr(bad, 1);
}

/* ------------------------------------------------------------------ */
/* Reductions                                                          */
/* ------------------------------------------------------------------ */

constexpr int kRedThreads = 256;
constexpr int kRedChunk = 4096;   /* elements per block */

__global__ void __launch_bounds__(kRedThreads) k_stats_part(const int64_t* __restrict__ d, int64_t k,
                                                            int64_t* __restrict__ part) {
    __shared__ int64_t s_sum[kRedThreads], s_max[kRedThreads], s_arg[kRedThreads], s_pos[kRedThreads],
        s_zero[kRedThreads];
    const int64_t base = (int64_t)blockIdx.x * kRedChunk;
    int64_t sum = 0, mx = INT64_MIN, arg = -1, pos = 0, zero = 0;
    for (int q = threadIdx.x; q < kRedChunk; q += kRedThreads) {
        const int64_t i = base + q;
        if (i >= k) break;
        const int64_t x = d[i];
        sum += x;
        pos += x > 0;
        zero += x == 0;
        if (x > mx) { mx = x; arg = i; }
    }
    s_sum[threadIdx.x] = sum; s_max[threadIdx.x] = mx; s_arg[threadIdx.x] = arg;
    s_pos[threadIdx.x] = pos; s_zero[threadIdx.x] = zero;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const int a = threadIdx.x, b = threadIdx.x + o;
            s_sum[a] += s_sum[b];
            s_pos[a] += s_pos[b];
            s_zero[a] += s_zero[b];
            if (s_max[b] > s_max[a] || (s_max[b] == s_max[a] && s_arg[b] >= 0 && (s_arg[a] < 0 || s_arg[b] < s_arg[a]))) {
                s_max[a] = s_max[b];
                s_arg[a] = s_arg[b];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int64_t* o = part + (int64_t)blockIdx.x * 5;
        o[0] = s_sum[0]; o[1] = s_max[0]; o[2] = s_arg[0]; o[3] = s_pos[0]; o[4] = s_zero[0];
    }
}

__global__ void __launch_bounds__(kRedThreads) k_stats_final(const int64_t* __restrict__ part, int64_t nb, int64_t k,
                                                             int64_t* __restrict__ out) {
    __shared__ int64_t s_sum[kRedThreads], s_max[kRedThreads], s_arg[kRedThreads], s_pos[kRedThreads],
        s_zero[kRedThreads];
    int64_t sum = 0, mx = INT64_MIN, arg = -1, pos = 0, zero = 0;
    for (int64_t b = threadIdx.x; b < nb; b += kRedThreads) {
        const int64_t* p = part + b * 5;
        sum += p[0];
        pos += p[3];
        zero += p[4];
        if (p[2] >= 0 && (p[1] > mx || (p[1] == mx && (arg < 0 || p[2] < arg)))) { mx = p[1]; arg = p[2]; }
    }
    s_sum[threadIdx.x] = sum; s_max[threadIdx.x] = mx; s_arg[threadIdx.x] = arg;
    s_pos[threadIdx.x] = pos; s_zero[threadIdx.x] = zero;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const int a = threadIdx.x, b = threadIdx.x + o;
            s_sum[a] += s_sum[b];
            s_pos[a] += s_pos[b];
            s_zero[a] += s_zero[b];
            if (s_arg[b] >= 0 && (s_max[b] > s_max[a] || (s_max[b] == s_max[a] && (s_arg[a] < 0 || s_arg[b] < s_arg[a])))) {
                s_max[a] = s_max[b];
                s_arg[a] = s_arg[b];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = k; out[1] = s_sum[0]; out[2] = s_max[0]; out[3] = s_arg[0]; out[4] = s_pos[0]; out[5] = s_zero[0];
    }
}

__global__ void __launch_bounds__(kRedThreads) k_pick_count(const int64_t* __restrict__ d, int64_t k, int32_t kind,
                                                            int64_t* __restrict__ cnt) {
    __shared__ int64_t s[kRedThreads];
    const int64_t base = (int64_t)blockIdx.x * kRedChunk;
    int64_t c = 0;
    for (int q = threadIdx.x; q < kRedChunk; q += kRedThreads) {
        const int64_t i = base + q;
        if (i >= k) break;
        const int64_t x = d[i];
        c += kind ? (x > 0) : (x == 0);
    }
    s[threadIdx.x] = c;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[blockIdx.x] = s[0];
}

/* One block: locate the chunk holding the rank-th match, then the element. */
__global__ void __launch_bounds__(1024) k_pick_find(const int64_t* __restrict__ d, int64_t k, int32_t kind,
                                                    const int64_t* __restrict__ cnt, int64_t nb, int64_t rank,
                                                    int64_t* __restrict__ out) {
    __shared__ int64_t s[1024];
    __shared__ int64_t s_chunk, s_before;
    const int t = threadIdx.x;
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b0 = (int64_t)t * per, b1 = b0 + per < nb ? b0 + per : nb;
    int64_t mine = 0;
    for (int64_t b = b0; b < b1; b++) mine += cnt[b];
    s[t] = mine;
    if (t == 0) { s_chunk = -1; s_before = 0; }
    __syncthreads();
    /* inclusive scan (Hillis-Steele) */
    for (int o = 1; o < 1024; o <<= 1) {
        int64_t x = t >= o ? s[t - o] : 0;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    const int64_t incl = s[t], excl = incl - mine;
    if (rank >= excl && rank < incl) {
        int64_t acc = excl;
        for (int64_t b = b0; b < b1; b++) {
            if (rank < acc + cnt[b]) { s_chunk = b; s_before = acc; break; }
            acc += cnt[b];
        }
    }
    __syncthreads();
    if (s_chunk < 0) {
        if (t == 0) out[0] = -1;
        return;
    }
    /* within the chunk: 4 elements per thread, block prefix over counts */
    const int64_t base = s_chunk * kRedChunk;
    int64_t local = 0;
    int flags[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int64_t i = base + (int64_t)t * 4 + q;
        const int64_t x = i < k ? d[i] : -1;
        flags[q] = i < k && (kind ? (x > 0) : (x == 0));
        local += flags[q];
    }
    __syncthreads();
    s[t] = local;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int64_t x = t >= o ? s[t - o] : 0;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    int64_t r = s_before + s[t] - local;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (flags[q]) {
            if (r == rank) out[0] = base + (int64_t)t * 4 + q;
            r++;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Disjoint acceptance                                                 */
/* ------------------------------------------------------------------ */

struct PosDelta {
    const int64_t* d;
    __host__ __device__ bool operator()(const int64_t& i) const { return d[i] > 0; }
};

__global__ void k_gather_keys(const int64_t* __restrict__ d, const int64_t* __restrict__ idx, int64_t m,
                              int64_t* __restrict__ keys) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) keys[i] = d[idx[i]];
}

constexpr int kAccThreads = 1024;

/* Sequential greedy in priority order, one block.  lock[] is a bitmap of
 * vertices in the closed neighbourhood of accepted swaps (LDS, n+1 bits). */
__global__ void __launch_bounds__(kAccThreads) k_accept_greedy(BmDevGraph g, int32_t* __restrict__ pi,
                                                               const int2* __restrict__ cand,
                                                               const int64_t* __restrict__ d,
                                                               const int64_t* __restrict__ order, int64_t npos,
                                                               int64_t max_accept, int64_t* __restrict__ acc_idx,
                                                               int64_t* __restrict__ meta, int4* __restrict__ moves) {
    extern __shared__ uint32_t lock[];
    __shared__ int s_first;
    __shared__ int32_t s_u, s_v;
    __shared__ int64_t s_nacc, s_gain;
    const int t = threadIdx.x;
    const int words = (g.n + 32) / 32;
    for (int i = t; i < words; i += kAccThreads) lock[i] = 0;
    if (t == 0) { s_nacc = 0; s_gain = 0; }
    __syncthreads();
    for (int64_t base = 0; base < npos; base += kAccThreads) {
        if (s_nacc >= max_accept) break;
        const int64_t i = base + t;
        const bool valid = i < npos;
        int64_t ci = -1;
        int32_t u = 0, v = 0;
        if (valid) {
            ci = order[i];
            const int2 p = cand[ci];
            u = p.x;
            v = p.y;
        }
        int start = 0;
        for (;;) {
            if (t == 0) s_first = kAccThreads;
            __syncthreads();
            const bool free_ = valid && t >= start && u != v && !((lock[u >> 5] >> (u & 31)) & 1u) &&
                               !((lock[v >> 5] >> (v & 31)) & 1u);
            if (free_) atomicMin(&s_first, t);
            __syncthreads();
            const int first = s_first;
            if (first >= kAccThreads) break;
            if (t == first) {
                acc_idx[s_nacc] = ci;
                s_nacc += 1;
                s_gain += d[ci];
                s_u = u;
                s_v = v;
            }
            __syncthreads();
            const int32_t au = s_u, av = s_v;
            if (t == 0) {
                atomicOr(&lock[au >> 5], 1u << (au & 31));
                atomicOr(&lock[av >> 5], 1u << (av & 31));
            }
            const int32_t ends[2] = {au, av};
            for (int q = 0; q < 2; q++) {
                const int32_t x = ends[q];
                for (int64_t e = g.m_out_ptr[x] + t; e < g.m_out_ptr[x + 1]; e += kAccThreads) {
                    const int32_t y = g.m_out[e].x;
                    atomicOr(&lock[y >> 5], 1u << (y & 31));
                }
                for (int64_t e = g.m_in_ptr[x] + t; e < g.m_in_ptr[x + 1]; e += kAccThreads) {
                    const int32_t y = g.m_in[e].x;
                    atomicOr(&lock[y >> 5], 1u << (y & 31));
                }
            }
            __syncthreads();
            start = first + 1;
            if (s_nacc >= max_accept) break;
        }
        __syncthreads();
    }
    __syncthreads();
    const int64_t nacc = s_nacc;
    for (int64_t a = t; a < nacc; a += kAccThreads) {
        const int2 p = cand[acc_idx[a]];
        const int32_t fu = pi[p.x], fv = pi[p.y];
        if (moves) {
            moves[2 * a] = make_int4(p.x, fu, fv, 0);
            moves[2 * a + 1] = make_int4(p.y, fv, fu, 0);
        }
        pi[p.x] = fv;
        pi[p.y] = fu;
    }
    if (t == 0) { meta[0] = nacc; meta[1] = s_gain; meta[2] = 2 * nacc; }
}

__global__ void k_apply_swap(int32_t* pi, int32_t u, int32_t v) {
    const int32_t t = pi[u];
    pi[u] = pi[v];
    pi[v] = t;
}

/* apply_permutation (src/graph-re-alignment-gpu-3-swap.cu:925-956). */
__global__ void k_apply_perm3(int32_t* pi, int32_t a, int32_t b, int32_t c, int32_t p) {
    const int32_t f[3] = {pi[a], pi[b], pi[c]};
    pi[a] = f[perm_of(p, 0)];
    pi[b] = f[perm_of(p, 1)];
    pi[c] = f[perm_of(p, 2)];
}

/* ---- u-major uniform stream (bm_umajor_first / bm_umajor_partner) ---- */

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t count, int32_t v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = v;
}

/* counts[u] += #{c : first vertex of draw c is u}; a block-private histogram
 * in LDS (n+1 ints) when it fits, merged with one global atomic per bin. */
__global__ void __launch_bounds__(1024) k_umajor_hist(int32_t n, uint64_t seed, int64_t k, int32_t* __restrict__ counts,
                                                     int use_lds) {
    extern __shared__ int32_t h[];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (use_lds) {
        for (int32_t x = threadIdx.x; x <= n; x += blockDim.x) h[x] = 0;
        __syncthreads();
        for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < k; c += stride)
            atomicAdd(&h[bm_umajor_first(seed, (uint64_t)c, n)], 1);
        __syncthreads();
        for (int32_t x = threadIdx.x; x <= n; x += blockDim.x)
            if (h[x]) atomicAdd(&counts[x], h[x]);
    } else {
        for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < k; c += stride)
            atomicAdd(&counts[bm_umajor_first(seed, (uint64_t)c, n)], 1);
    }
}

/* off[0] = off[1] = 0, off[u + 1] = off[u] + counts[u] (u = 1..n): one block,
 * each thread a contiguous run, then a block scan of the run sums. */
__global__ void __launch_bounds__(1024) k_umajor_offsets(int32_t n, const int32_t* __restrict__ counts,
                                                        int64_t* __restrict__ off) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = ((int64_t)n + 1023) / 1024;
    const int64_t a = 1 + (int64_t)t * per, b = a + per < (int64_t)n + 1 ? a + per : (int64_t)n + 1;
    int64_t sum = 0;
    for (int64_t x = a; x < b; x++) sum += counts[x];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int64_t y = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += y;
        __syncthreads();
    }
    int64_t run = t ? part[t - 1] : 0;
    if (t == 0) {
        off[0] = 0;
        off[1] = 0;
    }
    for (int64_t x = a; x < b; x++) {
        run += counts[x];
        off[x + 1] = run;
    }
}

__global__ void k_gen_umajor(int32_t n, uint64_t seed, const int64_t* __restrict__ off, int64_t k,
                             int2* __restrict__ cand) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += stride) {
        const int32_t u = umajor_owner(off, n, i);
        cand[i] = make_int2(u, bm_umajor_partner(seed, (uint64_t)i, u, n));
    }
}

/* a-major triples: entry i of the stream (seed) is (a, b, c, 0), a the
 * owner of i in the offsets of the first-vertex draws. */
__global__ void k_gen_amajor(int32_t n, uint64_t seed, const int64_t* __restrict__ off, int64_t k,
                             int4* __restrict__ cand) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += stride) {
        const int32_t a = umajor_owner(off, n, i);
        int32_t b, c;
        bm_amajor_partners(seed, (uint64_t)i, a, n, &b, &c);
        cand[i] = make_int4(a, b, c, 0);
    }
}

int grid_for(int64_t work, int threads) {
    int64_t b = (work + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (int)b;
}

}  // namespace

/* ------------------------------------------------------------------ */
/* Launchers                                                           */
/* ------------------------------------------------------------------ */

static hipError_t set_lds(const void* fn, size_t bytes) {
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t bm_launch_gen_pairs(int32_t n, uint64_t seed, int64_t k, int2* cand, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_pairs, dim3(grid_for(k, 256)), dim3(256), 0, s, n, seed, k, cand);
    return hipGetLastError();
}

hipError_t bm_launch_gen_triples(int32_t n, uint64_t seed, int64_t k, int4* cand, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_triples, dim3(grid_for(k, 256)), dim3(256), 0, s, n, seed, k, cand);
    return hipGetLastError();
}

hipError_t bm_launch_gen_changed_pairs(int32_t n, int32_t c, const int32_t* reorder, int64_t P, int2* cand,
                                       hipStream_t s) {
    if (P <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_changed_pairs, dim3(grid_for(P, 256)), dim3(256), 0, s, n, c, reorder, P, cand);
    return hipGetLastError();
}

hipError_t bm_launch_gen_changed_triples(int32_t n, int32_t c, const int32_t* reorder, int64_t T, int4* cand,
                                         hipStream_t s) {
    if (T <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_changed_triples, dim3(grid_for(T, 256)), dim3(256), 0, s, n, c, reorder, T, cand);
    return hipGetLastError();
}

hipError_t bm_launch_check_pairs(int32_t n, const int2* cand, int64_t k, int32_t* bad, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_check_pairs, dim3(grid_for(k, 256)), dim3(256), 0, s, n, cand, k, bad);
    return hipGetLastError();
}

hipError_t bm_launch_check_triples(int32_t n, const int4* cand, int64_t k, int32_t* bad, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_check_triples, dim3(grid_for(k, 256)), dim3(256), 0, s, n, cand, k, bad);
    return hipGetLastError();
}

int64_t bm_stats_blocks(int64_t k) { return k <= 0 ? 1 : (k + kRedChunk - 1) / kRedChunk; }

hipError_t bm_launch_stats(const int64_t* d, int64_t k, int64_t* part, int64_t* out6, hipStream_t s) {
    const int64_t nb = bm_stats_blocks(k);
    hipLaunchKernelGGL(k_stats_part, dim3((unsigned)nb), dim3(kRedThreads), 0, s, d, k, part);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(kRedThreads), 0, s, part, nb, k, out6);
    return hipGetLastError();
}

hipError_t bm_launch_pick(const int64_t* d, int64_t k, int32_t kind, int64_t rank, int64_t* cnt, int64_t* out,
                          hipStream_t s) {
    const int64_t nb = bm_stats_blocks(k);
    hipLaunchKernelGGL(k_pick_count, dim3((unsigned)nb), dim3(kRedThreads), 0, s, d, k, kind, cnt);
    hipLaunchKernelGGL(k_pick_find, dim3(1), dim3(1024), 0, s, d, k, kind, cnt, nb, rank, out);
    return hipGetLastError();
}

/* Scratch layout: [idx k][keys k][sorted keys k][sorted idx k][count 1][cub temp ...] */
static size_t cub_temp_bytes(int64_t k) {
    size_t a = 0, b = 0;
    hipcub::CountingInputIterator<int64_t> it(0);
    PosDelta pred{nullptr};
    (void)hipcub::DeviceSelect::If(nullptr, a, it, (int64_t*)nullptr, (int64_t*)nullptr, (int)k, pred);
    (void)hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b, (const int64_t*)nullptr, (int64_t*)nullptr,
                                                 (const int64_t*)nullptr, (int64_t*)nullptr, (int)k);
    return a > b ? a : b;
}

size_t bm_accept_scratch_bytes(int64_t k) {
    if (k < 1) k = 1;
    return (size_t)k * 4 * sizeof(int64_t) + 256 + cub_temp_bytes(k) + 256;
}

hipError_t bm_launch_accept(const BmDevGraph& g, int32_t* pi_rw, const int2* cand, const int64_t* d, int64_t k,
                            int64_t max_accept, void* scratch, size_t scratch_bytes, int64_t* out_acc_idx,
                            int64_t* out_meta, int4* out_moves, hipStream_t s) {
    if (k <= 0) {
        hipError_t e = hipMemsetAsync(out_meta, 0, 3 * sizeof(int64_t), s);
        return e;
    }
    char* p = (char*)scratch;
    int64_t* idx = (int64_t*)p;
    int64_t* keys = idx + k;
    int64_t* skeys = keys + k;
    int64_t* sidx = skeys + k;
    int64_t* cnt = sidx + k;
    char* temp = (char*)(cnt) + 256;
    size_t temp_bytes = scratch_bytes - ((size_t)(temp - p));
    hipcub::CountingInputIterator<int64_t> it(0);
    PosDelta pred{d};
    size_t tb = temp_bytes;
    hipError_t e = hipcub::DeviceSelect::If(temp, tb, it, idx, cnt, (int)k, pred, s);
    if (e != hipSuccess) return e;
    int64_t npos = 0;
    e = hipMemcpyAsync(&npos, cnt, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if (npos > 0) {
        size_t need = 0;
        e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, need, (const int64_t*)nullptr, (int64_t*)nullptr,
                                                         (const int64_t*)nullptr, (int64_t*)nullptr, (int)npos, 0,
                                                         64, s);
        if (e != hipSuccess) return e;
        if (need > temp_bytes) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_gather_keys, dim3(grid_for(npos, 256)), dim3(256), 0, s, d, idx, npos, keys);
        tb = temp_bytes;
        e = hipcub::DeviceRadixSort::SortPairsDescending(temp, tb, keys, skeys, idx, sidx, (int)npos, 0, 64, s);
        if (e != hipSuccess) return e;
    }
    const size_t lds = (size_t)((g.n + 32) / 32) * sizeof(uint32_t);
    e = set_lds((const void*)k_accept_greedy, lds);
    if (e != hipSuccess) return e;
    if (max_accept <= 0) max_accept = INT64_MAX;
    hipLaunchKernelGGL(k_accept_greedy, dim3(1), dim3(kAccThreads), lds, s, g, pi_rw, cand, d, sidx, npos,
                       max_accept, out_acc_idx, out_meta, out_moves);
    return hipGetLastError();
}

hipError_t bm_launch_apply_swap(int32_t* pi, int32_t u, int32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_apply_swap, dim3(1), dim3(1), 0, s, pi, u, v);
    return hipGetLastError();
}

hipError_t bm_launch_apply_perm3(int32_t* pi, int32_t a, int32_t b, int32_t c, int32_t p, hipStream_t s) {
    hipLaunchKernelGGL(k_apply_perm3, dim3(1), dim3(1), 0, s, pi, a, b, c, p);
    return hipGetLastError();
}

hipError_t bm_launch_umajor_offsets(int32_t n, uint64_t seed, int64_t k, int32_t* counts, int64_t* off,
                                    int32_t num_cus, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for((int64_t)n + 1, 256)), dim3(256), 0, s, counts, (int64_t)n + 1, 0);
    const size_t lds = ((size_t)n + 1) * sizeof(int32_t);
    const int use_lds = lds <= 150 * 1024 ? 1 : 0;
    if (use_lds) {
        hipError_t e = hipFuncSetAttribute((const void*)k_umajor_hist, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
    }
    int64_t blocks = (k + 1023) / 1024;
    const int64_t cap = (int64_t)(num_cus > 0 ? num_cus : 256);
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    if (k > 0)
        hipLaunchKernelGGL(k_umajor_hist, dim3((unsigned)blocks), dim3(1024), use_lds ? lds : 0, s, n, seed, k, counts,
                           use_lds);
    hipLaunchKernelGGL(k_umajor_offsets, dim3(1), dim3(1024), 0, s, n, counts, off);
    return hipGetLastError();
}

hipError_t bm_launch_gen_amajor(int32_t n, uint64_t seed, const int64_t* off, int64_t k, int4* cand, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_amajor, dim3(grid_for(k, 256)), dim3(256), 0, s, n, seed, off, k, cand);
    return hipGetLastError();
}

hipError_t bm_launch_gen_umajor(int32_t n, uint64_t seed, const int64_t* off, int64_t k, int2* cand, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_umajor, dim3(grid_for(k, 256)), dim3(256), 0, s, n, seed, off, k, cand);
    return hipGetLastError();
}
