/*
 * gfx950 kernels for exact swap / 3-permutation deltas, the full score and
 * deterministic disjoint acceptance.
 *
 * Work decomposition: one 64-lane wavefront per candidate, persistent grid
 * (each wave strides over the candidate index space).  For a candidate the
 * wave
 *   1. reads pi of the candidate's vertices (wave-uniform, scalar loads);
 *   2. stages the hash tables of the female rows it will probe into its own
 *      slice of LDS (coalesced 16-byte copies), or leaves a table in global
 *      memory when it does not fit (hub rows; then probes hit L2);
 *   3. streams the male out-lists (then in-lists) of the candidate's vertices
 *      with coalesced 8-byte {neighbour, weight} loads, gathers pi[neighbour]
 *      and probes the staged tables;
 *   4. reduces the per-lane int64 partial deltas across the wave.
 * The reference kernels scan a dense (n+1)^2 int16 matrix row per vertex
 * instead (src/graph-re-alignment-gpu-swap.cu:198-233), O(n) per pair.
 *
 * Exact contribution of a male edge (w) mapped onto female vertices (a,b):
 *   c = min(w, w_F(a,b)) if F has a->b, else 0   (src/score-matching.c:579-590)
 * A probe returns BM_ABSENT (INT32_MIN, rejected as a weight at load) when
 * the female edge does not exist.
 */
#include "bm_device.h"
#include "bm_rng.h"

#include <hipcub/hipcub.hpp>

#define BM_ABSENT INT32_MIN
#define BM_WAVE 64

namespace {

__device__ __forceinline__ uint32_t slot_hash(int32_t key, uint32_t bits) {
    return ((uint32_t)key * 0x9E3779B1u) >> (32u - bits);
}

__device__ __forceinline__ int32_t contrib(int32_t w, int32_t fw) {
    return fw == BM_ABSENT ? 0 : (w < fw ? w : fw);
}

__device__ __forceinline__ int32_t uni(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint32_t uniu(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ int64_t uni64(int64_t x) {
    uint32_t lo = uniu((uint32_t)x), hi = uniu((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

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

/* A female row table: either staged in this wave's LDS slice at `off`, or
 * probed in global memory.  All fields are wave-uniform. */
struct Tab {
    const int2* g;
    uint32_t off;
    uint32_t bits;
    uint32_t in_lds;
};

__device__ __forceinline__ Tab tab_of(const uint64_t* __restrict__ desc, const int2* __restrict__ tab,
                                      int32_t row) {
    Tab t;
    uint64_t d = desc[row];
    uint32_t lo = uniu((uint32_t)d), hi = uniu((uint32_t)(d >> 32));
    d = ((uint64_t)hi << 32) | lo;
    t.bits = (uint32_t)(d & 63u);
    t.g = tab + (d >> 6);
    t.off = 0;
    t.in_lds = 0;
    return t;
}

__device__ __forceinline__ uint32_t tab_slots(const Tab& t) { return t.bits ? (1u << t.bits) : 0u; }

__device__ __forceinline__ int32_t probe(const int2* lds, const Tab& t, int32_t key) {
    if (t.bits == 0) return BM_ABSENT;
    const uint32_t mask = (1u << t.bits) - 1u;
    uint32_t s = slot_hash(key, t.bits);
    if (t.in_lds) {
        const int2* T = lds + t.off;
        for (;;) {
            int2 e = T[s];
            if (e.x == 0) return BM_ABSENT;
            if (e.x == key) return e.y;
            s = (s + 1u) & mask;
        }
    } else {
        const int2* __restrict__ T = t.g;
        for (;;) {
            int2 e = T[s];
            if (e.x == 0) return BM_ABSENT;
            if (e.x == key) return e.y;
            s = (s + 1u) & mask;
        }
    }
}

/* Copy a table into LDS with 16-byte moves (sizes are even, offsets even). */
__device__ __forceinline__ void stage(int2* lds, const Tab& t, int lane) {
    if (!t.in_lds) return;
    const int4* __restrict__ src = reinterpret_cast<const int4*>(t.g);
    int4* dst = reinterpret_cast<int4*>(lds + t.off);
    const uint32_t n4 = tab_slots(t) >> 1;
    for (uint32_t i = lane; i < n4; i += BM_WAVE) dst[i] = src[i];
}

/* Place up to N tables into a slice of `cap` slots: smallest first, each
 * table goes to LDS if it still fits. */
template <int N>
__device__ __forceinline__ void place(Tab* t, uint32_t cap) {
    uint32_t order[N];
#pragma unroll
    for (int i = 0; i < N; i++) order[i] = i;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j + 1 < N - i; j++) {
            uint32_t a = order[j], b = order[j + 1];
            uint32_t sa = 0, sb = 0;
#pragma unroll
            for (int q = 0; q < N; q++) {
                if ((uint32_t)q == a) sa = tab_slots(t[q]);
                if ((uint32_t)q == b) sb = tab_slots(t[q]);
            }
            if (sb < sa) { order[j] = b; order[j + 1] = a; }
        }
    uint32_t used = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
#pragma unroll
        for (int q = 0; q < N; q++) {
            if ((uint32_t)q != order[i]) continue;
            uint32_t s = tab_slots(t[q]);
            if (s && used + s <= cap) {
                t[q].in_lds = 1;
                t[q].off = used;
                used += s;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Swap deltas                                                         */
/* ------------------------------------------------------------------ */

constexpr int kUnroll = 4;

/* Out-edges of k (old female row To, new row Tn).  An edge to u or v uses
 * the post-swap female of that endpoint; this covers the u<->v edges and
 * self-loops exactly. */
__device__ __forceinline__ int64_t swap_scan_out(const int2* __restrict__ adj, const int32_t* __restrict__ pi,
                                                 int64_t beg, int64_t end, const int2* lds, const Tab& To,
                                                 const Tab& Tn, int32_t u, int32_t v, int32_t f1, int32_t f2,
                                                 int lane) {
    int64_t acc = 0;
    for (int64_t e0 = beg + lane; e0 < end; e0 += (int64_t)BM_WAVE * kUnroll) {
        int2 ed[kUnroll];
        int32_t gx[kUnroll];
#pragma unroll
        for (int j = 0; j < kUnroll; j++) {
            int64_t e = e0 + (int64_t)BM_WAVE * j;
            ed[j] = e < end ? adj[e] : make_int2(0, 0);
        }
#pragma unroll
        for (int j = 0; j < kUnroll; j++) gx[j] = pi[ed[j].x];
#pragma unroll
        for (int j = 0; j < kUnroll; j++) {
            const int32_t x = ed[j].x, w = ed[j].y;
            if (x == 0) continue;
            int32_t go = gx[j], gn = gx[j];
            if (x == u) { go = f1; gn = f2; }
            else if (x == v) { go = f2; gn = f1; }
            acc += (int64_t)contrib(w, probe(lds, Tn, gn)) - (int64_t)contrib(w, probe(lds, To, go));
        }
    }
    return acc;
}

/* In-edges of k; edges from u or v were counted by their out-scan. */
__device__ __forceinline__ int64_t swap_scan_in(const int2* __restrict__ adj, const int32_t* __restrict__ pi,
                                                int64_t beg, int64_t end, const int2* lds, const Tab& To,
                                                const Tab& Tn, int32_t u, int32_t v, int lane) {
    int64_t acc = 0;
    for (int64_t e0 = beg + lane; e0 < end; e0 += (int64_t)BM_WAVE * kUnroll) {
        int2 ed[kUnroll];
        int32_t gx[kUnroll];
#pragma unroll
        for (int j = 0; j < kUnroll; j++) {
            int64_t e = e0 + (int64_t)BM_WAVE * j;
            ed[j] = e < end ? adj[e] : make_int2(0, 0);
        }
#pragma unroll
        for (int j = 0; j < kUnroll; j++) gx[j] = pi[ed[j].x];
#pragma unroll
        for (int j = 0; j < kUnroll; j++) {
            const int32_t x = ed[j].x, w = ed[j].y;
            if (x == 0 || x == u || x == v) continue;
            acc += (int64_t)contrib(w, probe(lds, Tn, gx[j])) - (int64_t)contrib(w, probe(lds, To, gx[j]));
        }
    }
    return acc;
}

template <int MODE>
__global__ void __launch_bounds__(512) k_swap(BmDevGraph g, const int2* __restrict__ cand, uint64_t seed,
                                             int64_t k, int64_t* __restrict__ delta, uint32_t cap) {
    extern __shared__ int4 smem4[];
    const int lane = threadIdx.x & (BM_WAVE - 1);
    const int wave = threadIdx.x / BM_WAVE;
    const int wpb = blockDim.x / BM_WAVE;
    int2* lds = reinterpret_cast<int2*>(smem4) + (size_t)wave * cap;
    const int64_t nw = (int64_t)gridDim.x * wpb;
    for (int64_t c = (int64_t)blockIdx.x * wpb + wave; c < k; c += nw) {
        int32_t u, v;
        if (MODE == BM_CAND_LIST) {
            int2 p = cand[c];
            u = p.x;
            v = p.y;
        } else {
            bm_uniform_pair(seed, (uint64_t)c, g.n, &u, &v);
        }
        u = uni(u);
        v = uni(v);
        if (u == v) {
            if (lane == 0) delta[c] = 0;
            continue;
        }
        const int32_t f1 = uni(g.pi[u]), f2 = uni(g.pi[v]);
        int64_t acc = 0;
        {
            Tab t[2] = {tab_of(g.f_out_desc, g.f_out_tab, f1), tab_of(g.f_out_desc, g.f_out_tab, f2)};
            place<2>(t, cap);
            wave_sync();
            stage(lds, t[0], lane);
            stage(lds, t[1], lane);
            wave_sync();
            const int64_t ub = uni64(g.m_out_ptr[u]), ue = uni64(g.m_out_ptr[u + 1]);
            const int64_t vb = uni64(g.m_out_ptr[v]), ve = uni64(g.m_out_ptr[v + 1]);
            acc += swap_scan_out(g.m_out, g.pi, ub, ue, lds, t[0], t[1], u, v, f1, f2, lane);
            acc += swap_scan_out(g.m_out, g.pi, vb, ve, lds, t[1], t[0], u, v, f1, f2, lane);
        }
        {
            Tab t[2] = {tab_of(g.f_in_desc, g.f_in_tab, f1), tab_of(g.f_in_desc, g.f_in_tab, f2)};
            place<2>(t, cap);
            wave_sync();
            stage(lds, t[0], lane);
            stage(lds, t[1], lane);
            wave_sync();
            const int64_t ub = uni64(g.m_in_ptr[u]), ue = uni64(g.m_in_ptr[u + 1]);
            const int64_t vb = uni64(g.m_in_ptr[v]), ve = uni64(g.m_in_ptr[v + 1]);
            acc += swap_scan_in(g.m_in, g.pi, ub, ue, lds, t[0], t[1], u, v, lane);
            acc += swap_scan_in(g.m_in, g.pi, vb, ve, lds, t[1], t[0], u, v, lane);
        }
        acc = wave_sum64(acc);
        if (lane == 0) delta[c] = acc;
    }
}

/* ------------------------------------------------------------------ */
/* 3-vertex permutation deltas                                         */
/* ------------------------------------------------------------------ */

/* New female index (0..2) of position i under permutation p (1..5), in the
 * numbering of calculateThreeNodeSwapKernel's table
 * (src/graph-re-alignment-gpu-3-swap.cu:167-174). */
__host__ __device__ __forceinline__ int perm_of(int p, int i) {
    /* p: 1 (0,2,1)  2 (1,0,2)  3 (1,2,0)  4 (2,0,1)  5 (2,1,0) */
    switch (p) {
        case 1: return i == 0 ? 0 : (i == 1 ? 2 : 1);
        case 2: return i == 0 ? 1 : (i == 1 ? 0 : 2);
        case 3: return i == 0 ? 1 : (i == 1 ? 2 : 0);
        case 4: return i == 0 ? 2 : (i == 1 ? 0 : 1);
        case 5: return i == 0 ? 2 : (i == 1 ? 1 : 0);
        default: return i;
    }
}

template <typename T>
__device__ __forceinline__ T sel3(int j, const T& a, const T& b, const T& c) {
    return j == 0 ? a : (j == 1 ? b : c);
}

struct Rot3Acc {
    int64_t s[3][3];   /* s[i][j]: external contributions of position i if it maps to female j */
    int64_t in[6];     /* internal-edge delta per permutation 1..5 (index 0 unused) */
};

__device__ __forceinline__ void rot3_scan_out(const int2* __restrict__ adj, const int32_t* __restrict__ pi,
                                              int64_t beg, int64_t end, const int2* lds, const Tab* T,
                                              const int32_t* vtx, const int32_t* f, int i, Rot3Acc& A,
                                              int lane) {
    int64_t a0 = 0, a1 = 0, a2 = 0;
    for (int64_t e0 = beg + lane; e0 < end; e0 += (int64_t)BM_WAVE * 2) {
        int2 ed[2];
        int32_t gx[2];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int64_t e = e0 + (int64_t)BM_WAVE * j;
            ed[j] = e < end ? adj[e] : make_int2(0, 0);
        }
#pragma unroll
        for (int j = 0; j < 2; j++) gx[j] = pi[ed[j].x];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int32_t x = ed[j].x, w = ed[j].y;
            if (x == 0) continue;
            const int jx = x == vtx[0] ? 0 : (x == vtx[1] ? 1 : (x == vtx[2] ? 2 : -1));
            if (jx < 0) {
                a0 += contrib(w, probe(lds, T[0], gx[j]));
                a1 += contrib(w, probe(lds, T[1], gx[j]));
                a2 += contrib(w, probe(lds, T[2], gx[j]));
            } else {
                const int64_t old = contrib(w, probe(lds, sel3(i, T[0], T[1], T[2]), sel3(jx, f[0], f[1], f[2])));
#pragma unroll
                for (int p = 1; p <= 5; p++) {
                    const int ni = perm_of(p, i), nj = perm_of(p, jx);
                    A.in[p] += (int64_t)contrib(w, probe(lds, sel3(ni, T[0], T[1], T[2]), sel3(nj, f[0], f[1], f[2]))) - old;
                }
            }
        }
    }
    A.s[i][0] += a0;
    A.s[i][1] += a1;
    A.s[i][2] += a2;
}

__device__ __forceinline__ void rot3_scan_in(const int2* __restrict__ adj, const int32_t* __restrict__ pi,
                                             int64_t beg, int64_t end, const int2* lds, const Tab* T,
                                             const int32_t* vtx, int i, Rot3Acc& A, int lane) {
    int64_t a0 = 0, a1 = 0, a2 = 0;
    for (int64_t e0 = beg + lane; e0 < end; e0 += (int64_t)BM_WAVE * 2) {
        int2 ed[2];
        int32_t gx[2];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int64_t e = e0 + (int64_t)BM_WAVE * j;
            ed[j] = e < end ? adj[e] : make_int2(0, 0);
        }
#pragma unroll
        for (int j = 0; j < 2; j++) gx[j] = pi[ed[j].x];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int32_t x = ed[j].x, w = ed[j].y;
            if (x == 0 || x == vtx[0] || x == vtx[1] || x == vtx[2]) continue;
            a0 += contrib(w, probe(lds, T[0], gx[j]));
            a1 += contrib(w, probe(lds, T[1], gx[j]));
            a2 += contrib(w, probe(lds, T[2], gx[j]));
        }
    }
    A.s[i][0] += a0;
    A.s[i][1] += a1;
    A.s[i][2] += a2;
}

__global__ void __launch_bounds__(512) k_rot3(BmDevGraph g, int4* __restrict__ cand, int64_t k,
                                             int64_t* __restrict__ best, uint8_t* __restrict__ perm,
                                             int write_perm_into_cand, uint32_t cap) {
    extern __shared__ int4 smem4[];
    const int lane = threadIdx.x & (BM_WAVE - 1);
    const int wave = threadIdx.x / BM_WAVE;
    const int wpb = blockDim.x / BM_WAVE;
    int2* lds = reinterpret_cast<int2*>(smem4) + (size_t)wave * cap;
    const int64_t nw = (int64_t)gridDim.x * wpb;
    for (int64_t c = (int64_t)blockIdx.x * wpb + wave; c < k; c += nw) {
        const int4 q = cand[c];
        int32_t vtx[3] = {uni(q.x), uni(q.y), uni(q.z)};
        int32_t f[3] = {uni(g.pi[vtx[0]]), uni(g.pi[vtx[1]]), uni(g.pi[vtx[2]])};
        Rot3Acc A;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) A.s[i][j] = 0;
#pragma unroll
        for (int p = 0; p < 6; p++) A.in[p] = 0;
        {
            Tab T[3] = {tab_of(g.f_out_desc, g.f_out_tab, f[0]), tab_of(g.f_out_desc, g.f_out_tab, f[1]),
                        tab_of(g.f_out_desc, g.f_out_tab, f[2])};
            place<3>(T, cap);
            wave_sync();
            stage(lds, T[0], lane);
            stage(lds, T[1], lane);
            stage(lds, T[2], lane);
            wave_sync();
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const int64_t b = uni64(g.m_out_ptr[vtx[i]]), e = uni64(g.m_out_ptr[vtx[i] + 1]);
                rot3_scan_out(g.m_out, g.pi, b, e, lds, T, vtx, f, i, A, lane);
            }
        }
        {
            Tab T[3] = {tab_of(g.f_in_desc, g.f_in_tab, f[0]), tab_of(g.f_in_desc, g.f_in_tab, f[1]),
                        tab_of(g.f_in_desc, g.f_in_tab, f[2])};
            place<3>(T, cap);
            wave_sync();
            stage(lds, T[0], lane);
            stage(lds, T[1], lane);
            stage(lds, T[2], lane);
            wave_sync();
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const int64_t b = uni64(g.m_in_ptr[vtx[i]]), e = uni64(g.m_in_ptr[vtx[i] + 1]);
                rot3_scan_in(g.m_in, g.pi, b, e, lds, T, vtx, i, A, lane);
            }
        }
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) A.s[i][j] = wave_sum64(A.s[i][j]);
#pragma unroll
        for (int p = 1; p <= 5; p++) A.in[p] = wave_sum64(A.in[p]);
        if (lane == 0) {
            int64_t bd = 0;
            int bp = 0;
#pragma unroll
            for (int p = 1; p <= 5; p++) {
                int64_t d = A.in[p];
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const int ni = perm_of(p, i);
                    d += sel3(ni, A.s[i][0], A.s[i][1], A.s[i][2]) - A.s[i][i];
                }
                if (d > bd) { bd = d; bp = p; }
            }
            best[c] = bd;
            if (perm) perm[c] = (uint8_t)bp;
            if (write_perm_into_cand) cand[c].w = bp;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Full score                                                          */
/* ------------------------------------------------------------------ */

__global__ void __launch_bounds__(512) k_score(BmDevGraph g, int64_t* __restrict__ acc_out, uint32_t cap) {
    extern __shared__ int4 smem4[];
    const int lane = threadIdx.x & (BM_WAVE - 1);
    const int wave = threadIdx.x / BM_WAVE;
    const int wpb = blockDim.x / BM_WAVE;
    int2* lds = reinterpret_cast<int2*>(smem4) + (size_t)wave * cap;
    const int64_t nw = (int64_t)gridDim.x * wpb;
    int64_t acc = 0;
    for (int64_t uu = (int64_t)blockIdx.x * wpb + wave + 1; uu <= g.n; uu += nw) {
        const int32_t u = (int32_t)uu;
        const int64_t b = uni64(g.m_out_ptr[u]), e = uni64(g.m_out_ptr[u + 1]);
        if (b == e) continue;
        Tab t[1] = {tab_of(g.f_out_desc, g.f_out_tab, uni(g.pi[u]))};
        place<1>(t, cap);
        wave_sync();
        stage(lds, t[0], lane);
        wave_sync();
        for (int64_t e0 = b + lane; e0 < e; e0 += BM_WAVE) {
            const int2 ed = g.m_out[e0];
            acc += contrib(ed.y, probe(lds, t[0], g.pi[ed.x]));
        }
    }
    acc = wave_sum64(acc);
    if (lane == 0 && acc != 0) atomicAdd(reinterpret_cast<unsigned long long*>(acc_out), (unsigned long long)acc);
}

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
                                                               int64_t* __restrict__ meta) {
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
        const int32_t fu = pi[p.x];
        pi[p.x] = pi[p.y];
        pi[p.y] = fu;
    }
    if (t == 0) { meta[0] = nacc; meta[1] = s_gain; }
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

hipError_t bm_launch_swap(const BmDevGraph& g, const BmLaunch& L, int mode, const int2* cand, uint64_t seed,
                          int64_t k, int64_t* delta, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    const int threads = L.waves_per_block * BM_WAVE;
    const size_t lds = (size_t)L.waves_per_block * (size_t)L.lds_slots * sizeof(int2);
    const int64_t waves_needed = k;
    int64_t blocks = (waves_needed + L.waves_per_block - 1) / L.waves_per_block;
    if (blocks > L.grid_blocks) blocks = L.grid_blocks;
    if (mode == BM_CAND_LIST) {
        hipError_t e = set_lds((const void*)k_swap<BM_CAND_LIST>, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_swap<BM_CAND_LIST>, dim3((unsigned)blocks), dim3(threads), lds, s, g, cand, seed, k,
                           delta, (uint32_t)L.lds_slots);
    } else {
        hipError_t e = set_lds((const void*)k_swap<BM_CAND_UNIFORM>, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_swap<BM_CAND_UNIFORM>, dim3((unsigned)blocks), dim3(threads), lds, s, g, cand, seed,
                           k, delta, (uint32_t)L.lds_slots);
    }
    return hipGetLastError();
}

hipError_t bm_launch_rot3(const BmDevGraph& g, const BmLaunch& L, const int4* cand, int64_t k, int64_t* best,
                          uint8_t* perm, int4* cand_out_perm, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    const int threads = L.waves_per_block * BM_WAVE;
    const size_t lds = (size_t)L.waves_per_block * (size_t)L.lds_slots * sizeof(int2);
    int64_t blocks = (k + L.waves_per_block - 1) / L.waves_per_block;
    if (blocks > L.grid_blocks) blocks = L.grid_blocks;
    hipError_t e = set_lds((const void*)k_rot3, lds);
    if (e != hipSuccess) return e;
    /* cand_out_perm, when given, is the same buffer as cand */
    hipLaunchKernelGGL(k_rot3, dim3((unsigned)blocks), dim3(threads), lds, s, g, const_cast<int4*>(cand), k, best,
                       perm, cand_out_perm ? 1 : 0, (uint32_t)L.lds_slots);
    return hipGetLastError();
}

hipError_t bm_launch_score(const BmDevGraph& g, const BmLaunch& L, int64_t* acc, hipStream_t s) {
    const int threads = L.waves_per_block * BM_WAVE;
    const size_t lds = (size_t)L.waves_per_block * (size_t)L.lds_slots * sizeof(int2);
    int64_t blocks = ((int64_t)g.n + L.waves_per_block - 1) / L.waves_per_block;
    if (blocks > L.grid_blocks) blocks = L.grid_blocks;
    if (blocks < 1) blocks = 1;
    hipError_t e = set_lds((const void*)k_score, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_score, dim3((unsigned)blocks), dim3(threads), lds, s, g, acc, (uint32_t)L.lds_slots);
    return hipGetLastError();
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
                            int64_t* out_meta, hipStream_t s) {
    if (k <= 0) {
        hipError_t e = hipMemsetAsync(out_meta, 0, 2 * sizeof(int64_t), s);
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
                       max_accept, out_acc_idx, out_meta);
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
