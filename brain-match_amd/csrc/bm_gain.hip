/*
 * Placement-gain table: exact deltas from an (n+1) x (n+1) int32 table in HBM.
 *
 *   G(u, f) = sum over male edges u->x (x != u) of c(w, f, pi(x))
 *           + sum over male edges x->u (x != u) of c(w, pi(x), f)
 *
 * i.e. the score of u's incident edges (self-loops excluded) if u alone were
 * placed on female f.  Row 0 and column 0 stay zero (unmatched).  For a move
 * that reassigns a set T of male vertices from females f_i to g_i:
 *   delta = sum_{i in T} (G(i, g_i) - G(i, f_i))
 *         + sum over male edges e = (i -> j), i != j, both in T, of
 *             c(w_e, g_i, g_j) - c(w_e, g_i, f_j) - c(w_e, f_i, g_j) + c(w_e, f_i, f_j)
 *         + sum over self-loops (i -> i), i in T, of c(w, g_i, g_i) - c(w, f_i, f_i)
 * (the G terms count an edge inside T against the other endpoint's old
 * female; the second line replaces that by the exact term).  Swaps are |T| = 2,
 * the 3-vertex permutations |T| = 3.  All terms are exact integers, so the
 * result equals S(pi') - S(pi) bit for bit.
 *
 * The reference keeps a related n x n int16 table of swap deltas on the CPU
 * (src/best-swap/score-lookup.c:246-293) and patches it after each swap
 * (:129-174).  Here the table is per placement: a move changes only the rows
 * of the moved vertices' neighbours, and the table is built and patched on the
 * device.
 *
 * Build: one block per male row; a column chunk of the row is accumulated in
 * LDS with atomic adds and written out coalesced.  A row's work is the list of
 * (male edge, female-row entry) pairs, flattened per wave (64 edges at a time,
 * prefix sum of their female-row lengths) so lanes stay busy for short and
 * long female rows alike.  Lookups: one thread per candidate.  Patches after
 * moves: the same flattening, with global integer atomic adds.
 */
#include "bm_device.h"
#include "bm_devutil.h"
#include "bm_rng.h"

using namespace bmd;

namespace {

constexpr int kBuildThreads = 512;
constexpr int kBuildWaves = kBuildThreads / BM_WAVE;

__device__ __forceinline__ int32_t probe_g(const int2* __restrict__ T, uint32_t bits, int32_t key) {
    if (!bits || key == 0) return BM_ABSENT;
    const uint32_t mask = (1u << bits) - 1u;
    uint32_t s = slot_hash(key, bits);
    for (;;) {
        const int2 e = T[s];
        if (e.x == 0) return BM_ABSENT;
        if (e.x == key) return e.y;
        s = (s + 1u) & mask;
    }
}

/* w_F(a, b) or BM_ABSENT. */
__device__ __forceinline__ int32_t female_w(const BmDevGraph& g, int32_t a, int32_t b) {
    if (a == 0 || b == 0) return BM_ABSENT;
    const uint64_t d = g.f_out_desc[a];
    return probe_g(g.f_out_tab + (d >> 6), (uint32_t)(d & 63u), b);
}

/* w_M(a, b) or BM_ABSENT. */
__device__ __forceinline__ int32_t male_w(const BmGain& G, int32_t a, int32_t b) {
    const uint64_t d = G.m_out_desc[a];
    return probe_g(G.m_out_tab + (d >> 6), (uint32_t)(d & 63u), b);
}

/* c(w, a, b) for a male weight that may be absent. */
__device__ __forceinline__ int64_t cm(const BmDevGraph& g, int32_t w, int32_t a, int32_t b) {
    if (w == BM_ABSENT) return 0;
    return contrib(w, female_w(g, a, b));
}

/* Exact correction of an edge i -> j (i != j) inside the moved set. */
__device__ __forceinline__ int64_t edge_corr(const BmDevGraph& g, int32_t w, int32_t fi, int32_t fj, int32_t gi,
                                             int32_t gj) {
    if (w == BM_ABSENT) return 0;
    return cm(g, w, gi, gj) - cm(g, w, gi, fj) - cm(g, w, fi, gj) + cm(g, w, fi, fj);
}

/* Per-wave scratch for the flattened (edge, female entry) loops. */
struct WaveScratch {
    int32_t incl[BM_WAVE];
    int32_t w[BM_WAVE];
    int32_t y[BM_WAVE];   /* per-edge target row (patch kernel) */
    int64_t lo[BM_WAVE];
};

/* One wave's batch: lane l holds a male edge of weight w (have = false for no
 * edge) whose female row is frow[rb, re).  Every entry {f, fw} with f in
 * [col0, col0 + ncols) contributes min(w, fw) at column f - col0, through
 * add(owner lane, column, value).  Rows are sorted by key, so the column
 * window is located by binary search.  y is stored for the callback. */
template <typename Add>
__device__ __forceinline__ void wave_flat(bool have, int32_t w, int32_t y, const int2* __restrict__ frow, int64_t rb,
                                          int64_t re, int32_t col0, int32_t ncols, bool windowed, WaveScratch* S,
                                          int lane, Add add) {
    int64_t lo = rb, hi = re;
    if (have && windowed) {
        int64_t a = rb, b = re;
        while (a < b) {
            const int64_t m = (a + b) >> 1;
            if (frow[m].x < col0) a = m + 1; else b = m;
        }
        lo = a;
        b = re;
        while (a < b) {
            const int64_t m = (a + b) >> 1;
            if (frow[m].x < col0 + ncols) a = m + 1; else b = m;
        }
        hi = a;
    }
    int32_t incl = have ? (int32_t)(hi - lo) : 0;
#pragma unroll
    for (int o = 1; o < BM_WAVE; o <<= 1) {
        const int32_t t = __shfl_up(incl, o, BM_WAVE);
        if (lane >= o) incl += t;
    }
    const int32_t total = __shfl(incl, BM_WAVE - 1, BM_WAVE);
    S->incl[lane] = incl;
    S->w[lane] = w;
    S->y[lane] = y;
    S->lo[lane] = lo;
    wave_sync();
    for (int32_t base = 0; base < total; base += BM_WAVE) {
        const int32_t t = base + lane;
        if (t < total) {
            int32_t a = 0, b = BM_WAVE - 1;
            while (a < b) {
                const int32_t m = (a + b) >> 1;
                if (S->incl[m] > t) b = m; else a = m + 1;
            }
            const int32_t before = a ? S->incl[a - 1] : 0;
            const int2 ent = frow[S->lo[a] + (t - before)];
            const int32_t wv = S->w[a];
            add(S->y[a], ent.x - col0, wv < ent.y ? wv : ent.y);
        }
    }
    wave_sync();
}

__global__ void __launch_bounds__(kBuildThreads) k_gain_build(BmDevGraph g, BmGain G, int32_t col0, int32_t ncols) {
    extern __shared__ int32_t row[];
    __shared__ WaveScratch scratch[kBuildWaves];
    const int tid = threadIdx.x;
    const int lane = tid & (BM_WAVE - 1);
    const int wave = tid / BM_WAVE;
    const bool windowed = !(col0 == 0 && (int64_t)ncols >= G.stride);
    WaveScratch* S = &scratch[wave];
    auto add = [&](int32_t, int32_t c, int32_t v) { atomicAdd(&row[c], v); };
    for (int32_t u = blockIdx.x + 1; u <= g.n; u += gridDim.x) {
        for (int32_t c = tid; c < ncols; c += kBuildThreads) row[c] = 0;
        __syncthreads();
        /* out-edges u -> x: female edges f -> pi(x), i.e. the in-row (CSC) of pi(x) */
        {
            const int64_t b = g.m_out_ptr[u], e = g.m_out_ptr[u + 1];
            for (int64_t e0 = b + (int64_t)wave * BM_WAVE; e0 < e; e0 += (int64_t)kBuildWaves * BM_WAVE) {
                const int64_t i = e0 + lane;
                bool have = i < e;
                int2 ed = have ? g.m_out[i] : make_int2(0, 0);
                have = have && ed.x != u;
                const int32_t gx = have ? g.pi[ed.x] : 0;
                have = have && gx != 0;
                const int64_t rb = have ? g.f_in_ptr[gx] : 0, re = have ? g.f_in_ptr[gx + 1] : 0;
                wave_flat(have, ed.y, 0, g.f_in, rb, re, col0, ncols, windowed, S, lane, add);
            }
        }
        /* in-edges x -> u: female edges pi(x) -> f, i.e. the out-row (CSR) of pi(x) */
        {
            const int64_t b = g.m_in_ptr[u], e = g.m_in_ptr[u + 1];
            for (int64_t e0 = b + (int64_t)wave * BM_WAVE; e0 < e; e0 += (int64_t)kBuildWaves * BM_WAVE) {
                const int64_t i = e0 + lane;
                bool have = i < e;
                int2 ed = have ? g.m_in[i] : make_int2(0, 0);
                have = have && ed.x != u;
                const int32_t gx = have ? g.pi[ed.x] : 0;
                have = have && gx != 0;
                const int64_t rb = have ? g.f_out_ptr[gx] : 0, re = have ? g.f_out_ptr[gx + 1] : 0;
                wave_flat(have, ed.y, 0, g.f_out, rb, re, col0, ncols, windowed, S, lane, add);
            }
        }
        __syncthreads();
        int32_t* out = G.G + (int64_t)u * G.stride + col0;
        for (int32_t c = tid; c < ncols; c += kBuildThreads) out[c] = row[c];
        __syncthreads();
    }
}

/* Patch the rows of the neighbours of each moved vertex x (pi(x): old -> new).
 * moves[i] = {x, old, new, 0}, *count moves, one block per move.  An in-edge
 * y -> x contributes c(w, f, pi(x)) to G(y, f) over the female in-row of pi(x);
 * an out-edge x -> y contributes c(w, pi(x), f) over the female out-row.  The
 * old female's terms are subtracted and the new one's added; self-loops are
 * not in the table.  Updates of different moves commute (integer atomics). */
__global__ void __launch_bounds__(kBuildThreads) k_gain_update(BmDevGraph g, BmGain G, const int4* __restrict__ moves,
                                                               const int64_t* __restrict__ count) {
    __shared__ WaveScratch scratch[kBuildWaves];
    const int tid = threadIdx.x;
    const int lane = tid & (BM_WAVE - 1);
    const int wave = tid / BM_WAVE;
    WaveScratch* S = &scratch[wave];
    const int64_t nm = *count;
    for (int64_t mi = blockIdx.x; mi < nm; mi += gridDim.x) {
        const int4 mv = moves[mi];
        const int32_t x = mv.x;
        if (mv.y == mv.z) continue;
        for (int side = 0; side < 2; side++) {
            const int32_t gx = side ? mv.z : mv.y;
            if (gx == 0) continue;
            const int32_t sign = side ? 1 : -1;
            auto add = [&](int32_t y, int32_t c, int32_t v) {
                atomicAdd(&G.G[(int64_t)y * G.stride + c], sign * v);
            };
            for (int dir = 0; dir < 2; dir++) {
                const int64_t* mptr = dir ? g.m_out_ptr : g.m_in_ptr;
                const int2* madj = dir ? g.m_out : g.m_in;
                const int64_t* fptr = dir ? g.f_out_ptr : g.f_in_ptr;
                const int2* fadj = dir ? g.f_out : g.f_in;
                const int64_t b = mptr[x], e = mptr[x + 1];
                const int64_t rb = fptr[gx], re = fptr[gx + 1];
                for (int64_t e0 = b + (int64_t)wave * BM_WAVE; e0 < e; e0 += (int64_t)kBuildWaves * BM_WAVE) {
                    const int64_t i = e0 + lane;
                    bool have = i < e;
                    const int2 ed = have ? madj[i] : make_int2(0, 0);
                    have = have && ed.x != x;
                    wave_flat(have, ed.y, ed.x, fadj, rb, re, 0, 0, false, S, lane, add);
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Lookups                                                             */
/* ------------------------------------------------------------------ */

__device__ __forceinline__ int64_t gain_swap_delta(const BmDevGraph& g, const BmGain& G, int32_t u, int32_t v) {
    if (u == v) return 0;
    const int32_t f1 = g.pi[u], f2 = g.pi[v];
    const int32_t* Gu = G.G + (int64_t)u * G.stride;
    const int32_t* Gv = G.G + (int64_t)v * G.stride;
    int64_t d = (int64_t)Gu[f2] - Gu[f1] + Gv[f1] - Gv[f2];
    const int32_t w1 = male_w(G, u, v), w2 = male_w(G, v, u);
    d += edge_corr(g, w1, f1, f2, f2, f1);     /* u -> v */
    d += edge_corr(g, w2, f2, f1, f1, f2);     /* v -> u */
    const int32_t su = G.m_self[u], sv = G.m_self[v];
    if (su != BM_ABSENT) d += contrib(su, G.f_self[f2]) - (int64_t)contrib(su, G.f_self[f1]);
    if (sv != BM_ABSENT) d += contrib(sv, G.f_self[f1]) - (int64_t)contrib(sv, G.f_self[f2]);
    return d;
}

template <int MODE>
__global__ void __launch_bounds__(256) k_gain_swap(BmDevGraph g, BmGain G, const int2* __restrict__ cand,
                                                   uint64_t seed, int64_t k, int64_t* __restrict__ delta) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < k; c += stride) {
        int32_t u, v;
        if (MODE == BM_CAND_LIST) {
            const int2 p = cand[c];
            u = p.x;
            v = p.y;
        } else {
            bm_uniform_pair(seed, (uint64_t)c, g.n, &u, &v);
        }
        delta[c] = gain_swap_delta(g, G, u, v);
    }
}

__global__ void __launch_bounds__(256) k_gain_rot3(BmDevGraph g, BmGain G, int4* __restrict__ cand, int64_t k,
                                                   int64_t* __restrict__ best, uint8_t* __restrict__ perm,
                                                   int write_perm_into_cand) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < k; c += stride) {
        const int4 q = cand[c];
        const int32_t vtx[3] = {q.x, q.y, q.z};
        const int32_t f[3] = {g.pi[q.x], g.pi[q.y], g.pi[q.z]};
        int64_t Gv[3][3];   /* Gv[i][j] = G(vtx[i], f[j]) */
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int32_t* Gr = G.G + (int64_t)vtx[i] * G.stride;
#pragma unroll
            for (int j = 0; j < 3; j++) Gv[i][j] = Gr[f[j]];
        }
        int32_t w[3][3];    /* w[i][j] = w_M(vtx[i], vtx[j]) for i != j */
        int32_t s[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            s[i] = G.m_self[vtx[i]];
#pragma unroll
            for (int j = 0; j < 3; j++) w[i][j] = i == j ? BM_ABSENT : male_w(G, vtx[i], vtx[j]);
        }
        int64_t bd = 0;
        int bp = 0;
#pragma unroll
        for (int p = 1; p <= 5; p++) {
            int32_t ng[3];
#pragma unroll
            for (int i = 0; i < 3; i++) ng[i] = f[perm_of(p, i)];
            int64_t d = 0;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                d += sel3(perm_of(p, i), Gv[i][0], Gv[i][1], Gv[i][2]) - Gv[i][i];
                if (s[i] != BM_ABSENT) d += contrib(s[i], G.f_self[ng[i]]) - (int64_t)contrib(s[i], G.f_self[f[i]]);
#pragma unroll
                for (int j = 0; j < 3; j++)
                    if (i != j) d += edge_corr(g, w[i][j], f[i], f[j], ng[i], ng[j]);
            }
            if (d > bd) { bd = d; bp = p; }
        }
        best[c] = bd;
        if (perm) perm[c] = (uint8_t)bp;
        if (write_perm_into_cand) cand[c].w = bp;
    }
}

__global__ void k_moves_from_swap(const int32_t* __restrict__ pi_before, int32_t u, int32_t v, int4* moves,
                                  int64_t* count) {
    const int32_t f1 = pi_before[u], f2 = pi_before[v];
    moves[0] = make_int4(u, f1, f2, 0);
    moves[1] = make_int4(v, f2, f1, 0);
    *count = 2;
}

__global__ void k_moves_from_perm3(const int32_t* __restrict__ pi_before, int32_t a, int32_t b, int32_t c, int32_t p,
                                   int4* moves, int64_t* count) {
    const int32_t v[3] = {a, b, c};
    const int32_t f[3] = {pi_before[a], pi_before[b], pi_before[c]};
    for (int i = 0; i < 3; i++) moves[i] = make_int4(v[i], f[i], f[perm_of(p, i)], 0);
    *count = 3;
}

}  // namespace

/* ------------------------------------------------------------------ */
/* Launchers                                                           */
/* ------------------------------------------------------------------ */

int32_t bm_gain_chunk_cols(int64_t stride) {
    const int64_t cap = 36 * 1024;   /* 144 KiB of int32 per block */
    return (int32_t)(stride < cap ? stride : cap);
}

hipError_t bm_launch_gain_build(const BmDevGraph& g, const BmGain& G, int32_t num_cus, hipStream_t s) {
    const int32_t chunk = bm_gain_chunk_cols(G.stride);
    const size_t lds = (size_t)chunk * sizeof(int32_t);
    hipError_t e = hipFuncSetAttribute((const void*)k_gain_build, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_gain_build, kBuildThreads, lds);
    if (e != hipSuccess) return e;
    if (occ < 1) occ = 1;
    int64_t blocks = (int64_t)occ * (num_cus > 0 ? num_cus : 256);
    if (blocks > g.n) blocks = g.n;
    if (blocks < 1) blocks = 1;
    for (int64_t c0 = 0; c0 < G.stride; c0 += chunk) {
        const int32_t nc = (int32_t)(G.stride - c0 < chunk ? G.stride - c0 : chunk);
        hipLaunchKernelGGL(k_gain_build, dim3((unsigned)blocks), dim3(kBuildThreads), lds, s, g, G, (int32_t)c0, nc);
    }
    return hipGetLastError();
}

hipError_t bm_launch_gain_update(const BmDevGraph& g, const BmGain& G, const int4* moves, const int64_t* count,
                                 int64_t max_moves, hipStream_t s) {
    if (max_moves <= 0) return hipSuccess;
    int64_t blocks = max_moves < 4096 ? max_moves : 4096;
    hipLaunchKernelGGL(k_gain_update, dim3((unsigned)blocks), dim3(kBuildThreads), 0, s, g, G, moves, count);
    return hipGetLastError();
}

hipError_t bm_launch_gain_swap(const BmDevGraph& g, const BmGain& G, int mode, const int2* cand, uint64_t seed,
                               int64_t k, int64_t* delta, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    int64_t blocks = (k + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (mode == BM_CAND_LIST)
        hipLaunchKernelGGL(k_gain_swap<BM_CAND_LIST>, dim3((unsigned)blocks), dim3(256), 0, s, g, G, cand, seed, k,
                           delta);
    else
        hipLaunchKernelGGL(k_gain_swap<BM_CAND_UNIFORM>, dim3((unsigned)blocks), dim3(256), 0, s, g, G, cand, seed, k,
                           delta);
    return hipGetLastError();
}

hipError_t bm_launch_gain_rot3(const BmDevGraph& g, const BmGain& G, const int4* cand, int64_t k, int64_t* best,
                               uint8_t* perm, int4* cand_out_perm, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    int64_t blocks = (k + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_gain_rot3, dim3((unsigned)blocks), dim3(256), 0, s, g, G, const_cast<int4*>(cand), k, best,
                       perm, cand_out_perm ? 1 : 0);
    return hipGetLastError();
}

hipError_t bm_launch_moves_from_swap(const int32_t* pi_before, int32_t u, int32_t v, int4* moves, int64_t* count,
                                     hipStream_t s) {
    hipLaunchKernelGGL(k_moves_from_swap, dim3(1), dim3(1), 0, s, pi_before, u, v, moves, count);
    return hipGetLastError();
}

hipError_t bm_launch_moves_from_perm3(const int32_t* pi_before, int32_t a, int32_t b, int32_t c, int32_t p,
                                      int4* moves, int64_t* count, hipStream_t s) {
    hipLaunchKernelGGL(k_moves_from_perm3, dim3(1), dim3(1), 0, s, pi_before, a, b, c, p, moves, count);
    return hipGetLastError();
}
