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

constexpr int kBuildThreads = 1024;   /* build: 16 waves per row */
constexpr int kBU = 8;                 /* build: female-row entries per lane in flight */
constexpr int kPatchThreads = 512;
constexpr int kPatchWaves = kPatchThreads / BM_WAVE;

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

/* Female-row entry {f, w} in either layout: int2, or the packed word
 * (f << 15) | w used when ids are < 2^17 and weights in 0..32767. */
template <bool PACKED>
struct Row;
template <>
struct Row<false> {
    typedef int2 T;
    static __device__ __forceinline__ int32_t key(T e) { return e.x; }
    static __device__ __forceinline__ int32_t w(T e) { return e.y; }
    static __device__ __forceinline__ T none() { return make_int2(0, 0); }
};
template <>
struct Row<true> {
    typedef uint32_t T;
    static __device__ __forceinline__ int32_t key(T e) { return (int32_t)(e >> 15); }
    static __device__ __forceinline__ int32_t w(T e) { return (int32_t)(e & 0x7fffu); }
    static __device__ __forceinline__ T none() { return 0u; }
};

/* Row u: every male edge of u (out-edges against the female in-row of the
 * neighbour's image, in-edges against its out-row) is taken by one wave at a
 * time from a block-wide counter (so a long female row does not hold up the
 * block); the wave's lanes stride the row with kBU entries each in flight and
 * add min(w, fw) into the LDS copy of the row's column window. */
template <bool PACKED>
__global__ void __launch_bounds__(kBuildThreads) k_gain_build(BmDevGraph g, BmGain G, int32_t col0, int32_t ncols) {
    typedef Row<PACKED> R;
    typedef typename R::T E;
    extern __shared__ int32_t row[];
    __shared__ int32_t next_edge;
    const int tid = threadIdx.x;
    const int lane = tid & (BM_WAVE - 1);
    const bool windowed = !(col0 == 0 && (int64_t)ncols >= G.stride);
    const E* __restrict__ f_in_rows = PACKED ? (const E*)G.f_in_p : (const E*)g.f_in;
    const E* __restrict__ f_out_rows = PACKED ? (const E*)G.f_out_p : (const E*)g.f_out;
    for (int32_t u = blockIdx.x + 1; u <= g.n; u += gridDim.x) {
        for (int32_t c = tid; c < ncols; c += kBuildThreads) row[c] = 0;
        if (tid == 0) next_edge = 0;
        __syncthreads();
        const int64_t ob = g.m_out_ptr[u], nout = g.m_out_ptr[u + 1] - ob;
        const int64_t ib = g.m_in_ptr[u], L = nout + (g.m_in_ptr[u + 1] - ib);
        for (;;) {
            int32_t t = 0;
            if (lane == 0) t = atomicAdd(&next_edge, 1);
            t = uni(t);
            if (t >= L) break;
            const bool outp = t < nout;
            const int2 ed = outp ? g.m_out[ob + t] : g.m_in[ib + (t - nout)];
            const int32_t x = uni(ed.x), w = uni(ed.y);
            if (x == u) continue;                      /* self-loops are not in the table */
            const int32_t gx = uni(g.pi[x]);
            if (gx == 0) continue;
            const int64_t* __restrict__ fptr = outp ? g.f_in_ptr : g.f_out_ptr;
            const E* __restrict__ frow = outp ? f_in_rows : f_out_rows;
            int64_t rb = uni64(fptr[gx]), re = uni64(fptr[gx + 1]);
            if (windowed) {
                int64_t a = rb, b = re;
                while (a < b) {
                    const int64_t m = (a + b) >> 1;
                    if (R::key(frow[m]) < col0) a = m + 1; else b = m;
                }
                const int64_t lo = a;
                b = re;
                while (a < b) {
                    const int64_t m = (a + b) >> 1;
                    if (R::key(frow[m]) < col0 + ncols) a = m + 1; else b = m;
                }
                rb = lo;
                re = a;
            }
            for (int64_t j0 = rb; j0 < re; j0 += (int64_t)BM_WAVE * kBU) {
                E ent[kBU];
#pragma unroll
                for (int k = 0; k < kBU; k++) {
                    const int64_t j = j0 + lane + (int64_t)BM_WAVE * k;
                    ent[k] = j < re ? frow[j] : R::none();
                }
#pragma unroll
                for (int k = 0; k < kBU; k++) {
                    if (j0 + lane + (int64_t)BM_WAVE * k >= re) continue;
                    const int32_t f = R::key(ent[k]), fw = R::w(ent[k]);
                    atomicAdd(&row[f - col0], w < fw ? w : fw);
                }
            }
        }
        __syncthreads();
        int32_t* out = G.G + (int64_t)u * G.stride + col0;
        for (int32_t c = tid; c < ncols; c += kBuildThreads) out[c] = row[c];
        __syncthreads();
    }
}

__global__ void k_pack_rows(const int2* __restrict__ in, int64_t m, uint32_t* __restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride)
        out[i] = ((uint32_t)in[i].x << 15) | (uint32_t)in[i].y;
}

/* Patch the rows of the neighbours of each moved vertex x (pi(x): old -> new).
 * moves[i] = {x, old, new, 0}, *count moves, one block per move.  An in-edge
 * y -> x contributes c(w, f, pi(x)) to G(y, f) over the female in-row of pi(x);
 * an out-edge x -> y contributes c(w, pi(x), f) over the female out-row.  The
 * old female's terms are subtracted and the new one's added; self-loops are
 * not in the table.  Updates of different moves commute (integer atomics). */
__global__ void __launch_bounds__(kPatchThreads) k_gain_update(BmDevGraph g, BmGain G, const int4* __restrict__ moves,
                                                               const int64_t* __restrict__ count) {
    __shared__ WaveScratch scratch[kPatchWaves];
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
                for (int64_t e0 = b + (int64_t)wave * BM_WAVE; e0 < e; e0 += (int64_t)kPatchWaves * BM_WAVE) {
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
                                                   const int64_t* __restrict__ off, uint64_t seed, int64_t k,
                                                   int64_t* __restrict__ delta) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < k; c += stride) {
        int32_t u, v;
        if (MODE == BM_CAND_LIST) {
            const int2 p = cand[c];
            u = p.x;
            v = p.y;
        } else if (MODE == BM_CAND_UMAJOR) {
            u = umajor_owner(off, g.n, c);
            v = bm_umajor_partner(seed, (uint64_t)c, u, g.n);
        } else {
            bm_uniform_pair(seed, (uint64_t)c, g.n, &u, &v);
        }
        delta[c] = gain_swap_delta(g, G, u, v);
    }
}

/* Transposed table: GT(f, u) = G(u, f), 64 x 64 tiles through LDS. */
__global__ void __launch_bounds__(256) k_gain_transpose(const int32_t* __restrict__ Gt, int32_t* __restrict__ GT,
                                                        int64_t stride) {
    __shared__ int32_t tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int64_t tiles = (stride + 63) / 64;
    for (int64_t t = blockIdx.x; t < tiles * tiles; t += gridDim.x) {
        const int64_t br = t / tiles, bc = t % tiles;
        for (int r = ty; r < 64; r += 4) {
            const int64_t u = br * 64 + r, f = bc * 64 + tx;
            tile[r][tx] = (u < stride && f < stride) ? Gt[u * stride + f] : 0;
        }
        __syncthreads();
        for (int r = ty; r < 64; r += 4) {
            const int64_t f = bc * 64 + r, u = br * 64 + tx;
            if (f < stride && u < stride) GT[f * stride + u] = tile[tx][r];
        }
        __syncthreads();
    }
}

/* Per vertex v: {pi(v), G(v, pi(v)), male self-loop weight of v, female
 * self-loop weight of pi(v)} (BM_ABSENT for a missing loop). */
__global__ void k_gain_pis(BmDevGraph g, BmGain G, int4* __restrict__ pis) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= g.n; v += stride) {
        const int32_t f = v ? g.pi[v] : 0;
        pis[v] = make_int4(f, v ? G.G[v * G.stride + f] : 0, v ? G.m_self[v] : BM_ABSENT,
                           f ? G.f_self[f] : BM_ABSENT);
    }
}

/* Four hash-table lookups with their probe sequences interleaved (one
 * round of loads for all unresolved keys at a time); key 0 or an empty
 * table gives BM_ABSENT. */
__device__ __forceinline__ void probe4(const uint64_t* desc, const int2* const* tab, const int32_t* key,
                                       int32_t* out) {
    const int2* T[4];
    uint32_t bits[4], s[4], pend = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        out[j] = BM_ABSENT;
        bits[j] = (uint32_t)(desc[j] & 63u);
        T[j] = tab[j] + (desc[j] >> 6);
        const bool live = bits[j] != 0 && key[j] != 0;
        s[j] = live ? slot_hash(key[j], bits[j]) : 0u;
        pend |= (live ? 1u : 0u) << j;
    }
    while (pend) {
        int2 e[4];
#pragma unroll
        for (int j = 0; j < 4; j++) e[j] = ((pend >> j) & 1u) ? T[j][s[j]] : make_int2(0, 0);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (!((pend >> j) & 1u)) continue;
            if (e[j].x == 0 || e[j].x == key[j]) {
                if (e[j].x != 0) out[j] = e[j].y;
                pend &= ~(1u << j);
            } else {
                s[j] = (s[j] + 1u) & ((1u << bits[j]) - 1u);
            }
        }
    }
}

/* u-major lookups: one block per first vertex u of the stream.  The block
 * stages row G(u, .) (indexed by the partner's female), row GT(pi(u), .) =
 * G(., pi(u)) (indexed by the partner) and u's out- and in-neighbour bitmaps
 * in LDS; a pair (u, v) then costs two LDS reads, one gather of v's packed
 * {pi, G(v, pi(v)), self-loop} and the delta store.  Pairs with an edge
 * between u and v (bitmap hit) add the exact edge correction through the
 * hash tables; self-loops add their exact terms. */
constexpr int kUmThreads = 1024;
constexpr int kUmStage = 8;    /* row words per thread in flight while staging */
constexpr int kUmPairs = 4;    /* pairs per thread in flight */

__global__ void __launch_bounds__(kUmThreads) k_gain_swap_umajor(BmDevGraph g, BmGain G, const int32_t* __restrict__ GT,
                                                                 const int4* __restrict__ pis,
                                                                 const int64_t* __restrict__ off, uint64_t seed,
                                                                 int64_t* __restrict__ delta) {
    extern __shared__ int32_t sm[];
    const int32_t n = g.n;
    const int64_t stride = G.stride;
    const int32_t words = (n + 1 + 31) >> 5;
    int32_t* rowG = sm;
    int32_t* rowT = sm + stride;
    uint32_t* adj = reinterpret_cast<uint32_t*>(sm + 2 * stride);   /* [0, words): u -> v, [words, 2 words): v -> u */
    const int tid = threadIdx.x;
    for (int32_t u = blockIdx.x + 1; u <= n; u += gridDim.x) {
        const int64_t i0 = off[u], i1 = off[u + 1];
        if (i0 == i1) continue;   /* block-uniform */
        const int4 pu = pis[u];
        const int32_t f1 = pu.x;
        const int32_t F11 = pu.w;                           /* w_F(f1, f1) */
        const uint64_t du = G.m_out_desc[u], df1 = f1 ? g.f_out_desc[f1] : 0;
        const int32_t* __restrict__ Gu = G.G + (int64_t)u * stride;
        const int32_t* __restrict__ Tu = GT + (int64_t)f1 * stride;
        for (int64_t c0 = tid; c0 < stride; c0 += (int64_t)kUmThreads * kUmStage) {
            int32_t a[kUmStage], t[kUmStage];
#pragma unroll
            for (int j = 0; j < kUmStage; j++) {
                const int64_t c = c0 + (int64_t)kUmThreads * j;
                a[j] = c < stride ? Gu[c] : 0;
                t[j] = c < stride ? Tu[c] : 0;
            }
#pragma unroll
            for (int j = 0; j < kUmStage; j++) {
                const int64_t c = c0 + (int64_t)kUmThreads * j;
                if (c < stride) {
                    rowG[c] = a[j];
                    rowT[c] = t[j];
                }
            }
        }
        for (int32_t w = tid; w < 2 * words; w += kUmThreads) adj[w] = 0;
        __syncthreads();
        for (int64_t e = g.m_out_ptr[u] + tid; e < g.m_out_ptr[u + 1]; e += kUmThreads) {
            const int32_t x = g.m_out[e].x;
            atomicOr(&adj[x >> 5], 1u << (x & 31));
        }
        for (int64_t e = g.m_in_ptr[u] + tid; e < g.m_in_ptr[u + 1]; e += kUmThreads) {
            const int32_t x = g.m_in[e].x;
            atomicOr(&adj[words + (x >> 5)], 1u << (x & 31));
        }
        __syncthreads();
        for (int64_t ib = i0 + tid; ib < i1; ib += (int64_t)kUmThreads * kUmPairs) {
            int32_t v[kUmPairs];
            int4 pv[kUmPairs];
#pragma unroll
            for (int j = 0; j < kUmPairs; j++) {   /* all gathers in flight before any use */
                const int64_t i = ib + (int64_t)kUmThreads * j;
                v[j] = i < i1 ? bm_umajor_partner(seed, (uint64_t)i, u, n) : 0;
                pv[j] = i < i1 ? pis[v[j]] : make_int4(0, 0, BM_ABSENT, 0);
            }
#pragma unroll
            for (int j = 0; j < kUmPairs; j++) {
                const int64_t i = ib + (int64_t)kUmThreads * j;
                if (i >= i1) continue;
                const int32_t f2 = pv[j].x, x = v[j];
                const int32_t F22 = pv[j].w;                /* w_F(f2, f2) */
                int64_t d = (int64_t)rowG[f2] - pu.y + rowT[x] - pv[j].y;
                const bool uv = (adj[x >> 5] >> (x & 31)) & 1u;
                const bool vu = (adj[words + (x >> 5)] >> (x & 31)) & 1u;
                if (uv || vu) {
                    /* edges between u and v: w_M(u, v), w_M(v, u), w_F(f1, f2), w_F(f2, f1) */
                    const bool ff = f1 != 0 && f2 != 0;
                    const uint64_t desc[4] = {du, G.m_out_desc[x], df1, f2 ? g.f_out_desc[f2] : 0};
                    const int2* tab[4] = {G.m_out_tab, G.m_out_tab, g.f_out_tab, g.f_out_tab};
                    const int32_t key[4] = {uv ? x : 0, vu ? u : 0, ff ? f2 : 0, ff ? f1 : 0};
                    int32_t r[4];
                    probe4(desc, tab, key, r);
                    const int32_t F12 = r[2], F21 = r[3];
                    if (uv) {
                        const int32_t w = r[0];
                        d += (int64_t)contrib(w, F21) - contrib(w, F22) - contrib(w, F11) + contrib(w, F12);
                    }
                    if (vu) {
                        const int32_t w = r[1];
                        d += (int64_t)contrib(w, F12) - contrib(w, F11) - contrib(w, F22) + contrib(w, F21);
                    }
                }
                if (pu.z != BM_ABSENT) d += contrib(pu.z, F22) - (int64_t)contrib(pu.z, F11);
                if (pv[j].z != BM_ABSENT) d += contrib(pv[j].z, F11) - (int64_t)contrib(pv[j].z, F22);
                delta[i] = d;
            }
        }
        __syncthreads();
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
    const int64_t cap = 20000;   /* 80,000 bytes of int32: two blocks per CU (160 KiB of LDS) */
    return (int32_t)(stride < cap ? stride : cap);
}

hipError_t bm_launch_gain_build(const BmDevGraph& g, const BmGain& G, int32_t num_cus, hipStream_t s) {
    const int32_t chunk = bm_gain_chunk_cols(G.stride);
    const size_t lds = (size_t)chunk * sizeof(int32_t);
    const void* kern = G.f_out_p ? (const void*)k_gain_build<true> : (const void*)k_gain_build<false>;
    hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kBuildThreads, lds);
    if (e != hipSuccess) return e;
    if (occ < 1) occ = 1;
    int64_t blocks = (int64_t)occ * (num_cus > 0 ? num_cus : 256);
    if (blocks > g.n) blocks = g.n;
    if (blocks < 1) blocks = 1;
    for (int64_t c0 = 0; c0 < G.stride; c0 += chunk) {
        const int32_t nc = (int32_t)(G.stride - c0 < chunk ? G.stride - c0 : chunk);
        if (G.f_out_p)
            hipLaunchKernelGGL(k_gain_build<true>, dim3((unsigned)blocks), dim3(kBuildThreads), lds, s, g, G,
                               (int32_t)c0, nc);
        else
            hipLaunchKernelGGL(k_gain_build<false>, dim3((unsigned)blocks), dim3(kBuildThreads), lds, s, g, G,
                               (int32_t)c0, nc);
    }
    return hipGetLastError();
}

hipError_t bm_launch_gain_update(const BmDevGraph& g, const BmGain& G, const int4* moves, const int64_t* count,
                                 int64_t max_moves, hipStream_t s) {
    if (max_moves <= 0) return hipSuccess;
    int64_t blocks = max_moves < 4096 ? max_moves : 4096;
    hipLaunchKernelGGL(k_gain_update, dim3((unsigned)blocks), dim3(kPatchThreads), 0, s, g, G, moves, count);
    return hipGetLastError();
}

hipError_t bm_launch_gain_swap(const BmDevGraph& g, const BmGain& G, int mode, const int2* cand, const int64_t* off,
                               uint64_t seed, int64_t k, int64_t* delta, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    int64_t blocks = (k + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (mode == BM_CAND_LIST)
        hipLaunchKernelGGL(k_gain_swap<BM_CAND_LIST>, dim3((unsigned)blocks), dim3(256), 0, s, g, G, cand, off, seed,
                           k, delta);
    else if (mode == BM_CAND_UMAJOR)
        hipLaunchKernelGGL(k_gain_swap<BM_CAND_UMAJOR>, dim3((unsigned)blocks), dim3(256), 0, s, g, G, cand, off, seed,
                           k, delta);
    else
        hipLaunchKernelGGL(k_gain_swap<BM_CAND_UNIFORM>, dim3((unsigned)blocks), dim3(256), 0, s, g, G, cand, off,
                           seed, k, delta);
    return hipGetLastError();
}

size_t bm_gain_umajor_lds(int32_t n) { return 2 * ((size_t)n + 1) * 4 + 2 * (size_t)((n + 1 + 31) >> 5) * 4; }

hipError_t bm_launch_gain_transpose(const BmGain& G, int32_t* GT, int32_t num_cus, hipStream_t s) {
    const int64_t tiles = (G.stride + 63) / 64;
    int64_t blocks = tiles * tiles;
    const int64_t cap = (int64_t)(num_cus > 0 ? num_cus : 256) * 16;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_gain_transpose, dim3((unsigned)blocks), dim3(256), 0, s, G.G, GT, G.stride);
    return hipGetLastError();
}

hipError_t bm_launch_gain_pis(const BmDevGraph& g, const BmGain& G, int4* pis, hipStream_t s) {
    int64_t blocks = ((int64_t)g.n + 256) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_gain_pis, dim3((unsigned)blocks), dim3(256), 0, s, g, G, pis);
    return hipGetLastError();
}

hipError_t bm_launch_gain_swap_umajor(const BmDevGraph& g, const BmGain& G, const int32_t* GT, const int4* pis,
                                      const int64_t* off, uint64_t seed, int64_t* delta, int32_t num_cus,
                                      hipStream_t s) {
    const size_t lds = bm_gain_umajor_lds(g.n);
    hipError_t e = hipFuncSetAttribute((const void*)k_gain_swap_umajor, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_gain_swap_umajor, kUmThreads, lds);
    if (e != hipSuccess) return e;
    if (occ < 1) occ = 1;
    int64_t blocks = (int64_t)occ * (num_cus > 0 ? num_cus : 256);
    if (blocks > g.n) blocks = g.n;
    hipLaunchKernelGGL(k_gain_swap_umajor, dim3((unsigned)blocks), dim3(kUmThreads), lds, s, g, G, GT, pis, off, seed,
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

hipError_t bm_launch_pack_rows(const int2* in, int64_t m, uint32_t* out, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    int64_t blocks = (m + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_pack_rows, dim3((unsigned)blocks), dim3(256), 0, s, in, m, out);
    return hipGetLastError();
}
