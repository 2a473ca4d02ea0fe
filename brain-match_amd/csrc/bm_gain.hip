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
 * Build: a work list first gives every male edge the female row it expands
 * (one lane per edge, so the dependent loads of all edges overlap).  The
 * longest female rows (hubs adjacent to a large share of the vertices, which
 * carry most of the expanded entries) are also kept as dense 16-bit column
 * arrays: a tile pass sums their edges for 16 male rows at a time with
 * packed 16-bit min/add, one load of a dense row serving the 16 rows, and
 * writes the sums to the table.  Then one block per male row starts a column
 * chunk of the row in LDS from those sums (or zero) and adds the other edges
 * with atomic adds, its waves taking the row's edges one at a time with the
 * next edge's entry already loading, and writes the chunk out coalesced.  Lookups: one
 * thread per candidate, or one block per first vertex of a u-major stream.
 * Patches after moves: (male edge, female-row entry) pairs flattened per wave
 * (64 edges at a time, prefix sum of their female-row lengths), with global
 * integer atomic adds.
 */
#include "bm_device.h"
#include "bm_devutil.h"
#include "bm_rng.h"

using namespace bmd;

namespace {

constexpr int kBuildThreads = 1024;   /* build: 16 waves per row */
constexpr int kDescChunk = BM_DESC_CHUNK;   /* work-list slots per wave of k_gain_edge_desc */
constexpr int kDescLong = 1024;             /* rows longer than this lead their work-list chunk */
constexpr int kPatchThreads = 512;
constexpr int kTileRows = 16;       /* dense tile kernel: male rows per block */
constexpr int kTileThreads = 256;
constexpr int kPatchWaves = kPatchThreads / BM_WAVE;

/* Word of the dense adjacency bitmap holding the pair {a, b}: bit (max & 31)
 * of row min(a, b).  The bitmap is symmetric, set when a -> b or b -> a is a
 * male edge, so one load answers both directions. */
__device__ __forceinline__ uint32_t adj_word(const BmGain& G, int32_t a, int32_t b) {
    const int32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return G.m_adj[(int64_t)lo * G.m_adj_words + (hi >> 5)];
}

/* False when the dense adjacency bitmap shows no male edge between a and b;
 * true when it shows one (either direction) or there is no bitmap. */
__device__ __forceinline__ bool male_maybe(const BmGain& G, int32_t a, int32_t b) {
    return !G.m_adj || ((adj_word(G, a, b) >> ((a < b ? b : a) & 31)) & 1u);
}

/* K hash lookups at once: key[k] in table (T[k], bits[k]) where go[k], else
 * BM_ABSENT without a load.  Every first slot is loaded before any collision
 * chain is followed, so the K lookups cost one memory latency unless a key
 * sits past its home slot. */
template <int K>
__device__ __forceinline__ void probe_many(const int2* const* T, const uint32_t* bits, const int32_t* key,
                                           const bool* go, int32_t* out) {
    int2 e[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        e[k] = make_int2(0, 0);
        if (go[k] && bits[k] != 0u && key[k] != 0) e[k] = T[k][slot_hash(key[k], bits[k])];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (e[k].x != 0 && e[k].x != key[k]) {   /* past its home slot: walk the chain */
            const uint32_t mask = (1u << bits[k]) - 1u;
            uint32_t s = slot_hash(key[k], bits[k]);
            do {
                s = (s + 1u) & mask;
                e[k] = T[k][s];
            } while (e[k].x != 0 && e[k].x != key[k]);
        }
        out[k] = e[k].x != 0 ? e[k].y : BM_ABSENT;
    }
}

/* Weights inside a triple of male vertices v[0..2] with images f[0..2]:
 * w[i][j] = w_M(v[i], v[j]) for the pairs with have[i][j] (the others are
 * known absent and cost nothing) and, when any such pair exists, the nine
 * female weights F[x][y] = w_F(f[x], f[y]) (the diagonal is the female
 * self-loops fs[x]); BM_ABSENT elsewhere.  md / fd are the hash-table
 * descriptors of the rows of v[i] in M and f[x] in F (loaded by the caller
 * with its other gathers), so the up to 12 probes cost one memory latency.
 * v[2] = f[2] = 0 serves a pair. */
__device__ __forceinline__ void triple_weights(const BmDevGraph& g, const BmGain& G, const int32_t* v,
                                               const uint64_t* md, const int32_t* f, const uint64_t* fd,
                                               const int32_t* fs, const bool (*have)[3], int32_t (*w)[3],
                                               int32_t (*F)[3]) {
    bool any = false;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            w[i][j] = BM_ABSENT;
            F[i][j] = BM_ABSENT;
            if (i != j) any = any || have[i][j];
        }
    if (!any) return;
    const int2* T[12];
    uint32_t bits[12];
    int32_t key[12], out[12];
    bool go[12];
    int k = 0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            if (i == j) continue;
            T[k] = G.m_out_tab + (md[i] >> 6);
            bits[k] = (uint32_t)(md[i] & 63u);
            key[k] = v[j];
            go[k] = have[i][j];
            T[k + 6] = g.f_out_tab + (fd[i] >> 6);
            bits[k + 6] = (uint32_t)(fd[i] & 63u);
            key[k + 6] = f[j];
            go[k + 6] = f[i] != 0;
            k++;
        }
    probe_many<12>(T, bits, key, go, out);
    k = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        F[i][i] = fs[i];
#pragma unroll
        for (int j = 0; j < 3; j++)
            if (i != j) {
                w[i][j] = out[k];
                F[i][j] = out[k + 6];
                k++;
            }
    }
}

/* Female-row entry {f, w} in one of three layouts: int2; the packed word
 * (f << 15) | w (ids < 2^17, weights in 0..32767); or (w << 19) | 4 f (ids
 * < 2^17, weights in 0..8191), whose low bits are already the byte address
 * of column f in an LDS row at address 0.  addr() is that byte address,
 * load() reads entry i of a chunk through a buffer resource whose record
 * count is the chunk's length in bytes: reads past the chunk return zero
 * instead of touching memory, and every lane's loads share one offset
 * register. */
typedef int V2i __attribute__((ext_vector_type(2)));
constexpr int32_t kBufWord3 = 0x00020000;   /* gfx9 raw-buffer resource word 3 */

template <int LAYOUT>
struct Row;
template <>
struct Row<BM_ROWS_INT2> {
    typedef int2 T;
    static __device__ __forceinline__ int32_t key(T e) { return e.x; }
    static __device__ __forceinline__ uint32_t addr(T e) { return (uint32_t)e.x << 2; }
    static __device__ __forceinline__ int32_t w(T e) { return e.y; }
    static __device__ __forceinline__ T load(__amdgpu_buffer_rsrc_t r, int32_t i) {
        const V2i v = __builtin_amdgcn_raw_buffer_load_b64(r, i * 8, 0, 0);
        return make_int2(v.x, v.y);
    }
};
template <>
struct Row<BM_ROWS_P15> {
    typedef uint32_t T;
    static __device__ __forceinline__ int32_t key(T e) { return (int32_t)(e >> 15); }
    static __device__ __forceinline__ uint32_t addr(T e) { return (e >> 13) & 0x7fffcu; }
    static __device__ __forceinline__ int32_t w(T e) { return (int32_t)(e & 0x7fffu); }
    static __device__ __forceinline__ T load(__amdgpu_buffer_rsrc_t r, int32_t i) {
        return (T)__builtin_amdgcn_raw_buffer_load_b32(r, i * 4, 0, 0);
    }
};
template <>
struct Row<BM_ROWS_P19> {
    typedef uint32_t T;
    static __device__ __forceinline__ int32_t key(T e) { return (int32_t)((e & 0x7ffffu) >> 2); }
    static __device__ __forceinline__ uint32_t addr(T e) { return e & 0x7ffffu; }
    static __device__ __forceinline__ int32_t w(T e) { return (int32_t)(e >> 19); }
    static __device__ __forceinline__ T load(__amdgpu_buffer_rsrc_t r, int32_t i) {
        return (T)__builtin_amdgcn_raw_buffer_load_b32(r, i * 4, 0, 0);
    }
};

/* LDS integer add at a byte address (the build's row window sits at LDS
 * address 0: the kernel has no static LDS). */
__device__ __forceinline__ void lds_add(uint32_t byte_addr, int32_t v) {
    asm volatile("ds_add_u32 %0, %1" ::"v"(byte_addr), "v"(v) : "memory");
}

/* Column-window boundaries of the female rows: thread (r, k) finds the
 * first entry of row r whose column is >= k * cols by binary search (k =
 * nwin gives the row's length). */
__global__ void k_row_bounds(const int64_t* __restrict__ ptr, const int2* __restrict__ rows, int32_t n, int32_t cols,
                             int32_t nwin, int32_t* __restrict__ bnd) {
    const int64_t total = ((int64_t)n + 1) * (nwin + 1);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int32_t r = (int32_t)(i / (nwin + 1)), k = (int32_t)(i % (nwin + 1));
        const int64_t rb = ptr[r], re = ptr[r + 1];
        int64_t a = rb, b = re;
        if (k < nwin) {
            const int64_t c0 = (int64_t)k * cols;
            while (a < b) {
                const int64_t m = (a + b) >> 1;
                if (rows[m].x < c0) a = m + 1; else b = m;
            }
        } else {
            a = re;
        }
        bnd[i] = (int32_t)(a - rb);
    }
}

/* Work list of the build: for row u, at base(u) = m_out_ptr[u] +
 * m_in_ptr[u], one entry per male edge whose female row is sparse and has
 * entries in the column window (row_fill[u] of them): the female row the
 * edge expands (the in-row of the neighbour's image for an out-edge, its
 * out-row for an in-edge) clipped to the window (the row's window
 * boundaries are precomputed, k_row_bounds), as {start lo, start hi |
 * in-edge flag << 31, length, w}.
 * One wave per chunk of at most kDescChunk slots of one vertex's list (the
 * chunks are fixed by the male graph, so a hub's list is spread over many
 * waves), kDescChunk / 64 slots per lane with each level of the dependent
 * loads (edge, pi of the neighbour, row bounds) issued for all of them
 * together. */
template <int LAYOUT>
__global__ void __launch_bounds__(256) k_gain_edge_desc(BmDevGraph g, BmGain G, int32_t col0, int32_t ncols,
                                                        int32_t win) {
    constexpr int kJ = kDescChunk / BM_WAVE;   /* edge slots per lane, their loads in flight together */
    const int lane = threadIdx.x & (BM_WAVE - 1);
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / BM_WAVE;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / BM_WAVE;
    const bool windowed = !(col0 == 0 && (int64_t)ncols >= G.stride);
    for (int64_t ci = wave; ci < G.n_echunks; ci += nwaves) {
        const int2 ch = G.echunks[ci];   /* {u, first slot}: up to kDescChunk slots of u's list */
        const int32_t u = uni(ch.x);
        const int64_t ob = g.m_out_ptr[u], nout = g.m_out_ptr[u + 1] - ob;
        const int64_t ib = g.m_in_ptr[u], L = nout + (g.m_in_ptr[u + 1] - ib);
        const int64_t t0 = uni(ch.y), t1 = t0 + kDescChunk < L ? t0 + kDescChunk : L;
        int4* out = G.edesc + ob + ib;
        int2 ed[kJ];
        int32_t gx[kJ], lo32[kJ], hi32[kJ], di[kJ];
        int64_t rb[kJ], re[kJ];
#pragma unroll
        for (int j = 0; j < kJ; j++) {
            const int64_t t = t0 + lane + BM_WAVE * j;
            ed[j] = t < t1 ? (t < nout ? g.m_out[ob + t] : g.m_in[ib + (t - nout)]) : make_int2(0, 0);
        }
#pragma unroll
        for (int j = 0; j < kJ; j++)   /* self-loops and empty slots are not in the table */
            gx[j] = (ed[j].x != 0 && ed[j].x != u) ? g.pi[ed[j].x] : 0;
#pragma unroll
        for (int j = 0; j < kJ; j++) {
            const bool outp = t0 + lane + BM_WAVE * j < nout;
            const int64_t* fptr = outp ? g.f_in_ptr : g.f_out_ptr;
            rb[j] = gx[j] ? fptr[gx[j]] : 0;
            re[j] = gx[j] ? fptr[gx[j] + 1] : 0;
            di[j] = (G.ndense > 0 && gx[j]) ? (outp ? G.f_in_dense : G.f_out_dense)[gx[j]] : -1;
            if (windowed) {   /* the row's part inside the window, from its precomputed boundaries */
                const int32_t* bnd = (outp ? G.f_in_bnd : G.f_out_bnd) + (int64_t)gx[j] * (G.nwin + 1) + win;
                lo32[j] = gx[j] ? bnd[0] : 0;
                hi32[j] = gx[j] ? bnd[1] : 0;
            }
        }
        if (windowed) {
#pragma unroll
            for (int j = 0; j < kJ; j++) {
                re[j] = rb[j] + hi32[j];
                rb[j] += lo32[j];
            }
        }
        /* Only edges that expand a non-empty sparse row go on the list: a
         * chunk reserves places for them at the front of u's list with one
         * atomic on the row's fill count (the row walk takes that many), and
         * within the chunk rows longer than kDescLong entries go first (the
         * build takes entries in list order, so long rows start early and a
         * male row's last entries are short ones).  Dense rows' edges only
         * set their tile weight; self-loops, unmatched neighbours and rows
         * with no entry in the window are dropped. */
        uint64_t mlong[kJ], mshort[kJ];
        int32_t nlong = 0, nshort = 0;
#pragma unroll
        for (int j = 0; j < kJ; j++) {
            const bool keep = t0 + lane + BM_WAVE * j < t1 && di[j] < 0 && re[j] > rb[j];
            const bool lng = keep && re[j] - rb[j] > kDescLong;
            mlong[j] = __ballot(lng);
            mshort[j] = __ballot(keep && !lng);
            nlong += __popcll(mlong[j]);
            nshort += __popcll(mshort[j]);
        }
        int32_t first = 0;
        if (lane == 0 && nlong + nshort > 0) first = atomicAdd(&G.row_fill[u], nlong + nshort);
        first = uni(first);
        const uint64_t below = (1ull << lane) - 1ull;
        int32_t before_l = 0, before_s = 0;
#pragma unroll
        for (int j = 0; j < kJ; j++) {
            const int64_t t = t0 + lane + BM_WAVE * j;
            if (t < t1) {
                if (di[j] >= 0 && win == 0)   /* the dense tile kernel's weight for (u, dense row) */
                    G.dense_w[((int64_t)(u / kTileRows) * G.ndense + di[j]) * kTileRows + (u % kTileRows)] =
                        (uint32_t)ed[j].y | ((uint32_t)ed[j].y << 16);
                const bool lng = (mlong[j] >> lane) & 1ull, shrt = (mshort[j] >> lane) & 1ull;
                if (lng || shrt) {
                    const int32_t pos = first + (lng ? before_l + __popcll(mlong[j] & below)
                                                     : nlong + before_s + __popcll(mshort[j] & below));
                    const uint32_t side = t < nout ? 0u : BM_EDESC_IN;
                    const uint32_t hi = (uint32_t)((uint64_t)rb[j] >> 32) | side;
                    out[pos] = make_int4((int32_t)(uint32_t)rb[j], (int32_t)hi, (int32_t)(re[j] - rb[j]), ed[j].y);
                }
            }
            before_l += __popcll(mlong[j]);
            before_s += __popcll(mshort[j]);
        }
    }
}

/* Dense part of the table: G(u, f) restricted to u's edges whose female
 * row is dense, the sum over them of min(w, dense weight of column f).  One
 * block per work item, a tile of 16 male rows by 512 words (1,024 columns):
 * the items are many more than the CUs' block slots, so the dispatcher
 * evens out the CUs (one block per tile over every column left the CUs
 * holding five tiles a quarter longer than those holding four).  The work
 * list has put each such edge's weight, packed w | w << 16, at dense_w[tile]
 * [dense row][row in tile] (0 where the male row has no edge to that dense
 * row), so a dense row's 16 weights are one scalar load; a thread owns two
 * column pairs (a 32-bit word of every dense row each) and keeps their 16
 * rows' sums as packed 16-bit halves, one load of a dense-row word serving
 * 16 rows; a row whose weight is 0 is skipped (a scalar branch: about 40% of
 * a tile's rows do not use a given dense row).  The sums stay below
 * 2^16: each term min(w, fw) is at most fw, and the dense rows are chosen
 * so that every column's weights sum to at most 65535.  The result stays
 * packed: word w of the second half of G's row u holds columns 2w (low
 * half) and 2w + 1, and the row build unpacks it into its LDS row, adds the
 * other edges and overwrites the row.  In the second half a column window's
 * packed words are never under an earlier window's stores (window [c0, c1)
 * stores bytes [4 c0, 4 c1) of the row, and its packed words start at byte
 * 2 stride + 2 c0 >= 4 c1 of every window before it). */
typedef unsigned short Us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pk_min_add(uint32_t acc, uint32_t v, uint32_t wpk) {
    const Us2 m = __builtin_elementwise_min(__builtin_bit_cast(Us2, v), __builtin_bit_cast(Us2, wpk));
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(Us2, acc) + m);
}

/* Four packed weights read as constant data (a scalar load for a uniform
 * address). */
__device__ __forceinline__ uint4 load_weights(const uint32_t* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef __attribute__((address_space(4))) const uint4 ConstUint4;
    return *(const ConstUint4*)p;
#else
    return *(const uint4*)p;
#endif
}

/* One work item: the tile's words [wb, wb + 512) (restrict-qualified, so
 * the weight reads are known not to alias the table writes and become
 * scalar loads). */
__device__ __forceinline__ void dense_tile_words(const uint32_t* __restrict__ wt, const uint32_t* __restrict__ dense,
                                                 int32_t* __restrict__ gout, int64_t stride, int64_t hw, int32_t nwords,
                                                 int32_t nd, int32_t u0, int tid, int32_t wb) {
    /* the dense-row loop and everything in it are block-uniform (scalar
     * weight loads); a thread's words past the row are clamped to the
     * last word for loading and not stored */
    {
        const int32_t w0 = wb + tid, w1 = w0 + kTileThreads;
        const int32_t l0 = w0 < nwords ? w0 : nwords - 1, l1 = w1 < nwords ? w1 : nwords - 1;
        uint32_t a0[kTileRows], a1[kTileRows];
#pragma unroll
        for (int r = 0; r < kTileRows; r++) {
            a0[r] = 0u;
            a1[r] = 0u;
        }
        /* dense row k's two words are loaded one iteration ahead (row nd is
         * a zero row, so the look-ahead needs no clamp); byte offsets are
         * loop-invariant and the row pointer advances by one add */
        const uint32_t o0 = (uint32_t)l0 * 4u, o1 = (uint32_t)l1 * 4u;
        auto word = [](const uint32_t* row, uint32_t off) {
            return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(row) + off);
        };
        uint32_t v0 = word(dense, o0), v1 = word(dense, o1);
        const uint32_t* dn = dense;
        for (int32_t k = 0; k < nd; k++) {
            dn += hw;
            const uint32_t n0 = word(dn, o0), n1 = word(dn, o1);
            const uint32_t* wk = wt + k * kTileRows;
            const uint4 q0 = load_weights(wk), q1 = load_weights(wk + 4), q2 = load_weights(wk + 8),
                        q3 = load_weights(wk + 12);
            /* no test of the 16 weights together: nearly every tile uses every
             * dense row, and the scalar unit is the pass's busiest (the test
             * cost 3% of the build) */
            const uint32_t w[kTileRows] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                                           q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
#pragma unroll
            for (int r = 0; r < kTileRows; r++) {
                if (w[r] != 0u) {
                    a0[r] = pk_min_add(a0[r], v0, w[r]);
                    a1[r] = pk_min_add(a1[r], v1, w[r]);
                }
            }
            v0 = n0;
            v1 = n1;
        }
        /* every row of the tile is stored, packed: row 0's sums are zero
         * (vertex 0 has no edges), and rows past n land in the table's 16
         * padding rows */
        const bool s0 = w0 < nwords, s1 = w1 < nwords;
#pragma unroll
        for (int r = 0; r < kTileRows; r++) {
            uint32_t* out = reinterpret_cast<uint32_t*>(gout + (int64_t)(u0 + r) * stride + hw);
            if (s0) out[w0] = a0[r];
            if (s1) out[w1] = a1[r];
        }
    }
}

__global__ void __launch_bounds__(kTileThreads) k_gain_dense_tile(BmDevGraph g, BmGain G) {
    const int tid = threadIdx.x;
    const int32_t ntiles = (g.n + kTileRows) / kTileRows;   /* rows 0..n */
    const int64_t hw = G.stride >> 1;
    const int32_t nwords = (int32_t)hw;
    const int32_t nd = G.ndense;
    /* items column-block-major: the blocks running at one time share the
     * dense rows' 2 KB column blocks in L2 */
    const int32_t nwb = (nwords + 2 * kTileThreads - 1) / (2 * kTileThreads);
    for (int64_t item = blockIdx.x; item < (int64_t)ntiles * nwb; item += gridDim.x) {
        const int32_t tile = (int32_t)(item % ntiles), wbi = (int32_t)(item / ntiles);
        const int32_t u0 = tile * kTileRows;
        const uint32_t* wt = G.dense_w + (int64_t)tile * nd * kTileRows;
        dense_tile_words(wt, G.dense, G.G, G.stride, hw, nwords, nd, u0, tid, wbi * 2 * kTileThreads);
    }
}

/* Clears n words (the dense tile weights before the work list fills them). */
__global__ void __launch_bounds__(256) k_zero_words(uint32_t* __restrict__ p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

/* Row u: the edges of u's work list are taken one at a time by the block's
 * waves from an LDS counter (so a long female row does not hold up the
 * block).  A wave walks its edges' female rows in chunks of 64 x U entries
 * and keeps S chunks in flight: while the oldest chunk's min(w, fw) are
 * added into the LDS copy of the row's column window, the next S - 1 (of the
 * same edge or the next ones) are loading.  Work-list entries are read
 * through the constant address space (scalar loads), so taking a new edge
 * does not wait on the row loads in flight; the entry after next is always
 * loading. */
/* One work-list entry, read as constant data (a scalar load for a uniform
 * index).  The host pass of the compiler only type-checks kernel bodies. */
__device__ __forceinline__ int4 load_entry(const int4* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef __attribute__((address_space(4))) const int4 ConstInt4;
    return *(const ConstInt4*)p;
#else
    return *p;
#endif
}

template <int LAYOUT, int U, int S, bool WINDOWED>
__global__ void __launch_bounds__(kBuildThreads) k_gain_build(BmDevGraph g, BmGain G, int32_t col0_arg, int32_t ncols) {
    typedef Row<LAYOUT> R;
    constexpr bool PACKED = LAYOUT != BM_ROWS_INT2;
    typedef typename R::T E;
    /* the row window sits at LDS address 0, and a single window starts at
     * column 0, so a column's LDS address is just 4 f */
    extern __shared__ int32_t row[];
    int32_t& next_edge = row[ncols];
    const int32_t col0 = WINDOWED ? col0_arg : 0;
    const int tid = threadIdx.x;
    const int lane = tid & (BM_WAVE - 1);
    constexpr int32_t kStep = BM_WAVE * U;
    const E* __restrict__ f_in_rows = PACKED ? (const E*)G.f_in_p : (const E*)g.f_in;
    const E* __restrict__ f_out_rows = PACKED ? (const E*)G.f_out_p : (const E*)g.f_out;
    for (int32_t u = blockIdx.x + 1; u <= g.n; u += gridDim.x) {
        const int64_t base = g.m_out_ptr[u] + g.m_in_ptr[u];
        const int32_t L = uni(G.row_fill[u]);   /* the list's entries: edges of non-empty sparse rows */
        const int4* list = G.edesc + base;
        /* the row starts from its dense part or zero; k_gain_dense_tile
         * leaves the dense part packed in the row's second half, columns 2w
         * and 2w + 1 in word w (col0 and ncols are even) */
        if (G.ndense > 0) {
            const uint32_t* src =
                reinterpret_cast<const uint32_t*>(G.G + (int64_t)u * G.stride + (G.stride >> 1)) + (col0 >> 1);
            for (int32_t c = tid; c < (ncols >> 1); c += kBuildThreads) {
                const uint32_t v = src[c];
                reinterpret_cast<int2*>(row)[c] = make_int2((int32_t)(v & 0xffffu), (int32_t)(v >> 16));
            }
        } else {
            for (int32_t c = tid; c < ncols; c += kBuildThreads) row[c] = 0;
        }
        if (tid == 0) next_edge = 0;
        __syncthreads();
        /* the list index is clamped so the look-ahead load is unconditional
         * (its value is used only while the index is < L) */
        const int32_t last = L > 0 ? L - 1 : 0;
        auto grab = [&]() {
            int32_t x = 0;
            if (lane == 0) x = atomicAdd(&next_edge, 1);
            return uni(x);
        };
        /* scalar walk state: current edge (row pointer, length, weight, next
         * offset) and the prefetched next edge */
        int32_t t = grab();
        int4 d = load_entry(list + (t < last ? t : last));
        int32_t tn = grab();
        int4 dn = load_entry(list + (tn < last ? tn : last));
        const E* frow = f_in_rows;
        int32_t len = 0, w = 0, j0 = 0;
        auto decode = [&](const int4 en) {
            const uint32_t hi = (uint32_t)uni(en.y);
            const int64_t start =
                (int64_t)((uint64_t)(uint32_t)uni(en.x) | ((uint64_t)(hi & ~(BM_EDESC_IN | BM_EDESC_DENSE)) << 32));
            frow = ((hi & BM_EDESC_IN) ? f_out_rows : f_in_rows) + start;
            len = (hi & BM_EDESC_DENSE) ? 0 : uni(en.z);
            w = uni(en.w);
            j0 = 0;
        };
        if (t < L) decode(d);
        /* next chunk: pointer, entry count (0 = no more work), weight */
        auto next_chunk = [&](const E*& p, int32_t& cnt, int32_t& wt) {
            for (;;) {
                if (j0 < len) {
                    p = frow + j0;
                    cnt = len - j0 < kStep ? len - j0 : kStep;
                    wt = w;
                    j0 += kStep;
                    return;
                }
                if (t >= L || tn >= L) {
                    t = L;
                    p = f_in_rows;
                    cnt = 0;
                    wt = 0;
                    return;
                }
                t = tn;
                decode(dn);
                tn = grab();
                dn = load_entry(list + (tn < last ? tn : last));
            }
        };
        /* Chunk loads are unconditional (buffer loads bounded by the chunk)
         * and the loop is unrolled S times so the register sets rotate
         * without copies: both keep the compiler's load counting exact, so an
         * add waits only on its own chunk while the newer ones load. */
        auto load_chunk = [&](E (&r)[U], const E* p, int32_t cnt) {
            /* cnt = 0 gives a resource of no records: every load returns
             * zero without touching memory, so the load count never depends
             * on the branch taken */
            const __amdgpu_buffer_rsrc_t rs =
                __builtin_amdgcn_make_buffer_rsrc(const_cast<E*>(p), 0, cnt * (int32_t)sizeof(E), kBufWord3);
#pragma unroll
            for (int k = 0; k < U; k++) r[k] = R::load(rs, lane + BM_WAVE * k);
        };
        /* column f of the window is at LDS byte 4 (f - col0); a full chunk
         * needs no lane mask, the last chunk of an edge masks the lanes past
         * its end (their zero entries must not add min(w, 0) for a negative w) */
        const uint32_t shift = 4u * (uint32_t)col0;
        auto add_chunk = [&](const E (&r)[U], int32_t cnt, int32_t wt) {
            if (cnt == kStep) {
#pragma unroll
                for (int k = 0; k < U; k++) {
                    const int32_t fw = R::w(r[k]);
                    lds_add(WINDOWED ? R::addr(r[k]) - shift : R::addr(r[k]), wt < fw ? wt : fw);
                }
            } else {
#pragma unroll
                for (int k = 0; k < U; k++) {
                    if (lane + BM_WAVE * k < cnt) {
                        const int32_t fw = R::w(r[k]);
                        lds_add(WINDOWED ? R::addr(r[k]) - shift : R::addr(r[k]), wt < fw ? wt : fw);
                    }
                }
            }
        };
        /* named register sets (arrays of them cost registers) */
        E ra[U], rb[U];
        const E* pa;
        const E* pb;
        int32_t ca, wa, cb, wb;
        if constexpr (S == 2) {
            next_chunk(pa, ca, wa);
            load_chunk(ra, pa, ca);
            while (ca > 0) {
                next_chunk(pb, cb, wb);
                load_chunk(rb, pb, cb);
                add_chunk(ra, ca, wa);
                if (cb == 0) break;
                next_chunk(pa, ca, wa);
                load_chunk(ra, pa, ca);
                add_chunk(rb, cb, wb);
            }
        } else {
            static_assert(S == 4, "the build pipelines 2 or 4 chunks");
            E rc[U], rd[U];
            const E* pc;
            const E* pd;
            int32_t cc, wc, cd, wd;
            next_chunk(pa, ca, wa);
            load_chunk(ra, pa, ca);
            next_chunk(pb, cb, wb);
            load_chunk(rb, pb, cb);
            next_chunk(pc, cc, wc);
            load_chunk(rc, pc, cc);
            while (ca > 0) {
                next_chunk(pd, cd, wd);
                load_chunk(rd, pd, cd);
                add_chunk(ra, ca, wa);
                if (cb == 0) break;
                next_chunk(pa, ca, wa);
                load_chunk(ra, pa, ca);
                add_chunk(rb, cb, wb);
                if (cc == 0) break;
                next_chunk(pb, cb, wb);
                load_chunk(rb, pb, cb);
                add_chunk(rc, cc, wc);
                if (cd == 0) break;
                next_chunk(pc, cc, wc);
                load_chunk(rc, pc, cc);
                add_chunk(rd, cd, wd);
            }
        }
        /* the compiler does not count the inline-asm LDS adds: wait for them
         * before the barrier that publishes the row */
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        __syncthreads();
        /* the finished row streams to HBM without taking L2 lines from the
         * female rows that later rows re-read */
        int32_t* out = G.G + (int64_t)u * G.stride + col0;
        for (int32_t c = tid; c < ncols; c += kBuildThreads) __builtin_nontemporal_store(row[c], &out[c]);
        __syncthreads();
    }
}

/* Dense male adjacency: bit b of row a set iff M has the edge a -> b.  One
 * wave per vertex, lanes over its out-list. */
__global__ void __launch_bounds__(256) k_adj_bits(BmDevGraph g, uint32_t* __restrict__ bits, int64_t words) {
    const int lane = threadIdx.x & (BM_WAVE - 1);
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / BM_WAVE;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / BM_WAVE;
    for (int64_t a = wave + 1; a <= g.n; a += nwaves) {
        uint32_t* row = bits + a * words;
        for (int64_t e = g.m_out_ptr[a] + lane; e < g.m_out_ptr[a + 1]; e += BM_WAVE) {
            const int32_t b = g.m_out[e].x;   /* the pair's bit sits in the row of its smaller id */
            if (b >= a) atomicOr(&row[b >> 5], 1u << (b & 31));
            else atomicOr(&bits[(int64_t)b * words + (a >> 5)], 1u << (a & 31));
        }
    }
}

__global__ void k_pack_rows(const int2* __restrict__ in, int64_t m, int layout, uint32_t* __restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride)
        out[i] = layout == BM_ROWS_P19 ? ((uint32_t)in[i].y << 19) | ((uint32_t)in[i].x << 2)
                                       : ((uint32_t)in[i].x << 15) | (uint32_t)in[i].y;
}

/* Dense row k: one block over the entries of female row rows[k].x (its
 * out-row when rows[k].y, else its in-row), each weight stored at its
 * column. */
__global__ void __launch_bounds__(256) k_dense_fill(const int64_t* __restrict__ out_ptr, const int2* __restrict__ out_rows,
                                                    const int64_t* __restrict__ in_ptr, const int2* __restrict__ in_rows,
                                                    const int2* __restrict__ rows, int64_t stride,
                                                    uint16_t* __restrict__ pool) {
    const int2 r = rows[blockIdx.x];
    const int64_t* ptr = r.y ? out_ptr : in_ptr;
    const int2* adj = r.y ? out_rows : in_rows;
    uint16_t* dst = pool + (int64_t)blockIdx.x * stride;
    for (int64_t e = ptr[r.x] + threadIdx.x; e < ptr[r.x + 1]; e += blockDim.x) dst[adj[e].x] = (uint16_t)adj[e].y;
}

/* Patch the rows of the neighbours of each moved vertex x (pi(x): old -> new).
 * moves[i] = {x, old, new, 0}, *count moves.  An in-edge y -> x contributes
 * c(w, f, pi(x)) to G(y, f) over the female in-row of pi(x); an out-edge
 * x -> y contributes c(w, pi(x), f) over the female out-row.  The old
 * female's terms are subtracted and the new one's added; self-loops are not
 * in the table.  Updates of different moves commute (integer atomics).
 *
 * Work items: (move, side = old / new female, direction, group of 64 of x's
 * edges, chunk of 64 entries of the female row).  k_gain_patch_plan counts
 * each (move, side, direction)'s items and scans them into offsets; the
 * patch kernel's waves take items from the whole index space, so the many
 * items of one high-degree vertex spread over the device.  An item: lanes
 * hold 64 edges and 64 row entries, and each entry is applied for every edge
 * in turn (the edge's row y and weight read out of its lane as scalars). */
constexpr int kPlanThreads = 1024;

__device__ __forceinline__ int64_t patch_items(const BmDevGraph& g, int4 mv, int side, int dir) {
    const int32_t gx = side ? mv.z : mv.y;
    if (mv.y == mv.z || gx == 0 || mv.x == 0) return 0;
    const int64_t* mptr = dir ? g.m_out_ptr : g.m_in_ptr;
    const int64_t* fptr = dir ? g.f_out_ptr : g.f_in_ptr;
    const int64_t deg = mptr[mv.x + 1] - mptr[mv.x], len = fptr[gx + 1] - fptr[gx];
    return ((deg + BM_WAVE - 1) / BM_WAVE) * ((len + BM_WAVE - 1) / BM_WAVE);
}

/* One block: off[s] = items of the slots before s, s = 4 move + 2 side + dir,
 * off[4 nm] = all items. */
__global__ void __launch_bounds__(kPlanThreads) k_gain_patch_plan(BmDevGraph g, const int4* __restrict__ moves,
                                                                  const int64_t* __restrict__ count,
                                                                  int64_t* __restrict__ off) {
    __shared__ int64_t part[kPlanThreads];
    const int64_t slots = 4 * *count;
    const int t = threadIdx.x;
    const int64_t per = (slots + kPlanThreads - 1) / kPlanThreads;
    const int64_t s0 = t * per, s1 = s0 + per < slots ? s0 + per : slots;
    int64_t mine = 0;
    for (int64_t sl = s0; sl < s1; sl++) mine += patch_items(g, moves[sl >> 2], (int)((sl >> 1) & 1), (int)(sl & 1));
    part[t] = mine;
    __syncthreads();
    for (int o = 1; o < kPlanThreads; o <<= 1) {
        const int64_t add = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    int64_t run = t ? part[t - 1] : 0;
    for (int64_t sl = s0; sl < s1; sl++) {
        off[sl] = run;
        run += patch_items(g, moves[sl >> 2], (int)((sl >> 1) & 1), (int)(sl & 1));
    }
    if (t == kPlanThreads - 1) off[slots] = part[t];
}

__global__ void __launch_bounds__(kPatchThreads) k_gain_update(BmDevGraph g, BmGain G, const int4* __restrict__ moves,
                                                               const int64_t* __restrict__ count,
                                                               const int64_t* __restrict__ off,
                                                               int32_t* __restrict__ GT) {
    const int lane = threadIdx.x & (BM_WAVE - 1);
    const int64_t slots = 4 * *count;
    const int64_t total = off[slots];
    const int64_t nw = (int64_t)gridDim.x * kPatchWaves;
    for (int64_t item = (int64_t)blockIdx.x * kPatchWaves + threadIdx.x / BM_WAVE; item < total; item += nw) {
        /* the slot holding the item: the last s with off[s] <= item */
        int64_t lo = 0, hi = slots;   /* off[lo] <= item < off[hi] */
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (off[mid] <= item) lo = mid;
            else hi = mid;
        }
        const int4 mv = moves[lo >> 2];
        const int side = (int)((lo >> 1) & 1), dir = (int)(lo & 1);
        const int32_t x = mv.x, gx = side ? mv.z : mv.y;
        const int32_t sign = side ? 1 : -1;
        const int64_t* mptr = dir ? g.m_out_ptr : g.m_in_ptr;
        const int2* madj = dir ? g.m_out : g.m_in;
        const int64_t* fptr = dir ? g.f_out_ptr : g.f_in_ptr;
        const int2* fadj = dir ? g.f_out : g.f_in;
        const int64_t b = mptr[x], e = mptr[x + 1];
        const int64_t rb = fptr[gx], re = fptr[gx + 1];
        const int64_t nchunk = (re - rb + BM_WAVE - 1) / BM_WAVE;
        const int64_t local = item - off[lo];
        const int64_t e0 = b + (local / nchunk) * BM_WAVE;
        const int64_t i = e0 + lane;
        const int2 ed = i < e ? madj[i] : make_int2(0, 0);
        const int32_t yl = ed.x != x ? ed.x : 0;   /* 0: past the list or a self-loop */
        const int32_t nv = e - e0 < BM_WAVE ? (int32_t)(e - e0) : BM_WAVE;
        const int64_t ri = rb + (local % nchunk) * BM_WAVE + lane;
        const bool rin = ri < re;
        const int2 ent = rin ? fadj[ri] : make_int2(0, 0);
        for (int32_t k = 0; k < nv; k++) {
            const int32_t y = __builtin_amdgcn_readlane(yl, k);
            if (y == 0) continue;
            const int32_t w = __builtin_amdgcn_readlane(ed.y, k);
            if (rin) {
                const int32_t d = sign * (w < ent.y ? w : ent.y);
                atomicAdd(&G.G[(int64_t)y * G.stride + ent.x], d);
                if (GT) atomicAdd(&GT[(int64_t)ent.x * G.stride + y], d);   /* the transposed copy, when kept */
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Lookups                                                             */
/* ------------------------------------------------------------------ */

/* Swap delta of u != v from its four table entries guf2 = G(u, f2),
 * guf1 = G(u, f1), gvf1 = G(v, f1), gvf2 = G(v, f2), with e_uv = "a male
 * edge between u and v may exist" (either direction): an edge's weight and
 * the female weights its correction needs are hash-probed only then. */
__device__ __forceinline__ int64_t gain_swap_delta_from(const BmDevGraph& g, const BmGain& G, int32_t u, int32_t v,
                                                        int32_t f1, int32_t f2, bool e_uv, int64_t guf2, int64_t guf1,
                                                        int64_t gvf1, int64_t gvf2, int32_t su, int32_t sv,
                                                        int32_t fs1, int32_t fs2) {
    int64_t d = guf2 - guf1 + gvf1 - gvf2;
    const int32_t fs[3] = {fs1, fs2, BM_ABSENT};
    if (e_uv) {
        /* male edges u -> v and v -> u; their corrections need the four
         * female weights among f1 and f2 (rare: the hash-row descriptors are
         * loaded only here) */
        const bool have[3][3] = {{false, true, false}, {true, false, false}, {false, false, false}};
        const uint64_t md[3] = {G.m_out_desc[u], G.m_out_desc[v], 0ull};
        const uint64_t fd[3] = {g.f_out_desc[f1], g.f_out_desc[f2], 0ull};
        const int32_t vtx[3] = {u, v, 0}, f[3] = {f1, f2, 0};
        int32_t w[3][3], F[3][3];
        triple_weights(g, G, vtx, md, f, fd, fs, have, w, F);
        const int32_t w1 = w[0][1], w2 = w[1][0];
        /* u -> v lands on f2 -> f1; v -> u on f1 -> f2 */
        if (w1 != BM_ABSENT)
            d += (int64_t)contrib(w1, F[1][0]) - contrib(w1, F[1][1]) - contrib(w1, F[0][0]) + contrib(w1, F[0][1]);
        if (w2 != BM_ABSENT)
            d += (int64_t)contrib(w2, F[0][1]) - contrib(w2, F[0][0]) - contrib(w2, F[1][1]) + contrib(w2, F[1][0]);
    }
    if (su != BM_ABSENT) d += contrib(su, fs[1]) - (int64_t)contrib(su, fs[0]);
    if (sv != BM_ABSENT) d += contrib(sv, fs[0]) - (int64_t)contrib(sv, fs[1]);
    return d;
}

__device__ __forceinline__ int64_t gain_swap_delta(const BmDevGraph& g, const BmGain& G, int32_t u, int32_t v) {
    if (u == v) return 0;
    const int32_t f1 = g.pi[u], f2 = g.pi[v];
    const int32_t* Gu = G.G + (int64_t)u * G.stride;
    const int32_t* Gv = G.G + (int64_t)v * G.stride;
    return gain_swap_delta_from(g, G, u, v, f1, f2, male_maybe(G, u, v), Gu[f2], Gu[f1], Gv[f1], Gv[f2], G.m_self[u],
                                G.m_self[v], G.f_self[f1], G.f_self[f2]);
}

/* The same from the per-vertex words {pi, G(v, pi(v)), male self-loop,
 * female self-loop of pi(v)}: two random table reads per pair instead of
 * four. */
__device__ __forceinline__ int64_t gain_swap_delta_pis(const BmDevGraph& g, const BmGain& G, const int4* pis, int32_t u,
                                                       int32_t v) {
    if (u == v) return 0;
    const int4 pu = pis[u], pv = pis[v];
    return gain_swap_delta_from(g, G, u, v, pu.x, pv.x, male_maybe(G, u, v), G.G[(int64_t)u * G.stride + pv.x], pu.y,
                                G.G[(int64_t)v * G.stride + pu.x], pv.y, pu.z, pv.z, pu.w, pv.w);
}

template <int MODE, bool PIS>
__global__ void __launch_bounds__(256) k_gain_swap(BmDevGraph g, BmGain G, const int2* __restrict__ cand,
                                                   const int64_t* __restrict__ off, uint64_t seed, int64_t k,
                                                   int64_t* __restrict__ delta, const int4* __restrict__ pis) {
    if (MODE == BM_CAND_UMAJOR) {
        /* a u-major stream is read in one contiguous range per block, so a
         * thread's next entry (blockDim later) almost always has the same
         * first vertex or the next one: one offset load confirms it, and a
         * binary search is needed only after a long jump */
        const int64_t per = (k + gridDim.x - 1) / gridDim.x;
        const int64_t c0 = (int64_t)blockIdx.x * per, c1 = c0 + per < k ? c0 + per : k;
        int32_t u = 0;
        for (int64_t c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
            if (u == 0 || off[u + 1] <= c) {
                if (u != 0 && u < g.n && off[u + 2] > c) u++;
                else u = umajor_owner(off, g.n, c);
            }
            const int32_t v = bm_umajor_partner(seed, (uint64_t)c, u, g.n);
            delta[c] = PIS ? gain_swap_delta_pis(g, G, pis, u, v) : gain_swap_delta(g, G, u, v);
        }
        return;
    }
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
        delta[c] = PIS ? gain_swap_delta_pis(g, G, pis, u, v) : gain_swap_delta(g, G, u, v);
    }
}

/* Exact best swap over all n(n-1)/2 pairs u < v.  One block per first vertex
 * u (grid-stride over u); the block's threads take v = u+1..n in order, so
 * every per-partner read is coalesced: GT(f1, v) = G(v, f1) from the
 * transposed table, the packed per-vertex words {pi(v), G(v, pi(v)),
 * self-loops}, and the dense adjacency bitmap's row u (bit v); only G(u, f2)
 * is a gather, inside row u, which the block shares through the caches.  The
 * order is: larger delta first, then smaller u, then smaller v, a total
 * order, so the result does not depend on the launch shape.  Each block
 * writes its best as {delta lo, delta hi, u, v}; k_best_swap_final reduces
 * the blocks' bests. */
constexpr int kBestThreads = 256;
constexpr int kBestPairs = 4;   /* partners per thread in flight */

__device__ __forceinline__ bool best_before(int64_t d, int32_t u, int32_t v, int64_t bd, int32_t bu, int32_t bv) {
    return d > bd || (d == bd && (u < bu || (u == bu && v < bv)));
}

/* Block-wide reduction of (d, u, v) under best_before; thread 0 gets the result. */
__device__ __forceinline__ void best_block_reduce(int64_t& d, int32_t& u, int32_t& v) {
    __shared__ int64_t sd[kBestThreads / BM_WAVE];
    __shared__ int32_t su[kBestThreads / BM_WAVE], sv[kBestThreads / BM_WAVE];
    const int lane = threadIdx.x & (BM_WAVE - 1), wave = threadIdx.x / BM_WAVE;
#pragma unroll
    for (int o = BM_WAVE / 2; o > 0; o >>= 1) {
        const int64_t od = __shfl_down(d, o, BM_WAVE);
        const int32_t ou = __shfl_down(u, o, BM_WAVE), ov = __shfl_down(v, o, BM_WAVE);
        if (best_before(od, ou, ov, d, u, v)) { d = od; u = ou; v = ov; }
    }
    if (lane == 0) { sd[wave] = d; su[wave] = u; sv[wave] = v; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBestThreads / BM_WAVE; w++)
            if (best_before(sd[w], su[w], sv[w], d, u, v)) { d = sd[w]; u = su[w]; v = sv[w]; }
    }
}

/* The block's threads scan the partners v = u+1..n of u and keep their best
 * (d, u, v) under best_before. */
__device__ __forceinline__ void best_partner_scan(const BmDevGraph& g, const BmGain& G, const int32_t* __restrict__ GT,
                                                  const int4* __restrict__ pis, int32_t u, int64_t& bd, int32_t& bu,
                                                  int32_t& bv) {
    const int32_t n = g.n;
    const int4 pu = pis[u];
    const int32_t f1 = pu.x;
    const int32_t* Gu = G.G + (int64_t)u * G.stride;
    const int32_t* Tf1 = GT + (int64_t)f1 * G.stride;
    for (int32_t v0 = u + 1 + (int32_t)threadIdx.x; v0 <= n; v0 += kBestThreads * kBestPairs) {
        /* every load of kBestPairs partners in flight before any use */
        int4 pv[kBestPairs];
        int32_t tv[kBestPairs], gu[kBestPairs];
        bool mm[kBestPairs];
#pragma unroll
        for (int j = 0; j < kBestPairs; j++) {
            const int32_t v = v0 + kBestThreads * j;
            const bool in = v <= n;
            pv[j] = in ? pis[v] : make_int4(0, 0, BM_ABSENT, BM_ABSENT);
            tv[j] = in ? Tf1[v] : 0;
            mm[j] = in && male_maybe(G, u, v);
        }
#pragma unroll
        for (int j = 0; j < kBestPairs; j++) gu[j] = Gu[pv[j].x];
#pragma unroll
        for (int j = 0; j < kBestPairs; j++) {
            const int32_t v = v0 + kBestThreads * j;
            if (v > n) continue;
            int64_t d;
            if (mm[j]) {   /* a male edge between u and v may exist (rare): the full correction */
                d = gain_swap_delta_from(g, G, u, v, f1, pv[j].x, true, gu[j], pu.y, tv[j], pv[j].y, pu.z, pv[j].z,
                                         pu.w, pv[j].w);
            } else {       /* only self-loops, whose weights the per-vertex words carry */
                d = (int64_t)gu[j] - pu.y + tv[j] - pv[j].y;
                if (pu.z != BM_ABSENT) d += contrib(pu.z, pv[j].w) - (int64_t)contrib(pu.z, pu.w);
                if (pv[j].z != BM_ABSENT) d += contrib(pv[j].z, pu.w) - (int64_t)contrib(pv[j].z, pv[j].w);
            }
            if (best_before(d, u, v, bd, bu, bv)) { bd = d; bu = u; bv = v; }
        }
    }
}

__global__ void __launch_bounds__(kBestThreads) k_gain_best_swap(BmDevGraph g, BmGain G, const int32_t* __restrict__ GT,
                                                                 const int4* __restrict__ pis, int4* __restrict__ part) {
    int64_t bd = INT64_MIN;
    int32_t bu = INT32_MAX, bv = INT32_MAX;
    for (int32_t u = blockIdx.x + 1; u < g.n; u += gridDim.x) best_partner_scan(g, G, GT, pis, u, bd, bu, bv);
    best_block_reduce(bd, bu, bv);
    if (threadIdx.x == 0)
        part[blockIdx.x] = make_int4((int32_t)(uint32_t)(uint64_t)bd, (int32_t)(uint32_t)((uint64_t)bd >> 32), bu, bv);
}

/* Best partner of each u in u0 .. u0 + count - 1: block i takes u0 + i (and
 * every gridDim.x-th after it) and writes {delta lo, delta hi, u, v} for it
 * (delta INT64_MIN, v INT32_MAX when u has no partner v > u). */
__global__ void __launch_bounds__(kBestThreads) k_gain_best_partners(BmDevGraph g, BmGain G,
                                                                     const int32_t* __restrict__ GT,
                                                                     const int4* __restrict__ pis, int32_t u0,
                                                                     int32_t count, int4* __restrict__ out) {
    for (int32_t i = blockIdx.x; i < count; i += gridDim.x) {
        const int32_t u = u0 + i;
        int64_t bd = INT64_MIN;
        int32_t bu = u, bv = INT32_MAX;
        best_partner_scan(g, G, GT, pis, u, bd, bu, bv);
        best_block_reduce(bd, bu, bv);
        if (threadIdx.x == 0)
            out[i] = make_int4((int32_t)(uint32_t)(uint64_t)bd, (int32_t)(uint32_t)((uint64_t)bd >> 32), u, bv);
        __syncthreads();   /* the reduction's shared words are reused for the next u */
    }
}

__global__ void __launch_bounds__(kBestThreads) k_best_swap_final(const int4* __restrict__ part, int32_t nparts,
                                                                  int4* __restrict__ out) {
    int64_t bd = INT64_MIN;
    int32_t bu = INT32_MAX, bv = INT32_MAX;
    for (int32_t i = threadIdx.x; i < nparts; i += kBestThreads) {
        const int4 q = part[i];
        const int64_t d = (int64_t)(((uint64_t)(uint32_t)q.y << 32) | (uint64_t)(uint32_t)q.x);
        if (best_before(d, q.z, q.w, bd, bu, bv)) { bd = d; bu = q.z; bv = q.w; }
    }
    best_block_reduce(bd, bu, bv);
    if (threadIdx.x == 0)
        out[0] = make_int4((int32_t)(uint32_t)(uint64_t)bd, (int32_t)(uint32_t)((uint64_t)bd >> 32), bu, bv);
}

/* Transposed table: GT(f, u) = G(u, f) for f < nrows, with GT(f, u) = 0 in
 * the padding columns u >= nrows.  64 x 64 tiles through LDS; rows start on
 * 128-byte lines (stride a multiple of 32), so every global access is an
 * aligned 16-byte load or store. */
__global__ void __launch_bounds__(256) k_gain_transpose(const int32_t* __restrict__ Gt, int32_t* __restrict__ GT,
                                                        int64_t nrows, int64_t stride) {
    __shared__ int32_t tile[64][65];
    const int q = threadIdx.x & 15, r0 = threadIdx.x >> 4;    /* int4 column, first row */
    const int64_t ut = (stride + 63) / 64, ft = (nrows + 63) / 64;
    for (int64_t t = blockIdx.x; t < ut * ft; t += gridDim.x) {
        const int64_t bu = t / ft, bf = t % ft;   /* tile rows u = 64 bu.., columns f = 64 bf.. */
        const int64_t f0 = bf * 64 + 4 * q;
        int4 a[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int64_t u = bu * 64 + r0 + 16 * k;
            a[k] = (u < nrows && f0 < stride) ? *reinterpret_cast<const int4*>(Gt + u * stride + f0)
                                              : make_int4(0, 0, 0, 0);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int32_t* tr = &tile[r0 + 16 * k][4 * q];
            tr[0] = a[k].x;
            tr[1] = a[k].y;
            tr[2] = a[k].z;
            tr[3] = a[k].w;
        }
        __syncthreads();
        const int64_t u0 = bu * 64 + 4 * q;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int fl = r0 + 16 * k;
            const int64_t f = bf * 64 + fl;
            if (f < nrows && u0 < stride)
                *reinterpret_cast<int4*>(GT + f * stride + u0) =
                    make_int4(tile[4 * q][fl], tile[4 * q + 1][fl], tile[4 * q + 2][fl], tile[4 * q + 3][fl]);
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

/* u-major lookups: one block per first vertex u of the stream.  The block
 * stages row G(u, .) (indexed by the partner's female) and row GT(pi(u), .) =
 * G(., pi(u)) (indexed by the partner) in LDS with aligned 16-byte loads, and
 * four bitmaps over vertex ids with a prefix popcount per 64-bit word: u's
 * male out- and in-neighbours and f1 = pi(u)'s female out- and in-neighbours.
 * A pair (u, v) then costs two LDS reads, one gather of v's packed {pi,
 * G(v, pi(v)), self-loops}, two bitmap reads and the delta store.  For a pair
 * with an edge between u and v (about 2% at n = 18,524, mean degree 216) the
 * bit's rank is the entry's position in the sorted CSR row, so each needed
 * weight (w_M(u, v), w_M(v, u), w_F(f1, f2), w_F(f2, f1)) is one direct load;
 * self-loops add their exact terms. */
constexpr int kUmThreads = 1024;
constexpr size_t kLdsBytes = 160 * 1024;
constexpr int kUmStage = 5;    /* 16-byte row words per thread and row in flight while staging (5 x 1024 x 4
                                  * covers a row of up to 20,480 entries in one round) */
constexpr int kUmPairs = 4;    /* pairs per thread in flight */

typedef unsigned V4u __attribute__((ext_vector_type(4)));

/* Neighbour id of entry i of an {id, w} list of len entries, 0 past its end
 * (a bounded buffer load: nothing is read there). */
__device__ __forceinline__ int32_t list_id(const int2* base, int64_t len, int32_t i) {
    const int32_t bytes = len < (int64_t)(INT32_MAX / 8) ? (int32_t)len * 8 : (INT32_MAX / 8) * 8;
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<int2*>(base), 0, bytes, kBufWord3);
    return (int32_t)__builtin_amdgcn_raw_buffer_load_b32(r, i * 8, 0, 0);
}

/* Rank of x in a bitmap (uint32 words) with prefix counts per 64-bit word. */
template <typename P>
__device__ __forceinline__ int32_t bit_rank(const uint32_t* bm, const P* pre, int32_t x) {
    const int32_t w = x >> 6;
    const uint64_t word = (uint64_t)bm[2 * w] | ((uint64_t)bm[2 * w + 1] << 32);
    return (int32_t)pre[w] + __popcll(word & ((1ull << (x & 63)) - 1ull));
}

__device__ __forceinline__ bool bit_test(const uint32_t* bm, int32_t x) { return (bm[x >> 5] >> (x & 31)) & 1u; }

/* ROWS = false is the form for tables whose two rows do not fit in LDS
 * (n >= 18,976): only the bitmaps are staged (with 32-bit prefix counts, as
 * a hub's list may pass 65,535 entries), G(u, f2) is gathered from row u
 * (which the block's loads share in L2) and G(v, f1) from v's row, so a pair
 * costs one random table read besides its per-vertex words; no transposed
 * table is needed. */
template <bool ROWS, typename P>
__global__ void __launch_bounds__(kUmThreads) k_gain_swap_umajor(BmDevGraph g, BmGain G, const int32_t* __restrict__ GT,
                                                                 const int4* __restrict__ pis,
                                                                 const int64_t* __restrict__ off, uint64_t seed,
                                                                 int64_t* __restrict__ delta) {
    extern __shared__ int4 sm4[];
    const int32_t n = g.n;
    const int64_t stride = G.stride;                   /* multiple of 32 */
    const int32_t s4 = (int32_t)(stride >> 2);
    const int32_t w64 = (n + 1 + 63) >> 6;             /* prefix entries per bitmap */
    int32_t* rowG = reinterpret_cast<int32_t*>(sm4);
    int32_t* rowT = rowG + stride;
    uint32_t* bits = ROWS ? reinterpret_cast<uint32_t*>(rowT + stride)   /* 4 bitmaps of 2 * w64 words */
                          : reinterpret_cast<uint32_t*>(sm4);
    uint32_t* mo = bits;                               /* u -> v in M */
    uint32_t* mi = bits + 2 * w64;                     /* v -> u in M */
    uint32_t* fo = bits + 4 * w64;                     /* f1 -> f in F */
    uint32_t* fi = bits + 6 * w64;                     /* f -> f1 in F */
    P* pre = reinterpret_cast<P*>(bits + 8 * w64);     /* 4 x w64 prefix counts */
    const int tid = threadIdx.x;
    const int lane = tid & (BM_WAVE - 1), wave = tid / BM_WAVE;
    for (int32_t u = blockIdx.x + 1; u <= n; u += gridDim.x) {
        const int64_t i0 = off[u], i1 = off[u + 1];
        if (i0 == i1) continue;   /* block-uniform */
        const int4 pu = pis[u];
        const int32_t f1 = pu.x;
        const int32_t F11 = pu.w;                           /* w_F(f1, f1) */
        /* Staging.  The first kUmThreads entries of the four adjacency lists and
         * both rows are requested together as bounded buffer loads (zero past
         * the end), lists first, so the bitmap marks wait only on the list
         * loads while the rows are still arriving. */
        for (int32_t w = tid; w < 8 * w64; w += kUmThreads) bits[w] = 0;
        __syncthreads();
        const int64_t mob = g.m_out_ptr[u], moe = g.m_out_ptr[u + 1];
        const int64_t mib = g.m_in_ptr[u], mie = g.m_in_ptr[u + 1];
        const int64_t fob = f1 ? g.f_out_ptr[f1] : 0, foe = f1 ? g.f_out_ptr[f1 + 1] : 0;
        const int64_t fib = f1 ? g.f_in_ptr[f1] : 0, fie = f1 ? g.f_in_ptr[f1 + 1] : 0;
        const int32_t x_mo = list_id(g.m_out + mob, moe - mob, tid);
        const int32_t x_mi = list_id(g.m_in + mib, mie - mib, tid);
        const int32_t x_fo = list_id(g.f_out + fob, foe - fob, tid);
        const int32_t x_fi = list_id(g.f_in + fib, fie - fib, tid);
        V4u a[kUmStage], t[kUmStage];
        if constexpr (ROWS) {
            const __amdgpu_buffer_rsrc_t rG =
                __builtin_amdgcn_make_buffer_rsrc(G.G + (int64_t)u * stride, 0, (int32_t)(stride * 4), kBufWord3);
            const __amdgpu_buffer_rsrc_t rT = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<int32_t*>(GT) + (int64_t)f1 * stride, 0, (int32_t)(stride * 4), kBufWord3);
#pragma unroll
            for (int j = 0; j < kUmStage; j++) {
                a[j] = __builtin_amdgcn_raw_buffer_load_b128(rG, (tid + kUmThreads * j) * 16, 0, 0);
                t[j] = __builtin_amdgcn_raw_buffer_load_b128(rT, (tid + kUmThreads * j) * 16, 0, 0);
            }
        }
        if (x_mo) atomicOr(&mo[x_mo >> 5], 1u << (x_mo & 31));   /* ids are >= 1: 0 marks nothing */
        if (x_mi) atomicOr(&mi[x_mi >> 5], 1u << (x_mi & 31));
        if (x_fo) atomicOr(&fo[x_fo >> 5], 1u << (x_fo & 31));
        if (x_fi) atomicOr(&fi[x_fi >> 5], 1u << (x_fi & 31));
        for (int64_t e = mob + tid + kUmThreads; e < moe; e += kUmThreads) {   /* lists longer than the block */
            const int32_t x = g.m_out[e].x;
            atomicOr(&mo[x >> 5], 1u << (x & 31));
        }
        for (int64_t e = mib + tid + kUmThreads; e < mie; e += kUmThreads) {
            const int32_t x = g.m_in[e].x;
            atomicOr(&mi[x >> 5], 1u << (x & 31));
        }
        for (int64_t e = fob + tid + kUmThreads; e < foe; e += kUmThreads) {
            const int32_t y = g.f_out[e].x;
            atomicOr(&fo[y >> 5], 1u << (y & 31));
        }
        for (int64_t e = fib + tid + kUmThreads; e < fie; e += kUmThreads) {
            const int32_t y = g.f_in[e].x;
            atomicOr(&fi[y >> 5], 1u << (y & 31));
        }
        if constexpr (ROWS) {
            V4u* sG = reinterpret_cast<V4u*>(rowG);
            V4u* sT = reinterpret_cast<V4u*>(rowT);
#pragma unroll
            for (int j = 0; j < kUmStage; j++) {
                const int32_t c = tid + kUmThreads * j;
                if (c < s4) {
                    sG[c] = a[j];
                    sT[c] = t[j];
                }
            }
        }
        __syncthreads();
        if (wave < 4) {   /* exclusive prefix popcounts, one wave per bitmap */
            const uint32_t* bm = bits + 2 * w64 * wave;
            P* pr = pre + w64 * wave;
            const int32_t per = (w64 + BM_WAVE - 1) / BM_WAVE;
            const int32_t lo = lane * per, hi = lo + per < w64 ? lo + per : w64;
            int32_t cnt = 0;
            for (int32_t w = lo; w < hi; w++) cnt += __popc(bm[2 * w]) + __popc(bm[2 * w + 1]);
            int32_t incl = cnt;
#pragma unroll
            for (int o = 1; o < BM_WAVE; o <<= 1) {
                const int32_t t = __shfl_up(incl, o, BM_WAVE);
                if (lane >= o) incl += t;
            }
            int32_t run = incl - cnt;
            for (int32_t w = lo; w < hi; w++) {
                pr[w] = (P)run;
                run += __popc(bm[2 * w]) + __popc(bm[2 * w + 1]);
            }
        }
        __syncthreads();
        const int32_t* Gu = G.G + (int64_t)u * stride;
        for (int64_t ib = i0 + tid; ib < i1; ib += (int64_t)kUmThreads * kUmPairs) {
            int32_t v[kUmPairs], gu[kUmPairs], gt[kUmPairs];
            int4 pv[kUmPairs];
#pragma unroll
            for (int j = 0; j < kUmPairs; j++) {   /* all gathers in flight before any use */
                const int64_t i = ib + (int64_t)kUmThreads * j;
                v[j] = i < i1 ? bm_umajor_partner(seed, (uint64_t)i, u, n) : 0;
                pv[j] = i < i1 ? pis[v[j]] : make_int4(0, 0, BM_ABSENT, 0);
                if constexpr (!ROWS) gt[j] = G.G[(int64_t)v[j] * stride + f1];   /* G(v, f1); row 0 when past the range */
            }
            if constexpr (!ROWS) {
#pragma unroll
                for (int j = 0; j < kUmPairs; j++) gu[j] = Gu[pv[j].x];   /* G(u, f2) */
            }
#pragma unroll
            for (int j = 0; j < kUmPairs; j++) {
                const int64_t i = ib + (int64_t)kUmThreads * j;
                if (i >= i1) continue;
                const int32_t f2 = pv[j].x, x = v[j];
                const int32_t F22 = pv[j].w;                /* w_F(f2, f2) */
                const int32_t g_uf2 = ROWS ? rowG[f2] : gu[j], g_vf1 = ROWS ? rowT[x] : gt[j];
                int64_t d = (int64_t)g_uf2 - pu.y + g_vf1 - pv[j].y;
                const bool uv = bit_test(mo, x);
                const bool vu = bit_test(mi, x);
                if (uv || vu) {
                    /* edges between u and v: w_M(u, v), w_M(v, u), w_F(f1, f2), w_F(f2, f1);
                     * f1 = 0 leaves fo, fi empty and bit 0 is never set, so an
                     * unmatched endpoint reads no female weight */
                    const int32_t wuv = uv ? g.m_out[mob + bit_rank(mo, pre, x)].y : BM_ABSENT;
                    const int32_t wvu = vu ? g.m_in[mib + bit_rank(mi, pre + w64, x)].y : BM_ABSENT;
                    const int32_t F12 = bit_test(fo, f2) ? g.f_out[fob + bit_rank(fo, pre + 2 * w64, f2)].y : BM_ABSENT;
                    const int32_t F21 = bit_test(fi, f2) ? g.f_in[fib + bit_rank(fi, pre + 3 * w64, f2)].y : BM_ABSENT;
                    if (uv) d += (int64_t)contrib(wuv, F21) - contrib(wuv, F22) - contrib(wuv, F11) + contrib(wuv, F12);
                    if (vu) d += (int64_t)contrib(wvu, F12) - contrib(wvu, F11) - contrib(wvu, F22) + contrib(wvu, F21);
                }
                if (pu.z != BM_ABSENT) d += contrib(pu.z, F22) - (int64_t)contrib(pu.z, F11);
                if (pv[j].z != BM_ABSENT) d += contrib(pv[j].z, F11) - (int64_t)contrib(pv[j].z, F22);
                delta[i] = d;
            }
        }
        __syncthreads();
    }
}

/* Best of the five permutations of a triple from its nine gains Gv[i][j] =
 * G(vtx_i, f_j), the male self-loops s and female self-loops fs of f_0..f_2
 * (BM_ABSENT when missing), the male weights w[i][j] of edges inside the
 * triple and the female weights F[x][y] = w_F(f_x, f_y) (needed only where a
 * male edge exists).  An edge i -> j is corrected exactly: under permutation
 * p it contributes c(w, F(f_pi, f_pj)) where the gains counted
 * c(w, F(f_pi, f_j)) and c(w, F(f_i, f_pj)) and the old score
 * c(w, F(f_i, f_j)); all of these are among the nine F[x][y].  bd = the
 * largest positive delta (0 if none), bp its permutation (first p on ties). */
__device__ __forceinline__ void rot3_from_gains(const int64_t (*Gv)[3], const int32_t (*w)[3], const int32_t (*F)[3],
                                                const int32_t* s, const int32_t* fs, int64_t& bd, int& bp) {
    bd = 0;
    bp = 0;
#pragma unroll
    for (int p = 1; p <= 5; p++) {
        int64_t d = 0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int pi = perm_of(p, i);
            d += sel3(pi, Gv[i][0], Gv[i][1], Gv[i][2]) - Gv[i][i];
            if (s[i] != BM_ABSENT) d += contrib(s[i], fs[pi]) - (int64_t)contrib(s[i], fs[i]);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                if (i == j) continue;
                const int pj = perm_of(p, j);
                const int32_t x = w[i][j];
                if (x != BM_ABSENT)
                    d += (int64_t)contrib(x, F[pi][pj]) - contrib(x, F[pi][j]) - contrib(x, F[i][pj]) +
                         contrib(x, F[i][j]);
            }
        }
        if (d > bd) { bd = d; bp = p; }
    }
}

template <bool PIS>
__global__ void __launch_bounds__(256) k_gain_rot3(BmDevGraph g, BmGain G, int4* __restrict__ cand, int64_t k,
                                                   int64_t* __restrict__ best, uint8_t* __restrict__ perm,
                                                   int write_perm_into_cand, const int4* __restrict__ pis) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < k; c += stride) {
        const int4 q = cand[c];
        const int32_t vtx[3] = {q.x, q.y, q.z};
        int32_t f[3], s[3], fs[3];
        int64_t Gv[3][3];   /* Gv[i][j] = G(vtx[i], f[j]) */
        if (PIS) {   /* images, diagonal entries and self-loops from the per-vertex words */
            const int4 p[3] = {pis[vtx[0]], pis[vtx[1]], pis[vtx[2]]};
#pragma unroll
            for (int i = 0; i < 3; i++) {
                f[i] = p[i].x;
                Gv[i][i] = p[i].y;
                s[i] = p[i].z;
                fs[i] = p[i].w;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 3; i++) {
                f[i] = g.pi[vtx[i]];
                s[i] = G.m_self[vtx[i]];
                fs[i] = G.f_self[f[i]];
                Gv[i][i] = G.G[(int64_t)vtx[i] * G.stride + f[i]];
            }
        }
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int32_t* Gr = G.G + (int64_t)vtx[i] * G.stride;
#pragma unroll
            for (int j = 0; j < 3; j++)
                if (j != i) Gv[i][j] = Gr[f[j]];
        }
        const bool e01 = male_maybe(G, vtx[0], vtx[1]), e02 = male_maybe(G, vtx[0], vtx[2]),
                   e12 = male_maybe(G, vtx[1], vtx[2]);   /* either direction */
        int32_t w[3][3], F[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) w[i][j] = F[i][j] = BM_ABSENT;
        if (e01 || e02 || e12) {   /* the hash-row descriptors only for a triple with an edge */
            uint64_t md[3], fd[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                md[i] = G.m_out_desc[vtx[i]];
                fd[i] = g.f_out_desc[f[i]];
            }
            const bool have[3][3] = {{false, e01, e02}, {e01, false, e12}, {e02, e12, false}};
            triple_weights(g, G, vtx, md, f, fd, fs, have, w, F);
        }
        int64_t bd;
        int bp;
        rot3_from_gains(Gv, w, F, s, fs, bd, bp);
        best[c] = bd;
        if (perm) perm[c] = (uint8_t)bp;
        if (write_perm_into_cand) cand[c].w = bp;
    }
}

/* a-major 3-vertex lookups: one block per first vertex a of the stream.  The
 * block stages rows G(a, .) and GT(pi(a), .) in LDS as the u-major pair
 * kernel does, and a's male out- and in-neighbour bitmaps.  A triple (a, b, c)
 * then reads G(a, f_b), G(a, f_c), G(b, f_a), G(c, f_a) from LDS, the diagonal
 * terms and self-loops from the gathered per-vertex words of b and c, and
 * only G(b, f_c) and G(c, f_b) from the table in HBM; male edges between a
 * and b or c are tested in the bitmaps, between b and c in the dense
 * adjacency bitmap.  A triple with such an edge (about 7% at n = 18,524)
 * needs the edge's weight and the female weights its correction uses from
 * the hash tables, a chain of dependent loads that nearly every wave's batch
 * would wait for: it is appended instead, with its four staged gains and two
 * gathered ones, to the block's region of a queue in device memory, and
 * k_gain_rot3_edges finishes the queue with full waves after the launch
 * (5.67 -> 4.86 ms per 2^26 triples at n = 18,524).  A triple that finds the region full is
 * finished in place. */
constexpr int kR3Triples = 2;     /* triples per thread in flight */
constexpr int kR3EdgeSplit = 16;  /* k_gain_rot3_edges blocks per queue region */

__global__ void __launch_bounds__(kUmThreads) k_gain_rot3_amajor(BmDevGraph g, BmGain G, const int32_t* __restrict__ GT,
                                                                 const int4* __restrict__ pis,
                                                                 const int64_t* __restrict__ off, uint64_t seed,
                                                                 int64_t* __restrict__ best,
                                                                 uint8_t* __restrict__ perm, int4* __restrict__ queue,
                                                                 int64_t qper, int32_t* __restrict__ qcount) {
    extern __shared__ int4 sm4[];
    const int32_t n = g.n;
    const int64_t stride = G.stride;                   /* multiple of 32 */
    const int32_t s4 = (int32_t)(stride >> 2);
    const int32_t w32 = (n + 1 + 31) >> 5;
    int32_t* rowG = reinterpret_cast<int32_t*>(sm4);
    int32_t* rowT = rowG + stride;
    uint32_t* mo = reinterpret_cast<uint32_t*>(rowT + stride);   /* a -> v */
    uint32_t* mi = mo + w32;                                     /* v -> a */
    int32_t& qn = *reinterpret_cast<int32_t*>(mi + w32);         /* entries this block has queued */
    int4* myq = queue + (int64_t)blockIdx.x * qper * 3;          /* this block's region, 3 x int4 per entry */
    const int tid = threadIdx.x;
    if (tid == 0) qn = 0;
    __syncthreads();
    for (int32_t a = blockIdx.x + 1; a <= n; a += gridDim.x) {
        const int64_t i0 = off[a], i1 = off[a + 1];
        if (i0 == i1) continue;   /* block-uniform */
        const int4 pa = pis[a];
        const int32_t fa = pa.x;
        const uint64_t mda = G.m_out_desc[a], fda = g.f_out_desc[fa];
        for (int32_t w = tid; w < 2 * w32; w += kUmThreads) mo[w] = 0;
        __syncthreads();
        const int64_t mob = g.m_out_ptr[a], moe = g.m_out_ptr[a + 1];
        const int64_t mib = g.m_in_ptr[a], mie = g.m_in_ptr[a + 1];
        const int32_t x_mo = list_id(g.m_out + mob, moe - mob, tid);
        const int32_t x_mi = list_id(g.m_in + mib, mie - mib, tid);
        const __amdgpu_buffer_rsrc_t rG =
            __builtin_amdgcn_make_buffer_rsrc(G.G + (int64_t)a * stride, 0, (int32_t)(stride * 4), kBufWord3);
        const __amdgpu_buffer_rsrc_t rT =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<int32_t*>(GT) + (int64_t)fa * stride, 0, (int32_t)(stride * 4),
                                              kBufWord3);
        V4u ra[kUmStage], rt[kUmStage];
#pragma unroll
        for (int j = 0; j < kUmStage; j++) {
            ra[j] = __builtin_amdgcn_raw_buffer_load_b128(rG, (tid + kUmThreads * j) * 16, 0, 0);
            rt[j] = __builtin_amdgcn_raw_buffer_load_b128(rT, (tid + kUmThreads * j) * 16, 0, 0);
        }
        if (x_mo) atomicOr(&mo[x_mo >> 5], 1u << (x_mo & 31));
        if (x_mi) atomicOr(&mi[x_mi >> 5], 1u << (x_mi & 31));
        for (int64_t e = mob + tid + kUmThreads; e < moe; e += kUmThreads) {
            const int32_t x = g.m_out[e].x;
            atomicOr(&mo[x >> 5], 1u << (x & 31));
        }
        for (int64_t e = mib + tid + kUmThreads; e < mie; e += kUmThreads) {
            const int32_t x = g.m_in[e].x;
            atomicOr(&mi[x >> 5], 1u << (x & 31));
        }
        V4u* sG = reinterpret_cast<V4u*>(rowG);
        V4u* sT = reinterpret_cast<V4u*>(rowT);
#pragma unroll
        for (int j = 0; j < kUmStage; j++) {
            const int32_t c = tid + kUmThreads * j;
            if (c < s4) {
                sG[c] = ra[j];
                sT[c] = rt[j];
            }
        }
        __syncthreads();
        /* the five permutations' best for entry i from its gathered values;
         * the edge corrections (descriptor loads and hash probes) only when
         * edges says so */
        auto finish = [&](int64_t i, int32_t b, int32_t c, const int4& pb, const int4& pc, uint32_t wbc, int32_t gbc,
                          int32_t gcb, bool edges) {
            const int32_t f[3] = {fa, pb.x, pc.x};
            const int64_t Gv[3][3] = {{pa.y, rowG[f[1]], rowG[f[2]]}, {rowT[b], pb.y, gbc}, {rowT[c], gcb, pc.y}};
            const int32_t sv[3] = {pa.z, pb.z, pc.z};
            const int32_t fs[3] = {pa.w, pb.w, pc.w};
            int32_t w[3][3], F[3][3];
#pragma unroll
            for (int x = 0; x < 3; x++)
#pragma unroll
                for (int y = 0; y < 3; y++) {
                    w[x][y] = BM_ABSENT;
                    F[x][y] = x == y ? fs[x] : BM_ABSENT;
                }
            if (edges) {
                const bool ab = (mo[b >> 5] >> (b & 31)) & 1u, ba = (mi[b >> 5] >> (b & 31)) & 1u;
                const bool ac = (mo[c >> 5] >> (c & 31)) & 1u, ca = (mi[c >> 5] >> (c & 31)) & 1u;
                const bool bc = (wbc >> ((b < c ? c : b) & 31)) & 1u;   /* either direction */
                const int32_t vtx[3] = {a, b, c};
                const bool have[3][3] = {{false, ab, ac}, {ba, false, bc}, {ca, bc, false}};
                const uint64_t md[3] = {mda, G.m_out_desc[b], G.m_out_desc[c]};
                const uint64_t fd[3] = {fda, g.f_out_desc[pb.x], g.f_out_desc[pc.x]};
                triple_weights(g, G, vtx, md, f, fd, fs, have, w, F);
            }
            int64_t bd;
            int bp;
            rot3_from_gains(Gv, w, F, sv, fs, bd, bp);
            best[i] = bd;
            perm[i] = (uint8_t)bp;
        };
        auto has_edge = [&](int32_t b, int32_t c, uint32_t wbc) {
            return (((mo[b >> 5] | mi[b >> 5]) >> (b & 31)) & 1u) || (((mo[c >> 5] | mi[c >> 5]) >> (c & 31)) & 1u) ||
                   ((wbc >> ((b < c ? c : b) & 31)) & 1u);
        };
        for (int64_t ib = i0 + tid; ib < i1; ib += (int64_t)kUmThreads * kR3Triples) {
            int32_t vb[kR3Triples], vc[kR3Triples];
            int4 pb[kR3Triples], pc[kR3Triples];
            uint32_t wbc[kR3Triples];   /* dense-bitmap word of the pair {b, c} */
            int32_t gbc[kR3Triples], gcb[kR3Triples];
#pragma unroll
            for (int j = 0; j < kR3Triples; j++) {   /* every gather of a level in flight before any use */
                const int64_t i = ib + (int64_t)kUmThreads * j;
                vb[j] = vc[j] = 0;
                if (i < i1) bm_amajor_partners(seed, (uint64_t)i, a, n, &vb[j], &vc[j]);
                pb[j] = i < i1 ? pis[vb[j]] : make_int4(0, 0, BM_ABSENT, BM_ABSENT);
                pc[j] = i < i1 ? pis[vc[j]] : make_int4(0, 0, BM_ABSENT, BM_ABSENT);
                wbc[j] = G.m_adj ? adj_word(G, vb[j], vc[j]) : ~0u;
            }
#pragma unroll
            for (int j = 0; j < kR3Triples; j++) {
                gbc[j] = G.G[(int64_t)vb[j] * stride + pc[j].x];   /* G(b, f_c) */
                gcb[j] = G.G[(int64_t)vc[j] * stride + pb[j].x];   /* G(c, f_b) */
            }
#pragma unroll
            for (int j = 0; j < kR3Triples; j++) {
                const int64_t i = ib + (int64_t)kUmThreads * j;
                if (i >= i1) continue;
                const int32_t b = vb[j], c = vc[j];
                const bool edges = has_edge(b, c, wbc[j]);
                if (edges && qper > 0) {
                    const int32_t slot = atomicAdd(&qn, 1);
                    if (slot < qper) {
                        const bool ab = (mo[b >> 5] >> (b & 31)) & 1u, ba = (mi[b >> 5] >> (b & 31)) & 1u;
                        const bool ac = (mo[c >> 5] >> (c & 31)) & 1u, ca = (mi[c >> 5] >> (c & 31)) & 1u;
                        const bool bc = (wbc[j] >> ((b < c ? c : b) & 31)) & 1u;
                        const int32_t flags = (ab ? 1 : 0) | (ba ? 2 : 0) | (ac ? 4 : 0) | (ca ? 8 : 0) | (bc ? 16 : 0);
                        int4* e = myq + (int64_t)slot * 3;
                        e[0] = make_int4((int32_t)(uint32_t)(uint64_t)i, (int32_t)((uint64_t)i >> 32), a, b);
                        e[1] = make_int4(c, rowG[pb[j].x], rowG[pc[j].x], rowT[b]);
                        e[2] = make_int4(rowT[c], gbc[j], gcb[j], flags);
                        continue;
                    }
                }
                finish(i, b, c, pb[j], pc[j], wbc[j], gbc[j], gcb[j], edges);
            }
        }
        __syncthreads();
    }
    if (tid == 0) qcount[blockIdx.x] = (int32_t)(qn < qper ? qn : qper);
}

/* The queued triples of k_gain_rot3_amajor: kR3EdgeSplit blocks per queue
 * region.  An entry holds the triple's stream index, its vertices, the six
 * off-diagonal gains and its edge flags (a->b, b->a, a->c, c->a, b-c); the
 * per-vertex words give the images, diagonal gains and self-loops, and
 * triple_weights the edges' weights and the female weights around them. */
__global__ void __launch_bounds__(256) k_gain_rot3_edges(BmDevGraph g, BmGain G, const int4* __restrict__ pis,
                                                         const int4* __restrict__ queue, int64_t qper,
                                                         const int32_t* __restrict__ qcount, int64_t* __restrict__ best,
                                                         uint8_t* __restrict__ perm) {
    const int32_t r = blockIdx.x / kR3EdgeSplit, part = blockIdx.x % kR3EdgeSplit;
    const int32_t cnt = qcount[r];
    const int4* q = queue + (int64_t)r * qper * 3;
    for (int32_t t = part * 256 + threadIdx.x; t < cnt; t += kR3EdgeSplit * 256) {
        const int4 e0 = q[(int64_t)t * 3], e1 = q[(int64_t)t * 3 + 1], e2 = q[(int64_t)t * 3 + 2];
        const int64_t i = (int64_t)(((uint64_t)(uint32_t)e0.y << 32) | (uint32_t)e0.x);
        const int32_t vtx[3] = {e0.z, e0.w, e1.x};
        const int4 p0 = pis[vtx[0]], p1 = pis[vtx[1]], p2 = pis[vtx[2]];
        const int32_t f[3] = {p0.x, p1.x, p2.x};
        const int64_t Gv[3][3] = {{p0.y, e1.y, e1.z}, {e1.w, p1.y, e2.y}, {e2.x, e2.z, p2.y}};
        const int32_t sv[3] = {p0.z, p1.z, p2.z};
        const int32_t fs[3] = {p0.w, p1.w, p2.w};
        const int32_t fl = e2.w;
        const bool have[3][3] = {{false, (fl & 1) != 0, (fl & 4) != 0},
                                 {(fl & 2) != 0, false, (fl & 16) != 0},
                                 {(fl & 8) != 0, (fl & 16) != 0, false}};
        const uint64_t md[3] = {G.m_out_desc[vtx[0]], G.m_out_desc[vtx[1]], G.m_out_desc[vtx[2]]};
        const uint64_t fd[3] = {g.f_out_desc[f[0]], g.f_out_desc[f[1]], g.f_out_desc[f[2]]};
        int32_t w[3][3], F[3][3];
        triple_weights(g, G, vtx, md, f, fd, fs, have, w, F);
        int64_t bd;
        int bp;
        rot3_from_gains(Gv, w, F, sv, fs, bd, bp);
        best[i] = bd;
        perm[i] = (uint8_t)bp;
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
    /* at most 40,768 columns (163,076 bytes of int32 with the edge counter:
     * one block per CU in 160 KiB of LDS), split evenly into multiples of 32.
     * A row of up to 20,384 columns still gets two blocks per CU; a wider
     * table is built in fewer, wider windows at one block per CU, which at
     * n = 100,000 beats more windows at two blocks per CU (three windows of
     * 33,344 columns: 81.8 ms; five of 20,032: 93.5 ms), since every window
     * walks the whole work list again. */
    const int64_t cap = 40768;
    const int64_t chunks = (stride + cap - 1) / cap;
    const int64_t w = ((stride + chunks - 1) / chunks + 31) & ~(int64_t)31;
    return (int32_t)(w < stride ? w : stride);
}

hipError_t bm_launch_dense_fill(const int64_t* out_ptr, const int2* out_rows, const int64_t* in_ptr,
                                const int2* in_rows, const int2* rows, int32_t ndense, int64_t stride,
                                uint16_t* pool, hipStream_t s) {
    if (ndense <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dense_fill, dim3((unsigned)ndense), dim3(256), 0, s, out_ptr, out_rows, in_ptr, in_rows, rows,
                       stride, pool);
    return hipGetLastError();
}

hipError_t bm_launch_row_bounds(const int64_t* ptr, const int2* rows, int32_t n, int32_t cols, int32_t nwin,
                                int32_t* bnd, hipStream_t s) {
    if (n < 0 || cols <= 0 || nwin < 1) return hipErrorInvalidValue;
    const int64_t total = ((int64_t)n + 1) * (nwin + 1);
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_row_bounds, dim3((unsigned)blocks), dim3(256), 0, s, ptr, rows, n, cols, nwin, bnd);
    return hipGetLastError();
}

template <int LAYOUT, int U>
static hipError_t launch_build_rows(const BmDevGraph& g, const BmGain& G, int32_t num_cus, hipStream_t s) {
    constexpr int S = U == 4 ? 4 : 2;   /* chunks in flight per wave */
    const int32_t chunk = bm_gain_chunk_cols(G.stride);
    const bool windowed = chunk < G.stride;
    const size_t lds = ((size_t)chunk + 1) * sizeof(int32_t);   /* row window + edge counter */
    const void* kern =
        windowed ? (const void*)k_gain_build<LAYOUT, U, S, true> : (const void*)k_gain_build<LAYOUT, U, S, false>;
    hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kBuildThreads, lds);
    if (e != hipSuccess) return e;
    if (occ < 1) occ = 1;
    const int64_t cus = num_cus > 0 ? num_cus : 256;
    int64_t blocks = (int64_t)occ * cus;
    if (blocks > g.n) blocks = g.n;
    if (blocks < 1) blocks = 1;
    int64_t dblocks = (G.n_echunks + 3) / 4;       /* 4 waves (chunks) per 256-thread block */
    if (dblocks > cus * 64) dblocks = cus * 64;
    if (dblocks < 1) dblocks = 1;
    if (windowed && (!G.f_in_bnd || !G.f_out_bnd || G.nwin != (int32_t)((G.stride + chunk - 1) / chunk)))
        return hipErrorInvalidValue;
    if (G.ndense > 0) {   /* the work list sets the dense tile weights that are in use */
        const int64_t nw = (int64_t)((g.n + kTileRows) / kTileRows) * G.ndense * kTileRows;
        int64_t zb = (nw + 255) / 256;
        if (zb > cus * 8) zb = cus * 8;
        hipLaunchKernelGGL(k_zero_words, dim3((unsigned)zb), dim3(256), 0, s, G.dense_w, nw);
    }
    const int64_t fb = ((int64_t)g.n + 1 + 255) / 256;
    for (int64_t c0 = 0, win = 0; c0 < G.stride; c0 += chunk, win++) {
        const int32_t nc = (int32_t)(G.stride - c0 < chunk ? G.stride - c0 : chunk);
        /* each window's work list counts its rows' entries afresh */
        hipLaunchKernelGGL(k_zero_words, dim3((unsigned)fb), dim3(256), 0, s, reinterpret_cast<uint32_t*>(G.row_fill),
                           (int64_t)g.n + 1);
        hipLaunchKernelGGL(k_gain_edge_desc<LAYOUT>, dim3((unsigned)dblocks), dim3(256), 0, s, g, G, (int32_t)c0, nc,
                           (int32_t)win);
        if (win == 0 && G.ndense > 0) {   /* every column of the dense part, from the first work list */
            const int64_t nwb = (G.stride / 2 + 2 * kTileThreads - 1) / (2 * kTileThreads);
            int64_t tblocks = (g.n + kTileRows) / kTileRows * nwb;   /* one block per work item */
            if (tblocks > cus * 128) tblocks = cus * 128;
            hipLaunchKernelGGL(k_gain_dense_tile, dim3((unsigned)tblocks), dim3(kTileThreads), 0, s, g, G);
        }
        if (windowed)
            hipLaunchKernelGGL((k_gain_build<LAYOUT, U, S, true>), dim3((unsigned)blocks), dim3(kBuildThreads), lds, s, g,
                               G, (int32_t)c0, nc);
        else
            hipLaunchKernelGGL((k_gain_build<LAYOUT, U, S, false>), dim3((unsigned)blocks), dim3(kBuildThreads), lds, s, g,
                               G, (int32_t)c0, nc);
    }
    return hipGetLastError();
}

template <int U>
static hipError_t launch_build_layout(const BmDevGraph& g, const BmGain& G, int32_t num_cus, hipStream_t s) {
    switch (G.row_layout) {
    case BM_ROWS_P19: return launch_build_rows<BM_ROWS_P19, U>(g, G, num_cus, s);
    case BM_ROWS_P15: return launch_build_rows<BM_ROWS_P15, U>(g, G, num_cus, s);
    default: return launch_build_rows<BM_ROWS_INT2, U>(g, G, num_cus, s);
    }
}

hipError_t bm_launch_gain_build(const BmDevGraph& g, const BmGain& G, int32_t num_cus, hipStream_t s) {
    if (!G.edesc) return hipErrorInvalidValue;
    if (G.row_layout != BM_ROWS_INT2 && !G.f_out_p) return hipErrorInvalidValue;
    switch (G.build_unroll) {
    case 4: return launch_build_layout<4>(g, G, num_cus, s);
    case 16: return launch_build_layout<16>(g, G, num_cus, s);
    case 32: return launch_build_layout<32>(g, G, num_cus, s);
    default: return launch_build_layout<8>(g, G, num_cus, s);
    }
}

hipError_t bm_launch_gain_update(const BmDevGraph& g, const BmGain& G, int32_t* GT, const int4* moves, const int64_t* count,
                                 int64_t max_moves, int64_t* plan, int32_t num_cus, hipStream_t s) {
    if (max_moves <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gain_patch_plan, dim3(1), dim3(kPlanThreads), 0, s, g, moves, count, plan);
    /* persistent grid: every wave walks the item space */
    const int64_t blocks = (int64_t)(num_cus > 0 ? num_cus : 256) * 4;
    hipLaunchKernelGGL(k_gain_update, dim3((unsigned)blocks), dim3(kPatchThreads), 0, s, g, G, moves, count, plan, GT);
    return hipGetLastError();
}

hipError_t bm_launch_gain_swap(const BmDevGraph& g, const BmGain& G, int mode, const int2* cand, const int64_t* off,
                               uint64_t seed, int64_t k, int64_t* delta, const int4* pis, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    int64_t blocks = (k + 255) / 256;
    if (blocks > 65536) blocks = 65536;
#define BM_GAIN_SWAP_LAUNCH(M)                                                                                        \
    if (pis)                                                                                                          \
        hipLaunchKernelGGL((k_gain_swap<M, true>), dim3((unsigned)blocks), dim3(256), 0, s, g, G, cand, off, seed, k, \
                           delta, pis);                                                                               \
    else                                                                                                              \
        hipLaunchKernelGGL((k_gain_swap<M, false>), dim3((unsigned)blocks), dim3(256), 0, s, g, G, cand, off, seed,  \
                           k, delta, pis);
    if (mode == BM_CAND_LIST) {
        BM_GAIN_SWAP_LAUNCH(BM_CAND_LIST)
    } else if (mode == BM_CAND_UMAJOR) {
        BM_GAIN_SWAP_LAUNCH(BM_CAND_UMAJOR)
    } else {
        BM_GAIN_SWAP_LAUNCH(BM_CAND_UNIFORM)
    }
#undef BM_GAIN_SWAP_LAUNCH
    return hipGetLastError();
}

int64_t bm_gain_stride(int32_t n) { return (((int64_t)n + 1) + 31) & ~(int64_t)31; }

size_t bm_gain_rot3_amajor_lds(int32_t n) {
    return 2 * (size_t)bm_gain_stride(n) * 4 + 2 * (size_t)((n + 1 + 31) >> 5) * 4 + 4;   /* rows, 2 bitmaps, count */
}

size_t bm_gain_umajor_lds(int32_t n) {
    const size_t w64 = ((size_t)n + 1 + 63) >> 6;
    return 2 * (size_t)bm_gain_stride(n) * 4 + 4 * w64 * 8 + 4 * w64 * 2;   /* rows, 4 bitmaps, 4 prefix arrays */
}

size_t bm_gain_umajor_bitmap_lds(int32_t n) {
    const size_t w64 = ((size_t)n + 1 + 63) >> 6;
    return 4 * w64 * 8 + 4 * w64 * 4;   /* 4 bitmaps, 4 arrays of 32-bit prefix counts */
}

hipError_t bm_launch_gain_transpose(const BmGain& G, int32_t nrows, int32_t* GT, int32_t num_cus, hipStream_t s) {
    if (G.stride & 31) return hipErrorInvalidValue;   /* 16-byte accesses need 128-byte rows */
    int64_t blocks = ((G.stride + 63) / 64) * (((int64_t)nrows + 63) / 64);
    const int64_t cap = (int64_t)(num_cus > 0 ? num_cus : 256) * 8;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_gain_transpose, dim3((unsigned)blocks), dim3(256), 0, s, G.G, GT, (int64_t)nrows, G.stride);
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
    /* rows staged when they fit (GT required), else bitmaps only */
    const bool rows = bm_gain_umajor_lds(g.n) <= kLdsBytes;
    const size_t lds = rows ? bm_gain_umajor_lds(g.n) : bm_gain_umajor_bitmap_lds(g.n);
    if (lds > kLdsBytes || (rows && !GT)) return hipErrorInvalidValue;
    const void* kern = rows ? (const void*)k_gain_swap_umajor<true, uint16_t> : (const void*)k_gain_swap_umajor<false, uint32_t>;
    hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kUmThreads, lds);
    if (e != hipSuccess) return e;
    if (occ < 1) occ = 1;
    int64_t blocks = (int64_t)occ * (num_cus > 0 ? num_cus : 256);
    if (blocks > g.n) blocks = g.n;
    if (rows)
        hipLaunchKernelGGL((k_gain_swap_umajor<true, uint16_t>), dim3((unsigned)blocks), dim3(kUmThreads), lds, s, g, G,
                           GT, pis, off, seed, delta);
    else
        hipLaunchKernelGGL((k_gain_swap_umajor<false, uint32_t>), dim3((unsigned)blocks), dim3(kUmThreads), lds, s, g, G,
                           GT, pis, off, seed, delta);
    return hipGetLastError();
}

hipError_t bm_launch_gain_best_swap(const BmDevGraph& g, const BmGain& G, const int32_t* GT, const int4* pis,
                                    int4* part, int32_t max_parts, int4* out, int32_t num_cus, hipStream_t s) {
    int64_t blocks = (int64_t)8 * (num_cus > 0 ? num_cus : 256);
    if (blocks > max_parts) blocks = max_parts;
    if (blocks > (int64_t)g.n - 1) blocks = (int64_t)g.n - 1;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_gain_best_swap, dim3((unsigned)blocks), dim3(kBestThreads), 0, s, g, G, GT, pis, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_best_swap_final, dim3(1), dim3(kBestThreads), 0, s, part, (int32_t)blocks, out);
    return hipGetLastError();
}

hipError_t bm_launch_gain_best_partners(const BmDevGraph& g, const BmGain& G, const int32_t* GT, const int4* pis,
                                        int32_t u0, int32_t count, int4* out, int32_t num_cus, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (u0 < 1 || (int64_t)u0 + count - 1 > g.n) return hipErrorInvalidValue;
    int64_t blocks = count;
    const int64_t cap = (int64_t)8 * (num_cus > 0 ? num_cus : 256);
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_gain_best_partners, dim3((unsigned)blocks), dim3(kBestThreads), 0, s, g, G, GT, pis, u0, count,
                       out);
    return hipGetLastError();
}

hipError_t bm_launch_gain_rot3_amajor(const BmDevGraph& g, const BmGain& G, const int32_t* GT, const int4* pis,
                                      const int64_t* off, uint64_t seed, int64_t* best, uint8_t* perm, int4* queue,
                                      int64_t queue_entries, int32_t* queue_counts, int32_t queue_regions,
                                      int32_t num_cus, hipStream_t s) {
    const size_t lds = bm_gain_rot3_amajor_lds(g.n);
    if (lds > kLdsBytes) return hipErrorInvalidValue;
    const void* kern = (const void*)k_gain_rot3_amajor;
    hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kUmThreads, lds);
    if (e != hipSuccess) return e;
    if (occ < 1) occ = 1;
    int64_t blocks = (int64_t)occ * (num_cus > 0 ? num_cus : 256);
    if (blocks > g.n) blocks = g.n;
    if (blocks > queue_regions) blocks = queue_regions;
    if (blocks < 1 || !queue_counts) return hipErrorInvalidValue;
    const int64_t qper = queue ? queue_entries / blocks : 0;   /* entries per block region */
    hipLaunchKernelGGL(k_gain_rot3_amajor, dim3((unsigned)blocks), dim3(kUmThreads), lds, s, g, G, GT, pis, off, seed,
                       best, perm, queue, qper, queue_counts);
    if (qper > 0)
        hipLaunchKernelGGL(k_gain_rot3_edges, dim3((unsigned)(blocks * kR3EdgeSplit)), dim3(256), 0, s, g, G, pis,
                           queue, qper, queue_counts, best, perm);
    return hipGetLastError();
}

hipError_t bm_launch_gain_rot3(const BmDevGraph& g, const BmGain& G, const int4* cand, int64_t k, int64_t* best,
                               uint8_t* perm, int4* cand_out_perm, const int4* pis, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    int64_t blocks = (k + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (pis)
        hipLaunchKernelGGL(k_gain_rot3<true>, dim3((unsigned)blocks), dim3(256), 0, s, g, G, const_cast<int4*>(cand), k,
                           best, perm, cand_out_perm ? 1 : 0, pis);
    else
        hipLaunchKernelGGL(k_gain_rot3<false>, dim3((unsigned)blocks), dim3(256), 0, s, g, G, const_cast<int4*>(cand), k,
                           best, perm, cand_out_perm ? 1 : 0, pis);
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

hipError_t bm_launch_adj_bits(const BmDevGraph& g, uint32_t* bits, int64_t words, int32_t num_cus, hipStream_t s) {
    int64_t blocks = ((int64_t)g.n + 4) / 4;
    const int64_t cap = (int64_t)(num_cus > 0 ? num_cus : 256) * 32;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_adj_bits, dim3((unsigned)blocks), dim3(256), 0, s, g, bits, words);
    return hipGetLastError();
}

hipError_t bm_launch_pack_rows(const int2* in, int64_t m, int layout, uint32_t* out, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    if (layout != BM_ROWS_P15 && layout != BM_ROWS_P19) return hipErrorInvalidValue;
    int64_t blocks = (m + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_pack_rows, dim3((unsigned)blocks), dim3(256), 0, s, in, m, layout, out);
    return hipGetLastError();
}
