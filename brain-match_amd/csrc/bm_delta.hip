/*
 * gfx950 delta kernels: exact swap deltas, exact 3-vertex permutation deltas
 * and the full score.
 *
 * One 64-lane wavefront per candidate, persistent grid.  For a candidate the
 * wave
 *   1. reads pi of the candidate's vertices and the list bounds it needs
 *      (wave-uniform values: scalar loads);
 *   2. builds, in its own slice of LDS, an open-addressing hash table of each
 *      female row it will probe (out-rows of the candidate's female images
 *      for the out-phase, in-rows for the in-phase), by streaming the sorted
 *      row with coalesced 8-byte loads and inserting with LDS compare-and-swap
 *      (load factor <= 3/4).  A row whose table does not fit the slice is
 *      probed in its precomputed global table instead (hub rows, L2 hits);
 *   3. streams the male lists of the candidate's vertices as one index space,
 *      eight entries per lane in flight, gathers pi[neighbour], probes;
 *   4. reduces the per-lane int64 partial deltas across the wave.
 * LDS slots are 4 bytes ((key << 15) | weight) when the female graph has
 * n < 2^17 and weights in 0..32767 ("compact"), 8 bytes otherwise.
 *
 * The reference kernels scan a dense (n+1)^2 int16 matrix row per vertex,
 * O(n) per pair (src/graph-re-alignment-gpu-swap.cu:198-233,
 * src/graph-re-alignment-gpu-3-swap.cu:186-248); here the work is
 * O(deg) per candidate and exact for any int32 weights.
 */
#include <type_traits>

#include "bm_detexp.h"
#include "bm_device.h"
#include "bm_devutil.h"
#include "bm_rng.h"

using namespace bmd;

namespace {

constexpr int kU = 8;       /* list entries per lane in flight */
constexpr int kU3 = 4;      /* same for the 3-vertex kernel (more live registers) */

struct Tab {
    const int2* gt;    /* precomputed global table of the row */
    uint32_t gbits;    /* its log2 size; 0 = empty row */
    uint32_t lbits;    /* log2 size of the LDS table built from the row */
    uint32_t off;      /* slot offset inside the wave's LDS slice */
    uint32_t in_lds;
    int64_t rb, re;    /* row range in the female CSR/CSC */
};

__device__ __forceinline__ uint32_t lds_bits_for(int64_t deg) {
    if (deg <= 0) return 0;
    const int64_t need = (deg * 4 + 2) / 3;   /* ceil(4/3 deg) slots */
    /* the smallest b >= 2 with 2^b >= need */
    return need > 4 ? 64u - (uint32_t)__clzll((unsigned long long)(need - 1)) : 2u;
}

__device__ __forceinline__ Tab make_tab(const uint64_t* __restrict__ desc, const int2* __restrict__ gtab,
                                        const int64_t* __restrict__ ptr, int32_t row) {
    Tab t;
    const uint64_t d = (uint64_t)uni64((int64_t)desc[row]);
    t.gbits = (uint32_t)(d & 63u);
    t.gt = gtab + (d >> 6);
    t.rb = uni64(ptr[row]);
    t.re = uni64(ptr[row + 1]);
    t.lbits = lds_bits_for(t.re - t.rb);
    t.off = 0;
    t.in_lds = 0;
    return t;
}

/* Smallest tables first into the slice; the rest stay global. */
template <int N>
__device__ __forceinline__ void place(Tab* t, uint32_t cap) {
    uint32_t used = 0;
    bool done[N];
#pragma unroll
    for (int i = 0; i < N; i++) done[i] = false;
#pragma unroll
    for (int r = 0; r < N; r++) {
        int pick = -1;
        uint32_t best = 0xffffffffu;
#pragma unroll
        for (int i = 0; i < N; i++) {
            const uint32_t s = t[i].lbits ? (1u << t[i].lbits) : 0u;
            if (!done[i] && s < best) { best = s; pick = i; }
        }
#pragma unroll
        for (int i = 0; i < N; i++) {
            if (i != pick) continue;
            done[i] = true;
            const uint32_t s = t[i].lbits ? (1u << t[i].lbits) : 0u;
            if (s && used + s <= cap) {
                t[i].in_lds = 1;
                t[i].off = used;
                used += s;
            }
        }
    }
}

template <bool COMPACT>
struct Slot;
template <>
struct Slot<true> {
    typedef uint32_t T;
};
template <>
struct Slot<false> {
    typedef unsigned long long T;
};

/* LDS tables are linear-probing tables whose home slots are the first slots
 * of 16-byte buckets (4 compact or 2 wide slots): a key's probe sequence
 * starts at its home bucket, so a lookup reads a whole bucket per round and
 * most lookups, present or absent, end in the first round. */
template <bool COMPACT>
struct Bucket {
    static constexpr uint32_t kLog = COMPACT ? 2u : 1u;   /* log2 slots per 16-byte bucket */
    static constexpr uint32_t kSlots = 1u << kLog;
};

/* First slot of key's home bucket in a table of 2^bits slots (bits >= 2). */
template <bool COMPACT>
__device__ __forceinline__ uint32_t lds_home(int32_t key, uint32_t bits) {
    const uint32_t nb = bits - Bucket<COMPACT>::kLog;
    return nb ? slot_hash(key, nb) << Bucket<COMPACT>::kLog : 0u;
}

/* One bucket of a lookup: w = key's weight if the bucket holds it; stop =
 * the lookup ends here (key found, or an empty slot shows it is absent). */
template <bool COMPACT>
__device__ __forceinline__ void bucket_find(const typename Slot<COMPACT>::T* T, uint32_t s, int32_t key, int32_t& w,
                                            bool& stop) {
    const uint4 v = *reinterpret_cast<const uint4*>(T + s);
    if (COMPACT) {
        const uint32_t k = (uint32_t)key;
        const bool h0 = (v.x >> 15) == k, h1 = (v.y >> 15) == k, h2 = (v.z >> 15) == k, h3 = (v.w >> 15) == k;
        const uint32_t e = h0 ? v.x : (h1 ? v.y : (h2 ? v.z : v.w));
        const bool hit = h0 || h1 || h2 || h3;
        w = hit ? (int32_t)(e & 0x7fffu) : BM_ABSENT;
        stop = hit || v.x == 0u || v.y == 0u || v.z == 0u || v.w == 0u;
    } else {   /* slots {key, w} in (x, y) and (z, w) */
        const bool h0 = (int32_t)v.x == key, h1 = (int32_t)v.z == key;
        w = h0 ? (int32_t)v.y : (h1 ? (int32_t)v.w : BM_ABSENT);
        stop = h0 || h1 || v.x == 0u || v.z == 0u;
    }
}

template <bool COMPACT>
__device__ __forceinline__ int32_t probe_lds(const typename Slot<COMPACT>::T* T, uint32_t bits, int32_t key) {
    const uint32_t mask = (1u << bits) - 1u;
    uint32_t s = lds_home<COMPACT>(key, bits);
    for (;;) {
        int32_t w;
        bool stop;
        bucket_find<COMPACT>(T, s, key, w, stop);
        if (stop) return w;
        s = (s + Bucket<COMPACT>::kSlots) & mask;
    }
}

__device__ __forceinline__ int32_t probe_glb(const int2* __restrict__ T, uint32_t bits, int32_t key) {
    const uint32_t mask = (1u << bits) - 1u;
    uint32_t s = slot_hash(key, bits);
    for (;;) {
        const int2 e = T[s];
        if (e.x == 0) return BM_ABSENT;
        if (e.x == key) return e.y;
        s = (s + 1u) & mask;
    }
}

template <bool COMPACT>
__device__ __forceinline__ int32_t look(const typename Slot<COMPACT>::T* lds, const Tab& t, int32_t key) {
    if (t.in_lds) return probe_lds<COMPACT>(lds + t.off, t.lbits, key);
    if (t.gbits) return probe_glb(t.gt, t.gbits, key);
    return BM_ABSENT;
}

/* U lookups in one table with their probe sequences interleaved: every
 * round issues the next slot of all unresolved keys before checking any, so
 * the latency of a round is paid once for all U keys.  key 0 (an unmatched
 * neighbour, or an inactive item) resolves to BM_ABSENT without a probe. */
template <bool COMPACT, int U>
__device__ __forceinline__ void look_many(const typename Slot<COMPACT>::T* lds, const Tab& t, const int32_t* key,
                                          int32_t* out) {
    typedef typename Slot<COMPACT>::T S;
#pragma unroll
    for (int j = 0; j < U; j++) out[j] = BM_ABSENT;
    if (!t.in_lds && !t.gbits) return;
    const uint32_t bits = t.in_lds ? t.lbits : t.gbits;
    const uint32_t mask = (1u << bits) - 1u;
    uint32_t s[U];
    uint32_t pend = 0;
#pragma unroll
    for (int j = 0; j < U; j++) {
        s[j] = key[j] != 0 ? (t.in_lds ? lds_home<COMPACT>(key[j], bits) : slot_hash(key[j], bits)) : 0u;
        pend |= (key[j] != 0 ? 1u : 0u) << j;
    }
    if (t.in_lds) {
        const S* T = lds + t.off;
        /* first round (the home bucket) for every key at once, with no
         * per-key branch (a key 0 reads bucket 0 and is ignored) */
        pend = 0;
#pragma unroll
        for (int j = 0; j < U; j++) {
            int32_t w;
            bool stop;
            bucket_find<COMPACT>(T, s[j], key[j], w, stop);
            if (key[j] != 0) out[j] = w;
            pend |= (key[j] != 0 && !stop ? 1u : 0u) << j;
            s[j] = (s[j] + Bucket<COMPACT>::kSlots) & mask;
        }
        while (pend) {
#pragma unroll
            for (int j = 0; j < U; j++) {
                if (!((pend >> j) & 1u)) continue;
                int32_t w;
                bool stop;
                bucket_find<COMPACT>(T, s[j], key[j], w, stop);
                out[j] = w;
                if (stop) pend &= ~(1u << j);
                else s[j] = (s[j] + Bucket<COMPACT>::kSlots) & mask;
            }
        }
    } else {
        const int2* __restrict__ T = t.gt;
        {   /* first round for every key at once (a key 0 reads slot 0 and is ignored) */
            int2 e[U];
#pragma unroll
            for (int j = 0; j < U; j++) e[j] = T[s[j]];
            pend = 0;
#pragma unroll
            for (int j = 0; j < U; j++) {
                const bool hit = e[j].x == key[j];
                if (key[j] != 0 && hit) out[j] = e[j].y;
                pend |= (key[j] != 0 && !hit && e[j].x != 0 ? 1u : 0u) << j;
                s[j] = (s[j] + 1u) & mask;
            }
            if (!__ballot(pend != 0)) return;
        }
        while (pend) {
            int2 e[U];
#pragma unroll
            for (int j = 0; j < U; j++) e[j] = ((pend >> j) & 1u) ? T[s[j]] : make_int2(0, 0);
#pragma unroll
            for (int j = 0; j < U; j++) {
                if (!((pend >> j) & 1u)) continue;
                if (e[j].x == 0 || e[j].x == key[j]) {
                    if (e[j].x != 0) out[j] = e[j].y;
                    pend &= ~(1u << j);
                } else {
                    s[j] = (s[j] + 1u) & mask;
                }
            }
        }
    }
}

/* Clear the LDS tables of t[0..N) and insert their rows (one index space
 * over the N rows, U entries per lane in flight).  Keys are unique per row,
 * so a compare-and-swap only ever races with a different key. */
template <bool COMPACT, int N, int U = kU>
__device__ __forceinline__ void build(typename Slot<COMPACT>::T* lds, const Tab* t, const int2* __restrict__ rows,
                                      int lane) {
    typedef typename Slot<COMPACT>::T S;
    wave_sync();
#pragma unroll
    for (int q = 0; q < N; q++) {
        if (!t[q].in_lds) continue;
        uint4* z = reinterpret_cast<uint4*>(lds + t[q].off);
        const uint32_t n16 = ((1u << t[q].lbits) * (uint32_t)sizeof(S)) >> 4;
        for (uint32_t i = lane; i < n16; i += BM_WAVE) z[i] = make_uint4(0, 0, 0, 0);
    }
    wave_sync();
    /* positions in 32 bits: the tables fit the LDS slice, so the rows are short */
    int32_t len[N], tot = 0;
#pragma unroll
    for (int q = 0; q < N; q++) {
        len[q] = t[q].in_lds ? (int32_t)(t[q].re - t[q].rb) : 0;
        tot += len[q];
    }
    for (int32_t base = 0; base < tot; base += BM_WAVE * U) {
        int2 kv[U];
        int which[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            int32_t i = base + lane + BM_WAVE * j;
            which[j] = -1;
            kv[j] = make_int2(0, 0);
            if (i < tot) {
#pragma unroll
                for (int q = 0; q < N; q++) {
                    if (which[j] < 0 && i < len[q]) {
                        which[j] = q;
                        kv[j] = rows[t[q].rb + i];
                    } else if (which[j] < 0) {
                        i -= len[q];
                    }
                }
            }
        }
        uint32_t slot[U], toff[U], msk[U], pend = 0;
        S packed[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            uint32_t off = 0, bits = 2;
#pragma unroll
            for (int q = 0; q < N; q++)
                if (which[j] == q) { off = t[q].off; bits = t[q].lbits; }
            msk[j] = (1u << bits) - 1u;
            toff[j] = off;
            slot[j] = lds_home<COMPACT>(kv[j].x, bits);
            if (COMPACT)
                packed[j] = ((uint32_t)kv[j].x << 15) | (uint32_t)kv[j].y;
            else
                packed[j] = ((unsigned long long)(uint32_t)kv[j].y << 32) | (uint32_t)kv[j].x;
            pend |= (which[j] >= 0 ? 1u : 0u) << j;
        }
        /* interleaved inserts: all pending compare-and-swaps issued per
         * round.  The first round reads each entry's home bucket and aims at
         * its first empty slot (slots before it stay taken, so the probe
         * sequence stays unbroken); after it, only entries that found the
         * bucket full or lost a race go on, slot by slot. */
        {
            uint32_t first[U];
#pragma unroll
            for (int j = 0; j < U; j++) {
                const uint4 v = *reinterpret_cast<const uint4*>(lds + toff[j] + slot[j]);
                if (COMPACT)
                    first[j] = v.x == 0u ? 0u : (v.y == 0u ? 1u : (v.z == 0u ? 2u : (v.w == 0u ? 3u : 4u)));
                else
                    first[j] = v.x == 0u ? 0u : (v.z == 0u ? 1u : 2u);
            }
            S old[U];
#pragma unroll
            for (int j = 0; j < U; j++) {
                const bool go = which[j] >= 0 && first[j] < Bucket<COMPACT>::kSlots;
                slot[j] = (slot[j] + (go ? first[j] : Bucket<COMPACT>::kSlots - 1u)) & msk[j];
                old[j] = go ? atomicCAS(&lds[toff[j] + slot[j]], (S)0, packed[j]) : (S)1;
            }
            pend = 0;
#pragma unroll
            for (int j = 0; j < U; j++) {
                pend |= (which[j] >= 0 && old[j] != (S)0 ? 1u : 0u) << j;
                slot[j] = (slot[j] + 1u) & msk[j];
            }
        }
        while (pend) {
            S old[U];
#pragma unroll
            for (int j = 0; j < U; j++)
                old[j] = ((pend >> j) & 1u) ? atomicCAS(&lds[toff[j] + slot[j]], (S)0, packed[j]) : (S)0;
#pragma unroll
            for (int j = 0; j < U; j++) {
                if (!((pend >> j) & 1u)) continue;
                if (old[j] == (S)0) {
                    pend &= ~(1u << j);
                } else {
                    slot[j] = (slot[j] + 1u) & msk[j];
                }
            }
        }
    }
    wave_sync();
}

/* ------------------------------------------------------------------ */
/* Swap deltas                                                         */
/* ------------------------------------------------------------------ */

/* One phase over the lists of u (L0 entries at b0) and v (L1 at b1).
 * A = female row of f1 = pi(u), B = row of f2 = pi(v).
 * Out-phase: an edge to u or v uses that endpoint's post-swap female (this
 * covers u<->v edges and self-loops).  In-phase: edges from u or v were
 * counted by the out-phase and are skipped.  Compact graphs sum a lane's
 * contributions in 32 bits: each is within +-65534 and a lane takes at most
 * 2^12 + U edges (deduplicated lists of a graph with n < 2^17 hold fewer
 * than 2^18 entries together). */
template <bool COMPACT, bool OUTP, int U>
__device__ __forceinline__ int64_t swap_scan(const int2* __restrict__ adj, const int32_t* __restrict__ pi,
                                             int64_t b0, int64_t L0, int64_t b1, int64_t L1,
                                             const typename Slot<COMPACT>::T* lds, const Tab& A, const Tab& B,
                                             int32_t u, int32_t v, int32_t f1, int32_t f2, int lane) {
    typedef typename std::conditional<COMPACT, int32_t, int64_t>::type Acc;
    Acc acc = 0;
    /* list positions in 32 bits (a list is shorter than 2^31) */
    const int32_t M0 = (int32_t)L0, M = (int32_t)(L0 + L1);
    for (int32_t base = 0; base < M; base += BM_WAVE * U) {
        int2 ed[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            const int32_t i = base + lane + BM_WAVE * j;
            ed[j] = i < M ? adj[i < M0 ? b0 + i : b1 + (i - M0)] : make_int2(0, 0);
        }
        int32_t gx[U];
#pragma unroll
        for (int j = 0; j < U; j++) gx[j] = pi[ed[j].x];
        int32_t kA[U], kB[U], rA[U], rB[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            const int32_t x = ed[j].x;
            const bool side = base + lane + BM_WAVE * j >= M0;  /* false: list of u */
            int32_t go = gx[j], gn = gx[j];
            if (OUTP) {
                if (x == u) { go = f1; gn = f2; }
                else if (x == v) { go = f2; gn = f1; }
            } else if (x == u || x == v) {
                go = gn = 0;  /* counted by the out-phase */
            }
            if (x == 0) go = gn = 0;
            kA[j] = side ? gn : go;
            kB[j] = side ? go : gn;
        }
        look_many<COMPACT, U>(lds, A, kA, rA);
        look_many<COMPACT, U>(lds, B, kB, rB);
#pragma unroll
        for (int j = 0; j < U; j++) {
            const bool side = base + lane + BM_WAVE * j >= M0;
            const int32_t w = ed[j].y;
            const int32_t cA = contrib(w, rA[j]);
            const int32_t cB = contrib(w, rB[j]);
            acc += side ? (Acc)cA - (Acc)cB : (Acc)cB - (Acc)cA;
        }
    }
    return (int64_t)acc;
}

/* Exact delta of swapping the images of u != v (f1 = pi(u), f2 = pi(v)) in
 * the matching g.pi, computed by the whole wave; every lane returns it. */
template <bool COMPACT, int U>
__device__ __forceinline__ int64_t swap_delta_wave(const BmDevGraph& g, int32_t u, int32_t v, int32_t f1,
                                                   int32_t f2, typename Slot<COMPACT>::T* lds, uint32_t cap,
                                                   int lane) {
    const int64_t ou0 = uni64(g.m_out_ptr[u]), ou1 = uni64(g.m_out_ptr[u + 1]);
    const int64_t ov0 = uni64(g.m_out_ptr[v]), ov1 = uni64(g.m_out_ptr[v + 1]);
    const int64_t iu0 = uni64(g.m_in_ptr[u]), iu1 = uni64(g.m_in_ptr[u + 1]);
    const int64_t iv0 = uni64(g.m_in_ptr[v]), iv1 = uni64(g.m_in_ptr[v + 1]);
    int64_t acc;
    {
        Tab t[2] = {make_tab(g.f_out_desc, g.f_out_tab, g.f_out_ptr, f1),
                    make_tab(g.f_out_desc, g.f_out_tab, g.f_out_ptr, f2)};
        place<2>(t, cap);
        build<COMPACT, 2, U>(lds, t, g.f_out, lane);
        acc = swap_scan<COMPACT, true, U>(g.m_out, g.pi, ou0, ou1 - ou0, ov0, ov1 - ov0, lds, t[0], t[1], u, v, f1,
                                       f2, lane);
    }
    {
        Tab t[2] = {make_tab(g.f_in_desc, g.f_in_tab, g.f_in_ptr, f1),
                    make_tab(g.f_in_desc, g.f_in_tab, g.f_in_ptr, f2)};
        place<2>(t, cap);
        build<COMPACT, 2, U>(lds, t, g.f_in, lane);
        acc += swap_scan<COMPACT, false, U>(g.m_in, g.pi, iu0, iu1 - iu0, iv0, iv1 - iv0, lds, t[0], t[1], u, v,
                                         f1, f2, lane);
    }
    return wave_sum64(acc);
}

template <int MODE, bool COMPACT, int U>
__global__ void __launch_bounds__(512) k_swap(BmDevGraph g, const int2* __restrict__ cand,
                                             const int64_t* __restrict__ off, uint64_t seed, int64_t k,
                                             int64_t* __restrict__ delta, uint32_t cap) {
    typedef typename Slot<COMPACT>::T S;
    extern __shared__ uint4 smem[];
    const int lane = threadIdx.x & (BM_WAVE - 1);
    const int wave = threadIdx.x / BM_WAVE;
    const int wpb = blockDim.x / BM_WAVE;
    S* lds = reinterpret_cast<S*>(smem) + (size_t)wave * cap;
    const int64_t nw = (int64_t)gridDim.x * wpb;
    for (int64_t c = (int64_t)blockIdx.x * wpb + wave; c < k; c += nw) {
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
        u = uni(u);
        v = uni(v);
        if (u == v) {
            if (lane == 0) delta[c] = 0;
            continue;
        }
        const int32_t f1 = uni(g.pi[u]), f2 = uni(g.pi[v]);
        const int64_t acc = swap_delta_wave<COMPACT, U>(g, u, v, f1, f2, lds, cap, lane);
        if (lane == 0) delta[c] = acc;
    }
}

/* ------------------------------------------------------------------ */
/* 3-vertex permutation deltas                                         */
/* ------------------------------------------------------------------ */

struct Rot3Acc {
    int64_t s[3][3];   /* s[i][j]: external contributions of position i if it maps to female j */
    int64_t in[6];     /* internal-edge delta per permutation 1..5 (index 0 unused) */
};

/* One phase over the three lists (positions 0,1,2).  External edges add
 * their contribution under each of the three female images; edges inside
 * the triple (self-loops included) are summed per permutation in the
 * out-phase and skipped in the in-phase. */
template <bool COMPACT, bool OUTP, int U3>
__device__ __forceinline__ void rot3_scan(const int2* __restrict__ adj, const int32_t* __restrict__ pi,
                                          const int64_t* b, const int64_t* len, const typename Slot<COMPACT>::T* lds,
                                          const Tab* T, const int32_t* vtx, const int32_t* f, Rot3Acc& A, int lane) {
    const int64_t L = len[0] + len[1] + len[2];
    for (int64_t base = 0; base < L; base += (int64_t)BM_WAVE * U3) {
        int2 ed[U3];
        int pos[U3];
#pragma unroll
        for (int j = 0; j < U3; j++) {
            int64_t i = base + lane + (int64_t)BM_WAVE * j;
            pos[j] = -1;
            ed[j] = make_int2(0, 0);
            if (i < L) {
                if (i < len[0]) { pos[j] = 0; ed[j] = adj[b[0] + i]; }
                else if (i < len[0] + len[1]) { pos[j] = 1; ed[j] = adj[b[1] + i - len[0]]; }
                else { pos[j] = 2; ed[j] = adj[b[2] + i - len[0] - len[1]]; }
            }
        }
        int32_t gx[U3];
#pragma unroll
        for (int j = 0; j < U3; j++) gx[j] = pi[ed[j].x];
        int32_t key[U3], r0[U3], r1[U3], r2[U3];
        int jxs[U3];
#pragma unroll
        for (int j = 0; j < U3; j++) {
            const int32_t x = ed[j].x;
            jxs[j] = x == 0 ? -2 : (x == vtx[0] ? 0 : (x == vtx[1] ? 1 : (x == vtx[2] ? 2 : -1)));
            key[j] = jxs[j] == -1 ? gx[j] : 0;   /* external edges only */
        }
        look_many<COMPACT, U3>(lds, T[0], key, r0);
        look_many<COMPACT, U3>(lds, T[1], key, r1);
        look_many<COMPACT, U3>(lds, T[2], key, r2);
#pragma unroll
        for (int j = 0; j < U3; j++) {
            const int32_t w = ed[j].y;
            const int i = pos[j];
            const int jx = jxs[j];
            if (jx == -1) {
                const int64_t c0 = contrib(w, r0[j]), c1 = contrib(w, r1[j]), c2 = contrib(w, r2[j]);
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    if (i != q) continue;
                    A.s[q][0] += c0;
                    A.s[q][1] += c1;
                    A.s[q][2] += c2;
                }
            } else if (OUTP && jx >= 0) {
                /* edge pos i -> pos jx, both inside the triple (rare) */
                const int64_t old = contrib(w, look<COMPACT>(lds, sel3(i, T[0], T[1], T[2]), sel3(jx, f[0], f[1], f[2])));
#pragma unroll
                for (int p = 1; p <= 5; p++) {
                    const int ni = perm_of(p, i), nj = perm_of(p, jx);
                    A.in[p] += (int64_t)contrib(w, look<COMPACT>(lds, sel3(ni, T[0], T[1], T[2]),
                                                                 sel3(nj, f[0], f[1], f[2]))) - old;
                }
            }
        }
    }
}

/* Exact deltas d[p], p = 1..5, of the five non-identity permutations of the
 * images f of the distinct vertices vtx in g.pi, computed by the whole wave;
 * every lane returns them (d[0] = 0). */
template <bool COMPACT, int U3 = kU3>
__device__ __forceinline__ void rot3_delta_wave(const BmDevGraph& g, const int32_t* vtx, const int32_t* f,
                                                typename Slot<COMPACT>::T* lds, uint32_t cap, int lane,
                                                int64_t* d) {
    Rot3Acc A;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) A.s[i][j] = 0;
#pragma unroll
    for (int p = 0; p < 6; p++) A.in[p] = 0;
    {
        Tab T[3] = {make_tab(g.f_out_desc, g.f_out_tab, g.f_out_ptr, f[0]),
                    make_tab(g.f_out_desc, g.f_out_tab, g.f_out_ptr, f[1]),
                    make_tab(g.f_out_desc, g.f_out_tab, g.f_out_ptr, f[2])};
        place<3>(T, cap);
        build<COMPACT, 3, U3>(lds, T, g.f_out, lane);
        int64_t b[3], len[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            b[i] = uni64(g.m_out_ptr[vtx[i]]);
            len[i] = uni64(g.m_out_ptr[vtx[i] + 1]) - b[i];
        }
        rot3_scan<COMPACT, true, U3>(g.m_out, g.pi, b, len, lds, T, vtx, f, A, lane);
    }
    {
        Tab T[3] = {make_tab(g.f_in_desc, g.f_in_tab, g.f_in_ptr, f[0]),
                    make_tab(g.f_in_desc, g.f_in_tab, g.f_in_ptr, f[1]),
                    make_tab(g.f_in_desc, g.f_in_tab, g.f_in_ptr, f[2])};
        place<3>(T, cap);
        build<COMPACT, 3, U3>(lds, T, g.f_in, lane);
        int64_t b[3], len[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            b[i] = uni64(g.m_in_ptr[vtx[i]]);
            len[i] = uni64(g.m_in_ptr[vtx[i] + 1]) - b[i];
        }
        rot3_scan<COMPACT, false, U3>(g.m_in, g.pi, b, len, lds, T, vtx, f, A, lane);
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) A.s[i][j] = wave_sum64(A.s[i][j]);
    d[0] = 0;
#pragma unroll
    for (int p = 1; p <= 5; p++) {
        int64_t x = wave_sum64(A.in[p]);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int ni = perm_of(p, i);
            x += sel3(ni, A.s[i][0], A.s[i][1], A.s[i][2]) - A.s[i][i];
        }
        d[p] = x;
    }
}

template <bool COMPACT, int U3>
__global__ void __launch_bounds__(512) k_rot3(BmDevGraph g, int4* __restrict__ cand, int64_t k,
                                             int64_t* __restrict__ best, uint8_t* __restrict__ perm,
                                             int write_perm_into_cand, uint32_t cap) {
    typedef typename Slot<COMPACT>::T S;
    extern __shared__ uint4 smem[];
    const int lane = threadIdx.x & (BM_WAVE - 1);
    const int wave = threadIdx.x / BM_WAVE;
    const int wpb = blockDim.x / BM_WAVE;
    S* lds = reinterpret_cast<S*>(smem) + (size_t)wave * cap;
    const int64_t nw = (int64_t)gridDim.x * wpb;
    for (int64_t c = (int64_t)blockIdx.x * wpb + wave; c < k; c += nw) {
        const int4 q = cand[c];
        const int32_t vtx[3] = {uni(q.x), uni(q.y), uni(q.z)};
        const int32_t f[3] = {uni(g.pi[vtx[0]]), uni(g.pi[vtx[1]]), uni(g.pi[vtx[2]])};
        int64_t d5[6];
        rot3_delta_wave<COMPACT, U3>(g, vtx, f, lds, cap, lane, d5);
        if (lane == 0) {
            int64_t bd = 0;
            int bp = 0;
#pragma unroll
            for (int p = 1; p <= 5; p++)
                if (d5[p] > bd) { bd = d5[p]; bp = p; }
            best[c] = bd;
            if (perm) perm[c] = (uint8_t)bp;
            if (write_perm_into_cand) cand[c].w = bp;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Full score                                                          */
/* ------------------------------------------------------------------ */

template <bool COMPACT>
__global__ void __launch_bounds__(512) k_score(BmDevGraph g, int64_t* __restrict__ acc_out, uint32_t cap) {
    typedef typename Slot<COMPACT>::T S;
    extern __shared__ uint4 smem[];
    const int lane = threadIdx.x & (BM_WAVE - 1);
    const int wave = threadIdx.x / BM_WAVE;
    const int wpb = blockDim.x / BM_WAVE;
    S* lds = reinterpret_cast<S*>(smem) + (size_t)wave * cap;
    const int64_t nw = (int64_t)gridDim.x * wpb;
    int64_t acc = 0;
    for (int64_t uu = (int64_t)blockIdx.x * wpb + wave + 1; uu <= g.n; uu += nw) {
        const int32_t u = (int32_t)uu;
        const int64_t b = uni64(g.m_out_ptr[u]), e = uni64(g.m_out_ptr[u + 1]);
        if (b == e) continue;
        Tab t[1] = {make_tab(g.f_out_desc, g.f_out_tab, g.f_out_ptr, uni(g.pi[u]))};
        if (t[0].re - t[0].rb < 4 * (e - b) + 64) place<1>(t, cap); /* build only when probes pay for it */
        build<COMPACT, 1>(lds, t, g.f_out, lane);
        for (int64_t base = b; base < e; base += (int64_t)BM_WAVE * kU) {
            int2 ed[kU];
#pragma unroll
            for (int j = 0; j < kU; j++) {
                const int64_t i = base + lane + (int64_t)BM_WAVE * j;
                ed[j] = i < e ? g.m_out[i] : make_int2(0, 0);
            }
            int32_t gx[kU];
#pragma unroll
            for (int j = 0; j < kU; j++) gx[j] = g.pi[ed[j].x];
            int32_t r[kU];
            look_many<COMPACT, kU>(lds, t[0], gx, r);
#pragma unroll
            for (int j = 0; j < kU; j++) acc += contrib(ed[j].y, r[j]);
        }
    }
    acc = wave_sum64(acc);
    if (lane == 0 && acc != 0) atomicAdd(reinterpret_cast<unsigned long long*>(acc_out), (unsigned long long)acc);
}

/* ------------------------------------------------------------------ */
/* Annealing chains                                                    */
/* ------------------------------------------------------------------ */

/* One wave per chain.  Step s of a chain with stream seed q:
 *   kind:  r_k = rand(q2, s); a 3-vertex move if lo32(r_k) < p_rot3 (n >= 3);
 *   pool:  with a proposal pool of P pairs for the chain's instance,
 *          r_l = rand(q4, s); the move is built on pool entry
 *          range(hi32(r_l), P) if lo32(r_l) < p_pool;
 *   swap:  (u, v) = that pool pair, else uniform pair s of stream q0;
 *          delta exact;
 *   3-move: (a, b) = that pool pair and c = bm_third_vertex(q3, s, n, a, b),
 *          else (a, b, c) = uniform triple s of stream q3; permutation
 *          p = 1 + range(hi32(r_k), 5), delta of that permutation exact;
 *   accept if delta >= 0, else if U < exp(delta * beta_s) with
 *          U = (rand(q1, s) >> 11) * 2^-53 and beta_s = beta0 * exp(s * lam),
 *          exp being the deterministic bm_det_exp;
 * with q_i = rand(q, i).  The best matching is kept lazily: a chain standing
 * on its best score copies the matching aside only when it accepts a move
 * that lowers the score (the best is the latest matching reaching the
 * maximum).  The Metropolis rule and the geometric schedule follow
 * simulated_annealing (src/best-swap/score-lookup.c:345-400); its delta
 * comes from a lookup table and its stream from xoshiro256**, here the
 * delta is recomputed exactly every step and the streams are counter-based.
 * Each chain also sums the male degrees of the vertices of its moves (the
 * byte model of the delta kernels) and counts its 3-vertex and pool moves. */
template <bool COMPACT, int U, bool ROT>
__global__ void __launch_bounds__(512) k_anneal(const BmDevGraph* __restrict__ insts, BmChain* __restrict__ chains,
                                               int64_t n_chains, int32_t* __restrict__ pool, int64_t steps,
                                               BmAnnealSched sc, int64_t* __restrict__ traj, uint32_t cap,
                                               const BmProposalPool* __restrict__ props) {
    typedef typename Slot<COMPACT>::T S;
    extern __shared__ uint4 smem[];
    const int lane = threadIdx.x & (BM_WAVE - 1);
    const int wave = threadIdx.x / BM_WAVE;
    const int wpb = blockDim.x / BM_WAVE;
    S* lds = reinterpret_cast<S*>(smem) + (size_t)wave * cap;
    const int64_t nw = (int64_t)gridDim.x * wpb;
    for (int64_t c = (int64_t)blockIdx.x * wpb + wave; c < n_chains; c += nw) {
        const BmChain ch = chains[c];
        BmDevGraph g = insts[uni(ch.inst)];
        const int32_t n = g.n;
        int32_t* pi = pool + uni64(ch.pi_off);
        int32_t* pib = pi + ((int64_t)n + 1);
        g.pi = pi;
        const uint64_t q = (uint64_t)uni64((int64_t)ch.seed);
        const uint64_t q0 = bm_rand_at(q, 0), q1 = bm_rand_at(q, 1), q2 = bm_rand_at(q, 2), q3 = bm_rand_at(q, 3);
        const uint64_t q4 = bm_rand_at(q, 4);
        int64_t score = uni64(ch.score), best = uni64(ch.best), nacc = uni64(ch.accepted);
        int at_best = uni(ch.at_best);
        const int64_t s0 = uni64(ch.step);
        /* the instance's proposal pool (count 0: every move uniform) */
        const int2* pp = nullptr;
        int64_t pcount = 0;
        if (sc.p_pool && props) {
            pp = reinterpret_cast<const int2*>(uni64((int64_t)reinterpret_cast<uintptr_t>(props[uni(ch.inst)].pairs)));
            pcount = uni64(props[uni(ch.inst)].count);
        }
        int64_t dsw = 0, drt = 0, nrt = 0, npool = 0;
        for (int64_t s = s0; s < s0 + steps; s++) {
            const uint64_t rk = bm_rand_at(q2, (uint64_t)s);
            /* ROT = false: swap-only chains (p_rot3 == 0), built without the
             * 3-vertex path and its registers */
            const bool rot = ROT && n >= 3 && (uint64_t)(uint32_t)rk < sc.p_rot3;
            /* a move from the pool: entry range(hi32, count) of draw s of stream q4 */
            int2 from = make_int2(0, 0);
            if (pcount > 0) {
                const uint64_t rl = bm_rand_at(q4, (uint64_t)s);
                if ((uint64_t)(uint32_t)rl < sc.p_pool) {
                    from = pp[(int64_t)(((uint64_t)(uint32_t)(rl >> 32) * (uint64_t)pcount) >> 32)];
                    from.x = uni(from.x);
                    from.y = uni(from.y);
                    npool++;
                }
            }
            int32_t vtx[3] = {0, 0, 0}, f[3] = {0, 0, 0};
            int64_t d;
            int p = 0;
            if (!rot) {
                if (from.x) {
                    vtx[0] = from.x;
                    vtx[1] = from.y;
                } else {
                    bm_uniform_pair(q0, (uint64_t)s, n, &vtx[0], &vtx[1]);
                }
                dsw += uni64((g.m_out_ptr[vtx[0] + 1] - g.m_out_ptr[vtx[0]]) + (g.m_in_ptr[vtx[0] + 1] - g.m_in_ptr[vtx[0]]) +
                             (g.m_out_ptr[vtx[1] + 1] - g.m_out_ptr[vtx[1]]) + (g.m_in_ptr[vtx[1] + 1] - g.m_in_ptr[vtx[1]]));
                f[0] = uni(pi[vtx[0]]);
                f[1] = uni(pi[vtx[1]]);
                d = swap_delta_wave<COMPACT, U>(g, vtx[0], vtx[1], f[0], f[1], lds, cap, lane);
            } else {
                if (from.x) {
                    vtx[0] = from.x;
                    vtx[1] = from.y;
                    vtx[2] = bm_third_vertex(q3, (uint64_t)s, n, from.x, from.y);
                } else {
                    bm_uniform_triple(q3, (uint64_t)s, n, &vtx[0], &vtx[1], &vtx[2]);
                }
                nrt++;
#pragma unroll
                for (int i = 0; i < 3; i++)
                    drt += uni64((g.m_out_ptr[vtx[i] + 1] - g.m_out_ptr[vtx[i]]) + (g.m_in_ptr[vtx[i] + 1] - g.m_in_ptr[vtx[i]]));
#pragma unroll
                for (int i = 0; i < 3; i++) f[i] = uni(pi[vtx[i]]);
                int64_t d5[6];
                rot3_delta_wave<COMPACT, U>(g, vtx, f, lds, cap, lane, d5);
                p = 1 + (int)bm_range32((uint32_t)(rk >> 32), 5u);
                d = d5[1];
#pragma unroll
                for (int j = 2; j <= 5; j++)
                    if (p == j) d = d5[j];
            }
            bool accept = d >= 0;
            if (!accept) {
                const double beta = sc.beta0 * bm_det_exp((double)s * sc.lam);
                const double x = (double)d * beta;
                const double r = (double)(bm_rand_at(q1, (uint64_t)s) >> 11) * 0x1p-53;
                accept = r < bm_det_exp(x);
            }
            if (!accept) continue;
            if (at_best && d < 0) {
                for (int64_t i = lane; i <= n; i += BM_WAVE) pib[i] = pi[i];
                at_best = 0;
            }
            if (lane == 0) {
                if (!rot) {
                    pi[vtx[0]] = f[1];
                    pi[vtx[1]] = f[0];
                } else {
#pragma unroll
                    for (int i = 0; i < 3; i++) pi[vtx[i]] = sel3(perm_of(p, i), f[0], f[1], f[2]);
                }
            }
            __threadfence_block();
            score += d;
            nacc++;
            if (score >= best) {
                best = score;
                at_best = 1;
            }
        }
        if (lane == 0) {
            BmChain o = ch;
            o.score = score;
            o.best = best;
            o.accepted = nacc;
            o.step = s0 + steps;
            o.at_best = at_best;
            o.deg_swap += dsw;
            o.deg_rot3 += drt;
            o.n_rot3 += nrt;
            o.n_learned += npool;
            if (traj && o.nseg < sc.traj_cap) traj[c * sc.traj_cap + o.nseg] = score;
            o.nseg++;
            chains[c] = o;
        }
    }
}

template <typename K>
hipError_t launch_delta(K kernel, const BmLaunch& L, int64_t work, size_t slot_bytes, hipStream_t s,
                        dim3* grid, dim3* block, size_t* lds) {
    (void)s;
    *block = dim3(L.waves_per_block * BM_WAVE);
    *lds = (size_t)L.waves_per_block * (size_t)L.lds_slots * slot_bytes;
    hipError_t e = hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)*lds);
    if (e != hipSuccess) return e;
    /* persistent grid: never more blocks than can be resident at once */
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)kernel, (int)block->x, *lds);
    if (e != hipSuccess) return e;
    int per_cu = L.blocks_per_cu < occ ? L.blocks_per_cu : occ;
    if (per_cu < 1) per_cu = 1;
    int64_t blocks = (work + L.waves_per_block - 1) / L.waves_per_block;
    const int64_t cap = (int64_t)per_cu * (L.num_cus > 0 ? L.num_cus : 256);
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    *grid = dim3((unsigned)blocks);
    return hipSuccess;
}

}  // namespace

size_t bm_slot_bytes(const BmDevGraph& g) { return g.compact ? 4 : 8; }

hipError_t bm_launch_swap(const BmDevGraph& g, const BmLaunch& L, int mode, const int2* cand, const int64_t* off,
                          uint64_t seed, int64_t k, int64_t* delta, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    dim3 grid, block;
    size_t lds;
    hipError_t e;
#define BM_SWAP_CASE(M, C, U)                                                                           \
    e = launch_delta(k_swap<M, C, U>, L, k, C ? 4 : 8, s, &grid, &block, &lds);                         \
    if (e != hipSuccess) return e;                                                                      \
    hipLaunchKernelGGL((k_swap<M, C, U>), grid, block, lds, s, g, cand, off, seed, k, delta, (uint32_t)L.lds_slots);
#define BM_SWAP_U(M, C)                             \
    if (L.unroll == 4) { BM_SWAP_CASE(M, C, 4) }     \
    else if (L.unroll == 2) { BM_SWAP_CASE(M, C, 2) } \
    else { BM_SWAP_CASE(M, C, 8) }
    if (mode == BM_CAND_LIST) {
        if (g.compact) { BM_SWAP_U(BM_CAND_LIST, true) } else { BM_SWAP_U(BM_CAND_LIST, false) }
    } else if (mode == BM_CAND_UMAJOR) {
        if (g.compact) { BM_SWAP_U(BM_CAND_UMAJOR, true) } else { BM_SWAP_U(BM_CAND_UMAJOR, false) }
    } else {
        if (g.compact) { BM_SWAP_U(BM_CAND_UNIFORM, true) } else { BM_SWAP_U(BM_CAND_UNIFORM, false) }
    }
#undef BM_SWAP_U
#undef BM_SWAP_CASE
    return hipGetLastError();
}

template <bool C, int U3>
static hipError_t launch_rot3_u(const BmDevGraph& g, const BmLaunch& L, const int4* cand, int64_t k, int64_t* best,
                                uint8_t* perm, int4* cand_out_perm, hipStream_t s) {
    dim3 grid, block;
    size_t lds;
    const hipError_t e = launch_delta(k_rot3<C, U3>, L, k, C ? 4 : 8, s, &grid, &block, &lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_rot3<C, U3>), grid, block, lds, s, g, const_cast<int4*>(cand), k, best, perm,
                       cand_out_perm ? 1 : 0, (uint32_t)L.lds_slots);
    return hipGetLastError();
}

/* 3-vertex kernel: 2 or 4 list entries per lane in flight (L.unroll). */
hipError_t bm_launch_rot3(const BmDevGraph& g, const BmLaunch& L, const int4* cand, int64_t k, int64_t* best,
                          uint8_t* perm, int4* cand_out_perm, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    if (g.compact)
        return L.unroll == 2 ? launch_rot3_u<true, 2>(g, L, cand, k, best, perm, cand_out_perm, s)
                             : launch_rot3_u<true, 4>(g, L, cand, k, best, perm, cand_out_perm, s);
    return L.unroll == 2 ? launch_rot3_u<false, 2>(g, L, cand, k, best, perm, cand_out_perm, s)
                         : launch_rot3_u<false, 4>(g, L, cand, k, best, perm, cand_out_perm, s);
}

/* Annealing chains: 2 or 4 list entries per lane in flight (L.unroll). */
hipError_t bm_anneal_grid(int compact, int rot, const BmLaunch& L, int64_t n_chains, dim3* grid, dim3* block,
                          size_t* lds) {
#define BM_ANNEAL_GRID(C, U, R) launch_delta(k_anneal<C, U, R>, L, n_chains, C ? 4 : 8, nullptr, grid, block, lds)
    if (L.unroll == 2) {
        if (compact) return rot ? BM_ANNEAL_GRID(true, 2, true) : BM_ANNEAL_GRID(true, 2, false);
        return rot ? BM_ANNEAL_GRID(false, 2, true) : BM_ANNEAL_GRID(false, 2, false);
    }
    if (compact) return rot ? BM_ANNEAL_GRID(true, 4, true) : BM_ANNEAL_GRID(true, 4, false);
    return rot ? BM_ANNEAL_GRID(false, 4, true) : BM_ANNEAL_GRID(false, 4, false);
#undef BM_ANNEAL_GRID
}

hipError_t bm_launch_anneal(const BmDevGraph* insts, int compact, BmChain* chains, int64_t n_chains,
                            int32_t* pool, int64_t steps, const BmAnnealSched& sc, int64_t* traj,
                            const BmProposalPool* proposals, const BmLaunch& L, dim3 grid, dim3 block, size_t lds,
                            hipStream_t s) {
    if (n_chains <= 0 || steps <= 0) return hipSuccess;
    const uint32_t cap = (uint32_t)L.lds_slots;
#define BM_ANNEAL_LAUNCH(C, U, R)                                                                                    \
    hipLaunchKernelGGL((k_anneal<C, U, R>), grid, block, lds, s, insts, chains, n_chains, pool, steps, sc, traj, cap, \
                       proposals)
    const bool rot = sc.p_rot3 != 0;
    if (L.unroll == 2) {
        if (compact) {
            if (rot) BM_ANNEAL_LAUNCH(true, 2, true); else BM_ANNEAL_LAUNCH(true, 2, false);
        } else {
            if (rot) BM_ANNEAL_LAUNCH(false, 2, true); else BM_ANNEAL_LAUNCH(false, 2, false);
        }
    } else {
        if (compact) {
            if (rot) BM_ANNEAL_LAUNCH(true, 4, true); else BM_ANNEAL_LAUNCH(true, 4, false);
        } else {
            if (rot) BM_ANNEAL_LAUNCH(false, 4, true); else BM_ANNEAL_LAUNCH(false, 4, false);
        }
    }
#undef BM_ANNEAL_LAUNCH
    return hipGetLastError();
}

hipError_t bm_launch_score(const BmDevGraph& g, const BmLaunch& L, int64_t* acc, hipStream_t s) {
    dim3 grid, block;
    size_t lds;
    hipError_t e;
    if (g.compact) {
        e = launch_delta(k_score<true>, L, g.n, 4, s, &grid, &block, &lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_score<true>, grid, block, lds, s, g, acc, (uint32_t)L.lds_slots);
    } else {
        e = launch_delta(k_score<false>, L, g.n, 8, s, &grid, &block, &lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_score<false>, grid, block, lds, s, g, acc, (uint32_t)L.lds_slots);
    }
    return hipGetLastError();
}
