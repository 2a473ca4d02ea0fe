/*
 * CPU oracle for the swap-delta path.  TEST INFRASTRUCTURE ONLY: it is loaded
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * never by the library or the tools.  Plain C, deliberately simple, sharing no
 * code with brain-match_amd/ (its graph store, lookups and enumerations are
 * written independently) so that agreement means something.
 *
 * Contents and what each follows / how it is pinned:
 *
 *  or_build / or_score          score-matching's semantics: duplicate (src,dst)
 *                               keeps the last weight once (src/score-matching.c:339-359),
 *                               contribution min(w_M, w_F) when the female edge
 *                               exists else 0, unmatched = 0 (:558-597).
 *                               Pinned against the compiled score-matching
 *                               binary (oracle/_ref) on fixtures and by the
 *                               closed form S(identity; M, M) = sum of weights.
 *  or_swap_delta / or_rot3      exact S(pi') - S(pi) over the edges incident to
 *                               the moved vertices.  Pinned by brute force
 *                               (or_score of the permuted matching) on small
 *                               graphs including self-loops and u<->v edges.
 *  or_quick_enum                the changed-first pair order of
 *                               calculateSwapDeltasKernelQuick, restated from
 *                               its index loops (src/graph-re-alignment-gpu-swap.cu:163-195).
 *  or_three_enum                the 3-section triple order of
 *                               calculateThreeNodeSwapKernel (src/graph-re-alignment-gpu-3-swap.cu:67-150).
 *  or_ref_dense_*               the reference kernels' arithmetic on dense
 *                               int16 matrices, quirks included (weights
 *                               stored as short, entries <= 0 ignored,
 *                               self-loops skipped): :198-246 of the swap tool
 *                               and :176-298 of the 3-swap tool.  Pinned:
 *                               equal to the exact delta on graphs without
 *                               self-loops / duplicates / weights above 32767.
 *  or_ref_lists_delta           calculate_swap_delta (src/graph-alignment-optimized.c:321-408)
 *                               on the reference's edge lists + dense int16
 *                               female matrix, self-loop quirk included.
 *                               Pinned against the compiled reference function
 *                               (oracle/_ref) on fixtures.
 *  or_accept_disjoint           sequential greedy disjoint acceptance
 *                               (delta desc, index asc; closed 1-hop
 *                               neighbourhood lock).  Pinned by the additivity
 *                               check: gain == brute-force score difference.
 *  or_best_swap / _climb        the exact maximum over all pairs u < v of
 *                               or_swap_delta (ties: smaller u, then smaller
 *                               v) and the steepest-ascent loop of the
 *                               score-lookup driver (src/best-swap/score-lookup.c:519-716)
 *                               with that maximum in place of its heap.
 *                               Pinned: the maximum equals the argmax of
 *                               brute-force recounts over every pair of small
 *                               graphs; the climb's score trace equals
 *                               or_score after every swap and ends with no
 *                               positive pair left.
 *  or_best_partner / or_sweep   the OpenMP tool's sweep (src/re-alignment-parallel.c:556-676):
 *                               best partner v > u of each u in turn (ties to
 *                               the smaller v), swapped when its delta is > 0;
 *                               a restart from the best matching perturbed by
 *                               k disjoint swaps after a pass without a swap
 *                               from the third pass on; the perturbation's
 *                               vertex stream as bm_sweep documents it.
 *                               Pinned: best partners equal the argmax of
 *                               brute-force recounts; the pass trace equals
 *                               or_score of the best matching; a pass that
 *                               applies no swap leaves no positive pair; a
 *                               perturbation moves exactly 2k distinct
 *                               vertices and is a permutation of the images.
 *  or_uniform_pair              the candidate stream definition; pinned by
 *                               comparing with pairs downloaded from the device.
 *  or_cpu_baseline_*            the reference CPU arithmetic (a4) driven by an
 *                               OpenMP loop with schedule(dynamic) like
 *                               src/re-alignment-parallel.c:601, used as the
 *                               timed CPU baseline ("port").
 *  or_gain_row                  a placement-gain table row from its
 *                               definition.  Pinned by full recounts: moving
 *                               u alone onto f changes the score by
 *                               G(u, f) - G(u, pi(u)) plus its self-loop
 *                               term, for every u and f of the quirky graph;
 *                               the diagonal sums to twice the score without
 *                               self-loops; every swap delta equals four
 *                               entries plus the corrections (tests).
 *  or_det_exp                   deterministic exp (Cody & Waite reduction,
 *                               Taylor polynomial).  Pinned against libm exp
 *                               to 2 ulp over [-700, 700].
 *  or_anneal_chain              replay of an annealing chain: the Metropolis
 *                               rule and geometric schedule of
 *                               simulated_annealing (src/best-swap/score-lookup.c:345-400)
 *                               over exact deltas, with moves optionally built
 *                               on pairs of a proposal pool.  Pinned by the score
 *                               invariant (running score == or_score of the
 *                               final matching), a brute-force step check, and
 *                               for pools: a one-pair pool moves only that
 *                               pair, p_pool = 0 gives the uniform chain.
 *  or_third_vertex              third vertex of a move on a pool pair;
 *                               pinned by distinctness and uniformity over
 *                               the other n - 2 vertices.
 *  or_twohop_set / _pairs       collect_neighborhood's 2-hop set
 *                               (src/refine-twohop-random.c:582-625) and the
 *                               K-smallest-key sample of it.  Pinned by set
 *                               equality with a Python BFS over the raw edge
 *                               list and by the sample's uniformity.
 *  or_umajor_pairs              the u-major uniform pair stream.  Pinned by
 *                               its law (first vertices = the uniform pair
 *                               stream's multiset, partners uniform) and by
 *                               equality with pairs downloaded from the device.
 *  or_amajor_triples            the a-major uniform triple stream.  Pinned by
 *                               its law (first vertices = the u-major first
 *                               vertices of the derived seed, the ordered
 *                               partner pairs uniform over the (n-1)(n-2)
 *                               choices, checked exhaustively at small n) and
 *                               by equality with triples downloaded from the
 *                               device.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Graph store: deduplicated edge arrays + sorted key index             */
/* ------------------------------------------------------------------ */

typedef struct {
    int64_t key;   /* (src << 32) | dst */
    int64_t idx;   /* input line index */
    int32_t w;
} rec_t;

typedef struct {
    int32_t n;
    int64_t m;
    int64_t* key;      /* sorted unique keys */
    int32_t* w;
    /* adjacency by source and by destination (edge ids into key/w) */
    int64_t* out_off;  /* n+2 */
    int64_t* out_e;
    int64_t* in_off;
    int64_t* in_e;
} og_t;

typedef struct or_graph {
    int32_t n;
    og_t M, F;
} or_graph;

static int cmp_rec(const void* a, const void* b) {
    const rec_t* x = (const rec_t*)a;
    const rec_t* y = (const rec_t*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

static int og_build(og_t* g, int32_t n, int64_t m, const int32_t* s, const int32_t* d, const int32_t* w) {
    memset(g, 0, sizeof *g);
    g->n = n;
    rec_t* r = (rec_t*)malloc(sizeof(rec_t) * (size_t)(m ? m : 1));
    if (!r) return -1;
    for (int64_t i = 0; i < m; i++) {
        if (s[i] < 1 || s[i] > n || d[i] < 1 || d[i] > n) { free(r); return -2; }
        r[i].key = ((int64_t)s[i] << 32) | (uint32_t)d[i];
        r[i].idx = i;
        r[i].w = w[i];
    }
    qsort(r, (size_t)m, sizeof(rec_t), cmp_rec);
    int64_t u = 0;
    for (int64_t i = 0; i < m; i++) {
        if (i + 1 < m && r[i + 1].key == r[i].key) continue; /* a later line wins */
        r[u++] = r[i];
    }
    g->m = u;
    g->key = (int64_t*)malloc(sizeof(int64_t) * (size_t)(u ? u : 1));
    g->w = (int32_t*)malloc(sizeof(int32_t) * (size_t)(u ? u : 1));
    g->out_off = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
    g->in_off = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
    g->out_e = (int64_t*)malloc(sizeof(int64_t) * (size_t)(u ? u : 1));
    g->in_e = (int64_t*)malloc(sizeof(int64_t) * (size_t)(u ? u : 1));
    for (int64_t i = 0; i < u; i++) {
        g->key[i] = r[i].key;
        g->w[i] = r[i].w;
        g->out_off[(r[i].key >> 32) + 1]++;
        g->in_off[(int32_t)(r[i].key & 0xffffffff) + 1]++;
    }
    free(r);
    for (int32_t v = 0; v <= n; v++) {
        g->out_off[v + 1] += g->out_off[v];
        g->in_off[v + 1] += g->in_off[v];
    }
    int64_t* po = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 2));
    int64_t* pi = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 2));
    memcpy(po, g->out_off, sizeof(int64_t) * ((size_t)n + 2));
    memcpy(pi, g->in_off, sizeof(int64_t) * ((size_t)n + 2));
    for (int64_t i = 0; i < u; i++) {
        g->out_e[po[g->key[i] >> 32]++] = i;
        g->in_e[pi[(int32_t)(g->key[i] & 0xffffffff)]++] = i;
    }
    free(po);
    free(pi);
    return 0;
}

static void og_free(og_t* g) {
    free(g->key);
    free(g->w);
    free(g->out_off);
    free(g->out_e);
    free(g->in_off);
    free(g->in_e);
}

static inline int32_t og_src(const og_t* g, int64_t e) { return (int32_t)(g->key[e] >> 32); }
static inline int32_t og_dst(const og_t* g, int64_t e) { return (int32_t)(g->key[e] & 0xffffffff); }

/* Female edge lookup by binary search; returns 1 and the weight if present. */
static inline int og_find(const og_t* g, int32_t a, int32_t b, int32_t* w) {
    if (a < 1 || b < 1) return 0;
    const int64_t k = ((int64_t)a << 32) | (uint32_t)b;
    int64_t lo = 0, hi = g->m;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (g->key[mid] < k) lo = mid + 1; else hi = mid;
    }
    if (lo < g->m && g->key[lo] == k) { *w = g->w[lo]; return 1; }
    return 0;
}

static inline int64_t contrib(const og_t* F, int32_t w, int32_t a, int32_t b) {
    int32_t fw;
    if (!og_find(F, a, b, &fw)) return 0;
    return w < fw ? w : fw;
}

or_graph* or_build(int32_t n, int64_t mm, const int32_t* ms, const int32_t* md, const int32_t* mw, int64_t mf,
                   const int32_t* fs, const int32_t* fd, const int32_t* fw) {
    or_graph* g = (or_graph*)calloc(1, sizeof(or_graph));
    if (!g) return NULL;
    g->n = n;
    if (og_build(&g->M, n, mm, ms, md, mw) != 0 || og_build(&g->F, n, mf, fs, fd, fw) != 0) {
        og_free(&g->M);
        og_free(&g->F);
        free(g);
        return NULL;
    }
    return g;
}

void or_free(or_graph* g) {
    if (!g) return;
    og_free(&g->M);
    og_free(&g->F);
    free(g);
}

int64_t or_num_edges(const or_graph* g, int female) { return female ? g->F.m : g->M.m; }

int64_t or_score(const or_graph* g, const int32_t* pi) {
    int64_t s = 0;
    for (int64_t e = 0; e < g->M.m; e++)
        s += contrib(&g->F, g->M.w[e], pi[og_src(&g->M, e)], pi[og_dst(&g->M, e)]);
    return s;
}

/* Sum over male edges incident to any vertex of set[0..k) of
 * contrib under pi_new minus contrib under pi_old.  Each edge once. */
static int64_t local_delta(const or_graph* g, const int32_t* pi_old, const int32_t* pi_new, const int32_t* set,
                           int k) {
    int64_t d = 0;
    for (int i = 0; i < k; i++) {
        const int32_t x = set[i];
        for (int64_t p = g->M.out_off[x]; p < g->M.out_off[x + 1]; p++) {
            const int64_t e = g->M.out_e[p];
            const int32_t y = og_dst(&g->M, e), w = g->M.w[e];
            d += contrib(&g->F, w, pi_new[x], pi_new[y]) - contrib(&g->F, w, pi_old[x], pi_old[y]);
        }
        for (int64_t p = g->M.in_off[x]; p < g->M.in_off[x + 1]; p++) {
            const int64_t e = g->M.in_e[p];
            const int32_t y = og_src(&g->M, e), w = g->M.w[e];
            int in_set = 0;
            for (int j = 0; j < k; j++) in_set |= (set[j] == y);
            if (in_set) continue; /* counted as an out-edge of y */
            d += contrib(&g->F, w, pi_new[y], pi_new[x]) - contrib(&g->F, w, pi_old[y], pi_old[x]);
        }
    }
    return d;
}

int64_t or_swap_delta(const or_graph* g, const int32_t* pi, int32_t u, int32_t v) {
    if (u == v) return 0;
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * ((size_t)g->n + 1));
    memcpy(q, pi, sizeof(int32_t) * ((size_t)g->n + 1));
    q[u] = pi[v];
    q[v] = pi[u];
    const int32_t set[2] = {u, v};
    int64_t d = local_delta(g, pi, q, set, 2);
    free(q);
    return d;
}

/* Same, without the per-call matching copy (for the timed loop). */
static int64_t swap_delta_fast(const or_graph* g, const int32_t* pi, int32_t u, int32_t v) {
    if (u == v) return 0;
    const int32_t f1 = pi[u], f2 = pi[v];
#define NEWPI(x) ((x) == u ? f2 : ((x) == v ? f1 : pi[(x)]))
    int64_t d = 0;
    const int32_t set[2] = {u, v};
    for (int i = 0; i < 2; i++) {
        const int32_t x = set[i];
        for (int64_t p = g->M.out_off[x]; p < g->M.out_off[x + 1]; p++) {
            const int64_t e = g->M.out_e[p];
            const int32_t y = og_dst(&g->M, e), w = g->M.w[e];
            d += contrib(&g->F, w, NEWPI(x), NEWPI(y)) - contrib(&g->F, w, pi[x], pi[y]);
        }
        for (int64_t p = g->M.in_off[x]; p < g->M.in_off[x + 1]; p++) {
            const int64_t e = g->M.in_e[p];
            const int32_t y = og_src(&g->M, e), w = g->M.w[e];
            if (y == u || y == v) continue;
            d += contrib(&g->F, w, pi[y], NEWPI(x)) - contrib(&g->F, w, pi[y], pi[x]);
        }
    }
#undef NEWPI
    return d;
}

void or_swap_deltas(const or_graph* g, const int32_t* pi, const int32_t* uv, int64_t k, int64_t* out) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < k; i++) out[i] = swap_delta_fast(g, pi, uv[2 * i], uv[2 * i + 1]);
}

int64_t or_swap_delta_brute(const or_graph* g, const int32_t* pi, int32_t u, int32_t v) {
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * ((size_t)g->n + 1));
    memcpy(q, pi, sizeof(int32_t) * ((size_t)g->n + 1));
    const int32_t t = q[u];
    q[u] = q[v];
    q[v] = t;
    int64_t d = or_score(g, q) - or_score(g, pi);
    free(q);
    return d;
}

/* Permutation table of calculateThreeNodeSwapKernel
 * (src/graph-re-alignment-gpu-3-swap.cu:167-174): new female of position i. */
static const int PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

/* All five deltas; brute=1 uses full recounts. */
void or_rot3_all(const or_graph* g, const int32_t* pi, int32_t a, int32_t b, int32_t c, int brute, int64_t* d5) {
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * ((size_t)g->n + 1));
    memcpy(q, pi, sizeof(int32_t) * ((size_t)g->n + 1));
    const int32_t v[3] = {a, b, c};
    const int32_t f[3] = {pi[a], pi[b], pi[c]};
    const int64_t base = brute ? or_score(g, pi) : 0;
    for (int p = 1; p <= 5; p++) {
        for (int i = 0; i < 3; i++) q[v[i]] = f[PERM[p][i]];
        d5[p - 1] = brute ? or_score(g, q) - base : local_delta(g, pi, q, v, 3);
        for (int i = 0; i < 3; i++) q[v[i]] = f[i];
    }
    free(q);
}

/* best = max(0, max_p delta_p), first strictly greater p, p = 0 if none. */
void or_rot3(const or_graph* g, const int32_t* pi, int32_t a, int32_t b, int32_t c, int64_t* best, int32_t* perm) {
    int64_t d5[5];
    or_rot3_all(g, pi, a, b, c, 0, d5);
    int64_t bd = 0;
    int32_t bp = 0;
    for (int p = 1; p <= 5; p++)
        if (d5[p - 1] > bd) { bd = d5[p - 1]; bp = p; }
    *best = bd;
    *perm = bp;
}

void or_rot3_many(const or_graph* g, const int32_t* pi, const int32_t* abc, int64_t k, int64_t* best, int32_t* perm) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < k; i++) or_rot3(g, pi, abc[3 * i], abc[3 * i + 1], abc[3 * i + 2], &best[i], &perm[i]);
}

/* ------------------------------------------------------------------ */
/* Reference enumerations                                              */
/* ------------------------------------------------------------------ */

/* reorder: changed ids ascending, then the remaining ids ascending (the
 * reference's sortArray + fillArray). */
static void make_reorder(int32_t n, const int32_t* changed, int32_t c, int32_t* re) {
    char* present = (char*)calloc((size_t)n + 1, 1);
    for (int32_t i = 0; i < c; i++) { re[i] = changed[i]; present[changed[i]] = 1; }
    /* bubble sort as sortArray */
    for (int32_t i = 0; i < c - 1; i++)
        for (int32_t j = 0; j < c - i - 1; j++)
            if (re[j] > re[j + 1]) { int32_t t = re[j]; re[j] = re[j + 1]; re[j + 1] = t; }
    int32_t p = c;
    for (int32_t x = 1; x <= n; x++)
        if (!present[x]) re[p++] = x;
    free(present);
}

/* Returns P; uv gets 2P ids.  Index loops as in the reference kernel. */
int64_t or_quick_enum(int32_t n, const int32_t* changed, int32_t c, int32_t* uv) {
    int32_t* re = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    make_reorder(n, changed, c, re);
    const int64_t nu = (int64_t)n - c;
    const int64_t cc = (int64_t)c * (c - 1) / 2;
    const int64_t P = cc + (int64_t)c * nu;
    for (int64_t idx0 = 0; idx0 < P; idx0++) {
        int64_t idx = idx0;
        int32_t a, b;
        if (idx < cc) {
            int64_t i = 0, temp = c - 1;
            while (idx >= temp) { idx -= temp; i++; temp--; }
            a = re[i];
            b = re[i + idx + 1];
        } else {
            idx -= cc;
            int64_t i = idx / nu, j = idx - i * nu;
            a = re[i];
            b = re[c + j];
        }
        uv[2 * idx0] = a;
        uv[2 * idx0 + 1] = b;
    }
    free(re);
    return P;
}

int64_t or_three_enum(int32_t n, const int32_t* changed, int32_t c, int32_t* abc) {
    int32_t* re = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    make_reorder(n, changed, c, re);
    const int64_t s1 = (int64_t)c * (c - 1) * (c - 2) / 6;
    const int64_t pc = (int64_t)c * (c - 1) / 2;
    const int64_t s2 = s1 + pc * ((int64_t)n - c);
    const int64_t rc = (int64_t)n - c - 1;
    const int64_t T = s2 + (rc > 0 ? (int64_t)c * rc : 0);
    for (int64_t idx0 = 0; idx0 < T; idx0++) {
        int64_t idx = idx0;
        int64_t i1 = 0, i2 = 0, i3 = 0;
        int32_t a, b, d;
        if (idx < s1) {
            int64_t nn = c, temp1 = (nn - 1) * (nn - 2) / 2;
            while (idx >= temp1) { idx -= temp1; i1++; nn--; temp1 = (nn - 1) * (nn - 2) / 2; }
            i2 = i1 + 1;
            temp1 = nn - 2;
            while (idx >= temp1) { idx -= temp1; i2++; temp1--; }
            i3 = i2 + idx + 1;
            a = re[i1]; b = re[i2]; d = re[i3];
        } else if (idx < s2) {
            idx -= s1;
            i3 = idx / pc;
            int64_t rem = idx % pc, temp = c - 1;
            while (rem >= temp) { rem -= temp; i1++; temp--; }
            i2 = i1 + rem + 1;
            a = re[i1]; b = re[i2]; d = re[c + i3];
        } else {
            idx -= s2;
            i1 = idx / rc;
            i2 = idx % rc;
            a = re[i1]; b = re[c + i2]; d = re[c + i2 + 1];
        }
        abc[3 * idx0] = a;
        abc[3 * idx0 + 1] = b;
        abc[3 * idx0 + 2] = d;
    }
    free(re);
    return T;
}

/* ------------------------------------------------------------------ */
/* The reference kernels' dense-matrix arithmetic (quirks included)    */
/* ------------------------------------------------------------------ */

typedef struct or_dense {
    int32_t n;
    int16_t* M;   /* (n+1)^2, row-major, M[from*(n+1)+to] */
    int16_t* F;
} or_dense;

/* adj_matrix[from][to] = weight for every line, last line wins, stored as short. */
or_dense* or_dense_build(int32_t n, int64_t mm, const int32_t* ms, const int32_t* md, const int32_t* mw, int64_t mf,
                         const int32_t* fs, const int32_t* fd, const int32_t* fw) {
    or_dense* D = (or_dense*)calloc(1, sizeof(or_dense));
    const size_t dim = (size_t)n + 1;
    D->n = n;
    D->M = (int16_t*)calloc(dim * dim, sizeof(int16_t));
    D->F = (int16_t*)calloc(dim * dim, sizeof(int16_t));
    if (!D->M || !D->F) { free(D->M); free(D->F); free(D); return NULL; }
    for (int64_t i = 0; i < mm; i++) D->M[(size_t)ms[i] * dim + (size_t)md[i]] = (int16_t)mw[i];
    for (int64_t i = 0; i < mf; i++) D->F[(size_t)fs[i] * dim + (size_t)fd[i]] = (int16_t)fw[i];
    return D;
}

void or_dense_free(or_dense* D) {
    if (!D) return;
    free(D->M);
    free(D->F);
    free(D);
}

static inline int imin(int a, int b) { return a < b ? a : b; }

/* Delta of one pair as calculateSwapDeltasKernelQuick computes it
 * (src/graph-re-alignment-gpu-swap.cu:198-246). */
int32_t or_ref_dense_swap(const or_dense* D, const int32_t* map, int32_t m1, int32_t m2) {
    const size_t dim = (size_t)D->n + 1;
    const int16_t* M = D->M;
    const int16_t* F = D->F;
    const int32_t f1 = map[m1], f2 = map[m2];
    int delta = 0;
    for (int32_t i = 1; i <= D->n; i++) {
        if (i == m1 || i == m2) continue;
        if (M[m1 * dim + i] > 0) {
            int w = M[m1 * dim + i], d = map[i];
            delta += imin(w, F[f2 * dim + d]) - imin(w, F[f1 * dim + d]);
        }
        if (M[i * dim + m1] > 0) {
            int w = M[i * dim + m1], s = map[i];
            delta += imin(w, F[s * dim + f2]) - imin(w, F[s * dim + f1]);
        }
        if (M[m2 * dim + i] > 0) {
            int w = M[m2 * dim + i], d = map[i];
            delta += imin(w, F[f1 * dim + d]) - imin(w, F[f2 * dim + d]);
        }
        if (M[i * dim + m2] > 0) {
            int w = M[i * dim + m2], s = map[i];
            delta += imin(w, F[s * dim + f1]) - imin(w, F[s * dim + f2]);
        }
    }
    int w12 = M[m1 * dim + m2];
    if (w12 > 0) delta += imin(w12, F[f2 * dim + f1]) - imin(w12, F[f1 * dim + f2]);
    int w21 = M[m2 * dim + m1];
    if (w21 > 0) delta += imin(w21, F[f1 * dim + f2]) - imin(w21, F[f2 * dim + f1]);
    return delta;
}

/* The whole Quick launch: deltas[P] and pairs[2P] in the kernel's order. */
int64_t or_ref_dense_quick(const or_dense* D, const int32_t* map, const int32_t* changed, int32_t c, int32_t* deltas,
                           int32_t* pairs) {
    const int64_t P = or_quick_enum(D->n, changed, c, pairs);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < P; i++) deltas[i] = or_ref_dense_swap(D, map, pairs[2 * i], pairs[2 * i + 1]);
    return P;
}

/* One triple as calculateThreeNodeSwapKernel computes it
 * (src/graph-re-alignment-gpu-3-swap.cu:152-298). */
void or_ref_dense_three(const or_dense* D, const int32_t* map, int32_t m1, int32_t m2, int32_t m3, int32_t* best,
                        int32_t* bestp) {
    const size_t dim = (size_t)D->n + 1;
    const int16_t* M = D->M;
    const int16_t* F = D->F;
    const int32_t f[3] = {map[m1], map[m2], map[m3]};
    const int32_t m[3] = {m1, m2, m3};
    int bd = 0, bp = 0;
    for (int p = 1; p < 6; p++) {
        int delta = 0;
        int nf[3];
        for (int q = 0; q < 3; q++) nf[q] = f[PERM[p][q]];
        for (int q = 0; q < 3; q++) {
            for (int32_t i = 1; i <= D->n; i++) {
                if (i == m1 || i == m2 || i == m3) continue;
                if (M[m[q] * dim + i] > 0) {
                    int w = M[m[q] * dim + i], d = map[i];
                    delta += imin(w, F[nf[q] * dim + d]) - imin(w, F[f[q] * dim + d]);
                }
                if (M[i * dim + m[q]] > 0) {
                    int w = M[i * dim + m[q]], s = map[i];
                    delta += imin(w, F[s * dim + nf[q]]) - imin(w, F[s * dim + f[q]]);
                }
            }
        }
        /* the six ordered pairs in the kernel's order: 12, 21, 23, 32, 31, 13 */
        static const int PR[6][2] = {{0, 1}, {1, 0}, {1, 2}, {2, 1}, {2, 0}, {0, 2}};
        for (int r = 0; r < 6; r++) {
            const int x = PR[r][0], y = PR[r][1];
            int w = M[m[x] * dim + m[y]];
            if (w > 0) delta += imin(w, F[nf[x] * dim + nf[y]]) - imin(w, F[f[x] * dim + f[y]]);
        }
        if (delta > bd) { bd = delta; bp = p; }
    }
    *best = bd;
    *bestp = bp;
}

int64_t or_ref_dense_three_launch(const or_dense* D, const int32_t* map, const int32_t* changed, int32_t c,
                                  int32_t* deltas, int32_t* trip4) {
    const int64_t cc = c, rest = (int64_t)D->n - c - 1;
    const int64_t T0 = cc * (cc - 1) * (cc - 2) / 6 + cc * (cc - 1) / 2 * ((int64_t)D->n - cc) + (rest > 0 ? cc * rest : 0);
    int32_t* abc = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(T0 ? T0 : 1));
    const int64_t T = or_three_enum(D->n, changed, c, abc);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < T; i++) {
        int32_t b, p;
        or_ref_dense_three(D, map, abc[3 * i], abc[3 * i + 1], abc[3 * i + 2], &b, &p);
        deltas[i] = b;
        trip4[4 * i] = abc[3 * i];
        trip4[4 * i + 1] = abc[3 * i + 1];
        trip4[4 * i + 2] = abc[3 * i + 2];
        trip4[4 * i + 3] = p;
    }
    free(abc);
    return T;
}

/* ------------------------------------------------------------------ */
/* calculate_swap_delta on edge lists + dense female matrix            */
/* ------------------------------------------------------------------ */

typedef struct or_lists {
    int32_t n;
    /* every input line appended (duplicates kept), as add_edge does */
    int64_t* out_off;
    int32_t* out_to;
    int32_t* out_w;
    int64_t* in_off;
    int32_t* in_from;
    int32_t* in_w;
    int16_t* M;   /* dense male matrix (for the direct-edge terms) */
    int16_t* F;   /* dense female matrix */
} or_lists;

or_lists* or_lists_build(int32_t n, int64_t mm, const int32_t* ms, const int32_t* md, const int32_t* mw, int64_t mf,
                         const int32_t* fs, const int32_t* fd, const int32_t* fw) {
    or_lists* L = (or_lists*)calloc(1, sizeof(or_lists));
    const size_t dim = (size_t)n + 1;
    L->n = n;
    L->out_off = (int64_t*)calloc(dim + 1, sizeof(int64_t));
    L->in_off = (int64_t*)calloc(dim + 1, sizeof(int64_t));
    L->out_to = (int32_t*)malloc(sizeof(int32_t) * (size_t)(mm ? mm : 1));
    L->out_w = (int32_t*)malloc(sizeof(int32_t) * (size_t)(mm ? mm : 1));
    L->in_from = (int32_t*)malloc(sizeof(int32_t) * (size_t)(mm ? mm : 1));
    L->in_w = (int32_t*)malloc(sizeof(int32_t) * (size_t)(mm ? mm : 1));
    L->M = (int16_t*)calloc(dim * dim, sizeof(int16_t));
    L->F = (int16_t*)calloc(dim * dim, sizeof(int16_t));
    if (!L->out_off || !L->in_off || !L->out_to || !L->out_w || !L->in_from || !L->in_w || !L->M || !L->F) return NULL;
    for (int64_t i = 0; i < mm; i++) { L->out_off[ms[i] + 1]++; L->in_off[md[i] + 1]++; }
    for (size_t v = 0; v < dim; v++) { L->out_off[v + 1] += L->out_off[v]; L->in_off[v + 1] += L->in_off[v]; }
    int64_t* po = (int64_t*)malloc(sizeof(int64_t) * (dim + 1));
    int64_t* pi = (int64_t*)malloc(sizeof(int64_t) * (dim + 1));
    memcpy(po, L->out_off, sizeof(int64_t) * (dim + 1));
    memcpy(pi, L->in_off, sizeof(int64_t) * (dim + 1));
    for (int64_t i = 0; i < mm; i++) {
        int64_t a = po[ms[i]]++, b = pi[md[i]]++;
        L->out_to[a] = md[i];
        L->out_w[a] = mw[i];
        L->in_from[b] = ms[i];
        L->in_w[b] = mw[i];
        L->M[(size_t)ms[i] * dim + (size_t)md[i]] = (int16_t)mw[i];
    }
    free(po);
    free(pi);
    for (int64_t i = 0; i < mf; i++) L->F[(size_t)fs[i] * dim + (size_t)fd[i]] = (int16_t)fw[i];
    return L;
}

void or_lists_free(or_lists* L) {
    if (!L) return;
    free(L->out_off); free(L->in_off); free(L->out_to); free(L->out_w);
    free(L->in_from); free(L->in_w); free(L->M); free(L->F);
    free(L);
}

/* calculate_swap_delta (src/graph-alignment-optimized.c:321-408). */
int32_t or_ref_lists_delta(const or_lists* L, const int32_t* map, int32_t m1, int32_t m2) {
    const size_t dim = (size_t)L->n + 1;
    const int16_t* F = L->F;
    const int32_t f1 = map[m1], f2 = map[m2];
    int delta = 0;
    for (int64_t p = L->out_off[m1]; p < L->out_off[m1 + 1]; p++) {
        const int32_t d = L->out_to[p];
        if (d == m2) continue;
        const int w = L->out_w[p], df = map[d];
        delta += imin(w, F[f2 * dim + df]) - imin(w, F[f1 * dim + df]);
    }
    for (int64_t p = L->in_off[m1]; p < L->in_off[m1 + 1]; p++) {
        const int32_t s = L->in_from[p];
        if (s == m2) continue;
        const int w = L->in_w[p], sf = map[s];
        delta += imin(w, F[sf * dim + f2]) - imin(w, F[sf * dim + f1]);
    }
    for (int64_t p = L->out_off[m2]; p < L->out_off[m2 + 1]; p++) {
        const int32_t d = L->out_to[p];
        if (d == m1) continue;
        const int w = L->out_w[p], df = map[d];
        delta += imin(w, F[f1 * dim + df]) - imin(w, F[f2 * dim + df]);
    }
    for (int64_t p = L->in_off[m2]; p < L->in_off[m2 + 1]; p++) {
        const int32_t s = L->in_from[p];
        if (s == m1) continue;
        const int w = L->in_w[p], sf = map[s];
        delta += imin(w, F[sf * dim + f1]) - imin(w, F[sf * dim + f2]);
    }
    const int w12 = L->M[m1 * dim + m2];
    if (w12 > 0) delta += imin(w12, F[f2 * dim + f1]) - imin(w12, F[f1 * dim + f2]);
    const int w21 = L->M[m2 * dim + m1];
    if (w21 > 0) delta += imin(w21, F[f1 * dim + f2]) - imin(w21, F[f2 * dim + f1]);
    return delta;
}

/* The timed CPU baseline: calculate_swap_delta over a candidate list with
 * an OpenMP dynamic schedule (the loop shape of src/re-alignment-parallel.c:601-622
 * without its per-delta critical section).  threads <= 0: OpenMP default. */
void or_ref_lists_deltas(const or_lists* L, const int32_t* map, const int32_t* uv, int64_t k, int32_t* out,
                         int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < k; i++) out[i] = or_ref_lists_delta(L, map, uv[2 * i], uv[2 * i + 1]);
}

int or_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* Candidate stream and disjoint acceptance                            */
/* ------------------------------------------------------------------ */

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Pair c of stream seed: u uniform in 1..n, v uniform in 1..n minus u. */
void or_uniform_pair(uint64_t seed, uint64_t c, int32_t n, int32_t* u, int32_t* v) {
    const uint64_t r = mix64(seed * 0x9E3779B97F4A7C15ULL + (c + 1) * 0xD1B54A32D192ED03ULL);
    const uint32_t a = (uint32_t)(((uint64_t)(uint32_t)r * (uint64_t)n) >> 32);
    uint32_t b = (uint32_t)(((uint64_t)(uint32_t)(r >> 32) * (uint64_t)(n - 1)) >> 32);
    if (b >= a) b++;
    *u = (int32_t)a + 1;
    *v = (int32_t)b + 1;
}

void or_uniform_pairs(uint64_t seed, int64_t k, int32_t n, int32_t* uv) {
    for (int64_t c = 0; c < k; c++) or_uniform_pair(seed, (uint64_t)c, n, &uv[2 * c], &uv[2 * c + 1]);
}

typedef struct { int64_t d; int64_t i; } pri_t;
static int cmp_pri(const void* a, const void* b) {
    const pri_t* x = (const pri_t*)a;
    const pri_t* y = (const pri_t*)b;
    if (x->d != y->d) return x->d > y->d ? -1 : 1;
    return x->i < y->i ? -1 : (x->i > y->i);
}

/* Candidates with delta > 0 in (delta desc, index asc) order; accept one
 * iff neither endpoint is locked; lock the closed neighbourhoods (in + out)
 * of both endpoints.  Applies the swaps to pi.  Returns the count; gain and
 * the accepted indices (capacity k) are written. */
int64_t or_accept_disjoint(const or_graph* g, int32_t* pi, const int32_t* uv, const int64_t* d, int64_t k,
                           int64_t max_accept, int64_t* accepted, int64_t* gain) {
    pri_t* q = (pri_t*)malloc(sizeof(pri_t) * (size_t)(k ? k : 1));
    int64_t np = 0;
    for (int64_t i = 0; i < k; i++)
        if (d[i] > 0) { q[np].d = d[i]; q[np].i = i; np++; }
    qsort(q, (size_t)np, sizeof(pri_t), cmp_pri);
    char* lock = (char*)calloc((size_t)g->n + 1, 1);
    int64_t na = 0, gsum = 0;
    if (max_accept <= 0) max_accept = INT64_MAX;
    for (int64_t t = 0; t < np && na < max_accept; t++) {
        const int32_t u = uv[2 * q[t].i], v = uv[2 * q[t].i + 1];
        if (u == v || lock[u] || lock[v]) continue;
        accepted[na++] = q[t].i;
        gsum += q[t].d;
        const int32_t ends[2] = {u, v};
        for (int s = 0; s < 2; s++) {
            const int32_t x = ends[s];
            lock[x] = 1;
            for (int64_t p = g->M.out_off[x]; p < g->M.out_off[x + 1]; p++) lock[og_dst(&g->M, g->M.out_e[p])] = 1;
            for (int64_t p = g->M.in_off[x]; p < g->M.in_off[x + 1]; p++) lock[og_src(&g->M, g->M.in_e[p])] = 1;
        }
    }
    for (int64_t a = 0; a < na; a++) {
        const int32_t u = uv[2 * accepted[a]], v = uv[2 * accepted[a] + 1];
        const int32_t t = pi[u];
        pi[u] = pi[v];
        pi[v] = t;
    }
    free(q);
    free(lock);
    *gain = gsum;
    return na;
}

/* CPU replay of the disjoint hill climb: per round, batch pairs of stream
 * seed + r, exact deltas, acceptance.  trace gets rounds+1 scores. */
void or_hillclimb(const or_graph* g, int32_t* pi, uint64_t seed, int64_t rounds, int64_t batch, int64_t* trace) {
    int32_t* uv = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)batch);
    int64_t* d = (int64_t*)malloc(sizeof(int64_t) * (size_t)batch);
    int64_t* acc = (int64_t*)malloc(sizeof(int64_t) * (size_t)batch);
    int64_t s = or_score(g, pi);
    trace[0] = s;
    for (int64_t r = 0; r < rounds; r++) {
        or_uniform_pairs(seed + (uint64_t)r, batch, g->n, uv);
        or_swap_deltas(g, pi, uv, batch, d);
        int64_t gain = 0;
        or_accept_disjoint(g, pi, uv, d, batch, 0, acc, &gain);
        s += gain;
        trace[r + 1] = s;
    }
    free(uv);
    free(d);
    free(acc);
}

/* Row u of the placement-gain table, straight from its definition:
 * row[f] = sum over male edges u->x (x != u) of c(w, f, pi(x))
 *        + sum over male edges x->u (x != u) of c(w, pi(x), f),   f = 0..n.
 * O(n * deg(u)) female lookups; for checking small graphs only. */
void or_gain_row(const or_graph* g, const int32_t* pi, int32_t u, int32_t* row) {
    for (int32_t f = 0; f <= g->n; f++) {
        int64_t s = 0;
        for (int64_t p = g->M.out_off[u]; p < g->M.out_off[u + 1]; p++) {
            const int64_t e = g->M.out_e[p];
            const int32_t x = og_dst(&g->M, e);
            if (x != u) s += contrib(&g->F, g->M.w[e], f, pi[x]);
        }
        for (int64_t p = g->M.in_off[u]; p < g->M.in_off[u + 1]; p++) {
            const int64_t e = g->M.in_e[p];
            const int32_t x = og_src(&g->M, e);
            if (x != u) s += contrib(&g->F, g->M.w[e], pi[x], f);
        }
        row[f] = (int32_t)s;
    }
}

/* ------------------------------------------------------------------ */
/* Annealing chains                                                    */
/* ------------------------------------------------------------------ */

static inline uint64_t rand_at(uint64_t seed, uint64_t c) {
    return mix64(seed * 0x9E3779B97F4A7C15ULL + (c + 1) * 0xD1B54A32D192ED03ULL);
}

static inline uint32_t range32(uint32_t bits, uint32_t range) {
    return (uint32_t)(((uint64_t)bits * (uint64_t)range) >> 32);
}

/* Triple c of stream seed: a uniform, b uniform != a, d uniform outside {a, b}. */
void or_uniform_triple(uint64_t seed, uint64_t c, int32_t n, int32_t* x, int32_t* y, int32_t* z) {
    const uint64_t r0 = rand_at(seed, 2 * c), r1 = rand_at(seed, 2 * c + 1);
    const uint32_t a = range32((uint32_t)r0, (uint32_t)n);
    uint32_t b = range32((uint32_t)(r0 >> 32), (uint32_t)(n - 1));
    uint32_t d = range32((uint32_t)r1, (uint32_t)(n - 2));
    if (b >= a) b++;
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    if (d >= lo) d++;
    if (d >= hi) d++;
    *x = (int32_t)a + 1;
    *y = (int32_t)b + 1;
    *z = (int32_t)d + 1;
}

/* exp(x) for |x| <= 700 from IEEE double +, -, * only (the file is compiled
 * with -ffp-contract=off): x = k ln2 + r with ln2 split in a 32-bit head
 * and a tail (Cody & Waite), exp(r) by the degree-13 Taylor polynomial in
 * Horner form, times 2^k.  The acceptance tests of the annealing chains use
 * this function so that a replay takes the device's decisions bit for bit;
 * it is pinned against libm's exp (<= 2 ulp) by the tests. */
double or_det_exp(double x) {
    if (x != x) return 1.0;
    if (x < -700.0) return 0.0;
    if (x > 700.0) return 1.0e304;
    const double t = x * 1.4426950408889634;
    const int k = (int)(t < 0.0 ? t - 0.5 : t + 0.5);
    const double kd = (double)k;
    const double r = (x - kd * 6.93147180369123816490e-01) - kd * 1.90821492927058770002e-10;
    static const double inv_fact[14] = {1.0,
                                        1.0,
                                        0.5,
                                        1.0 / 6.0,
                                        1.0 / 24.0,
                                        1.0 / 120.0,
                                        1.0 / 720.0,
                                        1.0 / 5040.0,
                                        1.0 / 40320.0,
                                        1.0 / 362880.0,
                                        1.0 / 3628800.0,
                                        1.0 / 39916800.0,
                                        1.0 / 479001600.0,
                                        1.0 / 6227020800.0};
    double p = inv_fact[13];
    for (int i = 12; i >= 0; i--) p = p * r + inv_fact[i];
    if (k < -1022) return 0.0;
    union {
        uint64_t u;
        double d;
    } two_k;
    two_k.u = (uint64_t)(k + 1023) << 52;
    return p * two_k.d;
}

/* Third vertex of a 3-vertex move on the pair (a, b): uniform over the
 * n - 2 others, from draw 2s + 1 of stream seed (the draw the uniform triple
 * takes its third vertex from), skipping a and b in increasing order. */
int32_t or_third_vertex(uint64_t seed, uint64_t s, int32_t n, int32_t a, int32_t b) {
    uint32_t d = range32((uint32_t)rand_at(seed, 2 * s + 1), (uint32_t)(n - 2));
    const uint32_t x = (uint32_t)(a < b ? a : b) - 1, y = (uint32_t)(a < b ? b : a) - 1;
    if (d >= x) d++;
    if (d >= y) d++;
    return (int32_t)d + 1;
}

/* Replays `steps` steps (from step index s0) of one annealing chain with
 * stream seed q: step s draws its move kind from rand(q2, s) (a 3-vertex move
 * when the low 32 bits are below p_rot3, a 2^-32 fraction, and n >= 3); with
 * a proposal pool of pool_k pairs, r = rand(q4, s) selects pool entry
 * (hi32(r) pool_k) >> 32 when lo32(r) < p_pool, and the move is that pair's
 * swap or, for a 3-vertex move, that pair plus third_vertex(q3, s); otherwise
 * the move is uniform pair s of stream q0 or uniform triple s of stream q3.
 * A 3-vertex move takes permutation 1 + range(hi32(rand(q2, s)), 5).  The
 * exact delta is accepted when delta >= 0 or (rand(q1, s) >> 11) * 2^-53 <
 * or_det_exp(delta * beta0 * or_det_exp(s * lam)), q_i = rand(q, i).
 * st = {score, best, accepted, at_best} is updated in place; pi and pib (the
 * latest matching reaching the best score, valid when at_best == 0) too. */
void or_anneal_chain(const or_graph* g, int32_t* pi, int32_t* pib, uint64_t q, int64_t s0, int64_t steps,
                     double beta0, double lam, uint64_t p_rot3, int64_t* st, const int32_t* pool_uv,
                     int64_t pool_k, uint64_t p_pool) {
    const int32_t n = g->n;
    const uint64_t q0 = rand_at(q, 0), q1 = rand_at(q, 1), q2 = rand_at(q, 2), q3 = rand_at(q, 3);
    const uint64_t q4 = rand_at(q, 4);
    int64_t score = st[0], best = st[1], nacc = st[2];
    int at_best = (int)st[3];
    for (int64_t s = s0; s < s0 + steps; s++) {
        const uint64_t rk = rand_at(q2, (uint64_t)s);
        const int rot = n >= 3 && (uint64_t)(uint32_t)rk < p_rot3;
        int32_t v[3];
        int64_t d;
        int p = 0;
        int pooled = 0;
        if (pool_k > 0) {
            const uint64_t r = rand_at(q4, (uint64_t)s);
            if ((uint64_t)(uint32_t)r < p_pool) {
                const int64_t j = (int64_t)(((r >> 32) * (uint64_t)pool_k) >> 32);
                v[0] = pool_uv[2 * j];
                v[1] = pool_uv[2 * j + 1];
                pooled = 1;
            }
        }
        if (!rot) {
            if (!pooled) or_uniform_pair(q0, (uint64_t)s, n, &v[0], &v[1]);
            d = or_swap_delta(g, pi, v[0], v[1]);
        } else {
            if (pooled)
                v[2] = or_third_vertex(q3, (uint64_t)s, n, v[0], v[1]);
            else
                or_uniform_triple(q3, (uint64_t)s, n, &v[0], &v[1], &v[2]);
            int64_t d5[5];
            or_rot3_all(g, pi, v[0], v[1], v[2], 0, d5);
            p = 1 + (int)range32((uint32_t)(rk >> 32), 5u);
            d = d5[p - 1];
        }
        int accept = d >= 0;
        if (!accept) {
            const double beta = beta0 * or_det_exp((double)s * lam);
            const double r = (double)(rand_at(q1, (uint64_t)s) >> 11) * (1.0 / 9007199254740992.0);
            accept = r < or_det_exp((double)d * beta);
        }
        if (!accept) continue;
        if (at_best && d < 0) {
            memcpy(pib, pi, sizeof(int32_t) * ((size_t)n + 1));
            at_best = 0;
        }
        if (!rot) {
            const int32_t t = pi[v[0]];
            pi[v[0]] = pi[v[1]];
            pi[v[1]] = t;
        } else {
            const int32_t f[3] = {pi[v[0]], pi[v[1]], pi[v[2]]};
            for (int i = 0; i < 3; i++) pi[v[i]] = f[PERM[p][i]];
        }
        score += d;
        nacc++;
        if (score >= best) {
            best = score;
            at_best = 1;
        }
    }
    st[0] = score;
    st[1] = best;
    st[2] = nacc;
    st[3] = at_best;
}

/* Stream seed of chain `index` of an annealing run seeded `seed`. */
uint64_t or_chain_seed(uint64_t seed, int64_t index) { return rand_at(seed, (uint64_t)index); }

void or_det_exp_many(const double* x, int64_t k, double* out) {
    for (int64_t i = 0; i < k; i++) out[i] = or_det_exp(x[i]);
}

/* ------------------------------------------------------------------ */
/* 2-hop proposals                                                     */
/* ------------------------------------------------------------------ */

/* collect_neighborhood (src/refine-twohop-random.c:582-625) in the male
 * graph: out-neighbours of v and theirs, in-neighbours of v and theirs, v
 * excluded; ascending ids into out (n entries).  Returns the count. */
int32_t or_twohop_set(const or_graph* g, int32_t v, int32_t* out) {
    const og_t* M = &g->M;
    char* mark = (char*)calloc((size_t)g->n + 1, 1);
    for (int64_t p = M->out_off[v]; p < M->out_off[v + 1]; p++) {
        const int32_t x = og_dst(M, M->out_e[p]);
        mark[x] = 1;
        for (int64_t q = M->out_off[x]; q < M->out_off[x + 1]; q++) mark[og_dst(M, M->out_e[q])] = 1;
    }
    for (int64_t p = M->in_off[v]; p < M->in_off[v + 1]; p++) {
        const int32_t x = og_src(M, M->in_e[p]);
        mark[x] = 1;
        for (int64_t q = M->in_off[x]; q < M->in_off[x + 1]; q++) mark[og_src(M, M->in_e[q])] = 1;
    }
    mark[v] = 0;
    int32_t c = 0;
    for (int32_t x = 1; x <= g->n; x++)
        if (mark[x]) out[c++] = x;
    free(mark);
    return c;
}

typedef struct { uint32_t key; int32_t x; } kx_t;
static int cmp_kx(const void* a, const void* b) {
    const kx_t* p = (const kx_t*)a;
    const kx_t* q = (const kx_t*)b;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    return (p->x > q->x) - (p->x < q->x);
}
static int cmp_i32(const void* a, const void* b) {
    const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* Proposal pairs of a 2-hop batch: anchor a = anchors[a] or drawn from
 * rand(seed, 2a); its K targets are the set members with the K smallest
 * (hi32(rand(rand(seed, 2a+1), x)), x), ascending x, padded with (v, v). */
void or_twohop_pairs(const or_graph* g, const int32_t* anchors, int32_t n_anchors, int32_t K, uint64_t seed,
                     int32_t* pairs) {
    int32_t* set = (int32_t*)malloc(sizeof(int32_t) * ((size_t)g->n + 1));
    kx_t* kx = (kx_t*)malloc(sizeof(kx_t) * ((size_t)g->n + 1));
    int32_t* take = (int32_t*)malloc(sizeof(int32_t) * ((size_t)g->n + 1));
    for (int32_t a = 0; a < n_anchors; a++) {
        const int32_t v = anchors ? anchors[a]
                                  : 1 + (int32_t)range32((uint32_t)rand_at(seed, 2 * (uint64_t)a), (uint32_t)g->n);
        const uint64_t q = rand_at(seed, 2 * (uint64_t)a + 1);
        const int32_t S = or_twohop_set(g, v, set);
        int32_t t = 0;
        if (S <= K) {
            memcpy(take, set, sizeof(int32_t) * (size_t)S);
            t = S;
        } else {
            for (int32_t i = 0; i < S; i++) {
                kx[i].key = (uint32_t)(rand_at(q, (uint64_t)set[i]) >> 32);
                kx[i].x = set[i];
            }
            qsort(kx, (size_t)S, sizeof(kx_t), cmp_kx);
            for (int32_t i = 0; i < K; i++) take[i] = kx[i].x;
            t = K;
            qsort(take, (size_t)t, sizeof(int32_t), cmp_i32);
        }
        for (int32_t j = 0; j < K; j++) {
            pairs[2 * ((int64_t)a * K + j)] = v;
            pairs[2 * ((int64_t)a * K + j) + 1] = j < t ? take[j] : v;
        }
    }
    free(set);
    free(kx);
    free(take);
}

/* u-major uniform pairs: draw c picks its first vertex 1 + range(lo32(rand(seed, c)), n);
 * entries are grouped by first vertex in ascending order, entry i of vertex u
 * pairs it with 1 + range(hi32(rand(mix(seed ^ K), i)), n - 1), skipping u.
 * Writes k pairs; returns 0. */
int or_umajor_pairs(uint64_t seed, int64_t k, int32_t n, int32_t* uv) {
    int64_t* cnt = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
    for (int64_t c = 0; c < k; c++) cnt[1 + range32((uint32_t)rand_at(seed, (uint64_t)c), (uint32_t)n)]++;
    const uint64_t q = mix64(seed ^ 0x6A09E667F3BCC909ULL);
    int64_t i = 0;
    for (int32_t u = 1; u <= n; u++)
        for (int64_t j = 0; j < cnt[u]; j++, i++) {
            const uint64_t r = rand_at(q, (uint64_t)i);
            int32_t b = (int32_t)range32((uint32_t)(r >> 32), (uint32_t)(n - 1)) + 1;
            uv[2 * i] = u;
            uv[2 * i + 1] = b >= u ? b + 1 : b;
        }
    free(cnt);
    return 0;
}

/* Seed of the first-vertex draws of the a-major triple stream. */
uint64_t or_amajor_seed(uint64_t seed) { return mix64(seed ^ 0x3C6EF372FE94F82BULL); }

/* a-major uniform triples: the first vertices are those of the u-major
 * stream of or_amajor_seed(seed), grouped ascending; entry i of vertex a
 * takes r = rand(mix(seed ^ K'), i): b = 1 + range(lo32 r, n - 1) skipping a,
 * c = 1 + range(hi32 r, n - 2) skipping a and b in ascending order.  Writes
 * k triples (3 ints each); returns 0. */
int or_amajor_triples(uint64_t seed, int64_t k, int32_t n, int32_t* abc) {
    int64_t* cnt = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
    const uint64_t s1 = or_amajor_seed(seed);
    for (int64_t c = 0; c < k; c++) cnt[1 + range32((uint32_t)rand_at(s1, (uint64_t)c), (uint32_t)n)]++;
    const uint64_t q = mix64(seed ^ 0xA54FF53A5F1D36F1ULL);
    int64_t i = 0;
    for (int32_t a = 1; a <= n; a++)
        for (int64_t j = 0; j < cnt[a]; j++, i++) {
            const uint64_t r = rand_at(q, (uint64_t)i);
            uint32_t b = range32((uint32_t)r, (uint32_t)(n - 1)) + 1u;
            if (b >= (uint32_t)a) b++;
            uint32_t c = range32((uint32_t)(r >> 32), (uint32_t)(n - 2)) + 1u;
            const uint32_t lo = (uint32_t)a < b ? (uint32_t)a : b;
            const uint32_t hi = (uint32_t)a < b ? b : (uint32_t)a;
            if (c >= lo) c++;
            if (c >= hi) c++;
            abc[3 * i] = a;
            abc[3 * i + 1] = (int32_t)b;
            abc[3 * i + 2] = (int32_t)c;
        }
    free(cnt);
    return 0;
}

/* Pairs at the given entry indices of the u-major stream (seed, k). */
void or_umajor_pairs_at(uint64_t seed, int64_t k, int32_t n, const int64_t* idx, int64_t cnt, int32_t* uv) {
    int64_t* off = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
    for (int64_t c = 0; c < k; c++) off[1 + range32((uint32_t)rand_at(seed, (uint64_t)c), (uint32_t)n)]++;
    int64_t run = 0;
    for (int32_t u = 1; u <= n; u++) {
        const int64_t x = off[u];
        off[u] = run;
        run += x;
    }
    off[n + 1] = run;
    const uint64_t q = mix64(seed ^ 0x6A09E667F3BCC909ULL);
    for (int64_t j = 0; j < cnt; j++) {
        const int64_t i = idx[j];
        int32_t lo = 1, hi = n;
        while (lo < hi) {
            const int32_t m = (lo + hi) / 2;
            if (off[m + 1] > i) hi = m; else lo = m + 1;
        }
        const uint64_t r = rand_at(q, (uint64_t)i);
        const int32_t b = (int32_t)range32((uint32_t)(r >> 32), (uint32_t)(n - 1)) + 1;
        uv[2 * j] = lo;
        uv[2 * j + 1] = b >= lo ? b + 1 : b;
    }
    free(off);
}

/* Best swap over all pairs u < v: largest delta, then smaller u, then smaller
 * v; (0, 0, 0) when n < 2. */
void or_best_swap(const or_graph* g, const int32_t* pi, int64_t* delta, int32_t* bu, int32_t* bv) {
    const int32_t n = g->n;
    int64_t best = 0;
    int32_t ru = 0, rv = 0;
    int have = 0;
#pragma omp parallel
    {
        int64_t tb = 0;
        int32_t tu = 0, tv = 0;
        int th = 0;
#pragma omp for schedule(dynamic, 16)
        for (int32_t u = 1; u < n; u++)
            for (int32_t v = u + 1; v <= n; v++) {
                const int64_t d = swap_delta_fast(g, pi, u, v);
                if (!th || d > tb || (d == tb && (u < tu || (u == tu && v < tv)))) {
                    tb = d;
                    tu = u;
                    tv = v;
                    th = 1;
                }
            }
#pragma omp critical
        if (th && (!have || tb > best || (tb == best && (tu < ru || (tu == ru && tv < rv))))) {
            best = tb;
            ru = tu;
            rv = tv;
            have = 1;
        }
    }
    *delta = best;
    *bu = ru;
    *bv = rv;
}

/* Best partner v > u of u: the largest or_swap_delta, ties to the smaller v
 * (INT64_MIN and 0 when u = n). */
void or_best_partner(const or_graph* g, const int32_t* pi, int32_t u, int64_t* delta, int32_t* bv) {
    int64_t best = INT64_MIN;
    int32_t rv = 0;
    for (int32_t v = u + 1; v <= g->n; v++) {
        const int64_t d = swap_delta_fast(g, pi, u, v);
        if (rv == 0 || d > best) {
            best = d;
            rv = v;
        }
    }
    *delta = best;
    *bv = rv;
}

static uint64_t or_rand_at(uint64_t seed, uint64_t c) {
    return mix64(seed * 0x9E3779B97F4A7C15ULL + (c + 1) * 0xD1B54A32D192ED03ULL);
}

/* Perturbation p of seed: 2k distinct vertices of 1..n from the stream
 * or_rand_at(or_rand_at(seed, p), j) (vertex 1 + high 32 bits x n >> 32,
 * repeats skipped), the i-th image swapped with the (i + k)-th.  sel gets the
 * 2k vertices. */
void or_sweep_perturb(int32_t* pi, int32_t n, uint64_t seed, uint64_t p, int32_t k, int32_t* sel) {
    if (k <= 0) return;
    uint8_t* used = (uint8_t*)calloc((size_t)n + 1, 1);
    const uint64_t st = or_rand_at(seed, p);
    int32_t cnt = 0;
    for (uint64_t j = 0; cnt < 2 * k; j++) {
        const uint32_t hi = (uint32_t)(or_rand_at(st, j) >> 32);
        const int32_t x = 1 + (int32_t)(((uint64_t)hi * (uint64_t)(uint32_t)n) >> 32);
        if (!used[x]) {
            used[x] = 1;
            sel[cnt++] = x;
        }
    }
    for (int32_t i = 0; i < k; i++) {
        const int32_t t = pi[sel[i]];
        pi[sel[i]] = pi[sel[i + k]];
        pi[sel[i + k]] = t;
    }
    free(used);
}

/* The sweep over max_passes passes (restarts included).  pi: start in, best
 * out.  trace[p] = best score after pass p; out = {best score, swaps applied,
 * restarts, pairs evaluated}. */
void or_sweep(const or_graph* g, int32_t* pi, uint64_t seed, int64_t max_passes, int32_t initial_swaps,
              int32_t restart_swaps, int64_t* trace, int64_t* out) {
    const int32_t n = g->n;
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t* sel = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    memcpy(cur, pi, sizeof(int32_t) * ((size_t)n + 1));
    uint64_t np = 0;
    or_sweep_perturb(cur, n, seed, np++, initial_swaps, sel);
    int64_t score = or_score(g, cur), best = score, accepted = 0, restarts = 0, evals = 0;
    memcpy(pi, cur, sizeof(int32_t) * ((size_t)n + 1));
    int64_t pass = 0, improvements = 0, last = 0;
    for (int64_t passes = 0; passes < max_passes; passes++) {
        for (int32_t u = 1; u <= n; u++) {
            int64_t d;
            int32_t v;
            or_best_partner(g, cur, u, &d, &v);
            evals += n - u;
            if (d > 0) {
                const int32_t t = cur[u];
                cur[u] = cur[v];
                cur[v] = t;
                score += d;
                improvements++;
                accepted++;
                if (score > best) {
                    best = score;
                    memcpy(pi, cur, sizeof(int32_t) * ((size_t)n + 1));
                }
            }
        }
        pass++;
        trace[passes] = best;
        if (pass > 2 && improvements == last) {
            memcpy(cur, pi, sizeof(int32_t) * ((size_t)n + 1));
            or_sweep_perturb(cur, n, seed, np++, restart_swaps, sel);
            score = or_score(g, cur);
            improvements = last = pass = 0;
            restarts++;
        }
        if (last < improvements) last = improvements;
    }
    out[0] = best;
    out[1] = accepted;
    out[2] = restarts;
    out[3] = evals;
    free(cur);
    free(sel);
}

/* Steepest ascent: apply the best pair while its delta is > 0, at most
 * max_iters times.  trace[0..k] = running score; returns k (swaps applied). */
int64_t or_best_swap_climb(const or_graph* g, int32_t* pi, int64_t max_iters, int64_t* trace) {
    int64_t s = or_score(g, pi), k = 0;
    trace[0] = s;
    while (k < max_iters) {
        int64_t d;
        int32_t u, v;
        or_best_swap(g, pi, &d, &u, &v);
        if (d <= 0) break;
        const int32_t t = pi[u];
        pi[u] = pi[v];
        pi[v] = t;
        s += d;
        trace[++k] = s;
    }
    return k;
}
