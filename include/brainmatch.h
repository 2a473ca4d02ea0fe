/*
 * brainmatch.h -- C ABI of the MI355X (gfx950) exact swap-delta engine.
 *
 * Problem.  Two directed integer-weighted graphs on vertices 1..n, male M and
 * female F, and a matching pi: 1..n -> 0..n (0 = unmatched).  The alignment
 * score is
 *     S(pi) = sum over male edges (u,v,w) of  c(w, pi(u), pi(v)),
 *     c(w, a, b) = min(w, w_F(a,b)) if F has the edge a->b, else 0,
 * which is what score-matching computes (src/score-matching.c:558-597).
 * Every delta returned here is the exact integer S(pi') - S(pi), int64.
 *
 * Conventions for every entry point:
 *   - return 0 (BM_OK) on success, a negative BM_ERR_* code on failure; the
 *     message is then available from bm_last_error(ctx) (bm_last_error(NULL)
 *     for a failed bm_create);
 *   - host pointers are caller-owned and only read/written during the call;
 *   - a context owns one HIP stream and its device buffers; it is not
 *     thread-safe (one host thread per context), like the single host thread
 *     that drives the reference kernels (src/graph-re-alignment-gpu-swap.cu:1002-1350);
 *   - calls that return host data synchronise the context stream.
 *
 * The reference has no library boundary: its tools launch
 * calculateSwapDeltasKernelQuick / calculateThreeNodeSwapKernel directly with
 * cudaMemcpy before and after (src/graph-re-alignment-gpu-swap.cu:1164-1195,
 * src/graph-re-alignment-gpu-3-swap.cu:1128-1195).  Each entry point below
 * names the reference code it replaces.
 */
#ifndef BRAINMATCH_H
#define BRAINMATCH_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BM_OK 0
#define BM_ERR_ARG (-1)      /* invalid argument (range, size, null) */
#define BM_ERR_HIP (-2)      /* HIP runtime failure */
#define BM_ERR_NOMEM (-3)    /* host or device allocation failed */
#define BM_ERR_IO (-4)       /* file could not be opened/parsed */
#define BM_ERR_STATE (-5)    /* call not valid in the current state */
#define BM_ERR_NODEV (-6)    /* no HIP device */

#define BM_API_VERSION 1

typedef struct bm_ctx bm_ctx;
typedef struct bm_batch bm_batch;

/* One graph as an edge list.  Vertex ids in 1..n.  Duplicate (src,dst) lines
 * keep the LAST weight once, as score-matching's add_edge does
 * (src/score-matching.c:339-359).  Weights are int32; INT32_MIN is rejected. */
typedef struct bm_edges {
    int64_t m;
    const int32_t* src;
    const int32_t* dst;
    const int32_t* w;
} bm_edges;

/* Edge list owned by the library (from the CSV loader or the generator). */
typedef struct bm_edge_list {
    int64_t m;
    int32_t n;          /* largest vertex id seen */
    int32_t* src;
    int32_t* dst;
    int32_t* w;
} bm_edge_list;

/* ------------------------------------------------------------------ */
/* Library / device                                                    */
/* ------------------------------------------------------------------ */

int bm_api_version(void);
const char* bm_last_error(const bm_ctx* ctx);

/* ------------------------------------------------------------------ */
/* Files (host only; no device needed)                                 */
/* ------------------------------------------------------------------ */

/* Graph CSV: one header line, then "src,dst,weight" lines; ids may carry an
 * 'm'/'f' prefix.  Replaces load_graph_from_csv
 * (src/graph-re-alignment-gpu-swap.cu:1389-1433) with score-matching's parse
 * semantics (src/score-matching.c:361-488). */
int bm_load_graph_csv(const char* path, bm_edge_list* out);
int bm_write_graph_csv(const char* path, const bm_edge_list* g);
void bm_edge_list_free(bm_edge_list* g);

/* Matching CSV: header, then "male,female" lines, generic or m/f-prefixed.
 * pi must hold n+1 ints; unlisted vertices get 0.  Replaces
 * load_benchmark_mapping (src/graph-re-alignment-gpu-swap.cu:699-746).
 * Returns BM_ERR_ARG if an id exceeds n. */
int bm_load_mapping_csv(const char* path, int32_t n, int32_t* pi, int64_t* n_lines);
/* Writes "Male Node ID,Female Node ID" then "m<u>,f<pi[u]>" for pi[u] != 0.
 * Replaces write_mapping (src/graph-re-alignment-gpu-swap.cu:681-696). */
int bm_write_mapping_csv(const char* path, int32_t n, const int32_t* pi);
/* Converts a matching file between the two text forms, keeping line order
 * and the id text: to_prefixed = 1 reads "male,female" lines (digits only)
 * and writes "Male Node ID,Female Node ID" then "m<male>,f<female>", as
 * transform-solution (src/transform-solution.c:63-92); to_prefixed = 0
 * reads "m<male>,f<female>" lines and writes "Male ID,Female ID" then
 * "<male>,<female>", as convert-VNC-matching (src/convert-VNC-matching.c:53-73).
 * The first input line is a header and is skipped; lines are read 1,023
 * bytes at a time, as the reference reads them.  A line that does not match
 * stops the conversion with BM_ERR_IO ("Error on line <k>: ..." in
 * bm_last_error(NULL)), the lines before it written.  *n_lines (may be NULL)
 * = lines read, header included. */
int bm_convert_mapping_csv(const char* in_path, const char* out_path, int to_prefixed, int64_t* n_lines);

/* ------------------------------------------------------------------ */
/* Synthetic workloads (host only)                                     */
/* ------------------------------------------------------------------ */

#define BM_WEIGHTS_GEOMETRIC 1   /* geometric(p=0.3)+1: failures-before-success + 1, >= 1 */
#define BM_WEIGHTS_LOGNORMAL 2   /* max(1, round(exp(0.5 + 1.2 Z))) */

typedef struct bm_synth_params {
    int32_t n;              /* vertices per graph */
    double mean_out_degree; /* target mean out-degree of the male graph */
    double alpha;           /* power-law exponent of out-degrees (2.1) */
    int32_t weights;        /* BM_WEIGHTS_* */
    double modify_frac;     /* fraction of female edges modified (0.2) */
    double shuffle_frac;    /* fraction of pi0 entries shuffled among themselves */
    uint64_t seed;
} bm_synth_params;

/* Male graph, female graph (male relabelled by a hidden permutation P, then
 * modify_frac of edges half reweighted / half rewired), the hidden P (n+1
 * ints, P[0]=0) and pi0 = P with shuffle_frac of its entries shuffled.
 * No self-loops, no duplicate (src,dst), weights >= 1. */
int bm_synth_pair(const bm_synth_params* p, bm_edge_list* male, bm_edge_list* female,
                  int32_t* perm_true, int32_t* pi0);

/* Deterministic greedy initial matching in the manner of greedy-match
 * (src/greedy-match.c:213-530): features (in, out, 2-hop reach, 2-hop
 * co-reach), both sides sorted by feature total descending, then each male
 * vertex takes the unmatched female vertex with the smallest L1 feature
 * distance.  Ties are broken by the smaller vertex id (the reference's qsort
 * leaves them to libc).  pi gets n+1 ints. */
int bm_greedy_match(int32_t n, const bm_edges* male, const bm_edges* female, int32_t* pi);
/* The same greedy matching computed on the device from a context's resident
 * graphs (bm_create's edges): the 2-hop counts with an LDS bitmap per vertex,
 * a radix sort of both vertex lists, and one 1024-thread block for the greedy
 * walk (the O(n^2) loop of src/greedy-match.c:495-522).  pi gets n+1 ints;
 * the context matching is not changed.  BM_ERR_ARG when the two graphs have
 * different numbers of vertices with an incident edge. */
int bm_greedy_match_device(bm_ctx* ctx, int32_t* pi);

/* ------------------------------------------------------------------ */
/* Context: graphs + matching resident in HBM                          */
/* ------------------------------------------------------------------ */

/* Builds male CSR/CSC and the per-row female hash tables, copies them to
 * device `device`, and sets pi = identity-free zero matching.  Replaces the
 * dense (n+1)^2 short matrices and their 1.37 GB H2D copy
 * (src/graph-re-alignment-gpu-swap.cu:1040-1062). */
int bm_create(bm_ctx** out, int32_t n, const bm_edges* male, const bm_edges* female,
              int32_t device);
void bm_destroy(bm_ctx* ctx);

int32_t bm_num_vertices(const bm_ctx* ctx);
int64_t bm_num_edges(const bm_ctx* ctx, int female);

/* Per-vertex degree arrays of the male (female=0) or female (female=1)
 * graph, n+1 entries each (index 0 = 0). */
int bm_degrees(const bm_ctx* ctx, int female, int32_t* out_deg, int32_t* in_deg);

/* Replace the matching: pi has n+1 entries, pi[0] ignored, values in 0..n.
 * Replaces the per-iteration cudaMemcpy of the mapping
 * (src/graph-re-alignment-gpu-swap.cu:1167). */
int bm_set_mapping(bm_ctx* ctx, const int32_t* pi);
int bm_get_mapping(bm_ctx* ctx, int32_t* pi);

/* Full score of the current matching on the device.  Replaces
 * calculate_alignment_score (src/graph-re-alignment-gpu-swap.cu:563-578);
 * equals score-matching (src/score-matching.c:558-597) exactly, in int64. */
int bm_score(bm_ctx* ctx, int64_t* score);

/* ------------------------------------------------------------------ */
/* Host-buffer convenience calls (upload, evaluate, download)          */
/* ------------------------------------------------------------------ */

/* delta[i] = S(pi o (u_i v_i)) - S(pi) for uv = {u0,v0,u1,v1,...}, k pairs;
 * u == v gives 0.  Replaces calculate_swap_delta
 * (src/graph-alignment-optimized.c:321-408) for a whole list. */
int bm_swap_deltas(bm_ctx* ctx, const int32_t* uv, int64_t k, int64_t* delta);

/* The reference's changed-first enumeration: with changed[0..c) (any order;
 * sorted internally like sortArray, src/graph-re-alignment-gpu-swap.cu:296-307)
 * and the remaining vertices ascending (fillArray, :983-999), the
 * P = c(c-1)/2 + c(n-c) pairs of calculateSwapDeltasKernelQuick
 * (:141-250) in its output order.  delta gets P entries, uv 2P (may be NULL).
 * Returns P through *count. */
int bm_swap_deltas_changed(bm_ctx* ctx, const int32_t* changed, int32_t c,
                           int64_t* delta, int32_t* uv, int64_t* count);

/* For each triple (a,b,c) of distinct vertices: the best of the five
 * non-identity reassignments of (pi(a),pi(b),pi(c)), numbered as in
 * calculateThreeNodeSwapKernel (src/graph-re-alignment-gpu-3-swap.cu:167-174):
 *   1 (f1,f3,f2)  2 (f2,f1,f3)  3 (f2,f3,f1)  4 (f3,f1,f2)  5 (f3,f2,f1),
 * best = max(0, max_p delta_p) with the first strictly greater p winning and
 * perm 0 when none is positive (:176-177, :294-297). */
int bm_rot3_deltas(bm_ctx* ctx, const int32_t* abc, int64_t k, int64_t* best, uint8_t* perm);

/* The reference's 3-section triple enumeration (:67-150) over changed[0..c)
 * (sorted internally) and the remaining vertices ascending:
 * T = C(c,3) + C(c,2)(n-c) + c(n-c-1).  abcp gets (a,b,c,perm) per triple like
 * node_triplets (may be NULL).  Returns T through *count. */
int bm_rot3_deltas_changed(bm_ctx* ctx, const int32_t* changed, int32_t c,
                           int64_t* best, int32_t* abcp, int64_t* count);

/* Deterministic disjoint acceptance over a host candidate list: candidates
 * with delta > 0 are taken in order (delta desc, index asc); one is accepted
 * iff neither endpoint lies in the closed 1-hop neighbourhood (in- and
 * out-edges of M) of an endpoint of an already accepted swap.  Accepted swaps
 * are applied to the device matching and their deltas (which are additive by
 * construction) summed.  accepted[] gets candidate indices in acceptance
 * order (capacity k, may be NULL).  No reference counterpart is exact: the
 * reference accepts one swap per launch with a full recount
 * (src/graph-re-alignment-gpu-swap.cu:1220-1249). */
int bm_accept_disjoint(bm_ctx* ctx, const int32_t* uv, const int64_t* delta, int64_t k,
                       int32_t* accepted, int64_t* n_accepted, int64_t* gain);

/* Apply one swap / one 3-permutation to the device matching
 * (src/graph-re-alignment-gpu-3-swap.cu:925-956 for apply_permutation). */
int bm_apply_swap(bm_ctx* ctx, int32_t u, int32_t v);
int bm_apply_perm3(bm_ctx* ctx, int32_t a, int32_t b, int32_t c, int32_t perm);

/* ------------------------------------------------------------------ */
/* Device-resident batches (no host round trip per launch)             */
/* ------------------------------------------------------------------ */

typedef struct bm_stats {
    int64_t count;      /* candidates in the batch */
    int64_t sum;        /* sum of deltas (int64) */
    int64_t max;        /* largest delta (INT64_MIN for an empty batch) */
    int64_t argmax;     /* first index attaining max (-1 if empty) */
    int64_t n_pos;      /* deltas > 0 */
    int64_t n_zero;     /* deltas == 0 */
} bm_stats;

int bm_batch_create(bm_ctx* ctx, int64_t capacity, bm_batch** out);
void bm_batch_destroy(bm_batch* b);
int64_t bm_batch_size(const bm_batch* b);

/* Fill the batch with candidates. */
int bm_batch_set_pairs(bm_batch* b, const int32_t* uv, int64_t k);
int bm_batch_set_triples(bm_batch* b, const int32_t* abc, int64_t k);
int bm_batch_gen_pairs(bm_batch* b, uint64_t seed, int64_t k);      /* bm_uniform_pair stream */
int bm_batch_gen_triples(bm_batch* b, uint64_t seed, int64_t k);    /* bm_uniform_triple stream */
int bm_batch_gen_changed_pairs(bm_batch* b, const int32_t* changed, int32_t c);
int bm_batch_gen_changed_triples(bm_batch* b, const int32_t* changed, int32_t c);

/* 2-hop proposals.  The 2-hop set of v is collect_neighborhood's
 * (src/refine-twohop-random.c:582-625): out-neighbours, their out-neighbours,
 * in-neighbours and their in-neighbours in the male graph, minus v.
 * bm_twohop_set writes it in ascending order (out: n entries).
 * bm_batch_gen_twohop fills n_anchors * per_anchor pairs: anchor a is
 * anchors[a] (or, with anchors NULL, 1 + range(lo32(bm_rand_at(seed, 2a)), n))
 * and its per_anchor slots hold (anchor, x) for the per_anchor members x of
 * its set with the smallest hi32(bm_rand_at(bm_rand_at(seed, 2a+1), x)) (ties:
 * smaller x), in ascending x, i.e. a uniform sample without replacement as
 * the reference's partial shuffle (:686-691) draws it; slots beyond the set
 * size hold (anchor, anchor), delta 0.  Needs (n+1)/8 bytes of LDS per
 * block (n <= 1.2M). */
int bm_twohop_set(bm_ctx* ctx, int32_t v, int32_t* out, int32_t* count);
int bm_batch_gen_twohop(bm_batch* b, const int32_t* anchors, int32_t n_anchors, int32_t per_anchor, uint64_t seed);

/* Evaluate.  eval_swaps_uniform generates pair c of stream `seed` inside the
 * delta kernel (nothing is read for the candidate) and writes the deltas. */
int bm_batch_eval_swaps(bm_batch* b);
int bm_batch_eval_swaps_uniform(bm_batch* b, uint64_t seed, int64_t k);
int bm_batch_eval_rot3(bm_batch* b);

/* u-major uniform pairs (bm_umajor_first / bm_umajor_partner in
 * brain-match_amd/csrc/bm_rng.h): k independent uniform pairs listed grouped
 * by their first vertex, so an evaluator can visit every first vertex once.
 * The reference's changed-first enumeration is grouped the same way (all
 * partners of one changed vertex together, src/graph-re-alignment-gpu-swap.cu:163-195).
 * gen_umajor writes the pairs into the batch; eval_swaps_umajor evaluates the
 * same stream without materialising it.  With the gain table enabled and
 * n <= 18,975, one block per first vertex stages the two table rows it needs
 * and its adjacency bitmaps in LDS (the table's transpose is kept next to it,
 * doubling the table's memory); up to n of about 218,000 it stages only the
 * bitmaps and gathers both table entries (no transpose); beyond that each
 * pair is evaluated on its own. */
int bm_batch_gen_umajor(bm_batch* b, uint64_t seed, int64_t k);
int bm_batch_eval_swaps_umajor(bm_batch* b, uint64_t seed, int64_t k);

/* a-major uniform triples (bm_amajor_seed / bm_amajor_partners in
 * brain-match_amd/csrc/bm_rng.h): k independent uniform ordered triples of
 * distinct vertices (the law of bm_batch_gen_triples) listed grouped by their
 * first vertex.  gen_amajor writes them into the batch; eval_rot3_amajor
 * evaluates the best 3-vertex permutation of each (as bm_batch_eval_rot3) and,
 * with the gain table enabled and n <= 19,839, without materialising them:
 * one block per first vertex a stages the rows G(a, .) and GT(pi(a), .), so a
 * triple reads two table entries from HBM instead of nine.  Deltas and
 * permutations are read with bm_batch_get (cand is not available after the
 * staged evaluation). */
int bm_batch_gen_amajor(bm_batch* b, uint64_t seed, int64_t k);
int bm_batch_eval_rot3_amajor(bm_batch* b, uint64_t seed, int64_t k);

/* Read back a slice [first, first+count).  Any output pointer may be NULL.
 * cand gets 2 ints per pair batch / 4 ints (a,b,c,perm) per triple batch. */
int bm_batch_get(bm_batch* b, int64_t first, int64_t count, int64_t* delta,
                 uint8_t* perm, int32_t* cand);

/* Device reductions over the batch's deltas. */
int bm_batch_stats(bm_batch* b, bm_stats* out);
/* Index of the rank-th (0-based, index order) candidate with delta > 0
 * (kind = 1) or delta == 0 (kind = 0); -1 if there is none. */
int bm_batch_pick(bm_batch* b, int32_t kind, int64_t rank, int64_t* index);
/* Disjoint acceptance (as bm_accept_disjoint) over a pair batch, applied to
 * the context matching.  max_accept <= 0 means unlimited. */
int bm_batch_accept_disjoint(bm_batch* b, int64_t max_accept, int64_t* n_accepted, int64_t* gain);

/* ------------------------------------------------------------------ */
/* Placement-gain table (second evaluation strategy, same results)     */
/* ------------------------------------------------------------------ */

/* G(u, f) = score of u's incident male edges (self-loops excluded) if u alone
 * were placed on female f, an int32 table of n+1 rows in HBM, each row padded
 * to a multiple of 32 entries (1.37 GB at n = 18,524; bm_gain_table_bytes).  While enabled, every swap / 3-permutation evaluation reads it
 * (four table entries plus an exact correction for edges inside the moved set
 * and self-loops) instead of scanning the candidate's edges; bm_set_mapping
 * rebuilds it and bm_apply_swap, bm_apply_perm3 and the disjoint acceptance
 * patch the rows of the moved vertices' neighbours.  Results are identical to
 * the scan kernels.  The reference's closest relative is the n x n int16
 * swap-delta table of src/best-swap/score-lookup.c:246-293 (patched after each
 * swap, :129-174).  Fails with BM_ERR_NOMEM when the table does not fit the
 * device and with BM_ERR_ARG when a vertex's incident weights could overflow
 * int32. */
int bm_gain_enable(bm_ctx* ctx, int on);
int bm_gain_enabled(const bm_ctx* ctx);
/* Number of female rows whose part of the table the build sums as dense
 * 16-bit column arrays, 16 male rows at a time, instead of expanding them
 * edge by edge: the longest rows of at least a quarter of the columns, at
 * most 128 and at most 65535 / the largest female weight; only for graphs
 * whose weights fit 16 bits and whose male weights are all >= 0, whose
 * table rows fit one LDS column window (n + 1 <= 40,768), and only while no
 * two male vertices share a female.  0 when the table is off or no
 * row qualifies.  An introspection call with no counterpart in the
 * reference: the table it describes is the one above. */
int bm_gain_dense_rows(const bm_ctx* ctx);
int bm_gain_rebuild(bm_ctx* ctx);
int64_t bm_gain_table_bytes(int32_t n);
/* Copy row u (n+1 entries) of the table to the host. */
int bm_gain_row(bm_ctx* ctx, int32_t u, int32_t* row);
/* Exact best swap over all n(n-1)/2 pairs u < v of the context matching:
 * the largest delta, ties to the smaller u, then the smaller v.  Replaces the n x n int16 delta table and max-heap of
 * src/best-swap/score-lookup.c:246-293 (create_score_lookup) and :612-620
 * (extract_max): recomputed from the table on every call, so no entry is
 * stale and no delta is truncated to int16.  Needs the table (BM_ERR_ARG
 * otherwise) and, once, a transposed copy of it (bm_gain_table_bytes more). */
int bm_best_swap(bm_ctx* ctx, int64_t* delta, int32_t* u, int32_t* v);
/* Best partner of every u in u0 .. u0 + count - 1 under the context
 * matching: delta[i] = the largest exact delta of swapping u = u0 + i with a
 * v in u+1..n, v[i] = that v (ties to the smaller v); delta[i] = INT64_MIN and
 * v[i] = 0 for u = n.  Replaces the inner loop of the OpenMP tool's sweep
 * (src/re-alignment-parallel.c:598-621: the largest calculate_swap_delta
 * over node_m2 > node_m1, the first found among equal ones).  Needs the
 * placement-gain table (BM_ERR_ARG otherwise) and its transposed copy. */
int bm_best_partners(bm_ctx* ctx, int32_t u0, int32_t count, int64_t* delta, int32_t* v);

/* ------------------------------------------------------------------ */
/* Measurement                                                         */
/* ------------------------------------------------------------------ */

#define BM_KERNEL_SWAP 0
#define BM_KERNEL_ROT3 1
#define BM_KERNEL_SCORE 2
#define BM_KERNEL_GAIN_BUILD 3
#define BM_KERNEL_GAIN_SWAP 4
#define BM_KERNEL_GAIN_ROT3 5
#define BM_KERNEL_GAIN_UPDATE 6
#define BM_KERNEL_AUX 7          /* u-major stream offsets, table transpose */
#define BM_KERNEL_COUNT 8

/* When enabled, every delta/score launch is bracketed by HIP events on the
 * context stream; bm_timing_get synchronises and returns the number of
 * launches and their summed duration in milliseconds. */
int bm_timing_enable(bm_ctx* ctx, int on);
int bm_timing_get(bm_ctx* ctx, int32_t kernel, int64_t* launches, double* total_ms);
int bm_timing_reset(bm_ctx* ctx);
int bm_sync(bm_ctx* ctx);

/* Algorithmic bytes of a batch (per-candidate byte model, summed exactly from
 * the degree arrays): pairs 48 + 28*D(u,v) (+8 if read from memory), triples
 * 69 + 36*D3 (+12 if read from memory). */
int bm_batch_algo_bytes(bm_batch* b, int32_t generated, int64_t* bytes, int64_t* sum_degree);
int bm_uniform_algo_bytes(bm_ctx* ctx, uint64_t seed, int64_t k, int64_t* bytes, int64_t* sum_degree);
/* The same sum over the u-major stream (bm_batch_eval_swaps_umajor). */
int bm_umajor_algo_bytes(bm_ctx* ctx, uint64_t seed, int64_t k, int64_t* bytes, int64_t* sum_degree);

/* Launch-shape knobs (autotuner).  waves_per_block in {1,2,4,8},
 * lds_slots per wave (table slots of 4 bytes for graphs with n < 2^17 and
 * female weights in 0..32767, else 8 bytes; a multiple of 4), blocks_per_cu
 * >= 1 (capped at launch by the occupancy the kernel reaches). */
typedef struct bm_tuning {
    int32_t waves_per_block;
    int32_t lds_slots;
    int32_t blocks_per_cu;
    int32_t group_lanes;  /* lanes cooperating on one candidate (64) */
    int32_t unroll;       /* list entries per lane in flight: 2, 4 or 8 (swap kernel),
                             2 or 4 (3-vertex kernel, annealing chains);
                             female-row entries per lane, 4, 8, 16 or 32 (table build,
                             which uses only this field) */
} bm_tuning;
/* Until bm_set_tuning is called for it, the swap (3-vertex) kernel runs a
 * batch of fewer than 32,768 candidates in 4-wave blocks with 8 (4) list
 * entries per lane, where one candidate's latency sets the time, and a larger
 * batch in the shape bm_get_tuning reports. */
int bm_get_tuning(const bm_ctx* ctx, int32_t kernel, bm_tuning* t);
int bm_set_tuning(bm_ctx* ctx, int32_t kernel, const bm_tuning* t);

/* Device information string ("gfx950, 256 CUs, ...") for logs. */
int bm_device_info(bm_ctx* ctx, char* buf, size_t len);

/* ------------------------------------------------------------------ */
/* Search drivers (host loop over the device calls above)              */
/* ------------------------------------------------------------------ */

#define BM_SEARCH_SWAP_WEIGHTED 0  /* graph-re-alignment-gpu-swap: 100-101 score-weighted pairs */
#define BM_SEARCH_SWAP_UNIFORM 1   /* graph-re-alignment-gpu: 3-5 uniform pairs */
#define BM_SEARCH_ROT3 2           /* graph-re-alignment-gpu-3-swap: 3-5 uniform pairs, triples */

typedef struct bm_search_opts {
    int32_t mode;             /* BM_SEARCH_* */
    uint64_t seed;
    double time_limit_s;      /* <= 0: until *stop_flag or max_iters */
    int64_t max_iters;        /* <= 0: unlimited (kernel launches, as the reference's iter) */
    const char* out_path;     /* mapping written on every new best (NULL: none) */
    const char* sync_path;    /* adopted if better every sync_interval_s (NULL: none) */
    double sync_interval_s;
    int64_t verify_every;     /* device recount every this many accepted moves (0: never) */
    int32_t verbose;          /* 0 quiet, 1 [INFO] lines like the reference */
    volatile int* stop_flag;  /* polled between launches; nonzero stops */
    double restart_gap;       /* restart from best when (best-cur)/best exceeds this (0.002) */
} bm_search_opts;

typedef struct bm_search_result {
    int64_t initial_score;
    int64_t best_score;
    int64_t final_score;      /* score of the matching left on the device */
    int64_t iterations;       /* delta-kernel launches */
    int64_t accepted;         /* moves applied (positive and zero) */
    int64_t evaluations;      /* candidates evaluated */
    int64_t verify_failures;  /* running score != device recount (must stay 0) */
    double seconds;
    int64_t fallback_rounds;  /* bm_climb, learned proposal with fallback: rounds that took uniform
                                 candidates (0 elsewhere) */
    int64_t restarts;         /* bm_sweep: restarts from a perturbed best matching (0 elsewhere) */
} bm_search_result;

/* The perturb -> evaluate -> accept loop of optimize_mapping
 * (src/graph-re-alignment-gpu-swap.cu:1002-1350, src/graph-re-alignment-gpu-3-swap.cu:960-1361,
 * src/graph-re-alignment-gpu.cu:869), with the evaluation, the argmax and the
 * random positive/zero picks on the device and the running score kept from
 * the exact deltas instead of a full recount per move.  Starts from the
 * context matching; best_pi (n+1 ints, may be NULL) receives the best. */
int bm_search(bm_ctx* ctx, const bm_search_opts* opts, int32_t* best_pi, bm_search_result* res);

/* Greedy disjoint-swap hill climbing: for r in 0..rounds-1, evaluate `batch`
 * uniform pairs of stream (seed + r), accept a disjoint set (as
 * bm_batch_accept_disjoint) and apply it.  score_trace (rounds+1 entries, may
 * be NULL) gets the running score before round 0 and after each round; with
 * verify != 0 every round is checked against a device recount. */
int bm_hillclimb(bm_ctx* ctx, uint64_t seed, int64_t rounds, int64_t batch, int32_t verify,
                 int64_t* score_trace, bm_search_result* res);
/* Steepest-ascent swaps, the main loop of the score-lookup driver's
 * optimize_mapping (src/best-swap/score-lookup.c:519-716) without its heap:
 * each iteration applies bm_best_swap's pair while its delta is > 0, at most
 * max_iters times.  score_trace (max_iters+1 entries, may be NULL) gets the
 * running score before the first swap and after each; verify_every > 0
 * checks it against a device recount every that many swaps and at the end.
 * res->iterations = swaps applied; a stop with a positive pair left is a
 * max_iters stop.  Needs the placement-gain table. */
int bm_best_swap_climb(bm_ctx* ctx, int64_t max_iters, int64_t verify_every, int64_t* score_trace,
                       bm_search_result* res);
/* Options of bm_sweep.  The random disjoint swaps of perturbation p (0 =
 * before the first pass, 1, 2, ... = the restarts) pick 2k distinct vertices
 * of 1..n from the stream bm_rand_at(bm_rand_at(seed, p), j), j = 0, 1, ...
 * (vertex 1 + the high 32 bits of each draw times n, >> 32; repeats skipped)
 * and swap the images of the i-th and (i + k)-th. */
typedef struct bm_sweep_opts {
    uint64_t seed;
    int32_t initial_swaps;    /* random disjoint swaps before the first pass (the tool: 50) */
    int32_t restart_swaps;    /* random disjoint swaps of the best matching at a restart (the tool: 200) */
    int32_t batch;            /* vertices whose best partners one launch finds (0: 512) */
    int32_t verify;           /* recount the score after every applied swap, as the tool does */
    int64_t max_passes;       /* passes over the vertices, restarts included (0: no limit) */
    double time_limit_s;      /* checked between launches (0: no limit); one of the two limits is required */
} bm_sweep_opts;
/* The OpenMP tool's optimisation loop, optimize_mapping of
 * src/re-alignment-parallel.c:556-676: after a random perturbation, passes
 * over node_m1 = 1..n, each swapping node_m1 with its best partner
 * (bm_best_partners) when that delta is > 0 and keeping the best matching;
 * from the third pass on, a pass without an applied swap restarts from the
 * best matching perturbed by restart_swaps random swaps.  Each vertex is
 * evaluated under the matching left by the swaps before it, as in the tool:
 * a launch finds the best partners of `batch` consecutive vertices, the first
 * of them with a positive delta is swapped, and the next launch starts after
 * it.  Differences from the tool, all deliberate: its perturbation draws
 * rand() % n (so it can swap vertex 0's slot and never vertex n) and is seeded
 * by the clock; ties among best partners go to the smaller v (its OpenMP
 * critical section keeps whichever thread reports first); it runs until
 * killed.  On return the context holds the best matching (res->final_score =
 * res->best_score), best_pi (n+1 ints, may be NULL) receives it, pass_trace
 * (max_passes entries, may be NULL; only with max_passes > 0) the best score
 * after each pass; res->iterations = passes, res->accepted = swaps applied,
 * res->evaluations = pairs evaluated, res->restarts.  Needs the
 * placement-gain table. */
int bm_sweep(bm_ctx* ctx, const bm_sweep_opts* opts, int32_t* best_pi, int64_t* pass_trace, bm_search_result* res);
/* The same climb over 2-hop proposals: round r evaluates
 * bm_batch_gen_twohop(NULL, anchors, per_anchor, seed + r) (refine-twohop-random's
 * proposal, src/refine-twohop-random.c:676-693, with 10 anchors x 1000
 * targets per thread there). */
int bm_hillclimb_twohop(bm_ctx* ctx, uint64_t seed, int64_t rounds, int32_t anchors, int32_t per_anchor,
                        int32_t verify, int64_t* score_trace, bm_search_result* res);

/* ------------------------------------------------------------------ */
/* Learned swap proposal                                               */
/* ------------------------------------------------------------------ */

/* A 16 -> 128 -> 128 -> 1 perceptron (ReLU) in BF16 with FP32 accumulation on
 * the matrix cores, trained on the GPU with its own forward and backward
 * kernels on (candidate, exact delta) samples that the context's delta
 * kernels produce, regressing t = sign(delta) log(1 + |delta|).  A pair's 16
 * inputs are 8 per-vertex features of u then of v under the current matching:
 * log1p out-degree, log1p in-degree, log1p weight sum, carried score / (weight
 * sum + 1), fraction of incident edges contributing 0, log1p out- and
 * in-degree of the female image, log1p carried score.  It only ranks
 * candidates: every proposed swap is still judged by the exact delta kernels.
 * The reference has no learned component; its proposals are uniform,
 * score-weighted (src/graph-re-alignment-gpu-swap.cu:816-887) or 2-hop
 * (src/refine-twohop-random.c:676-693).  Parameter layout (float):
 * W1 [16][128], b1 [128], W2 [128][128], b2 [128], w3 [128], b3. */
typedef struct bm_mlp bm_mlp;
#define BM_MLP_FEATURES 16
#define BM_MLP_HIDDEN 128
int32_t bm_mlp_num_params(void);                     /* 18,817 */
int bm_mlp_create(bm_ctx* ctx, uint64_t seed, bm_mlp** out);   /* He-normal init, zero biases */
void bm_mlp_destroy(bm_mlp* m);
int bm_mlp_get_params(bm_mlp* m, float* params);
int bm_mlp_set_params(bm_mlp* m, const float* params);   /* also resets the optimiser */
/* Recompute the per-vertex features for the context's current matching. */
int bm_mlp_refresh_features(bm_mlp* m);
int bm_mlp_vertex_features(bm_mlp* m, float* out);       /* (n+1) x 8 floats */
int bm_mlp_score_pairs(bm_mlp* m, const int32_t* uv, int64_t k, float* out);
/* Gradient of mean (y - target)^2 over the k pairs (no update); mse may be NULL. */
int bm_mlp_gradient(bm_mlp* m, const int32_t* uv, const float* target, int64_t k, float* grad, float* mse);
/* `steps` Adam steps (lr, betas 0.9 / 0.999) on batches of uniform pairs of
 * stream bm_rand_at(seed, step) labelled by the context's exact deltas;
 * mse_trace (steps floats, may be NULL) gets each batch's loss. */
int bm_mlp_train(bm_mlp* m, uint64_t seed, int64_t steps, int32_t batch, float lr, float* mse_trace);
/* One Adam step on the pairs of an evaluated pair batch and their exact
 * deltas (the candidates a search has just sent to the delta kernels), with
 * the features of the current matching (refresh them before the batch is
 * ranked or evaluated).  mse (may be NULL) gets the batch loss before the
 * step.  BM_ERR_STATE when the batch holds no evaluated pairs. */
int bm_mlp_train_batch(bm_mlp* m, bm_batch* b, float lr, float* mse);
/* The `top` best-scored of `pool` uniform pairs of stream `seed` into the
 * batch as a pair list (descending score). */
int bm_batch_gen_learned(bm_batch* b, bm_mlp* m, uint64_t seed, int64_t pool, int64_t top);

/* bm_batch_gen_learned without any device allocation: the pool's scores
 * and the select state live in a fixed device array of the library (one per
 * device, 4.25 MiB, held until the process ends; its address is taken in
 * bm_mlp_create), so the call can run inside HIP graph capture or a serving
 * loop that must not allocate.  Limits: pool <= BM_NOALLOC_POOL, top <= 8192
 * (BM_ERR_ARG beyond).  The same pairs as bm_batch_gen_learned.  Calls that
 * overlap on one device race on the shared array.  No counterpart in the
 * reference: it is for callers that cannot allocate. */
#define BM_NOALLOC_POOL (1 << 20)
int bm_batch_gen_learned_noalloc(bm_batch* b, bm_mlp* m, uint64_t seed, int64_t pool, int64_t top);

#define BM_PROPOSAL_UNIFORM 0
#define BM_PROPOSAL_TWOHOP 1
#define BM_PROPOSAL_LEARNED 2

typedef struct bm_climb_opts {
    int32_t proposal;         /* BM_PROPOSAL_* */
    uint64_t seed;
    double time_limit_s;      /* <= 0: none (then max_rounds > 0) */
    int64_t max_rounds;       /* <= 0: none */
    int64_t per_round;        /* pairs evaluated exactly per round */
    int64_t pool;             /* learned: uniform pairs ranked per round */
    int32_t per_anchor;       /* 2-hop: targets per anchor */
    bm_mlp* mlp;              /* learned */
    int32_t refresh_every;    /* learned: rounds between feature refreshes */
    int32_t verify;           /* device recount after every round */
    int32_t retrain_every;    /* learned: one bm_mlp_train_batch step on a learned round's evaluated
                                 candidates every this many learned rounds (0: never) */
    float retrain_lr;         /* learned: Adam step size of those steps (> 0 when retrain_every > 0) */
    int32_t fallback;         /* learned: nonzero lets each round take uniform candidates instead when
                                 uniform rounds have recently gained more exact score per second */
} bm_climb_opts;

/* Disjoint-swap climb with one proposal source, for comparing proposals at
 * equal wall clock (see bm_batch_accept_disjoint for the acceptance).  With
 * the learned proposal and fallback set, each source keeps a moving average
 * (weight 1/8 per round) of the exact gain per second of its rounds; a round
 * uses the source with the larger average, every 16th round the other one,
 * and the first two rounds one of each.  res->evaluations counts every
 * candidate sent to the delta kernels; res->accepted every applied swap. */
int bm_climb(bm_ctx* ctx, const bm_climb_opts* opts, int64_t* trace_score, double* trace_time, int64_t trace_cap,
             bm_search_result* res);

/* ------------------------------------------------------------------ */
/* Annealing chains (many independent chains, batched over instances)  */
/* ------------------------------------------------------------------ */

/* Simulated annealing as simulated_annealing / probabilistic_swaps in
 * src/best-swap/score-lookup.c:300-400: a uniform random move is accepted
 * when delta >= 0 or when U < exp(delta / T_s), with T_s = t0 * decay^s.
 * Differences: the delta is recomputed exactly every step (no int16 lookup
 * table to patch), moves are swaps or, with probability p_rot3, one of the
 * five permutations of a uniform vertex triple, the random streams are
 * counter-based per chain, and exp is a deterministic function with the same
 * bits on the host and the device, so that any chain can be replayed exactly
 * on the CPU.  One chain per wavefront; all chains of all instances advance
 * in one kernel launch, replayed from a captured HIP graph. */
typedef struct bm_anneal bm_anneal;

typedef struct bm_anneal_params {
    uint64_t seed;
    int32_t chains_per_instance;  /* >= 1 */
    double t0;                    /* initial temperature, > 0 (50) */
    double decay;                 /* per-step factor in (0, 1] (0.9999) */
    double p_rot3;                /* probability of a 3-vertex move in [0, 1] */
    int32_t traj_cap;             /* score samples kept per chain (one per launch) */
    double p_pool;                /* probability in [0, 1] that a move is built on a pair of its
                                     instance's proposal pool (when that pool is not empty) */
} bm_anneal_params;

/* One instance per context (all on the same device); every chain of instance
 * i starts from context i's current matching.  Chain j of instance i has the
 * global index i * chains_per_instance + j and the stream seed
 * bm_rand_at(seed, global index). */
int bm_anneal_create(bm_anneal** out, bm_ctx* const* ctxs, int32_t n_instances, const bm_anneal_params* p);
void bm_anneal_destroy(bm_anneal* a);
int64_t bm_anneal_num_chains(const bm_anneal* a);
/* Advance every chain by steps_per_launch steps, `launches` times (one
 * trajectory sample per chain per launch). */
int bm_anneal_run(bm_anneal* a, int64_t steps_per_launch, int32_t launches);
/* Per-chain state; every array has num_chains entries and may be NULL. */
int bm_anneal_state(bm_anneal* a, int64_t* score, int64_t* best, int64_t* steps, int64_t* accepted);
/* Current (which = 0) or best (which = 1) matching of a chain (n+1 ints). */
int bm_anneal_mapping(bm_anneal* a, int64_t chain, int32_t which, int32_t* pi);
/* Trajectory: traj[chain * traj_cap + t] = score after launch t (t < count). */
int bm_anneal_trajectory(bm_anneal* a, int64_t* traj, int32_t* count);
/* Device recount of every chain's current matching; *mismatches gets the
 * number of chains whose running score differs (must be 0). */
int bm_anneal_verify(bm_anneal* a, int64_t* mismatches);
/* Launches run and their summed duration (HIP events around each launch). */
int bm_anneal_timing(bm_anneal* a, int64_t* launches, double* total_ms);
/* Algorithmic bytes of the chain steps run so far, from the byte model of
 * the swap and triple kernels (48 + 28 D per swap step, 69 + 36 D3 per
 * 3-vertex step, D summed by the chains from the male degrees). */
int bm_anneal_algo_bytes(bm_anneal* a, int64_t* bytes, int64_t* steps);
/* Per chain: 3-vertex moves and moves built on a pool pair, taken so far
 * (either array may be NULL). */
int bm_anneal_move_counts(bm_anneal* a, int64_t* rot3, int64_t* pooled);

/* Proposal pools.  With p_pool > 0, step s of a chain of instance i draws
 * r = bm_rand_at(bm_rand_at(chain seed, 4), s); when lo32(r) < round(p_pool
 * 2^32) and instance i's pool holds P > 0 pairs, the move is built on pool
 * entry (hi32(r) P) >> 32: the swap of that pair, or for a 3-vertex move
 * that pair plus a third vertex uniform over the others (bm_third_vertex in
 * brain-match_amd/csrc/bm_rng.h); every delta stays exact.  A pool can be
 * replaced between runs (bm_anneal_run), with no recapture.
 * bm_anneal_set_pool copies k host pairs (two distinct vertices in 1..n_i
 * each; k = 0 empties the pool); bm_anneal_pool_learned fills it with the
 * `top` best-ranked of `pool` uniform pairs of stream `seed` under a network
 * of instance i's context (as bm_batch_gen_learned, with the features of that
 * context's matching); bm_anneal_get_pool reads it back (uv may be NULL to
 * get only *k).  The reference's annealer draws uniform pairs only
 * (src/best-swap/score-lookup.c:345-400). */
int bm_anneal_set_pool(bm_anneal* a, int32_t inst, const int32_t* uv, int64_t k);
int bm_anneal_pool_learned(bm_anneal* a, int32_t inst, bm_mlp* m, uint64_t seed, int64_t pool, int64_t top);
int bm_anneal_get_pool(bm_anneal* a, int32_t inst, int32_t* uv, int64_t* k);
/* round(p_pool 2^32), the threshold the kernels compare lo32(r) with. */
uint64_t bm_anneal_p_pool(const bm_anneal* a);
/* Chains back to their state at creation. */
int bm_anneal_reset(bm_anneal* a);
/* Launch shape of the chain kernel (waves_per_block, lds_slots, blocks_per_cu
 * as in bm_set_tuning; the default is the first context's swap-kernel shape). */
int bm_anneal_get_tuning(const bm_anneal* a, bm_tuning* t);
int bm_anneal_set_tuning(bm_anneal* a, const bm_tuning* t);
/* Autotuner: times one launch of `steps` steps for each shape of a grid of
 * list entries per lane {2,4} x waves per block {1,2,4,8} x table slots per
 * wave x blocks per CU {4,16}, keeps the fastest and resets the chains.
 * trials (5 doubles per trial: waves, slots, blocks_per_cu, unroll, ms; up to
 * max_trials) may be NULL.  bm_anneal_set_tuning takes unroll 2 or 4. */
int bm_anneal_autotune(bm_anneal* a, int64_t steps, bm_tuning* best, double* trials, int32_t max_trials,
                       int32_t* n_trials);
/* Last error of an annealing object (or of the last failed create). */
const char* bm_anneal_last_error(const bm_anneal* a);
/* The schedule as the kernels use it, for exact replays: beta_s =
 * beta0 * exp(s * lam) with beta0 = 1 / t0 and lam = -ln(decay); a move is a
 * 3-vertex move when 32 random bits are below p_rot3 (round(p * 2^32)). */
int bm_anneal_schedule(const bm_anneal* a, double* beta0, double* lam, uint64_t* p_rot3);

#ifdef __cplusplus
}
#endif

#endif /* BRAINMATCH_H */
