/*
 * Internal interface between the C ABI (bm_api.hip) and the gfx950 kernels
 * (bm_kernels.hip).  Device pointers only; every launcher enqueues on the
 * given stream and returns the launch status.
 */
#ifndef BM_DEVICE_H
#define BM_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

/* Graph and matching as seen by the kernels (passed by value). */
struct BmDevGraph {
    int32_t n;
    const int64_t* m_out_ptr;   /* male CSR row offsets, n+2 */
    const int2* m_out;          /* {dst, w} */
    const int64_t* m_in_ptr;    /* male CSC column offsets, n+2 */
    const int2* m_in;           /* {src, w} */
    const uint64_t* f_out_desc; /* female out-row hash table descriptors, n+1 */
    const int2* f_out_tab;
    const uint64_t* f_in_desc;  /* female in-row hash table descriptors, n+1 */
    const int2* f_in_tab;
    const int64_t* f_out_ptr;   /* female CSR (sorted rows), n+2 */
    const int2* f_out;          /* {dst, w} */
    const int64_t* f_in_ptr;    /* female CSC, n+2 */
    const int2* f_in;           /* {src, w} */
    const int32_t* pi;          /* matching, n+1 */
    int32_t compact;            /* 1: LDS slots are (key << 15) | w (n < 2^17, weights 0..32767) */
};

/* Bytes per LDS table slot for this graph (4 compact, 8 otherwise). */
size_t bm_slot_bytes(const BmDevGraph& g);

/* Placement-gain table (bm_gain.hip). */
struct BmGain {
    int32_t* G;                 /* (n+1) x stride int32, row-major; row 0 / column 0 zero */
    int64_t stride;             /* row stride in int32: n + 1 rounded up to 32 (rows start on 128-byte lines) */
    const uint64_t* m_out_desc; /* male out-row hash tables (adjacency tests) */
    const int2* m_out_tab;
    const uint32_t* m_adj;      /* dense male adjacency, m_adj_words uint32 per row (n+1 rows), or NULL;
                                 * symmetric: bit max(a,b) of row min(a,b) is set when a -> b or b -> a */
    int64_t m_adj_words;
    const int32_t* m_self;      /* male self-loop weight per vertex or BM_ABSENT */
    const int32_t* f_self;      /* female self-loop weight per vertex or BM_ABSENT */
    const uint32_t* f_out_p;    /* female rows packed in row_layout when the graph is compact, else NULL */
    const uint32_t* f_in_p;
    int32_t row_layout;         /* BM_ROWS_*: layout of the rows the table build reads */
    int4* edesc;                /* build work list: one entry per male edge slot (2 x male edges + 1) */
    const int2* echunks;        /* {u, first slot} chunks of at most BM_DESC_CHUNK slots covering every list */
    int64_t n_echunks;
    int32_t build_unroll;       /* build: female-row entries per lane in flight (8, 16 or 32) */
    /* builds in column windows (nwin > 1): for female row r, entry k of
     * r * (nwin + 1) + [0, nwin] is the offset in the row of its first entry
     * with a column >= k * window width (nwin: the row's length); NULL when
     * the table is built in one window */
    const int32_t* f_in_bnd;
    const int32_t* f_out_bnd;
    int32_t nwin;
    /* Dense copies of the longest female rows (compact graphs whose male
     * weights are all >= 0): ndense rows of stride / 2 words and one zero row
     * after them, word i holding the weights of columns 2i (low half) and
     * 2i + 1 (high half), 0 where the row has no entry.  f_out_dense / f_in_dense give a female vertex's
     * row in it, or -1.  A separate pass sums the dense edges of 16 male
     * rows at a time and the row build starts from its result; their
     * work-list entries carry BM_EDESC_DENSE.  NULL / 0 when no row is
     * dense. */
    const uint32_t* dense;
    const int32_t* f_out_dense;
    const int32_t* f_in_dense;
    int32_t ndense;             /* every column's dense-row weights sum to <= 65535: 16-bit sums cannot overflow */
    uint32_t* dense_w;          /* per tile of 16 male rows: ndense x 16 packed weights (rebuilt by every build) */
    int32_t* row_fill;          /* per male row: work-list entries of the current window (n + 1, rebuilt) */
};

/* Work-list entry flags (high word of the start offset). */
#define BM_EDESC_IN 0x80000000u      /* in-edge: expands the female out-row */
#define BM_EDESC_DENSE 0x40000000u   /* dense row: {row in G.dense, flags, 0, w} */
#define BM_DENSE_MAX 128             /* at most this many dense female rows */

/* Row stride of the gain table and its transpose: n + 1 rounded up to a
 * multiple of 32 int32 (128 bytes), so every row starts on a cache line and
 * 16-byte accesses are aligned.  Both tables hold n + 1 rows; the padding
 * columns of G are zero. */
int64_t bm_gain_stride(int32_t n);

/* Female-row layouts read by the table build: {f, w} int2 rows; packed
 * (f << 15) | w (ids < 2^17, weights 0..32767); packed (w << 19) | 4 f (ids
 * < 2^17, weights 0..8191). */
#define BM_DESC_CHUNK 256   /* slots of one vertex's list per chunk of the work-list kernel */
#define BM_ROWS_INT2 0
#define BM_ROWS_P15 1
#define BM_ROWS_P19 2
/* Sets bit b of row a of the zeroed dense adjacency bitmap for every male
 * edge a -> b (words uint32 per row, rows 0..n). */
hipError_t bm_launch_adj_bits(const BmDevGraph& g, uint32_t* bits, int64_t words, int32_t num_cus, hipStream_t s);
/* Packs {f, w} rows into one of the packed layouts. */
hipError_t bm_launch_pack_rows(const int2* in, int64_t m, int layout, uint32_t* out, hipStream_t s);
/* Column-window width of the table build for a row stride. */
int32_t bm_gain_chunk_cols(int64_t stride);
/* Fills the zeroed dense rows: dense row k is female vertex rows[k].x's
 * out-row (rows[k].y = 1) or in-row (0); pool holds ndense x stride uint16. */
hipError_t bm_launch_dense_fill(const int64_t* out_ptr, const int2* out_rows, const int64_t* in_ptr,
                                const int2* in_rows, const int2* rows, int32_t ndense, int64_t stride,
                                uint16_t* pool, hipStream_t s);
/* Window boundaries of every female row for a build in nwin windows of
 * width cols (see BmGain::f_in_bnd): bnd[r * (nwin + 1) + k], rows 0..n. */
hipError_t bm_launch_row_bounds(const int64_t* ptr, const int2* rows, int32_t n, int32_t cols, int32_t nwin,
                                int32_t* bnd, hipStream_t s);

/* Table build: per column chunk, the work list (k_gain_edge_desc) then the
 * rows.  G.edesc must hold 2 x male edges + 1 entries. */
hipError_t bm_launch_gain_build(const BmDevGraph& g, const BmGain& G, int32_t num_cus, hipStream_t s);
/* Table patch after moves (moves[0..*count), *count <= max_moves); plan
 * holds 4 max_moves + 1 int64 of scratch for the work-item offsets; GT (may
 * be NULL) is a transposed copy patched alongside. */
hipError_t bm_launch_gain_update(const BmDevGraph& g, const BmGain& G, int32_t* GT, const int4* moves, const int64_t* count,
                                 int64_t max_moves, int64_t* plan, int32_t num_cus, hipStream_t s);
/* Per-pair table lookups; with pis (the per-vertex words, refreshed) the
 * diagonal entries and self-loops come from them. */
hipError_t bm_launch_gain_swap(const BmDevGraph& g, const BmGain& G, int mode, const int2* cand, const int64_t* off,
                               uint64_t seed, int64_t k, int64_t* delta, const int4* pis, hipStream_t s);
/* u-major evaluation: pis = {pi, G(v, pi(v)), self-loops} per vertex, off =
 * stream offsets.  With bm_gain_umajor_lds(n) <= 160 KiB both rows of a first
 * vertex are staged in LDS (GT = the transposed table, required); otherwise
 * only its bitmaps, bm_gain_umajor_bitmap_lds(n) bytes (GT unused), and the
 * launch fails if those do not fit either. */
size_t bm_gain_umajor_lds(int32_t n);
size_t bm_gain_umajor_bitmap_lds(int32_t n);
hipError_t bm_launch_gain_transpose(const BmGain& G, int32_t nrows, int32_t* GT, int32_t num_cus, hipStream_t s);
hipError_t bm_launch_gain_pis(const BmDevGraph& g, const BmGain& G, int4* pis, hipStream_t s);
/* a-major 3-vertex lookups (best delta and permutation per stream entry);
 * needs bm_gain_rot3_amajor_lds(n) bytes of LDS.  Triples with an edge are
 * queued (queue_entries entries of 3 x int4, split into one region per
 * block; queue_counts holds a count per region, at most queue_regions
 * blocks) and finished by a second kernel; queue = NULL finishes them in
 * place. */
size_t bm_gain_rot3_amajor_lds(int32_t n);
hipError_t bm_launch_gain_rot3_amajor(const BmDevGraph& g, const BmGain& G, const int32_t* GT, const int4* pis,
                                      const int64_t* off, uint64_t seed, int64_t* best, uint8_t* perm, int4* queue,
                                      int64_t queue_entries, int32_t* queue_counts, int32_t queue_regions,
                                      int32_t num_cus, hipStream_t s);
hipError_t bm_launch_gain_swap_umajor(const BmDevGraph& g, const BmGain& G, const int32_t* GT, const int4* pis,
                                      const int64_t* off, uint64_t seed, int64_t* delta, int32_t num_cus,
                                      hipStream_t s);
/* Exact best swap over all pairs u < v (larger delta, then smaller u, then
 * smaller v) into out[0] = {delta lo, delta hi, u, v}; u = INT32_MAX when
 * n < 2.  part holds max_parts per-block results. */
constexpr int32_t kBestSwapParts = 2048;
hipError_t bm_launch_gain_best_swap(const BmDevGraph& g, const BmGain& G, const int32_t* GT, const int4* pis,
                                    int4* part, int32_t max_parts, int4* out, int32_t num_cus, hipStream_t s);
/* Best partner of each u in u0 .. u0 + count - 1 over v = u+1..n: out[i] =
 * {delta lo, delta hi, u0 + i, v} (ties to the smaller v; delta INT64_MIN and
 * v INT32_MAX for u = n). */
hipError_t bm_launch_gain_best_partners(const BmDevGraph& g, const BmGain& G, const int32_t* GT, const int4* pis,
                                        int32_t u0, int32_t count, int4* out, int32_t num_cus, hipStream_t s);
hipError_t bm_launch_gain_rot3(const BmDevGraph& g, const BmGain& G, const int4* cand, int64_t k, int64_t* best,
                               uint8_t* perm, int4* cand_out_perm, const int4* pis, hipStream_t s);
hipError_t bm_launch_moves_from_swap(const int32_t* pi_before, int32_t u, int32_t v, int4* moves, int64_t* count,
                                     hipStream_t s);
hipError_t bm_launch_moves_from_perm3(const int32_t* pi_before, int32_t a, int32_t b, int32_t c, int32_t p,
                                      int4* moves, int64_t* count, hipStream_t s);

struct BmLaunch {
    int32_t waves_per_block;   /* 1, 2, 4 or 8 */
    int32_t lds_slots;         /* table slots per wave (8 bytes each) */
    int32_t blocks_per_cu;     /* persistent grid: blocks per CU (capped by occupancy) */
    int32_t num_cus;
    int32_t unroll;            /* list entries per lane in flight (2, 4 or 8) */
};

/* Candidate sources. */
enum BmCandMode { BM_CAND_LIST = 0, BM_CAND_UNIFORM = 1, BM_CAND_UMAJOR = 2 };

hipError_t bm_launch_swap(const BmDevGraph& g, const BmLaunch& L, int mode, const int2* cand,
                          const int64_t* off, uint64_t seed, int64_t k, int64_t* delta, hipStream_t s);

/* u-major stream (bm_umajor_first / bm_umajor_partner): counts (n+1 int32)
 * and offsets off[0..n+1] (int64) of k draws; pairs written as a list. */
hipError_t bm_launch_umajor_offsets(int32_t n, uint64_t seed, int64_t k, int32_t* counts, int64_t* off,
                                    int32_t num_cus, hipStream_t s);
hipError_t bm_launch_gen_umajor(int32_t n, uint64_t seed, const int64_t* off, int64_t k, int2* cand, hipStream_t s);
/* a-major triple stream entries 0..k-1 as {a, b, c, 0} (off from the
 * first-vertex draws of bm_amajor_seed(seed)). */
hipError_t bm_launch_gen_amajor(int32_t n, uint64_t seed, const int64_t* off, int64_t k, int4* cand, hipStream_t s);
hipError_t bm_launch_rot3(const BmDevGraph& g, const BmLaunch& L, const int4* cand, int64_t k,
                          int64_t* best, uint8_t* perm, int4* cand_out_perm, hipStream_t s);
hipError_t bm_launch_score(const BmDevGraph& g, const BmLaunch& L, int64_t* acc, hipStream_t s);

/* Annealing chains (k_anneal in bm_delta.hip): per-chain state in device memory. */
struct BmChain {
    int64_t score;      /* score of the current matching */
    int64_t best;       /* best score reached */
    int64_t step;       /* steps taken */
    int64_t accepted;   /* accepted moves */
    int64_t pi_off;     /* current matching at pool + pi_off, best-so-far copy right after it */
    uint64_t seed;      /* stream seed of the chain */
    int32_t inst;       /* graph (instance) index */
    int32_t at_best;    /* 1: the current matching is the best one (the copy is stale) */
    int32_t nseg;       /* launches run (trajectory entries) */
    int32_t pad;
    int64_t deg_swap;   /* sum over the swap moves taken of D(u, v) (byte model) */
    int64_t deg_rot3;   /* sum over the 3-vertex moves taken of D3 */
    int64_t n_rot3;     /* 3-vertex moves taken */
    int64_t n_learned;  /* moves drawn from the instance's proposal pool */
};

/* Proposal pool of one instance: the chains of that instance draw a fraction
 * of their moves from these pairs (see k_anneal). */
struct BmProposalPool {
    const int2* pairs;  /* count pairs u != v in 1..n */
    int64_t count;      /* 0: no pool */
};

struct BmAnnealSched {
    double beta0;       /* inverse initial temperature */
    double lam;         /* beta_s = beta0 * exp(s * lam) */
    uint64_t p_rot3;    /* probability of a 3-vertex move, in units of 2^-32 (0..2^32) */
    int32_t traj_cap;   /* trajectory entries kept per chain */
    int32_t pad;
    uint64_t p_pool;    /* probability that a move is drawn from the proposal pool, units of 2^-32 */
};

/* Persistent grid for n_chains chains (one wave each); computed once. */
hipError_t bm_anneal_grid(int compact, int rot, const BmLaunch& L, int64_t n_chains, dim3* grid, dim3* block, size_t* lds);
/* Advances every chain by `steps` steps.  insts: device array of graphs. */
hipError_t bm_launch_anneal(const BmDevGraph* insts, int compact, BmChain* chains, int64_t n_chains,
                            int32_t* pool, int64_t steps, const BmAnnealSched& sc, int64_t* traj,
                            const BmProposalPool* proposals, const BmLaunch& L, dim3 grid, dim3 block, size_t lds,
                            hipStream_t s);

hipError_t bm_launch_gen_pairs(int32_t n, uint64_t seed, int64_t k, int2* cand, hipStream_t s);
hipError_t bm_launch_gen_triples(int32_t n, uint64_t seed, int64_t k, int4* cand, hipStream_t s);
/* reorder: c changed ids ascending, then the rest ascending (n entries). */
hipError_t bm_launch_gen_changed_pairs(int32_t n, int32_t c, const int32_t* reorder, int64_t P,
                                       int2* cand, hipStream_t s);
hipError_t bm_launch_gen_changed_triples(int32_t n, int32_t c, const int32_t* reorder, int64_t T,
                                         int4* cand, hipStream_t s);

/* Validate candidate ids on the device: sets *bad (int32, zeroed by the
 * caller) to 1 if any id is outside 1..n. */
hipError_t bm_launch_check_pairs(int32_t n, const int2* cand, int64_t k, int32_t* bad, hipStream_t s);
hipError_t bm_launch_check_triples(int32_t n, const int4* cand, int64_t k, int32_t* bad, hipStream_t s);

/* Reductions over int64 deltas.  part must hold bm_stats_parts(k) * 6 int64. */
int64_t bm_stats_blocks(int64_t k);
hipError_t bm_launch_stats(const int64_t* d, int64_t k, int64_t* part, int64_t* out6, hipStream_t s);
/* rank-th index with delta > 0 (kind 1) or == 0 (kind 0); out = -1 if none.
 * cnt must hold bm_stats_blocks(k) int64. */
hipError_t bm_launch_pick(const int64_t* d, int64_t k, int32_t kind, int64_t rank, int64_t* cnt,
                          int64_t* out, hipStream_t s);

/* Disjoint acceptance.  Scratch sizes come from bm_accept_scratch_bytes. */
size_t bm_accept_scratch_bytes(int64_t k);
hipError_t bm_launch_accept(const BmDevGraph& g, int32_t* pi_rw, const int2* cand, const int64_t* d,
                            int64_t k, int64_t max_accept, void* scratch, size_t scratch_bytes,
                            int64_t* out_acc_idx /* k */, int64_t* out_meta /* [n_acc, gain, n_moves] */,
                            int4* out_moves /* 2k {x, old, new, 0}, may be NULL */, hipStream_t s);

/* 2-hop proposals (bm_twohop.hip).  K > 0: pairs[a*K + j] for anchor a
 * (anchors[a], or drawn from `seed` when anchors is NULL); K == 0: the whole
 * set of anchors[a] into out_list (n entries) and its size into out_count[a]
 * (n_anchors must be 1 then). */
size_t bm_twohop_lds_bytes(int32_t n);
hipError_t bm_launch_twohop(const BmDevGraph& g, const int32_t* anchors, uint64_t seed, int64_t n_anchors,
                            int32_t K, int2* pairs, int32_t* out_list, int32_t* out_count, int32_t num_cus,
                            hipStream_t s);

/* Learned proposal (bm_mlp.hip).  params/grad/adam_m/adam_v: fp32 in the
 * layout W1 [16][128], b1 [128], W2 [128][128], b2 [128], w3 [128], b3;
 * W1f/W2f/W2Tf: bf16 weights in MFMA fragment order (4 KB / 32 KB / 32 KB). */
struct BmMlpDev {
    float* params;
    float* grad;
    float* adam_m;
    float* adam_v;
    void* W1f;
    void* W2f;
    void* W2Tf;
};
int32_t bm_mlp_param_count(void);
/* Per-vertex features from edge chunks {x, out (1) or in (0), first, count}
 * of the male lists (at most bm_feature_chunk_edges() edges each); acc
 * holds (n+1) x 32 zeroed bytes of accumulators and is left zeroed. */
int32_t bm_feature_chunk_edges(void);
hipError_t bm_launch_vertex_features(const BmDevGraph& g, const int4* chunks, int64_t nchunks, void* acc, float* vf,
                                     int32_t num_cus, hipStream_t s);
hipError_t bm_launch_mlp_pack(const BmMlpDev& d, hipStream_t s);
/* Scores of k pairs (list, or uniform stream `seed` when list is NULL). */
hipError_t bm_launch_mlp_forward(const BmMlpDev& d, const float* vf, const int2* list, uint64_t seed, int32_t n,
                                 int64_t k, float* score, int32_t num_cus, hipStream_t s);
/* Gradient of the mean squared error over k pairs, added into d.grad; the
 * summed squared error is added into *loss.  act: bm_mlp_train_bytes(k). */
size_t bm_mlp_train_bytes(int64_t k);
hipError_t bm_launch_mlp_gradient(const BmMlpDev& d, const float* vf, const int2* list, uint64_t seed, int32_t n,
                                  int64_t k, const float* target, void* act, float* loss, int32_t num_cus,
                                  hipStream_t s);
hipError_t bm_launch_mlp_adam(const BmMlpDev& d, float lr, int64_t step, hipStream_t s);
hipError_t bm_launch_mlp_targets(const int64_t* delta, int64_t k, float* t, hipStream_t s);
/* Top `top` of `pool` uniform-stream pairs by score, as a pair list. */
size_t bm_mlp_rank_scratch(int64_t pool);
/* Scratch bytes of the select-only ranking of a pool (bm_launch_mlp_rank_select)
 * and the largest top it sorts. */
size_t bm_mlp_select_bytes(int64_t pool);
int64_t bm_mlp_select_max_top(void);
hipError_t bm_launch_mlp_rank_select(const BmMlpDev& d, const float* vf, uint64_t seed, int32_t n, int64_t pool,
                                     int64_t top, void* scratch, size_t scratch_bytes, int2* out, int32_t num_cus,
                                     hipStream_t s);
hipError_t bm_launch_mlp_rank(const BmMlpDev& d, const float* vf, uint64_t seed, int32_t n, int64_t pool, int64_t top,
                              void* scratch, size_t scratch_bytes, int2* out, int32_t num_cus, hipStream_t s);

/* Context internals for other host translation units (bm_api.hip). */
struct bm_ctx;
BmDevGraph bm_ctx_devgraph(const bm_ctx* c);
int bm_ctx_device(const bm_ctx* c);
int bm_ctx_num_cus(const bm_ctx* c);
hipStream_t bm_ctx_stream(const bm_ctx* c);
BmLaunch bm_ctx_launch(const bm_ctx* c, int kernel);
int bm_ctx_set_error(bm_ctx* c, int code, const char* msg);
struct bm_batch;
int2* bm_batch_pairs_dev(bm_batch* b);
int64_t* bm_batch_deltas_dev(bm_batch* b);
int64_t bm_batch_capacity(const bm_batch* b);
/* The context a network was created for. */
typedef struct bm_mlp bm_mlp;
bm_ctx* bm_mlp_context(const bm_mlp* m);
/* The `top` best-scored of `pool` uniform pairs of stream `seed` (as
 * bm_batch_gen_learned) into the device array out; synchronises the
 * context stream before returning. */
int bm_mlp_rank_into(bm_mlp* m, uint64_t seed, int64_t pool, int64_t top, int2* out);
/* Nonzero when the batch holds pairs whose exact deltas have been evaluated. */
int bm_batch_has_pair_deltas(const bm_batch* b);
/* The batch now holds k device-generated pairs (not yet evaluated). */
void bm_batch_mark_pairs(bm_batch* b, int64_t k);

hipError_t bm_launch_apply_swap(int32_t* pi, int32_t u, int32_t v, hipStream_t s);
hipError_t bm_launch_apply_perm3(int32_t* pi, int32_t a, int32_t b, int32_t c, int32_t p, hipStream_t s);

#endif
