/*
 * C ABI implementation (include/brainmatch.h): contexts, batches, launches,
 * timing.  Host logic only; the kernels are in bm_kernels.hip.
 */
#include "bm_device.h"
#include "bm_host.h"
#include "bm_rng.h"

#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

extern "C" const char* bm_global_error(void);

struct KTiming {
    int kernel;
    hipEvent_t a, b;
};

struct bm_ctx {
    int32_t n;
    int device;
    int num_cus;
    char arch[64];
    hipStream_t stream;
    bm_graph_host male;     /* host copy (degrees, neighbourhoods) */
    int64_t female_m;
    int32_t* f_out_deg;     /* female degrees, n+1 */
    int32_t* f_in_deg;
    /* device buffers */
    int64_t* d_m_out_ptr;
    int2* d_m_out;
    int64_t* d_m_in_ptr;
    int2* d_m_in;
    uint64_t* d_f_out_desc;
    int2* d_f_out_tab;
    uint64_t* d_f_in_desc;
    int2* d_f_in_tab;
    int64_t* d_f_out_ptr;   /* female CSR / CSC rows */
    int2* d_f_out;
    int64_t* d_f_in_ptr;
    int2* d_f_in;
    int32_t compact;
    int32_t row_layout;     /* BM_ROWS_P19 when female weights are in 0..8191, else BM_ROWS_P15 (compact graphs) */
    int32_t* d_pi;
    int64_t* d_small;       /* 16 int64 of scratch for scores / flags */
    int32_t* d_reorder;     /* n ints for the changed-first enumerations */
    BmLaunch tune[BM_KERNEL_COUNT];
    int32_t user_tuned[BM_KERNEL_COUNT];   /* bm_set_tuning was called: no batch-size choice */
    int32_t group_lanes[BM_KERNEL_COUNT];
    int timing_on;
    std::vector<KTiming>* pending;
    int64_t launches[BM_KERNEL_COUNT];
    double total_ms[BM_KERNEL_COUNT];
    bm_batch* work;         /* internal batch for the host-buffer calls */
    /* placement-gain table (bm_gain.hip) */
    int gain_on;
    int32_t* h_m_self;      /* host: self-loop weights or INT32_MIN */
    int32_t* h_f_self;
    int32_t max_abs_fw;     /* largest |female weight| */
    int32_t* d_G;
    uint64_t* d_m_out_desc;
    int2* d_m_out_tab;
    int32_t* d_m_self;
    int32_t* d_f_self;
    int4* d_moves;
    int64_t moves_cap;
    int64_t* d_patch_plan;  /* 4 moves_cap + 1 work-item offsets of a table patch */
    int64_t* d_nmoves;
    int64_t gain_version;   /* bumped by every build and patch of the table */
    /* u-major evaluation (bm_batch_eval_swaps_umajor) */
    int32_t* d_um_counts;   /* n+1 */
    int64_t* d_um_off;      /* n+2 */
    int32_t* d_GT;          /* transposed table, valid when gt_version == gain_version */
    int4* d_pis;            /* per vertex v: {pi(v), G(v, pi(v)), male self-loop weight of v, female
                               self-loop weight of pi(v)}, BM_ABSENT for a missing self-loop;
                               rewritten before every staged lookup */
    uint32_t* d_f_out_p;    /* packed female rows for the table build (compact graphs) */
    uint32_t* d_f_in_p;
    int4* d_edesc;          /* table-build work list, 2 x male edges entries */
    int2* d_echunks;        /* {u, first slot} chunks of the work list (fixed by the male graph) */
    uint32_t* d_m_adj;      /* dense male adjacency bitmap (NULL when it would exceed 2 GiB) */
    int64_t m_adj_words;
    int32_t* d_f_in_bnd;    /* column-window boundaries of the female rows (windowed builds only) */
    int32_t* d_f_out_bnd;
    int32_t nwin;
    uint32_t* d_dense;      /* dense copies of the longest female rows (BmGain::dense) */
    uint32_t* d_dense_w;    /* dense tile weights (BmGain::dense_w) */
    int32_t* d_row_fill;    /* work-list entries per male row (BmGain::row_fill) */
    int32_t* d_f_out_dense; /* female vertex -> its dense out-row, or -1 */
    int32_t* d_f_in_dense;
    int32_t ndense;
    int32_t male_nonneg;    /* every male weight is >= 0 */
    int32_t pi_distinct;    /* no two vertices share a nonzero image (moves keep this; bm_set_mapping checks it) */
    int64_t n_echunks;
    int64_t gt_version;
    int4* d_best;           /* best-swap result, then kBestSwapParts per-block results */
    int4* d_partners;       /* best partner per vertex of a bm_best_partners range */
    int4* d_r3q;            /* a-major 3-vertex lookups: queued triples with an edge (3 x int4 each) */
    int64_t r3q_entries;
    int32_t* d_r3q_counts;  /* per queue region (kR3QueueRegions) */
    int32_t partners_cap;
    char err[512];
};

struct bm_batch {
    bm_ctx* ctx;
    int64_t cap;
    int64_t k;
    int kind;               /* 0 empty, 1 pairs, 2 triples */
    int generated;          /* candidates came from a device generator */
    int evaluated;
    void* d_cand;           /* cap * 16 bytes (int2 pairs or int4 triples) */
    int64_t* d_delta;
    uint8_t* d_perm;
    int64_t* d_part;
    int64_t* d_out;         /* 16 int64 */
    void* d_acc_scratch;
    size_t acc_scratch_bytes;
    int64_t* d_acc_idx;
};

/* ------------------------------------------------------------------ */

static int fail(bm_ctx* c, int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    if (c)
        vsnprintf(c->err, sizeof c->err, fmt, ap);
    else {
        char buf[512];
        vsnprintf(buf, sizeof buf, fmt, ap);
        bm_set_global_error("%s", buf);
    }
    va_end(ap);
    return code;
}

#define HIPCHK(ctx, expr)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail((ctx), BM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                           \
    } while (0)

static BmDevGraph devgraph(const bm_ctx* c) {
    BmDevGraph g;
    g.n = c->n;
    g.m_out_ptr = c->d_m_out_ptr;
    g.m_out = c->d_m_out;
    g.m_in_ptr = c->d_m_in_ptr;
    g.m_in = c->d_m_in;
    g.f_out_desc = c->d_f_out_desc;
    g.f_out_tab = c->d_f_out_tab;
    g.f_in_desc = c->d_f_in_desc;
    g.f_in_tab = c->d_f_in_tab;
    g.f_out_ptr = c->d_f_out_ptr;
    g.f_out = c->d_f_out;
    g.f_in_ptr = c->d_f_in_ptr;
    g.f_in = c->d_f_in;
    g.pi = c->d_pi;
    g.compact = c->compact;
    return g;
}

static void timing_begin(bm_ctx* c, int kernel, KTiming* t) {
    t->kernel = -1;
    if (!c->timing_on) return;
    if (hipEventCreate(&t->a) != hipSuccess) return;
    if (hipEventCreate(&t->b) != hipSuccess) { (void)hipEventDestroy(t->a); return; }
    t->kernel = kernel;
    (void)hipEventRecord(t->a, c->stream);
}

static void timing_end(bm_ctx* c, KTiming* t) {
    if (t->kernel < 0) return;
    (void)hipEventRecord(t->b, c->stream);
    c->pending->push_back(*t);
}

static int timing_resolve(bm_ctx* c) {
    if (c->pending->empty()) return BM_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (KTiming& t : *c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            c->launches[t.kernel]++;
            c->total_ms[t.kernel] += ms;
        }
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    c->pending->clear();
    return BM_OK;
}

template <typename T>
static int dalloc(bm_ctx* c, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) return fail(c, BM_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", count * sizeof(T), hipGetErrorString(e));
    return BM_OK;
}

template <typename T>
static int h2d(bm_ctx* c, T* d, const void* h, size_t count) {
    if (count == 0) return BM_OK;
    HIPCHK(c, hipMemcpyAsync(d, h, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return BM_OK;
}

/* ------------------------------------------------------------------ */

extern "C" int bm_api_version(void) { return BM_API_VERSION; }

extern "C" const char* bm_last_error(const bm_ctx* ctx) { return ctx ? ctx->err : bm_global_error(); }

extern "C" int32_t bm_num_vertices(const bm_ctx* ctx) { return ctx ? ctx->n : 0; }

extern "C" int64_t bm_num_edges(const bm_ctx* ctx, int female) {
    if (!ctx) return 0;
    return female ? ctx->female_m : ctx->male.m;
}

static void free_ctx(bm_ctx* c) {
    if (!c) return;
    if (c->work) bm_batch_destroy(c->work);
    (void)hipFree(c->d_m_out_ptr);
    (void)hipFree(c->d_m_out);
    (void)hipFree(c->d_m_in_ptr);
    (void)hipFree(c->d_m_in);
    (void)hipFree(c->d_f_out_desc);
    (void)hipFree(c->d_f_out_tab);
    (void)hipFree(c->d_f_in_desc);
    (void)hipFree(c->d_f_in_tab);
    (void)hipFree(c->d_f_out_ptr);
    (void)hipFree(c->d_f_out);
    (void)hipFree(c->d_f_in_ptr);
    (void)hipFree(c->d_f_in);
    (void)hipFree(c->d_pi);
    (void)hipFree(c->d_small);
    (void)hipFree(c->d_reorder);
    if (c->pending) {
        for (KTiming& t : *c->pending) {
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
        delete c->pending;
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    (void)hipFree(c->d_G);
    (void)hipFree(c->d_m_out_desc);
    (void)hipFree(c->d_m_out_tab);
    (void)hipFree(c->d_m_self);
    (void)hipFree(c->d_f_self);
    (void)hipFree(c->d_moves);
    (void)hipFree(c->d_patch_plan);
    (void)hipFree(c->d_nmoves);
    (void)hipFree(c->d_um_counts);
    (void)hipFree(c->d_um_off);
    (void)hipFree(c->d_GT);
    (void)hipFree(c->d_pis);
    (void)hipFree(c->d_r3q);
    (void)hipFree(c->d_r3q_counts);
    (void)hipFree(c->d_f_out_p);
    (void)hipFree(c->d_f_in_p);
    (void)hipFree(c->d_edesc);
    (void)hipFree(c->d_echunks);
    (void)hipFree(c->d_f_in_bnd);
    (void)hipFree(c->d_f_out_bnd);
    (void)hipFree(c->d_dense);
    (void)hipFree(c->d_dense_w);
    (void)hipFree(c->d_row_fill);
    (void)hipFree(c->d_f_out_dense);
    (void)hipFree(c->d_f_in_dense);
    (void)hipFree(c->d_m_adj);
    (void)hipFree(c->d_best);
    (void)hipFree(c->d_partners);
    bm_graph_free(&c->male);
    free(c->f_out_deg);
    free(c->f_in_deg);
    free(c->h_m_self);
    free(c->h_f_self);
    free(c);
}

static BmGain devgain(const bm_ctx* c) {
    BmGain G;
    G.G = c->d_G;
    G.stride = bm_gain_stride(c->n);
    G.m_out_desc = c->d_m_out_desc;
    G.m_out_tab = c->d_m_out_tab;
    G.m_self = c->d_m_self;
    G.f_self = c->d_f_self;
    G.f_out_p = c->d_f_out_p;
    G.f_in_p = c->d_f_in_p;
    G.row_layout = c->d_f_out_p ? c->row_layout : BM_ROWS_INT2;
    G.edesc = c->d_edesc;
    G.echunks = c->d_echunks;
    G.n_echunks = c->n_echunks;
    G.f_in_bnd = c->d_f_in_bnd;
    G.f_out_bnd = c->d_f_out_bnd;
    G.nwin = c->nwin;
    G.m_adj = c->d_m_adj;
    G.m_adj_words = c->m_adj_words;
    G.build_unroll = c->tune[BM_KERNEL_GAIN_BUILD].unroll;
    /* a male row lists each dense female row at most once per side only
     * while images are distinct, which sizes the build's dense-edge list */
    const bool dense = c->ndense > 0 && c->pi_distinct;
    G.dense = dense ? c->d_dense : nullptr;
    G.f_out_dense = c->d_f_out_dense;
    G.f_in_dense = c->d_f_in_dense;
    G.ndense = dense ? c->ndense : 0;
    G.dense_w = c->d_dense_w;
    G.row_fill = c->d_row_fill;
    return G;
}

static int gain_rebuild(bm_ctx* c) {
    KTiming t;
    timing_begin(c, BM_KERNEL_GAIN_BUILD, &t);
    HIPCHK(c, bm_launch_gain_build(devgraph(c), devgain(c), c->num_cus, c->stream));
    timing_end(c, &t);
    c->gain_version++;
    return BM_OK;
}

/* Grow the device move list to hold `need` moves. */
static int gain_moves_cap(bm_ctx* c, int64_t need) {
    if (c->moves_cap >= need) return BM_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->d_moves);
    (void)hipFree(c->d_patch_plan);
    c->d_moves = nullptr;
    c->d_patch_plan = nullptr;
    c->moves_cap = 0;
    HIPCHK(c, hipMalloc((void**)&c->d_moves, (size_t)need * sizeof(int4)));
    HIPCHK(c, hipMalloc((void**)&c->d_patch_plan, (size_t)(4 * need + 1) * sizeof(int64_t)));
    c->moves_cap = need;
    return BM_OK;
}

/* Male degree (out + in) of vertex x. */
static int64_t male_degree(const bm_ctx* c, int32_t x) {
    return (c->male.out.ptr[x + 1] - c->male.out.ptr[x]) + (c->male.in.ptr[x + 1] - c->male.in.ptr[x]);
}

/* Whether a patch of moves of the vertices xs (k of them) should also patch
 * the transposed copy: its additions land in as many rows as the moved
 * vertices have male edges, one 4-byte atomic per row and female entry, so it
 * pays against a 0.54 ms transpose (n = 18,524) only for moves of vertices of
 * ordinary degree (together at most 8x the mean); hub moves leave the copy to
 * be rebuilt. */
static bool gt_patch_pays(const bm_ctx* c, const int32_t* xs, int k) {
    int64_t deg = 0;
    for (int i = 0; i < k; i++) deg += male_degree(c, xs[i]);
    const int64_t mean2 = c->n > 0 ? 2 * c->male.m / c->n : 0;
    return deg <= 8 * (mean2 > 1 ? mean2 : 1);
}

static int gain_patch(bm_ctx* c, const int64_t* d_count, int64_t max_moves, bool gt_allowed) {
    KTiming t;
    timing_begin(c, BM_KERNEL_GAIN_UPDATE, &t);
    if (max_moves > c->moves_cap) return fail(c, BM_ERR_STATE, "gain patch: %lld moves exceed the move list",
                                              (long long)max_moves);
    /* a transposed copy that is current is patched alongside (the same
     * additions at transposed positions) instead of being rebuilt later when
     * the caller found that cheaper (gt_patch_pays) */
    const bool gt = gt_allowed && c->d_GT && c->gt_version == c->gain_version;
    HIPCHK(c, bm_launch_gain_update(devgraph(c), devgain(c), gt ? c->d_GT : nullptr, c->d_moves, d_count, max_moves,
                                    c->d_patch_plan, c->num_cus, c->stream));
    timing_end(c, &t);
    c->gain_version++;
    if (gt) c->gt_version = c->gain_version;
    return BM_OK;
}

extern "C" void bm_destroy(bm_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_ctx(ctx);
}

/* Launch shape of the scan kernels for a batch of k candidates.  Unless the
 * caller set one (bm_set_tuning), a batch too small to fill the device's wave
 * slots (k < 32,768: at most four candidates per slot) runs in 4-wave blocks
 * with 8 list entries per lane -- there the slowest candidate's latency sets
 * the time (a climb round of 4,096 pairs: 1.0 vs 1.15 ms) -- and a larger
 * one in the default throughput shape. */
static BmLaunch scan_launch(const bm_ctx* c, int kernel, int64_t k);

static void default_tuning(bm_ctx* c) {
    const int cus = c->num_cus > 0 ? c->num_cus : 256;
    /* slots per wave; 32 waves per CU fit 160 KiB when slots * slot bytes <= 5 KiB */
    const int s = c->compact ? 1280 : 640;
    /* swap and 3-vertex kernels: the shape for large batches (see
     * scan_launch for small ones) -- 8-wave blocks, 2 list entries per lane:
     * occupancy over per-lane depth, 7.1e7 vs 4.9e7 pairs/s (4-wave blocks, 8
     * entries) at n = 18,524; 148.8 vs 156.1 ms per 2^22 triples */
    c->tune[BM_KERNEL_SWAP] = BmLaunch{8, s, 8, cus, 2};
    c->tune[BM_KERNEL_ROT3] = BmLaunch{8, 2 * s, 8, cus, 2};
    c->tune[BM_KERNEL_SCORE] = BmLaunch{4, s, 8, cus, 8};
    c->tune[BM_KERNEL_GAIN_BUILD] = BmLaunch{4, s, 8, cus, 4};   /* only unroll is used */
    for (int i = 0; i < BM_KERNEL_COUNT; i++) {
        c->group_lanes[i] = 64;
        c->user_tuned[i] = 0;
    }
}

static BmLaunch scan_launch(const bm_ctx* c, int kernel, int64_t k) {
    BmLaunch L = c->tune[kernel];
    if (c->user_tuned[kernel] || k >= 32768) return L;
    L.waves_per_block = 4;
    L.lds_slots = c->compact ? 1280 : 640;
    L.blocks_per_cu = 8;
    L.unroll = kernel == BM_KERNEL_ROT3 ? 4 : 8;
    return L;
}

extern "C" int bm_create(bm_ctx** out, int32_t n, const bm_edges* male, const bm_edges* female, int32_t device) {
    if (!out) return fail(nullptr, BM_ERR_ARG, "bm_create: out is NULL");
    *out = nullptr;
    if (n < 2) return fail(nullptr, BM_ERR_ARG, "bm_create: n must be >= 2 (got %d)", n);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(nullptr, BM_ERR_NODEV, "bm_create: no HIP device available");
    if (device < 0 || device >= ndev) return fail(nullptr, BM_ERR_ARG, "bm_create: device %d out of range", device);
    bm_ctx* c = (bm_ctx*)calloc(1, sizeof(bm_ctx));
    if (!c) return fail(nullptr, BM_ERR_NOMEM, "bm_create: out of host memory");
    c->n = n;
    c->device = device;
    c->pending = new std::vector<KTiming>();
    hipError_t he = hipSetDevice(device);
    if (he != hipSuccess) { free_ctx(c); return fail(nullptr, BM_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(he)); }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->num_cus = prop.multiProcessorCount;
        snprintf(c->arch, sizeof c->arch, "%s", prop.gcnArchName);
    }
    he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) { free_ctx(c); return fail(nullptr, BM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(he)); }

    const char* err = nullptr;
    int rc = bm_graph_build(n, male, &c->male, &err);
    if (rc != BM_OK) { free_ctx(c); return fail(nullptr, rc, "bm_create: male graph: %s", err); }
    bm_graph_host fg;
    rc = bm_graph_build(n, female, &fg, &err);
    if (rc != BM_OK) { free_ctx(c); return fail(nullptr, rc, "bm_create: female graph: %s", err); }
    c->female_m = fg.m;
    c->f_out_deg = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    c->f_in_deg = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    bm_rowhash_host ho, hi;
    memset(&ho, 0, sizeof ho);
    memset(&hi, 0, sizeof hi);
    if (!c->f_out_deg || !c->f_in_deg || bm_rowhash_build(&fg.out, &ho) != BM_OK ||
        bm_rowhash_build(&fg.in, &hi) != BM_OK) {
        bm_graph_free(&fg);
        bm_rowhash_free(&ho);
        bm_rowhash_free(&hi);
        free_ctx(c);
        return fail(nullptr, BM_ERR_NOMEM, "bm_create: out of host memory building female tables");
    }
    /* Compact LDS slots need keys below 2^17 and female weights in 0..32767;
     * compact kernels also sum a lane's contributions in 32 bits, which needs
     * male weights within +-32767 (every contribution is then within
     * +-32767, and a lane takes at most 2^12 + 8 edges of a list shorter
     * than 2^18). */
    c->compact = n < (1 << 17);
    for (int64_t e = 0; e < c->male.m && c->compact; e++)
        if (c->male.out.adj[e].y < -32767 || c->male.out.adj[e].y > 32767) c->compact = 0;
    c->male_nonneg = 1;
    for (int64_t e = 0; e < c->male.m; e++)
        if (c->male.out.adj[e].y < 0) c->male_nonneg = 0;
    for (int32_t r = 1; r <= n; r++) {
        c->f_out_deg[r] = (int32_t)(fg.out.ptr[r + 1] - fg.out.ptr[r]);
        c->f_in_deg[r] = (int32_t)(fg.in.ptr[r + 1] - fg.in.ptr[r]);
    }
    for (int64_t e = 0; e < fg.m && c->compact; e++)
        if (fg.out.adj[e].y < 0 || fg.out.adj[e].y > 32767) c->compact = 0;
    default_tuning(c);
    /* self-loop weights of both graphs and the largest |female weight| (gain table) */
    c->h_m_self = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    c->h_f_self = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    if (!c->h_m_self || !c->h_f_self) {
        bm_rowhash_free(&ho);
        bm_rowhash_free(&hi);
        bm_graph_free(&fg);
        free_ctx(c);
        return fail(nullptr, BM_ERR_NOMEM, "bm_create: out of host memory");
    }
    for (int32_t r = 0; r <= n; r++) c->h_m_self[r] = c->h_f_self[r] = INT32_MIN;
    for (int32_t r = 1; r <= n; r++) {
        for (int64_t e = c->male.out.ptr[r]; e < c->male.out.ptr[r + 1]; e++)
            if (c->male.out.adj[e].x == r) c->h_m_self[r] = c->male.out.adj[e].y;
        for (int64_t e = fg.out.ptr[r]; e < fg.out.ptr[r + 1]; e++)
            if (fg.out.adj[e].x == r) c->h_f_self[r] = fg.out.adj[e].y;
    }
    c->max_abs_fw = 0;
    for (int64_t e = 0; e < fg.m; e++) {
        const int32_t a = fg.out.adj[e].y < 0 ? -fg.out.adj[e].y : fg.out.adj[e].y;
        if (a > c->max_abs_fw) c->max_abs_fw = a;
    }

    const bm_graph_host& mg = c->male;
    rc = BM_OK;
    if (rc == BM_OK) rc = dalloc(c, &c->d_m_out_ptr, (size_t)n + 2);
    if (rc == BM_OK) rc = dalloc(c, &c->d_m_out, (size_t)mg.m);
    if (rc == BM_OK) rc = dalloc(c, &c->d_m_in_ptr, (size_t)n + 2);
    if (rc == BM_OK) rc = dalloc(c, &c->d_m_in, (size_t)mg.m);
    if (rc == BM_OK) rc = dalloc(c, &c->d_f_out_desc, (size_t)n + 1);
    if (rc == BM_OK) rc = dalloc(c, &c->d_f_out_tab, (size_t)(ho.slots ? ho.slots : 2));
    if (rc == BM_OK) rc = dalloc(c, &c->d_f_in_desc, (size_t)n + 1);
    if (rc == BM_OK) rc = dalloc(c, &c->d_f_in_tab, (size_t)(hi.slots ? hi.slots : 2));
    if (rc == BM_OK) rc = dalloc(c, &c->d_f_out_ptr, (size_t)n + 2);
    if (rc == BM_OK) rc = dalloc(c, &c->d_f_out, (size_t)fg.m);
    if (rc == BM_OK) rc = dalloc(c, &c->d_f_in_ptr, (size_t)n + 2);
    if (rc == BM_OK) rc = dalloc(c, &c->d_f_in, (size_t)fg.m);
    if (rc == BM_OK) rc = dalloc(c, &c->d_pi, (size_t)n + 1);
    if (rc == BM_OK) rc = dalloc(c, &c->d_small, 16);
    if (rc == BM_OK) rc = dalloc(c, &c->d_reorder, (size_t)n);
    if (rc == BM_OK) rc = h2d(c, c->d_m_out_ptr, mg.out.ptr, (size_t)n + 2);
    if (rc == BM_OK) rc = h2d(c, c->d_m_out, mg.out.adj, (size_t)mg.m);
    if (rc == BM_OK) rc = h2d(c, c->d_m_in_ptr, mg.in.ptr, (size_t)n + 2);
    if (rc == BM_OK) rc = h2d(c, c->d_m_in, mg.in.adj, (size_t)mg.m);
    if (rc == BM_OK) rc = h2d(c, c->d_f_out_desc, ho.desc, (size_t)n + 1);
    if (rc == BM_OK) rc = h2d(c, c->d_f_out_tab, ho.tab, (size_t)ho.slots);
    if (rc == BM_OK) rc = h2d(c, c->d_f_in_desc, hi.desc, (size_t)n + 1);
    if (rc == BM_OK) rc = h2d(c, c->d_f_in_tab, hi.tab, (size_t)hi.slots);
    if (rc == BM_OK) rc = h2d(c, c->d_f_out_ptr, fg.out.ptr, (size_t)n + 2);
    if (rc == BM_OK) rc = h2d(c, c->d_f_out, fg.out.adj, (size_t)fg.m);
    if (rc == BM_OK) rc = h2d(c, c->d_f_in_ptr, fg.in.ptr, (size_t)n + 2);
    if (rc == BM_OK) rc = h2d(c, c->d_f_in, fg.in.adj, (size_t)fg.m);
    c->pi_distinct = 1;   /* all unmatched */
    if (rc == BM_OK && hipMemsetAsync(c->d_pi, 0, ((size_t)n + 1) * sizeof(int32_t), c->stream) != hipSuccess)
        rc = fail(c, BM_ERR_HIP, "hipMemsetAsync failed");
    if (rc == BM_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, BM_ERR_HIP, "hipStreamSynchronize failed");
    bm_rowhash_free(&ho);
    bm_rowhash_free(&hi);
    bm_graph_free(&fg);
    if (rc != BM_OK) {
        bm_set_global_error("bm_create: %s", c->err);
        free_ctx(c);
        return rc;
    }
    *out = c;
    return BM_OK;
}

extern "C" int bm_degrees(const bm_ctx* ctx, int female, int32_t* out_deg, int32_t* in_deg) {
    if (!ctx) return BM_ERR_ARG;
    const int32_t n = ctx->n;
    for (int32_t r = 0; r <= n; r++) {
        if (female) {
            if (out_deg) out_deg[r] = r ? ctx->f_out_deg[r] : 0;
            if (in_deg) in_deg[r] = r ? ctx->f_in_deg[r] : 0;
        } else {
            if (out_deg) out_deg[r] = (int32_t)(ctx->male.out.ptr[r + 1] - ctx->male.out.ptr[r]);
            if (in_deg) in_deg[r] = (int32_t)(ctx->male.in.ptr[r + 1] - ctx->male.in.ptr[r]);
        }
    }
    return BM_OK;
}

extern "C" int bm_set_mapping(bm_ctx* c, const int32_t* pi) {
    if (!c || !pi) return fail(c, BM_ERR_ARG, "bm_set_mapping: null argument");
    for (int32_t i = 1; i <= c->n; i++)
        if (pi[i] < 0 || pi[i] > c->n)
            return fail(c, BM_ERR_ARG, "bm_set_mapping: pi[%d] = %d outside 0..%d", i, pi[i], c->n);
    {
        std::vector<uint8_t> seen((size_t)c->n + 1, 0);
        c->pi_distinct = 1;
        for (int32_t i = 1; i <= c->n && c->pi_distinct; i++) {
            if (pi[i] != 0 && seen[pi[i]]) c->pi_distinct = 0;
            seen[pi[i]] = 1;
        }
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->d_pi + 1, pi + 1, (size_t)c->n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    if (c->gain_on) {
        int rc = gain_rebuild(c);
        if (rc != BM_OK) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return BM_OK;
}

extern "C" int bm_get_mapping(bm_ctx* c, int32_t* pi) {
    if (!c || !pi) return fail(c, BM_ERR_ARG, "bm_get_mapping: null argument");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(pi, c->d_pi, ((size_t)c->n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    pi[0] = 0;
    return BM_OK;
}

extern "C" int bm_score(bm_ctx* c, int64_t* score) {
    if (!c || !score) return fail(c, BM_ERR_ARG, "bm_score: null argument");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemsetAsync(c->d_small, 0, sizeof(int64_t), c->stream));
    KTiming t;
    timing_begin(c, BM_KERNEL_SCORE, &t);
    HIPCHK(c, bm_launch_score(devgraph(c), c->tune[BM_KERNEL_SCORE], c->d_small, c->stream));
    timing_end(c, &t);
    HIPCHK(c, hipMemcpyAsync(score, c->d_small, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return BM_OK;
}

/* ------------------------------------------------------------------ */
/* Batches                                                             */
/* ------------------------------------------------------------------ */

extern "C" void bm_batch_destroy(bm_batch* b) {
    if (!b) return;
    if (b->ctx) {
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
    }
    (void)hipFree(b->d_cand);
    (void)hipFree(b->d_delta);
    (void)hipFree(b->d_perm);
    (void)hipFree(b->d_part);
    (void)hipFree(b->d_out);
    (void)hipFree(b->d_acc_scratch);
    (void)hipFree(b->d_acc_idx);
    free(b);
}

extern "C" int bm_batch_create(bm_ctx* c, int64_t capacity, bm_batch** out) {
    if (!c || !out || capacity < 1) return fail(c, BM_ERR_ARG, "bm_batch_create: invalid argument");
    if (capacity > ((int64_t)1 << 31) - 1)
        return fail(c, BM_ERR_ARG, "bm_batch_create: capacity %lld exceeds 2^31-1", (long long)capacity);
    *out = nullptr;
    HIPCHK(c, hipSetDevice(c->device));
    bm_batch* b = (bm_batch*)calloc(1, sizeof(bm_batch));
    if (!b) return fail(c, BM_ERR_NOMEM, "bm_batch_create: out of host memory");
    b->ctx = c;
    b->cap = capacity;
    int rc = BM_OK;
    const size_t nb = (size_t)bm_stats_blocks(capacity);
    if (rc == BM_OK) rc = dalloc(c, (int4**)&b->d_cand, (size_t)capacity);
    if (rc == BM_OK) rc = dalloc(c, &b->d_delta, (size_t)capacity);
    if (rc == BM_OK) rc = dalloc(c, &b->d_perm, (size_t)capacity);
    if (rc == BM_OK) rc = dalloc(c, &b->d_part, nb * 5);
    if (rc == BM_OK) rc = dalloc(c, &b->d_out, 16);
    if (rc != BM_OK) {
        bm_batch_destroy(b);
        return rc;
    }
    *out = b;
    return BM_OK;
}

extern "C" int64_t bm_batch_size(const bm_batch* b) { return b ? b->k : 0; }

extern "C" int bm_batch_set_pairs(bm_batch* b, const int32_t* uv, int64_t k) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (k < 0 || k > b->cap || (k > 0 && !uv)) return fail(c, BM_ERR_ARG, "bm_batch_set_pairs: k=%lld (capacity %lld)", (long long)k, (long long)b->cap);
    for (int64_t i = 0; i < 2 * k; i++)
        if (uv[i] < 1 || uv[i] > c->n)
            return fail(c, BM_ERR_ARG, "bm_batch_set_pairs: vertex %d at position %lld outside 1..%d", uv[i], (long long)i, c->n);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(b->d_cand, uv, (size_t)k * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    b->k = k;
    b->kind = 1;
    b->generated = 0;
    b->evaluated = 0;
    return BM_OK;
}

extern "C" int bm_batch_set_triples(bm_batch* b, const int32_t* abc, int64_t k) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (k < 0 || k > b->cap || (k > 0 && !abc)) return fail(c, BM_ERR_ARG, "bm_batch_set_triples: invalid k");
    std::vector<int4> tmp((size_t)k);
    for (int64_t i = 0; i < k; i++) {
        const int32_t x = abc[3 * i], y = abc[3 * i + 1], z = abc[3 * i + 2];
        if (x < 1 || x > c->n || y < 1 || y > c->n || z < 1 || z > c->n || x == y || x == z || y == z)
            return fail(c, BM_ERR_ARG, "bm_batch_set_triples: triple %lld = (%d,%d,%d) invalid (ids in 1..%d, distinct)",
                        (long long)i, x, y, z, c->n);
        tmp[(size_t)i] = make_int4(x, y, z, 0);
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(b->d_cand, tmp.data(), (size_t)k * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    b->k = k;
    b->kind = 2;
    b->generated = 0;
    b->evaluated = 0;
    return BM_OK;
}

extern "C" int bm_batch_gen_pairs(bm_batch* b, uint64_t seed, int64_t k) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (k < 0 || k > b->cap) return fail(c, BM_ERR_ARG, "bm_batch_gen_pairs: k=%lld exceeds capacity", (long long)k);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, bm_launch_gen_pairs(c->n, seed, k, (int2*)b->d_cand, c->stream));
    b->k = k;
    b->kind = 1;
    b->generated = 1;
    b->evaluated = 0;
    return BM_OK;
}

extern "C" int bm_batch_gen_triples(bm_batch* b, uint64_t seed, int64_t k) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (c->n < 3) return fail(c, BM_ERR_ARG, "bm_batch_gen_triples: n < 3");
    if (k < 0 || k > b->cap) return fail(c, BM_ERR_ARG, "bm_batch_gen_triples: k exceeds capacity");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, bm_launch_gen_triples(c->n, seed, k, (int4*)b->d_cand, c->stream));
    b->k = k;
    b->kind = 2;
    b->generated = 1;
    b->evaluated = 0;
    return BM_OK;
}

/* sortArray + fillArray (src/graph-re-alignment-gpu-swap.cu:296-307, :983-999):
 * changed ids ascending, then the other ids ascending. */
static int build_reorder(bm_ctx* c, const int32_t* changed, int32_t cnum, std::vector<int32_t>& re) {
    const int32_t n = c->n;
    if (cnum < 0 || cnum > n || (cnum > 0 && !changed)) return fail(c, BM_ERR_ARG, "changed list: invalid count %d", cnum);
    std::vector<uint8_t> present((size_t)n + 1, 0);
    re.clear();
    re.reserve((size_t)n);
    for (int32_t i = 0; i < cnum; i++) {
        const int32_t x = changed[i];
        if (x < 1 || x > n) return fail(c, BM_ERR_ARG, "changed list: vertex %d outside 1..%d", x, n);
        if (present[x]) return fail(c, BM_ERR_ARG, "changed list: vertex %d repeated", x);
        present[x] = 1;
        re.push_back(x);
    }
    std::sort(re.begin(), re.end());
    for (int32_t x = 1; x <= n; x++)
        if (!present[x]) re.push_back(x);
    return BM_OK;
}

extern "C" int bm_batch_gen_changed_pairs(bm_batch* b, const int32_t* changed, int32_t cnum) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    std::vector<int32_t> re;
    int rc = build_reorder(c, changed, cnum, re);
    if (rc != BM_OK) return rc;
    const int64_t P = (int64_t)cnum * (cnum - 1) / 2 + (int64_t)cnum * ((int64_t)c->n - cnum);
    if (P > b->cap) return fail(c, BM_ERR_ARG, "changed pairs: %lld pairs exceed batch capacity %lld", (long long)P, (long long)b->cap);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->d_reorder, re.data(), re.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, bm_launch_gen_changed_pairs(c->n, cnum, c->d_reorder, P, (int2*)b->d_cand, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  /* re goes out of scope */
    b->k = P;
    b->kind = 1;
    b->generated = 1;
    b->evaluated = 0;
    return BM_OK;
}

static int twohop_lds_ok(bm_ctx* c) {
    if (bm_twohop_lds_bytes(c->n) > 150 * 1024)
        return fail(c, BM_ERR_ARG, "2-hop sets: n=%d needs a %zu-byte LDS bitmap (limit 150 KiB)", c->n,
                    bm_twohop_lds_bytes(c->n));
    return BM_OK;
}

extern "C" int bm_twohop_set(bm_ctx* c, int32_t v, int32_t* out, int32_t* count) {
    if (!c || !out || !count) return fail(c, BM_ERR_ARG, "bm_twohop_set: null argument");
    if (v < 1 || v > c->n) return fail(c, BM_ERR_ARG, "bm_twohop_set: vertex %d outside 1..%d", v, c->n);
    int rc = twohop_lds_ok(c);
    if (rc != BM_OK) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    int32_t* d_v = reinterpret_cast<int32_t*>(c->d_small);          /* d_small[0] */
    int32_t* d_cnt = reinterpret_cast<int32_t*>(c->d_small + 1);    /* d_small[1] */
    HIPCHK(c, hipMemcpyAsync(d_v, &v, sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, bm_launch_twohop(devgraph(c), d_v, 0, 1, 0, nullptr, c->d_reorder, d_cnt, c->num_cus, c->stream));
    int32_t cnt = 0;
    HIPCHK(c, hipMemcpyAsync(&cnt, d_cnt, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (cnt < 0 || cnt > c->n) return fail(c, BM_ERR_STATE, "bm_twohop_set: bad count %d", cnt);
    if (cnt) HIPCHK(c, hipMemcpy(out, c->d_reorder, (size_t)cnt * sizeof(int32_t), hipMemcpyDeviceToHost));
    *count = cnt;
    return BM_OK;
}

extern "C" int bm_batch_gen_twohop(bm_batch* b, const int32_t* anchors, int32_t n_anchors, int32_t per_anchor,
                                   uint64_t seed) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (n_anchors < 0 || per_anchor < 1) return fail(c, BM_ERR_ARG, "bm_batch_gen_twohop: need n_anchors >= 0, per_anchor >= 1");
    const int64_t k = (int64_t)n_anchors * per_anchor;
    if (k > b->cap) return fail(c, BM_ERR_ARG, "bm_batch_gen_twohop: %lld pairs exceed batch capacity %lld", (long long)k, (long long)b->cap);
    int rc = twohop_lds_ok(c);
    if (rc != BM_OK) return rc;
    if (anchors)
        for (int32_t i = 0; i < n_anchors; i++)
            if (anchors[i] < 1 || anchors[i] > c->n)
                return fail(c, BM_ERR_ARG, "bm_batch_gen_twohop: anchor %d outside 1..%d", anchors[i], c->n);
    HIPCHK(c, hipSetDevice(c->device));
    /* anchors go to the upper half of the candidate buffer (pairs use the first 8 bytes of each 16) */
    int32_t* d_anch = nullptr;
    if (anchors && n_anchors) {
        d_anch = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(b->d_cand) + 8 * (size_t)b->cap);
        HIPCHK(c, hipMemcpyAsync(d_anch, anchors, sizeof(int32_t) * (size_t)n_anchors, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, bm_launch_twohop(devgraph(c), d_anch, seed, n_anchors, per_anchor, (int2*)b->d_cand, nullptr, nullptr,
                               c->num_cus, c->stream));
    if (anchors && n_anchors) HIPCHK(c, hipStreamSynchronize(c->stream));  /* host anchors may go */
    b->k = k;
    b->kind = 1;
    b->generated = 1;
    b->evaluated = 0;
    return BM_OK;
}

extern "C" int bm_batch_gen_changed_triples(bm_batch* b, const int32_t* changed, int32_t cnum) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (cnum > 4096) return fail(c, BM_ERR_ARG, "changed triples: c=%d above 4096", cnum);
    std::vector<int32_t> re;
    int rc = build_reorder(c, changed, cnum, re);
    if (rc != BM_OK) return rc;
    const int64_t cc = (int64_t)cnum;
    const int64_t nn = (int64_t)c->n;
    int64_t T = cc * (cc - 1) * (cc - 2) / 6 + cc * (cc - 1) / 2 * (nn - cc);
    if (nn - cc - 1 > 0) T += cc * (nn - cc - 1);
    if (T > b->cap) return fail(c, BM_ERR_ARG, "changed triples: %lld triples exceed batch capacity %lld", (long long)T, (long long)b->cap);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(c->d_reorder, re.data(), re.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, bm_launch_gen_changed_triples(c->n, cnum, c->d_reorder, T, (int4*)b->d_cand, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    b->k = T;
    b->kind = 2;
    b->generated = 1;
    b->evaluated = 0;
    return BM_OK;
}

/* The per-vertex words for per-candidate table lookups of k pairs or
 * triples: refreshed (one light launch) and used when the batch has at least
 * n candidates, where the random table reads saved (two per pair, three per
 * triple) outweigh it; NULL otherwise. */
static const int4* pair_pis(bm_ctx* c, int64_t k, int* rc) {
    *rc = BM_OK;
    if (k < c->n) return nullptr;
    if (!c->d_pis && (*rc = dalloc(c, &c->d_pis, (size_t)c->n + 1)) != BM_OK) return nullptr;
    if (bm_launch_gain_pis(devgraph(c), devgain(c), c->d_pis, c->stream) != hipSuccess) {
        *rc = fail(c, BM_ERR_HIP, "refreshing the per-vertex words failed");
        return nullptr;
    }
    return c->d_pis;
}

extern "C" int bm_batch_eval_swaps(bm_batch* b) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (b->kind != 1) return fail(c, BM_ERR_STATE, "bm_batch_eval_swaps: batch does not hold pairs");
    HIPCHK(c, hipSetDevice(c->device));
    KTiming t;
    if (c->gain_on) {
        timing_begin(c, BM_KERNEL_GAIN_SWAP, &t);
        int rc = BM_OK;
        const int4* pis = pair_pis(c, b->k, &rc);
        if (rc != BM_OK) {
            timing_end(c, &t);
            return rc;
        }
        HIPCHK(c, bm_launch_gain_swap(devgraph(c), devgain(c), BM_CAND_LIST, (const int2*)b->d_cand, nullptr, 0, b->k,
                                      b->d_delta, pis, c->stream));
    } else {
        timing_begin(c, BM_KERNEL_SWAP, &t);
        HIPCHK(c, bm_launch_swap(devgraph(c), scan_launch(c, BM_KERNEL_SWAP, b->k), BM_CAND_LIST, (const int2*)b->d_cand, nullptr, 0, b->k,
                                 b->d_delta, c->stream));
    }
    timing_end(c, &t);
    b->evaluated = 1;
    return BM_OK;
}

extern "C" int bm_batch_eval_swaps_uniform(bm_batch* b, uint64_t seed, int64_t k) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (k < 0 || k > b->cap) return fail(c, BM_ERR_ARG, "bm_batch_eval_swaps_uniform: k exceeds capacity");
    HIPCHK(c, hipSetDevice(c->device));
    KTiming t;
    if (c->gain_on) {
        timing_begin(c, BM_KERNEL_GAIN_SWAP, &t);
        int rc = BM_OK;
        const int4* pis = pair_pis(c, k, &rc);
        if (rc != BM_OK) {
            timing_end(c, &t);
            return rc;
        }
        HIPCHK(c, bm_launch_gain_swap(devgraph(c), devgain(c), BM_CAND_UNIFORM, nullptr, nullptr, seed, k, b->d_delta,
                                      pis, c->stream));
    } else {
        timing_begin(c, BM_KERNEL_SWAP, &t);
        HIPCHK(c, bm_launch_swap(devgraph(c), scan_launch(c, BM_KERNEL_SWAP, k), BM_CAND_UNIFORM, nullptr, nullptr, seed, k, b->d_delta,
                                 c->stream));
    }
    timing_end(c, &t);
    b->k = k;
    b->kind = 3; /* deltas only: candidates implicit in (seed, index) */
    b->generated = 1;
    b->evaluated = 1;
    return BM_OK;
}

/* Offsets of the u-major stream (seed, k) into c->d_um_off. */
static int umajor_offsets(bm_ctx* c, uint64_t seed, int64_t k) {
    if (!c->d_um_off) {
        int rc = dalloc(c, &c->d_um_counts, (size_t)c->n + 1);
        if (rc == BM_OK) rc = dalloc(c, &c->d_um_off, (size_t)c->n + 2);
        if (rc != BM_OK) return rc;
    }
    KTiming t;
    timing_begin(c, BM_KERNEL_AUX, &t);
    HIPCHK(c, bm_launch_umajor_offsets(c->n, seed, k, c->d_um_counts, c->d_um_off, c->num_cus, c->stream));
    timing_end(c, &t);
    return BM_OK;
}

/* The transposed table and the per-vertex words for the staged stream
 * lookups: GT is rebuilt only when the table changed since. */
static int gain_transposed(bm_ctx* c) {
    if (c->gt_version == c->gain_version) return BM_OK;
    if (!c->d_pis) {
        int rc = dalloc(c, &c->d_pis, (size_t)c->n + 1);
        if (rc != BM_OK) return rc;
    }
    if (!c->d_GT) {
        int rc = dalloc(c, &c->d_GT, (size_t)bm_gain_table_bytes(c->n) / sizeof(int32_t));
        if (rc != BM_OK) return rc;
    }
    KTiming t;
    timing_begin(c, BM_KERNEL_AUX, &t);
    HIPCHK(c, bm_launch_gain_transpose(devgain(c), c->n + 1, c->d_GT, c->num_cus, c->stream));
    timing_end(c, &t);
    c->gt_version = c->gain_version;
    return BM_OK;
}

extern "C" int bm_batch_gen_umajor(bm_batch* b, uint64_t seed, int64_t k) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (k < 0 || k > b->cap) return fail(c, BM_ERR_ARG, "bm_batch_gen_umajor: k exceeds capacity");
    HIPCHK(c, hipSetDevice(c->device));
    int rc = umajor_offsets(c, seed, k);
    if (rc != BM_OK) return rc;
    HIPCHK(c, bm_launch_gen_umajor(c->n, seed, c->d_um_off, k, (int2*)b->d_cand, c->stream));
    b->k = k;
    b->kind = 1;
    b->generated = 1;
    b->evaluated = 0;
    return BM_OK;
}

extern "C" int bm_batch_eval_swaps_umajor(bm_batch* b, uint64_t seed, int64_t k) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (k < 0 || k > b->cap) return fail(c, BM_ERR_ARG, "bm_batch_eval_swaps_umajor: k exceeds capacity");
    HIPCHK(c, hipSetDevice(c->device));
    int rc = umajor_offsets(c, seed, k);
    if (rc != BM_OK) return rc;
    KTiming t;
    if (c->gain_on) {
        /* rows and bitmaps staged per first vertex (n <= 18,975), bitmaps only
         * with the two table reads gathered (no transposed table), or per-pair
         * lookups when not even the bitmaps fit */
        const bool staged = bm_gain_umajor_lds(c->n) <= 160 * 1024;
        const bool bitmaps = !staged && bm_gain_umajor_bitmap_lds(c->n) <= 160 * 1024;
        if (staged && (rc = gain_transposed(c)) != BM_OK) return rc;
        if (bitmaps && !c->d_pis && (rc = dalloc(c, &c->d_pis, (size_t)c->n + 1)) != BM_OK) return rc;
        timing_begin(c, BM_KERNEL_GAIN_SWAP, &t);
        if (staged || bitmaps) {
            /* pi and the diagonal G(v, pi(v)) are cheap: refreshed every call */
            HIPCHK(c, bm_launch_gain_pis(devgraph(c), devgain(c), c->d_pis, c->stream));
            HIPCHK(c, bm_launch_gain_swap_umajor(devgraph(c), devgain(c), c->d_GT, c->d_pis, c->d_um_off, seed,
                                                 b->d_delta, c->num_cus, c->stream));
        } else {
            const int4* pis = pair_pis(c, k, &rc);
            if (rc != BM_OK) {
                timing_end(c, &t);
                return rc;
            }
            HIPCHK(c, bm_launch_gain_swap(devgraph(c), devgain(c), BM_CAND_UMAJOR, nullptr, c->d_um_off, seed, k,
                                          b->d_delta, pis, c->stream));
        }
    } else {
        timing_begin(c, BM_KERNEL_SWAP, &t);
        HIPCHK(c, bm_launch_swap(devgraph(c), scan_launch(c, BM_KERNEL_SWAP, k), BM_CAND_UMAJOR, nullptr, c->d_um_off, seed, k,
                                 b->d_delta, c->stream));
    }
    timing_end(c, &t);
    b->k = k;
    b->kind = 3; /* deltas only: candidates implicit in (seed, index) */
    b->generated = 1;
    b->evaluated = 1;
    return BM_OK;
}

extern "C" int bm_batch_eval_rot3(bm_batch* b) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (b->kind != 2) return fail(c, BM_ERR_STATE, "bm_batch_eval_rot3: batch does not hold triples");
    HIPCHK(c, hipSetDevice(c->device));
    KTiming t;
    if (c->gain_on) {
        timing_begin(c, BM_KERNEL_GAIN_ROT3, &t);
        int rc = BM_OK;
        const int4* pis = pair_pis(c, b->k, &rc);
        if (rc != BM_OK) {
            timing_end(c, &t);
            return rc;
        }
        HIPCHK(c, bm_launch_gain_rot3(devgraph(c), devgain(c), (const int4*)b->d_cand, b->k, b->d_delta, b->d_perm,
                                      (int4*)b->d_cand, pis, c->stream));
    } else {
        timing_begin(c, BM_KERNEL_ROT3, &t);
        HIPCHK(c, bm_launch_rot3(devgraph(c), scan_launch(c, BM_KERNEL_ROT3, b->k), (const int4*)b->d_cand, b->k, b->d_delta,
                                 b->d_perm, (int4*)b->d_cand, c->stream));
    }
    timing_end(c, &t);
    b->evaluated = 1;
    return BM_OK;
}

extern "C" int bm_batch_gen_amajor(bm_batch* b, uint64_t seed, int64_t k) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (k < 0 || k > b->cap) return fail(c, BM_ERR_ARG, "bm_batch_gen_amajor: k exceeds capacity");
    if (c->n < 3) return fail(c, BM_ERR_ARG, "bm_batch_gen_amajor: triples need n >= 3");
    HIPCHK(c, hipSetDevice(c->device));
    int rc = umajor_offsets(c, bm_amajor_seed(seed), k);
    if (rc != BM_OK) return rc;
    HIPCHK(c, bm_launch_gen_amajor(c->n, seed, c->d_um_off, k, (int4*)b->d_cand, c->stream));
    b->k = k;
    b->kind = 2;
    b->generated = 1;
    b->evaluated = 0;
    return BM_OK;
}

/* Regions of the a-major 3-vertex queue: at most one per block of that
 * launch (one block per CU there). */
static constexpr int32_t kR3QueueRegions = 4096;

extern "C" int bm_batch_eval_rot3_amajor(bm_batch* b, uint64_t seed, int64_t k) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (k < 0 || k > b->cap) return fail(c, BM_ERR_ARG, "bm_batch_eval_rot3_amajor: k exceeds capacity");
    if (c->n < 3) return fail(c, BM_ERR_ARG, "bm_batch_eval_rot3_amajor: triples need n >= 3");
    if (!c->gain_on || bm_gain_rot3_amajor_lds(c->n) > 160 * 1024) {
        /* no staged rows: materialise the stream and evaluate it as a list */
        int rc = bm_batch_gen_amajor(b, seed, k);
        return rc != BM_OK ? rc : bm_batch_eval_rot3(b);
    }
    HIPCHK(c, hipSetDevice(c->device));
    int rc = umajor_offsets(c, bm_amajor_seed(seed), k);
    if (rc == BM_OK) rc = gain_transposed(c);
    if (rc != BM_OK) return rc;
    /* the queue of triples with an edge: an eighth of the batch (about 7% of
     * the triples have one at n = 18,524; a block whose region fills
     * finishes the rest in place) */
    const int64_t want = k / 8 > 65536 ? k / 8 : 65536;
    if (!c->d_r3q_counts && (rc = dalloc(c, &c->d_r3q_counts, (size_t)kR3QueueRegions)) != BM_OK) return rc;
    if (c->r3q_entries < want) {
        (void)hipFree(c->d_r3q);
        c->d_r3q = nullptr;
        c->r3q_entries = 0;
        if ((rc = dalloc(c, &c->d_r3q, (size_t)want * 3)) != BM_OK) return rc;
        c->r3q_entries = want;
    }
    KTiming t;
    timing_begin(c, BM_KERNEL_GAIN_ROT3, &t);
    HIPCHK(c, bm_launch_gain_pis(devgraph(c), devgain(c), c->d_pis, c->stream));
    HIPCHK(c, bm_launch_gain_rot3_amajor(devgraph(c), devgain(c), c->d_GT, c->d_pis, c->d_um_off, seed, b->d_delta,
                                         b->d_perm, c->d_r3q, c->r3q_entries, c->d_r3q_counts, kR3QueueRegions,
                                         c->num_cus, c->stream));
    timing_end(c, &t);
    b->k = k;
    b->kind = 4; /* best deltas and permutations only: triples implicit in (seed, index) */
    b->generated = 1;
    b->evaluated = 1;
    return BM_OK;
}

extern "C" int bm_batch_get(bm_batch* b, int64_t first, int64_t count, int64_t* delta, uint8_t* perm, int32_t* cand) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (first < 0 || count < 0 || first + count > b->k) return fail(c, BM_ERR_ARG, "bm_batch_get: slice outside the batch");
    if (count == 0) return BM_OK;
    HIPCHK(c, hipSetDevice(c->device));
    if (delta) {
        if (!b->evaluated) return fail(c, BM_ERR_STATE, "bm_batch_get: batch not evaluated");
        HIPCHK(c, hipMemcpyAsync(delta, b->d_delta + first, (size_t)count * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    }
    if (perm) {
        if (b->kind != 2 && b->kind != 4) return fail(c, BM_ERR_STATE, "bm_batch_get: perm only for triple batches");
        HIPCHK(c, hipMemcpyAsync(perm, b->d_perm + first, (size_t)count, hipMemcpyDeviceToHost, c->stream));
    }
    if (cand) {
        if (b->kind == 1) {
            HIPCHK(c, hipMemcpyAsync(cand, (int2*)b->d_cand + first, (size_t)count * sizeof(int2), hipMemcpyDeviceToHost, c->stream));
        } else if (b->kind == 2) {
            HIPCHK(c, hipMemcpyAsync(cand, (int4*)b->d_cand + first, (size_t)count * sizeof(int4), hipMemcpyDeviceToHost, c->stream));
        } else if (b->kind == 3 || b->kind == 4) {
            return fail(c, BM_ERR_STATE, "bm_batch_get: stream-evaluated batches do not store candidates");
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return BM_OK;
}

extern "C" int bm_batch_stats(bm_batch* b, bm_stats* out) {
    if (!b || !out) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (!b->evaluated) return fail(c, BM_ERR_STATE, "bm_batch_stats: batch not evaluated");
    HIPCHK(c, hipSetDevice(c->device));
    int64_t h[6];
    if (b->k == 0) {
        out->count = 0; out->sum = 0; out->max = INT64_MIN; out->argmax = -1; out->n_pos = 0; out->n_zero = 0;
        return BM_OK;
    }
    HIPCHK(c, bm_launch_stats(b->d_delta, b->k, b->d_part, b->d_out, c->stream));
    HIPCHK(c, hipMemcpyAsync(h, b->d_out, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out->count = h[0]; out->sum = h[1]; out->max = h[2]; out->argmax = h[3]; out->n_pos = h[4]; out->n_zero = h[5];
    return BM_OK;
}

extern "C" int bm_batch_pick(bm_batch* b, int32_t kind, int64_t rank, int64_t* index) {
    if (!b || !index) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (!b->evaluated) return fail(c, BM_ERR_STATE, "bm_batch_pick: batch not evaluated");
    if (kind != 0 && kind != 1) return fail(c, BM_ERR_ARG, "bm_batch_pick: kind must be 0 or 1");
    *index = -1;
    if (rank < 0 || b->k == 0) return BM_OK;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, bm_launch_pick(b->d_delta, b->k, kind, rank, b->d_part, b->d_out, c->stream));
    HIPCHK(c, hipMemcpyAsync(index, b->d_out, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return BM_OK;
}

static int accept_impl(bm_batch* b, int64_t max_accept, int32_t* accepted, int64_t* n_acc, int64_t* gain) {
    bm_ctx* c = b->ctx;
    if (b->kind != 1 || !b->evaluated) return fail(c, BM_ERR_STATE, "accept: batch must hold evaluated pairs");
    if (c->n > 1200000) return fail(c, BM_ERR_ARG, "accept: n=%d above the LDS lock-bitmap limit 1200000", c->n);
    HIPCHK(c, hipSetDevice(c->device));
    if (!b->d_acc_scratch) {
        b->acc_scratch_bytes = bm_accept_scratch_bytes(b->cap);
        HIPCHK(c, hipMalloc(&b->d_acc_scratch, b->acc_scratch_bytes));
        HIPCHK(c, hipMalloc((void**)&b->d_acc_idx, (size_t)b->cap * sizeof(int64_t)));
    }
    if (c->gain_on) {
        int rc = gain_moves_cap(c, 2 * (b->k > 0 ? b->k : 1));
        if (rc != BM_OK) return rc;
    }
    HIPCHK(c, bm_launch_accept(devgraph(c), c->d_pi, (const int2*)b->d_cand, b->d_delta, b->k, max_accept,
                               b->d_acc_scratch, b->acc_scratch_bytes, b->d_acc_idx, b->d_out,
                               c->gain_on ? c->d_moves : nullptr, c->stream));
    if (c->gain_on && b->k > 0) {
        int rc = gain_patch(c, b->d_out + 2, 2 * b->k, false);
        if (rc != BM_OK) return rc;
    }
    int64_t meta[2];
    HIPCHK(c, hipMemcpyAsync(meta, b->d_out, sizeof meta, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (accepted && meta[0] > 0) {
        std::vector<int64_t> idx((size_t)meta[0]);
        HIPCHK(c, hipMemcpy(idx.data(), b->d_acc_idx, (size_t)meta[0] * sizeof(int64_t), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < meta[0]; i++) accepted[i] = (int32_t)idx[(size_t)i];
    }
    if (n_acc) *n_acc = meta[0];
    if (gain) *gain = meta[1];
    b->evaluated = 0; /* the matching changed: deltas are stale */
    return BM_OK;
}

extern "C" int bm_batch_accept_disjoint(bm_batch* b, int64_t max_accept, int64_t* n_accepted, int64_t* gain) {
    if (!b) return BM_ERR_ARG;
    return accept_impl(b, max_accept, nullptr, n_accepted, gain);
}

/* ------------------------------------------------------------------ */
/* Host-buffer convenience calls                                       */
/* ------------------------------------------------------------------ */

static int work_batch(bm_ctx* c, int64_t k, bm_batch** out) {
    if (k < 1) k = 1;
    if (c->work && c->work->cap >= k) { *out = c->work; return BM_OK; }
    if (c->work) { bm_batch_destroy(c->work); c->work = nullptr; }
    int rc = bm_batch_create(c, k, &c->work);
    if (rc != BM_OK) return rc;
    *out = c->work;
    return BM_OK;
}

extern "C" int bm_swap_deltas(bm_ctx* c, const int32_t* uv, int64_t k, int64_t* delta) {
    if (!c || k < 0 || (k > 0 && (!uv || !delta))) return fail(c, BM_ERR_ARG, "bm_swap_deltas: invalid argument");
    if (k == 0) return BM_OK;
    bm_batch* b;
    int rc = work_batch(c, k, &b);
    if (rc == BM_OK) rc = bm_batch_set_pairs(b, uv, k);
    if (rc == BM_OK) rc = bm_batch_eval_swaps(b);
    if (rc == BM_OK) rc = bm_batch_get(b, 0, k, delta, nullptr, nullptr);
    return rc;
}

extern "C" int bm_swap_deltas_changed(bm_ctx* c, const int32_t* changed, int32_t cnum, int64_t* delta, int32_t* uv,
                                      int64_t* count) {
    if (!c) return BM_ERR_ARG;
    const int64_t P = (int64_t)cnum * (cnum - 1) / 2 + (int64_t)cnum * ((int64_t)c->n - cnum);
    bm_batch* b;
    int rc = work_batch(c, P, &b);
    if (rc == BM_OK) rc = bm_batch_gen_changed_pairs(b, changed, cnum);
    if (rc == BM_OK) rc = bm_batch_eval_swaps(b);
    if (rc == BM_OK) rc = bm_batch_get(b, 0, P, delta, nullptr, uv);
    if (rc == BM_OK && count) *count = P;
    return rc;
}

extern "C" int bm_rot3_deltas(bm_ctx* c, const int32_t* abc, int64_t k, int64_t* best, uint8_t* perm) {
    if (!c || k < 0 || (k > 0 && (!abc || !best))) return fail(c, BM_ERR_ARG, "bm_rot3_deltas: invalid argument");
    if (k == 0) return BM_OK;
    bm_batch* b;
    int rc = work_batch(c, k, &b);
    if (rc == BM_OK) rc = bm_batch_set_triples(b, abc, k);
    if (rc == BM_OK) rc = bm_batch_eval_rot3(b);
    if (rc == BM_OK) rc = bm_batch_get(b, 0, k, best, perm, nullptr);
    return rc;
}

extern "C" int bm_rot3_deltas_changed(bm_ctx* c, const int32_t* changed, int32_t cnum, int64_t* best, int32_t* abcp,
                                      int64_t* count) {
    if (!c) return BM_ERR_ARG;
    const int64_t cc = cnum, nn = c->n;
    int64_t T = cc * (cc - 1) * (cc - 2) / 6 + cc * (cc - 1) / 2 * (nn - cc);
    if (nn - cc - 1 > 0) T += cc * (nn - cc - 1);
    bm_batch* b;
    int rc = work_batch(c, T, &b);
    if (rc == BM_OK) rc = bm_batch_gen_changed_triples(b, changed, cnum);
    if (rc == BM_OK) rc = bm_batch_eval_rot3(b);
    if (rc == BM_OK) rc = bm_batch_get(b, 0, T, best, nullptr, abcp);
    if (rc == BM_OK && count) *count = T;
    return rc;
}

extern "C" int bm_accept_disjoint(bm_ctx* c, const int32_t* uv, const int64_t* delta, int64_t k, int32_t* accepted,
                                  int64_t* n_accepted, int64_t* gain) {
    if (!c || k < 0 || (k > 0 && (!uv || !delta))) return fail(c, BM_ERR_ARG, "bm_accept_disjoint: invalid argument");
    bm_batch* b;
    int rc = work_batch(c, k, &b);
    if (rc != BM_OK) return rc;
    rc = bm_batch_set_pairs(b, uv, k);
    if (rc != BM_OK) return rc;
    if (k > 0) HIPCHK(c, hipMemcpyAsync(b->d_delta, delta, (size_t)k * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    b->evaluated = 1;
    return accept_impl(b, 0, accepted, n_accepted, gain);
}

extern "C" int bm_apply_swap(bm_ctx* c, int32_t u, int32_t v) {
    if (!c || u < 1 || u > c->n || v < 1 || v > c->n) return fail(c, BM_ERR_ARG, "bm_apply_swap: vertex outside 1..n");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->gain_on) {
        int rc = gain_moves_cap(c, 3);
        if (rc != BM_OK) return rc;
        HIPCHK(c, bm_launch_moves_from_swap(c->d_pi, u, v, c->d_moves, c->d_nmoves, c->stream));
    }
    HIPCHK(c, bm_launch_apply_swap(c->d_pi, u, v, c->stream));
    if (c->gain_on) {
        const int32_t xs[2] = {u, v};
        return gain_patch(c, c->d_nmoves, 2, gt_patch_pays(c, xs, 2));
    }
    return BM_OK;
}

/* Exact best swap over all n(n-1)/2 pairs from the placement-gain table and
 * its transpose (k_gain_best_swap): the maximum that the score-lookup
 * driver's heap holds (src/best-swap/score-lookup.c:246-293 fills the n x n
 * delta table, :612-620 extracts the maximum), computed afresh on every call,
 * so no entry is stale. */
extern "C" int bm_best_swap(bm_ctx* c, int64_t* delta, int32_t* u, int32_t* v) {
    if (!c || !delta || !u || !v) return fail(c, BM_ERR_ARG, "bm_best_swap: null argument");
    if (!c->gain_on) return fail(c, BM_ERR_ARG, "bm_best_swap: needs the placement-gain table (bm_gain_enable)");
    HIPCHK(c, hipSetDevice(c->device));
    int rc = gain_transposed(c);
    if (rc != BM_OK) return rc;
    if (!c->d_best && (rc = dalloc(c, &c->d_best, (size_t)kBestSwapParts + 1)) != BM_OK) return rc;
    KTiming t;
    timing_begin(c, BM_KERNEL_GAIN_SWAP, &t);
    HIPCHK(c, bm_launch_gain_pis(devgraph(c), devgain(c), c->d_pis, c->stream));
    HIPCHK(c, bm_launch_gain_best_swap(devgraph(c), devgain(c), c->d_GT, c->d_pis, c->d_best + 1, kBestSwapParts,
                                       c->d_best, c->num_cus, c->stream));
    timing_end(c, &t);
    int4 h;
    HIPCHK(c, hipMemcpyAsync(&h, c->d_best, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (h.z == INT32_MAX) {   /* n < 2: no pair */
        *delta = 0;
        *u = *v = 0;
        return BM_OK;
    }
    *delta = (int64_t)(((uint64_t)(uint32_t)h.y << 32) | (uint64_t)(uint32_t)h.x);
    *u = h.z;
    *v = h.w;
    return BM_OK;
}

/* Best partner of every u in u0 .. u0 + count - 1 over v = u+1..n (the
 * inner loop of the OpenMP tool's sweep, src/re-alignment-parallel.c:598-621,
 * which keeps the largest calculate_swap_delta over node_m2 > node_m1). */
extern "C" int bm_best_partners(bm_ctx* c, int32_t u0, int32_t count, int64_t* delta, int32_t* v) {
    if (!c || !delta || !v) return fail(c, BM_ERR_ARG, "bm_best_partners: null argument");
    if (count < 0 || (count > 0 && (u0 < 1 || (int64_t)u0 + count - 1 > c->n)))
        return fail(c, BM_ERR_ARG, "bm_best_partners: range %d + %d outside 1..%d", u0, count, c->n);
    if (!c->gain_on) return fail(c, BM_ERR_ARG, "bm_best_partners: needs the placement-gain table (bm_gain_enable)");
    if (count == 0) return BM_OK;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = gain_transposed(c);
    if (rc != BM_OK) return rc;
    if (c->partners_cap < count) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        (void)hipFree(c->d_partners);
        c->d_partners = nullptr;
        c->partners_cap = 0;
        if ((rc = dalloc(c, &c->d_partners, (size_t)count)) != BM_OK) return rc;
        c->partners_cap = count;
    }
    KTiming t;
    timing_begin(c, BM_KERNEL_GAIN_SWAP, &t);
    HIPCHK(c, bm_launch_gain_pis(devgraph(c), devgain(c), c->d_pis, c->stream));
    HIPCHK(c, bm_launch_gain_best_partners(devgraph(c), devgain(c), c->d_GT, c->d_pis, u0, count, c->d_partners,
                                           c->num_cus, c->stream));
    timing_end(c, &t);
    std::vector<int4> h((size_t)count);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_partners, (size_t)count * sizeof(int4), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int32_t i = 0; i < count; i++) {
        if (h[i].w == INT32_MAX) {   /* u = n: no partner */
            delta[i] = INT64_MIN;
            v[i] = 0;
        } else {
            delta[i] = (int64_t)(((uint64_t)(uint32_t)h[i].y << 32) | (uint64_t)(uint32_t)h[i].x);
            v[i] = h[i].w;
        }
    }
    return BM_OK;
}

extern "C" int bm_apply_perm3(bm_ctx* c, int32_t a, int32_t b, int32_t d, int32_t p) {
    if (!c || a < 1 || a > c->n || b < 1 || b > c->n || d < 1 || d > c->n || a == b || a == d || b == d || p < 0 || p > 5)
        return fail(c, BM_ERR_ARG, "bm_apply_perm3: invalid triple or permutation");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->gain_on) {
        int rc = gain_moves_cap(c, 3);
        if (rc != BM_OK) return rc;
        HIPCHK(c, bm_launch_moves_from_perm3(c->d_pi, a, b, d, p, c->d_moves, c->d_nmoves, c->stream));
    }
    HIPCHK(c, bm_launch_apply_perm3(c->d_pi, a, b, d, p, c->stream));
    if (c->gain_on) {
        const int32_t xs[3] = {a, b, d};
        return gain_patch(c, c->d_nmoves, 3, gt_patch_pays(c, xs, 3));
    }
    return BM_OK;
}

/* ------------------------------------------------------------------ */
/* Placement-gain table                                                */
/* ------------------------------------------------------------------ */

extern "C" int64_t bm_gain_table_bytes(int32_t n) {
    if (n < 1) return 0;
    return ((int64_t)n + 1) * bm_gain_stride(n) * (int64_t)sizeof(int32_t);
}

extern "C" int bm_gain_enabled(const bm_ctx* c) { return c ? c->gain_on : 0; }

extern "C" int bm_gain_dense_rows(const bm_ctx* c) { return c && c->gain_on && c->pi_distinct ? c->ndense : 0; }

/* Dense rows: female rows of at least stride / kDenseDiv entries. */
static constexpr int64_t kDenseDiv = 4;

extern "C" int bm_gain_enable(bm_ctx* c, int on) {
    if (!c) return BM_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!on) {
        if (!c->gain_on) return BM_OK;
        (void)hipFree(c->d_G);
        c->d_G = nullptr;
        (void)hipFree(c->d_GT);
        c->d_GT = nullptr;
        (void)hipFree(c->d_edesc);
        c->d_edesc = nullptr;
        c->gain_on = 0;
        c->gain_version++;
        return BM_OK;
    }
    if (c->gain_on) return BM_OK;
    /* every table entry must fit int32: |G(u,f)| <= sum over u's edges of max(|w|, max|w_F|) */
    const bm_graph_host& mg = c->male;
    for (int32_t u = 1; u <= c->n; u++) {
        int64_t bound = 0;
        for (int pass = 0; pass < 2; pass++) {
            const bm_csr_host& L = pass ? mg.in : mg.out;
            for (int64_t e = L.ptr[u]; e < L.ptr[u + 1]; e++) {
                const int64_t w = L.adj[e].y < 0 ? -(int64_t)L.adj[e].y : L.adj[e].y;
                bound += w > c->max_abs_fw ? w : c->max_abs_fw;
            }
        }
        if (bound > INT32_MAX)
            return fail(c, BM_ERR_ARG, "bm_gain_enable: vertex %d's incident weights can exceed int32 in the table", u);
    }
    const size_t bytes = (size_t)bm_gain_table_bytes(c->n);
    size_t free_b = 0, total_b = 0;
    HIPCHK(c, hipMemGetInfo(&free_b, &total_b));
    const size_t list_bytes = (size_t)(2 * c->male.m + 1) * sizeof(int4);
    if (bytes + list_bytes + ((size_t)1 << 30) > free_b)
        return fail(c, BM_ERR_NOMEM, "bm_gain_enable: table needs %zu bytes, %zu free on the device", bytes, free_b);
    if (!c->d_m_out_desc) {
        bm_rowhash_host hm;
        memset(&hm, 0, sizeof hm);
        if (bm_rowhash_build(&mg.out, &hm) != BM_OK) return fail(c, BM_ERR_NOMEM, "bm_gain_enable: out of host memory");
        int rc = dalloc(c, &c->d_m_out_desc, (size_t)c->n + 1);
        if (rc == BM_OK) rc = dalloc(c, &c->d_m_out_tab, (size_t)(hm.slots ? hm.slots : 2));
        if (rc == BM_OK) rc = dalloc(c, &c->d_m_self, (size_t)c->n + 1);
        if (rc == BM_OK) rc = dalloc(c, &c->d_f_self, (size_t)c->n + 1);
        if (rc == BM_OK) rc = dalloc(c, &c->d_nmoves, 1);
        if (rc == BM_OK) rc = h2d(c, c->d_m_out_desc, hm.desc, (size_t)c->n + 1);
        if (rc == BM_OK) rc = h2d(c, c->d_m_out_tab, hm.tab, (size_t)hm.slots);
        if (rc == BM_OK) rc = h2d(c, c->d_m_self, c->h_m_self, (size_t)c->n + 1);
        if (rc == BM_OK) rc = h2d(c, c->d_f_self, c->h_f_self, (size_t)c->n + 1);
        if (rc == BM_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, BM_ERR_HIP, "sync failed");
        bm_rowhash_free(&hm);
        if (rc != BM_OK) return rc;
    }
    if (c->compact && !c->d_f_out_p && c->female_m > 0) {
        int r2 = dalloc(c, &c->d_f_out_p, (size_t)c->female_m);
        if (r2 == BM_OK) r2 = dalloc(c, &c->d_f_in_p, (size_t)c->female_m);
        if (r2 != BM_OK) return r2;
        /* compact graphs have female weights in 0..32767; the narrower weight
         * field leaves the column's LDS byte address in the low bits */
        c->row_layout = c->max_abs_fw <= 8191 ? BM_ROWS_P19 : BM_ROWS_P15;
        HIPCHK(c, bm_launch_pack_rows(c->d_f_out, c->female_m, c->row_layout, c->d_f_out_p, c->stream));
        HIPCHK(c, bm_launch_pack_rows(c->d_f_in, c->female_m, c->row_layout, c->d_f_in_p, c->stream));
    }
    /* Dense copies of the longest female rows, summed 16 male rows at a time
     * in packed 16-bit registers: rows of at least stride / kDenseDiv
     * entries qualify, longest first (ties to the lower id), at most
     * BM_DENSE_MAX.  A dense row costs the tile pass about 10 us at
     * n = 18,524 whatever its length, so only the longest pay: 48 to 100
     * rows build in 3.80-3.88 ms there, 197 in 4.42 ms, 406 in 5.84 ms.
     * Exact when the graph is compact (female weights in 0..32767) and
     * every male weight is >= 0: an absent column then adds min(w, 0) = 0,
     * as a missing female edge does, and each term min(w, fw) is at most
     * fw, so a dense sum never exceeds its column's sum of dense-row
     * weights.  A row is taken only while every column's sum stays
     * <= 65535, which keeps the packed 16-bit sums exact.  The tile pass
     * leaves its sums packed in the second half of each row (half the
     * table's bytes written and read back by the row build), which also
     * serves a table built in column windows. */
    if (c->compact && c->male_nonneg && !c->d_f_out_dense && c->female_m > 0) {
        const int64_t stride = bm_gain_stride(c->n);
        const int32_t thr = (int32_t)(stride / kDenseDiv) > 1 ? (int32_t)(stride / kDenseDiv) : 1;
        std::vector<int2> cand;   /* {degree, side * (n + 1) + id}, side 0 = out-row */
        for (int32_t r = 1; r <= c->n; r++) {
            if (c->f_out_deg[r] >= thr) cand.push_back(make_int2(c->f_out_deg[r], r));
            if (c->f_in_deg[r] >= thr) cand.push_back(make_int2(c->f_in_deg[r], (c->n + 1) + r));
        }
        std::sort(cand.begin(), cand.end(), [](const int2& a, const int2& b) {
            return a.x != b.x ? a.x > b.x : a.y < b.y;
        });
        if (!cand.empty()) {   /* the column sums, from each candidate's entries {column, weight} */
            std::vector<int64_t> ptr_out((size_t)c->n + 2), ptr_in((size_t)c->n + 2);
            HIPCHK(c, hipMemcpy(ptr_out.data(), c->d_f_out_ptr, ptr_out.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
            HIPCHK(c, hipMemcpy(ptr_in.data(), c->d_f_in_ptr, ptr_in.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
            std::vector<int32_t> colsum((size_t)c->n + 1, 0);
            std::vector<int2> ent, kept;
            for (const int2& cd : cand) {
                if (kept.size() == (size_t)BM_DENSE_MAX) break;
                const bool in = cd.y > c->n;
                const int32_t r = in ? cd.y - (c->n + 1) : cd.y;
                const int64_t* p = in ? ptr_in.data() : ptr_out.data();
                ent.resize((size_t)(p[r + 1] - p[r]));
                HIPCHK(c, hipMemcpy(ent.data(), (in ? c->d_f_in : c->d_f_out) + p[r], ent.size() * sizeof(int2),
                                    hipMemcpyDeviceToHost));
                bool fits = true;
                for (const int2& e : ent) fits = fits && colsum[e.x] + e.y <= 65535;
                if (!fits) continue;
                for (const int2& e : ent) colsum[e.x] += e.y;
                kept.push_back(cd);
            }
            cand.swap(kept);
        }
        if (!cand.empty()) {
            std::vector<int32_t> hout((size_t)c->n + 1, -1), hin((size_t)c->n + 1, -1);
            std::vector<int2> rows(cand.size());
            for (size_t k = 0; k < cand.size(); k++) {
                const bool in = cand[k].y > c->n;
                const int32_t r = in ? cand[k].y - (c->n + 1) : cand[k].y;
                (in ? hin : hout)[r] = (int32_t)k;
                rows[k] = make_int2(r, in ? 0 : 1);
            }
            int2* d_rows = nullptr;
            int r2 = dalloc(c, &c->d_f_out_dense, (size_t)c->n + 1);
            if (r2 == BM_OK) r2 = dalloc(c, &c->d_f_in_dense, (size_t)c->n + 1);
            if (r2 == BM_OK) r2 = dalloc(c, &c->d_dense, (cand.size() + 1) * (size_t)(stride / 2));   /* + a zero row */
            const size_t wcount = (size_t)((c->n + 16) / 16) * cand.size() * 16;
            if (r2 == BM_OK) r2 = dalloc(c, &c->d_dense_w, wcount);
            if (r2 == BM_OK) r2 = dalloc(c, &d_rows, rows.size());
            if (r2 == BM_OK) r2 = h2d(c, c->d_f_out_dense, hout.data(), hout.size());
            if (r2 == BM_OK) r2 = h2d(c, c->d_f_in_dense, hin.data(), hin.size());
            if (r2 == BM_OK) r2 = h2d(c, d_rows, rows.data(), rows.size());
            if (r2 == BM_OK && hipMemsetAsync(c->d_dense, 0, (cand.size() + 1) * (size_t)stride * 2, c->stream) != hipSuccess)
                r2 = fail(c, BM_ERR_HIP, "bm_gain_enable: clearing the dense rows failed");
            if (r2 == BM_OK && bm_launch_dense_fill(c->d_f_out_ptr, c->d_f_out, c->d_f_in_ptr, c->d_f_in, d_rows,
                                                    (int32_t)rows.size(), stride, (uint16_t*)c->d_dense,
                                                    c->stream) != hipSuccess)
                r2 = fail(c, BM_ERR_HIP, "bm_gain_enable: filling the dense rows failed");
            if (r2 == BM_OK && hipStreamSynchronize(c->stream) != hipSuccess) r2 = fail(c, BM_ERR_HIP, "sync failed");
            (void)hipFree(d_rows);
            if (r2 != BM_OK) return r2;
            c->ndense = (int32_t)rows.size();
        }
    }
    /* one spare entry: the build's look-ahead reads entry base(u) of a row with no edges */
    int rc = dalloc(c, &c->d_edesc, (size_t)(2 * c->male.m + 1));
    if (rc != BM_OK) return rc;
    if (!c->d_row_fill && (rc = dalloc(c, &c->d_row_fill, (size_t)c->n + 1)) != BM_OK) return rc;
    if (!c->d_echunks) {
        std::vector<int2> ch;
        for (int32_t u = 1; u <= c->n; u++) {
            const int64_t L = (mg.out.ptr[u + 1] - mg.out.ptr[u]) + (mg.in.ptr[u + 1] - mg.in.ptr[u]);
            for (int64_t t = 0; t < L; t += BM_DESC_CHUNK) ch.push_back(make_int2(u, (int32_t)t));
        }
        c->n_echunks = (int64_t)ch.size();
        rc = dalloc(c, &c->d_echunks, ch.empty() ? (size_t)1 : ch.size());
        if (rc == BM_OK && !ch.empty()) rc = h2d(c, c->d_echunks, ch.data(), ch.size());
        if (rc == BM_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, BM_ERR_HIP, "sync failed");
        if (rc != BM_OK) return rc;
    }
    /* a table wider than one LDS window is built window by window; the
     * work list takes each female row's part in a window from boundaries
     * found once here (they depend only on the female graph) */
    {
        const int64_t stride = bm_gain_stride(c->n);
        const int32_t cols = bm_gain_chunk_cols(stride);
        const int32_t nwin = (int32_t)((stride + cols - 1) / cols);
        if (nwin > 1 && !c->d_f_in_bnd) {
            const size_t nb = ((size_t)c->n + 1) * (size_t)(nwin + 1);
            rc = dalloc(c, &c->d_f_in_bnd, nb);
            if (rc == BM_OK) rc = dalloc(c, &c->d_f_out_bnd, nb);
            if (rc != BM_OK) return rc;
            HIPCHK(c, bm_launch_row_bounds(c->d_f_in_ptr, c->d_f_in, c->n, cols, nwin, c->d_f_in_bnd, c->stream));
            HIPCHK(c, bm_launch_row_bounds(c->d_f_out_ptr, c->d_f_out, c->n, cols, nwin, c->d_f_out_bnd, c->stream));
        }
        c->nwin = nwin;
        /* a table built in column windows (long rows, one block per CU) walks
         * its rows in fewer, larger chunks: 8 entries per lane, two chunks in
         * flight (n = 100,000: 80.3 ms against 95.3 ms with 4 entries and four
         * chunks, which suit the single window of n = 18,524) */
        if (!c->user_tuned[BM_KERNEL_GAIN_BUILD]) c->tune[BM_KERNEL_GAIN_BUILD].unroll = nwin > 1 ? 8 : 4;
    }
    /* dense male adjacency for the table lookups' edge tests, when it fits
     * 2 GiB (43 MB at n = 18,524); otherwise they probe the hash tables */
    if (!c->d_m_adj) {
        const int64_t words = ((int64_t)c->n + 1 + 31) / 32;
        const size_t abytes = (size_t)((int64_t)c->n + 1) * (size_t)words * sizeof(uint32_t);
        if (abytes <= ((size_t)2 << 30) && bytes + abytes + ((size_t)1 << 30) <= free_b) {
            rc = dalloc(c, &c->d_m_adj, abytes / sizeof(uint32_t));
            if (rc != BM_OK) return rc;
            c->m_adj_words = words;
            HIPCHK(c, hipMemsetAsync(c->d_m_adj, 0, abytes, c->stream));
            HIPCHK(c, bm_launch_adj_bits(devgraph(c), c->d_m_adj, words, c->num_cus, c->stream));
        }
    }
    /* 16 rows of padding past row n: the dense tile pass stores whole tiles of
     * 16 male rows */
    rc = dalloc(c, &c->d_G, bytes / sizeof(int32_t) + (size_t)16 * (size_t)bm_gain_stride(c->n));
    if (rc != BM_OK) return rc;
    /* row 0 is never written by the build; clear it once */
    HIPCHK(c, hipMemsetAsync(c->d_G, 0, (size_t)bm_gain_stride(c->n) * sizeof(int32_t), c->stream));
    c->gain_on = 1;
    rc = gain_rebuild(c);
    if (rc != BM_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return BM_OK;
}

extern "C" int bm_gain_rebuild(bm_ctx* c) {
    if (!c) return BM_ERR_ARG;
    if (!c->gain_on) return fail(c, BM_ERR_STATE, "bm_gain_rebuild: gain table not enabled");
    HIPCHK(c, hipSetDevice(c->device));
    return gain_rebuild(c);
}

extern "C" int bm_gain_row(bm_ctx* c, int32_t u, int32_t* row) {
    if (!c || !row || u < 0 || u > c->n) return fail(c, BM_ERR_ARG, "bm_gain_row: invalid argument");
    if (!c->gain_on) return fail(c, BM_ERR_STATE, "bm_gain_row: gain table not enabled");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(row, c->d_G + (int64_t)u * bm_gain_stride(c->n), ((size_t)c->n + 1) * sizeof(int32_t),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return BM_OK;
}

/* ------------------------------------------------------------------ */
/* Measurement                                                         */
/* ------------------------------------------------------------------ */

extern "C" int bm_timing_enable(bm_ctx* c, int on) {
    if (!c) return BM_ERR_ARG;
    c->timing_on = on ? 1 : 0;
    return BM_OK;
}

extern "C" int bm_timing_get(bm_ctx* c, int32_t kernel, int64_t* launches, double* total_ms) {
    if (!c || kernel < 0 || kernel >= BM_KERNEL_COUNT) return fail(c, BM_ERR_ARG, "bm_timing_get: bad kernel id");
    HIPCHK(c, hipSetDevice(c->device));
    int rc = timing_resolve(c);
    if (rc != BM_OK) return rc;
    if (launches) *launches = c->launches[kernel];
    if (total_ms) *total_ms = c->total_ms[kernel];
    return BM_OK;
}

extern "C" int bm_timing_reset(bm_ctx* c) {
    if (!c) return BM_ERR_ARG;
    int rc = timing_resolve(c);
    if (rc != BM_OK) return rc;
    for (int i = 0; i < BM_KERNEL_COUNT; i++) { c->launches[i] = 0; c->total_ms[i] = 0; }
    return BM_OK;
}

extern "C" int bm_sync(bm_ctx* c) {
    if (!c) return BM_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return BM_OK;
}

static inline int64_t vdeg(const bm_graph_host& g, int32_t x) {
    return (g.out.ptr[x + 1] - g.out.ptr[x]) + (g.in.ptr[x + 1] - g.in.ptr[x]);
}

extern "C" int bm_batch_algo_bytes(bm_batch* b, int32_t generated, int64_t* bytes, int64_t* sum_degree) {
    if (!b) return BM_ERR_ARG;
    bm_ctx* c = b->ctx;
    if (b->kind != 1 && b->kind != 2) return fail(c, BM_ERR_STATE, "bm_batch_algo_bytes: batch holds no candidates");
    const int64_t k = b->k;
    const int w = b->kind == 1 ? 2 : 4;
    std::vector<int32_t> h((size_t)k * w);
    if (k) {
        HIPCHK(c, hipMemcpy(h.data(), b->d_cand, (size_t)k * w * sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    int64_t D = 0;
    for (int64_t i = 0; i < k; i++) {
        const int32_t* q = &h[(size_t)i * w];
        D += vdeg(c->male, q[0]) + vdeg(c->male, q[1]);
        if (w == 4) D += vdeg(c->male, q[2]);
    }
    int64_t B;
    if (b->kind == 1)
        B = k * (48 + (generated ? 0 : 8)) + 28 * D;
    else
        B = k * (69 + (generated ? 0 : 12)) + 36 * D;
    if (bytes) *bytes = B;
    if (sum_degree) *sum_degree = D;
    return BM_OK;
}

extern "C" int bm_get_tuning(const bm_ctx* c, int32_t kernel, bm_tuning* t) {
    if (!c || !t || kernel < 0 || kernel >= BM_KERNEL_COUNT) return BM_ERR_ARG;
    t->waves_per_block = c->tune[kernel].waves_per_block;
    t->lds_slots = c->tune[kernel].lds_slots;
    t->blocks_per_cu = c->tune[kernel].blocks_per_cu;
    t->group_lanes = c->group_lanes[kernel];
    t->unroll = c->tune[kernel].unroll;
    return BM_OK;
}

extern "C" int bm_set_tuning(bm_ctx* c, int32_t kernel, const bm_tuning* t) {
    if (!c || !t || kernel < 0 || kernel >= BM_KERNEL_COUNT) return BM_ERR_ARG;
    const int w = t->waves_per_block;
    if (w != 1 && w != 2 && w != 4 && w != 8) return fail(c, BM_ERR_ARG, "bm_set_tuning: waves_per_block must be 1,2,4,8");
    const size_t sb = c->compact ? 4 : 8;
    if (t->lds_slots < 4 || (t->lds_slots & 3) || (size_t)w * (size_t)t->lds_slots * sb > 160 * 1024)
        return fail(c, BM_ERR_ARG, "bm_set_tuning: lds_slots must be a multiple of 4 and waves*slots*slot bytes <= 160 KiB");
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > 64) return fail(c, BM_ERR_ARG, "bm_set_tuning: blocks_per_cu in 1..64");
    if (t->group_lanes != 64) return fail(c, BM_ERR_ARG, "bm_set_tuning: group_lanes other than 64 not built");
    const int cus = c->num_cus > 0 ? c->num_cus : 256;
    if (kernel == BM_KERNEL_GAIN_BUILD) {
        if (t->unroll != 4 && t->unroll != 8 && t->unroll != 16 && t->unroll != 32)
            return fail(c, BM_ERR_ARG, "bm_set_tuning: table-build unroll must be 4, 8, 16 or 32");
    } else if (t->unroll != 2 && t->unroll != 4 && t->unroll != 8) {
        return fail(c, BM_ERR_ARG, "bm_set_tuning: unroll must be 2, 4 or 8");
    }
    c->tune[kernel] = BmLaunch{w, t->lds_slots, t->blocks_per_cu, cus, t->unroll};
    c->group_lanes[kernel] = t->group_lanes;
    c->user_tuned[kernel] = 1;
    return BM_OK;
}

extern "C" int bm_device_info(bm_ctx* c, char* buf, size_t len) {
    if (!c || !buf || len == 0) return BM_ERR_ARG;
    snprintf(buf, len, "%s, %d CUs", c->arch, c->num_cus);
    return BM_OK;
}

extern "C" const bm_graph_host* bm_ctx_male(const bm_ctx* c) { return c ? &c->male : nullptr; }

BmDevGraph bm_ctx_devgraph(const bm_ctx* c) { return devgraph(c); }
int bm_ctx_set_error(bm_ctx* c, int code, const char* msg) { return fail(c, code, "%s", msg); }
int2* bm_batch_pairs_dev(bm_batch* b) { return (int2*)b->d_cand; }
int64_t* bm_batch_deltas_dev(bm_batch* b) { return b->d_delta; }
int64_t bm_batch_capacity(const bm_batch* b) { return b->cap; }
int bm_batch_has_pair_deltas(const bm_batch* b) { return b->kind == 1 && b->evaluated; }
void bm_batch_mark_pairs(bm_batch* b, int64_t k) {
    b->k = k;
    b->kind = 1;
    b->generated = 1;
    b->evaluated = 0;
}
int bm_ctx_device(const bm_ctx* c) { return c->device; }
int bm_ctx_num_cus(const bm_ctx* c) { return c->num_cus; }
hipStream_t bm_ctx_stream(const bm_ctx* c) { return c->stream; }
BmLaunch bm_ctx_launch(const bm_ctx* c, int kernel) { return c->tune[kernel]; }

extern "C" int bm_ctx_fail(bm_ctx* c, int code, const char* msg) { return fail(c, code, "%s", msg); }
