/*
 * Annealing chains (include/brainmatch.h, bm_anneal_*): host side.
 *
 * Layout in device memory:
 *   insts   one BmDevGraph per instance (the contexts' resident graphs, not
 *           copied);
 *   chains  one BmChain per chain (running score, best score, step and
 *           acceptance counters, stream seed, offset of its matchings);
 *   pool    per chain the current matching and the best-so-far copy,
 *           2 * (n_i + 1) int32 each, chains of one instance contiguous;
 *   traj    chains x traj_cap int64 score samples, one per launch;
 *   props   one BmProposalPool per instance, pointing at that instance's
 *           proposal pairs (a device buffer owned here, grown on demand).
 * A launch advances every chain by a fixed number of steps (k_anneal in
 * bm_delta.hip); it is captured once into a HIP graph and replayed.  The
 * chain state lives only on the device, so a replay needs no host input.
 */
#include "bm_device.h"
#include "bm_host.h"
#include "bm_rng.h"

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

extern "C" const char* bm_global_error(void);

struct bm_anneal {
    int device;
    hipStream_t stream;
    int32_t n_inst;
    int64_t per_inst;
    int64_t n_chains;
    std::vector<bm_ctx*> ctxs;
    std::vector<BmDevGraph> graphs;   /* host copies of the instance graphs */
    std::vector<int64_t> pi_off;      /* per chain */
    std::vector<BmChain> h_init;      /* chain states at creation (for bm_anneal_reset) */
    int32_t* d_init;                  /* each instance's initial matching, concatenated */
    std::vector<int64_t> init_off;    /* per instance offset into d_init */
    BmDevGraph* d_insts;
    BmChain* d_chains;
    int32_t* d_pool;
    int64_t* d_traj;
    int64_t* d_scores;                /* n_chains, for bm_anneal_verify */
    BmProposalPool* d_props;          /* n_inst pool descriptors */
    std::vector<BmProposalPool> h_props;
    std::vector<int2*> prop_buf;      /* per instance, capacity prop_cap */
    std::vector<int64_t> prop_cap;
    BmAnnealSched sc;
    BmLaunch L;
    int compact;
    dim3 grid, block;
    size_t lds;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int64_t graph_steps;
    int64_t launches;
    double total_ms;
    uint64_t seed;
    double p_rot3;
    char err[512];
};

static int afail(bm_anneal* a, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (a)
        snprintf(a->err, sizeof a->err, "%s", buf);
    else
        bm_set_global_error("%s", buf);
    return code;
}

#define AHIP(a, expr)                                                                                        \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            return afail((a), BM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                         __LINE__);                                                                          \
    } while (0)

static void free_anneal(bm_anneal* a) {
    if (!a) return;
    (void)hipSetDevice(a->device);
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    if (a->exec) (void)hipGraphExecDestroy(a->exec);
    if (a->graph) (void)hipGraphDestroy(a->graph);
    (void)hipFree(a->d_insts);
    (void)hipFree(a->d_chains);
    (void)hipFree(a->d_pool);
    (void)hipFree(a->d_traj);
    (void)hipFree(a->d_scores);
    (void)hipFree(a->d_init);
    (void)hipFree(a->d_props);
    for (int2* p : a->prop_buf) (void)hipFree(p);
    if (a->stream) (void)hipStreamDestroy(a->stream);
    delete a;
}

extern "C" void bm_anneal_destroy(bm_anneal* a) { free_anneal(a); }

extern "C" const char* bm_anneal_last_error(const bm_anneal* a) {
    return a ? a->err : bm_global_error();
}

extern "C" int64_t bm_anneal_num_chains(const bm_anneal* a) { return a ? a->n_chains : 0; }

extern "C" int bm_anneal_create(bm_anneal** out, bm_ctx* const* ctxs, int32_t n_inst, const bm_anneal_params* p) {
    if (!out) return afail(nullptr, BM_ERR_ARG, "bm_anneal_create: out is NULL");
    *out = nullptr;
    if (!ctxs || !p || n_inst < 1) return afail(nullptr, BM_ERR_ARG, "bm_anneal_create: need >= 1 context and params");
    if (p->chains_per_instance < 1) return afail(nullptr, BM_ERR_ARG, "bm_anneal_create: chains_per_instance must be >= 1");
    if (!(p->t0 > 0.0) || !isfinite(p->t0)) return afail(nullptr, BM_ERR_ARG, "bm_anneal_create: t0 must be finite and > 0");
    if (!(p->decay > 0.0 && p->decay <= 1.0)) return afail(nullptr, BM_ERR_ARG, "bm_anneal_create: decay must be in (0, 1]");
    if (!(p->p_rot3 >= 0.0 && p->p_rot3 <= 1.0)) return afail(nullptr, BM_ERR_ARG, "bm_anneal_create: p_rot3 must be in [0, 1]");
    if (p->traj_cap < 0) return afail(nullptr, BM_ERR_ARG, "bm_anneal_create: traj_cap must be >= 0");
    if (!(p->p_pool >= 0.0 && p->p_pool <= 1.0)) return afail(nullptr, BM_ERR_ARG, "bm_anneal_create: p_pool must be in [0, 1]");
    for (int32_t i = 0; i < n_inst; i++)
        if (!ctxs[i]) return afail(nullptr, BM_ERR_ARG, "bm_anneal_create: context %d is NULL", i);
    const int dev = bm_ctx_device(ctxs[0]);
    for (int32_t i = 1; i < n_inst; i++)
        if (bm_ctx_device(ctxs[i]) != dev)
            return afail(nullptr, BM_ERR_ARG, "bm_anneal_create: all contexts must be on one device");

    bm_anneal* a = new bm_anneal();
    memset(a->err, 0, sizeof a->err);
    a->device = dev;
    a->n_inst = n_inst;
    a->per_inst = p->chains_per_instance;
    a->n_chains = (int64_t)n_inst * p->chains_per_instance;
    a->seed = p->seed;
    a->p_rot3 = p->p_rot3;
    a->ctxs.assign(ctxs, ctxs + n_inst);
    int rc = BM_OK;
#define ACHK(expr)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) {                                                                       \
            rc = afail(nullptr, e_ == hipErrorOutOfMemory ? BM_ERR_NOMEM : BM_ERR_HIP, "%s failed: %s", \
                       #expr, hipGetErrorString(e_));                                                 \
            free_anneal(a);                                                                           \
            return rc;                                                                                \
        }                                                                                             \
    } while (0)
    ACHK(hipSetDevice(dev));
    ACHK(hipStreamCreate(&a->stream));

    /* initial scores: one device recount per instance, on its own stream */
    std::vector<int64_t> s0(n_inst);
    for (int32_t i = 0; i < n_inst; i++) {
        int r = bm_score(ctxs[i], &s0[i]);
        if (r != BM_OK) {
            rc = afail(nullptr, r, "bm_anneal_create: score of instance %d failed: %s", i, bm_last_error(ctxs[i]));
            free_anneal(a);
            return rc;
        }
    }

    a->compact = 1;
    for (int32_t i = 0; i < n_inst; i++) {
        a->graphs.push_back(bm_ctx_devgraph(ctxs[i]));
        a->compact &= a->graphs.back().compact ? 1 : 0;
    }
    /* launch shape: the swap kernel's tuning of the first context, the LDS
     * slice kept at the same byte size when the slot format widens */
    a->L = bm_ctx_launch(ctxs[0], BM_KERNEL_SWAP);
    if (!a->compact && a->graphs[0].compact) a->L.lds_slots /= 2;

    int64_t pool = 0;
    a->pi_off.resize((size_t)a->n_chains);
    std::vector<BmChain> h((size_t)a->n_chains);
    for (int32_t i = 0; i < n_inst; i++) {
        const int64_t len = (int64_t)a->graphs[i].n + 1;
        for (int64_t j = 0; j < a->per_inst; j++) {
            const int64_t c = (int64_t)i * a->per_inst + j;
            a->pi_off[c] = pool;
            BmChain& ch = h[c];
            memset(&ch, 0, sizeof ch);
            ch.score = s0[i];
            ch.best = s0[i];
            ch.pi_off = pool;
            ch.seed = bm_rand_at(p->seed, (uint64_t)c);
            ch.inst = i;
            ch.at_best = 1;
            pool += 2 * len;
        }
    }
    int64_t init_len = 0;
    for (int32_t i = 0; i < n_inst; i++) {
        a->init_off.push_back(init_len);
        init_len += (int64_t)a->graphs[i].n + 1;
    }
    ACHK(hipMalloc((void**)&a->d_init, sizeof(int32_t) * (size_t)init_len));
    for (int32_t i = 0; i < n_inst; i++)
        ACHK(hipMemcpyAsync(a->d_init + a->init_off[i], a->graphs[i].pi, sizeof(int32_t) * ((size_t)a->graphs[i].n + 1),
                            hipMemcpyDeviceToDevice, a->stream));
    a->h_init = h;
    ACHK(hipMalloc((void**)&a->d_insts, sizeof(BmDevGraph) * (size_t)n_inst));
    ACHK(hipMalloc((void**)&a->d_chains, sizeof(BmChain) * (size_t)a->n_chains));
    ACHK(hipMalloc((void**)&a->d_pool, sizeof(int32_t) * (size_t)pool));
    ACHK(hipMalloc((void**)&a->d_scores, sizeof(int64_t) * (size_t)a->n_chains));
    a->h_props.assign((size_t)n_inst, BmProposalPool{nullptr, 0});
    a->prop_buf.assign((size_t)n_inst, nullptr);
    a->prop_cap.assign((size_t)n_inst, 0);
    ACHK(hipMalloc((void**)&a->d_props, sizeof(BmProposalPool) * (size_t)n_inst));
    ACHK(hipMemcpyAsync(a->d_props, a->h_props.data(), sizeof(BmProposalPool) * (size_t)n_inst, hipMemcpyHostToDevice,
                        a->stream));
    if (p->traj_cap > 0) ACHK(hipMalloc((void**)&a->d_traj, sizeof(int64_t) * (size_t)a->n_chains * p->traj_cap));
    ACHK(hipMemcpyAsync(a->d_insts, a->graphs.data(), sizeof(BmDevGraph) * (size_t)n_inst, hipMemcpyHostToDevice,
                        a->stream));
    ACHK(hipMemcpyAsync(a->d_chains, h.data(), sizeof(BmChain) * (size_t)a->n_chains, hipMemcpyHostToDevice,
                        a->stream));
    for (int32_t i = 0; i < n_inst; i++) {
        const size_t bytes = sizeof(int32_t) * ((size_t)a->graphs[i].n + 1);
        for (int64_t j = 0; j < a->per_inst; j++) {
            const int64_t off = a->pi_off[(size_t)i * a->per_inst + j];
            ACHK(hipMemcpyAsync(a->d_pool + off, a->graphs[i].pi, bytes, hipMemcpyDeviceToDevice, a->stream));
            ACHK(hipMemcpyAsync(a->d_pool + off + a->graphs[i].n + 1, a->graphs[i].pi, bytes, hipMemcpyDeviceToDevice,
                                a->stream));
        }
    }
    a->sc.beta0 = 1.0 / p->t0;
    a->sc.lam = -log(p->decay);
    a->sc.p_rot3 = (uint64_t)llround(p->p_rot3 * 4294967296.0);
    a->sc.traj_cap = p->traj_cap;
    a->sc.pad = 0;
    a->sc.p_pool = (uint64_t)llround(p->p_pool * 4294967296.0);
    ACHK(bm_anneal_grid(a->compact, a->sc.p_rot3 != 0, a->L, a->n_chains, &a->grid, &a->block, &a->lds));
    ACHK(hipStreamSynchronize(a->stream));
#undef ACHK
    *out = a;
    return BM_OK;
}

extern "C" int bm_anneal_run(bm_anneal* a, int64_t steps, int32_t launches) {
    if (!a) return afail(nullptr, BM_ERR_ARG, "bm_anneal_run: NULL object");
    if (steps < 1 || launches < 0) return afail(a, BM_ERR_ARG, "bm_anneal_run: steps must be >= 1 and launches >= 0");
    AHIP(a, hipSetDevice(a->device));
    if (!a->exec || a->graph_steps != steps) {
        if (a->exec) (void)hipGraphExecDestroy(a->exec);
        if (a->graph) (void)hipGraphDestroy(a->graph);
        a->exec = nullptr;
        a->graph = nullptr;
        AHIP(a, hipStreamBeginCapture(a->stream, hipStreamCaptureModeThreadLocal));
        hipError_t le = bm_launch_anneal(a->d_insts, a->compact, a->d_chains, a->n_chains, a->d_pool, steps, a->sc,
                                         a->d_traj, a->d_props, a->L, a->grid, a->block, a->lds, a->stream);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(a->stream, &g);
        if (le != hipSuccess || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return afail(a, BM_ERR_HIP, "bm_anneal_run: capture failed: %s / %s", hipGetErrorString(le),
                         hipGetErrorString(ce));
        }
        a->graph = g;
        AHIP(a, hipGraphInstantiate(&a->exec, a->graph, nullptr, nullptr, 0));
        a->graph_steps = steps;
    }
    if (launches == 0) return BM_OK;
    std::vector<hipEvent_t> ev((size_t)launches + 1, nullptr);
    int rc = BM_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) { rc = afail(a, BM_ERR_HIP, "bm_anneal_run: hipEventCreate failed"); break; }
    if (rc == BM_OK) {
        hipError_t e = hipEventRecord(ev[0], a->stream);
        for (int32_t t = 0; t < launches && e == hipSuccess; t++) {
            e = hipGraphLaunch(a->exec, a->stream);
            if (e == hipSuccess) e = hipEventRecord(ev[(size_t)t + 1], a->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(a->stream);
        if (e != hipSuccess) rc = afail(a, BM_ERR_HIP, "bm_anneal_run: launch failed: %s", hipGetErrorString(e));
        if (rc == BM_OK) {
            for (int32_t t = 0; t < launches; t++) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, ev[t], ev[(size_t)t + 1]) == hipSuccess) a->total_ms += ms;
            }
            a->launches += launches;
        }
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

static int fetch_chains(bm_anneal* a, std::vector<BmChain>& h) {
    h.resize((size_t)a->n_chains);
    AHIP(a, hipSetDevice(a->device));
    AHIP(a, hipMemcpyAsync(h.data(), a->d_chains, sizeof(BmChain) * (size_t)a->n_chains, hipMemcpyDeviceToHost,
                           a->stream));
    AHIP(a, hipStreamSynchronize(a->stream));
    return BM_OK;
}

extern "C" int bm_anneal_state(bm_anneal* a, int64_t* score, int64_t* best, int64_t* steps, int64_t* accepted) {
    if (!a) return afail(nullptr, BM_ERR_ARG, "bm_anneal_state: NULL object");
    std::vector<BmChain> h;
    int rc = fetch_chains(a, h);
    if (rc != BM_OK) return rc;
    for (int64_t c = 0; c < a->n_chains; c++) {
        if (score) score[c] = h[c].score;
        if (best) best[c] = h[c].best;
        if (steps) steps[c] = h[c].step;
        if (accepted) accepted[c] = h[c].accepted;
    }
    return BM_OK;
}

extern "C" int bm_anneal_mapping(bm_anneal* a, int64_t chain, int32_t which, int32_t* pi) {
    if (!a || !pi) return afail(a, BM_ERR_ARG, "bm_anneal_mapping: null argument");
    if (chain < 0 || chain >= a->n_chains || (which != 0 && which != 1))
        return afail(a, BM_ERR_ARG, "bm_anneal_mapping: chain out of range or which not 0/1");
    std::vector<BmChain> h;
    int rc = fetch_chains(a, h);
    if (rc != BM_OK) return rc;
    const BmChain& ch = h[(size_t)chain];
    const int64_t len = (int64_t)a->graphs[ch.inst].n + 1;
    const int64_t off = ch.pi_off + ((which == 1 && !ch.at_best) ? len : 0);
    AHIP(a, hipMemcpyAsync(pi, a->d_pool + off, sizeof(int32_t) * (size_t)len, hipMemcpyDeviceToHost, a->stream));
    AHIP(a, hipStreamSynchronize(a->stream));
    return BM_OK;
}

extern "C" int bm_anneal_trajectory(bm_anneal* a, int64_t* traj, int32_t* count) {
    if (!a) return afail(nullptr, BM_ERR_ARG, "bm_anneal_trajectory: NULL object");
    std::vector<BmChain> h;
    int rc = fetch_chains(a, h);
    if (rc != BM_OK) return rc;
    const int32_t cnt = a->n_chains ? (h[0].nseg < a->sc.traj_cap ? h[0].nseg : a->sc.traj_cap) : 0;
    if (count) *count = cnt;
    if (traj && a->d_traj && a->sc.traj_cap > 0) {
        AHIP(a, hipMemcpyAsync(traj, a->d_traj, sizeof(int64_t) * (size_t)a->n_chains * a->sc.traj_cap,
                               hipMemcpyDeviceToHost, a->stream));
        AHIP(a, hipStreamSynchronize(a->stream));
    }
    return BM_OK;
}

extern "C" int bm_anneal_verify(bm_anneal* a, int64_t* mismatches) {
    if (!a || !mismatches) return afail(a, BM_ERR_ARG, "bm_anneal_verify: null argument");
    std::vector<BmChain> h;
    int rc = fetch_chains(a, h);
    if (rc != BM_OK) return rc;
    AHIP(a, hipMemsetAsync(a->d_scores, 0, sizeof(int64_t) * (size_t)a->n_chains, a->stream));
    for (int64_t c = 0; c < a->n_chains; c++) {
        BmDevGraph g = a->graphs[h[c].inst];
        g.pi = a->d_pool + h[c].pi_off;
        AHIP(a, bm_launch_score(g, bm_ctx_launch(a->ctxs[h[c].inst], BM_KERNEL_SCORE), a->d_scores + c, a->stream));
    }
    std::vector<int64_t> s((size_t)a->n_chains);
    AHIP(a, hipMemcpyAsync(s.data(), a->d_scores, sizeof(int64_t) * (size_t)a->n_chains, hipMemcpyDeviceToHost,
                           a->stream));
    AHIP(a, hipStreamSynchronize(a->stream));
    int64_t bad = 0;
    for (int64_t c = 0; c < a->n_chains; c++) bad += s[c] != h[c].score;
    *mismatches = bad;
    return BM_OK;
}

extern "C" int bm_anneal_timing(bm_anneal* a, int64_t* launches, double* total_ms) {
    if (!a) return afail(nullptr, BM_ERR_ARG, "bm_anneal_timing: NULL object");
    if (launches) *launches = a->launches;
    if (total_ms) *total_ms = a->total_ms;
    return BM_OK;
}

extern "C" int bm_anneal_algo_bytes(bm_anneal* a, int64_t* bytes, int64_t* steps) {
    if (!a) return afail(nullptr, BM_ERR_ARG, "bm_anneal_algo_bytes: NULL object");
    std::vector<BmChain> h;
    int rc = fetch_chains(a, h);
    if (rc != BM_OK) return rc;
    /* swap model 48 + 28 D per swap step, triple model 69 + 36 D3 per
     * 3-vertex step, from the degree sums the chains keep */
    int64_t B = 0, S = 0;
    for (int64_t c = 0; c < a->n_chains; c++) {
        const BmChain& ch = h[(size_t)c];
        B += 48 * (ch.step - ch.n_rot3) + 28 * ch.deg_swap + 69 * ch.n_rot3 + 36 * ch.deg_rot3;
        S += ch.step;
    }
    if (bytes) *bytes = B;
    if (steps) *steps = S;
    return BM_OK;
}

extern "C" int bm_anneal_move_counts(bm_anneal* a, int64_t* rot3, int64_t* pooled) {
    if (!a) return afail(nullptr, BM_ERR_ARG, "bm_anneal_move_counts: NULL object");
    std::vector<BmChain> h;
    int rc = fetch_chains(a, h);
    if (rc != BM_OK) return rc;
    for (int64_t c = 0; c < a->n_chains; c++) {
        if (rot3) rot3[c] = h[(size_t)c].n_rot3;
        if (pooled) pooled[c] = h[(size_t)c].n_learned;
    }
    return BM_OK;
}

/* Device buffer of instance inst's pool holding at least k pairs. */
static int pool_reserve(bm_anneal* a, int32_t inst, int64_t k) {
    if (a->prop_cap[(size_t)inst] >= k) return BM_OK;
    AHIP(a, hipStreamSynchronize(a->stream));
    (void)hipFree(a->prop_buf[(size_t)inst]);
    a->prop_buf[(size_t)inst] = nullptr;
    a->prop_cap[(size_t)inst] = 0;
    a->h_props[(size_t)inst] = BmProposalPool{nullptr, 0};
    hipError_t e = hipMalloc((void**)&a->prop_buf[(size_t)inst], sizeof(int2) * (size_t)k);
    if (e != hipSuccess)
        return afail(a, e == hipErrorOutOfMemory ? BM_ERR_NOMEM : BM_ERR_HIP, "bm_anneal pool: hipMalloc failed: %s",
                     hipGetErrorString(e));
    a->prop_cap[(size_t)inst] = k;
    return BM_OK;
}

static int pool_publish(bm_anneal* a, int32_t inst, int64_t k) {
    a->h_props[(size_t)inst] = BmProposalPool{k ? a->prop_buf[(size_t)inst] : nullptr, k};
    AHIP(a, hipMemcpyAsync(a->d_props + inst, &a->h_props[(size_t)inst], sizeof(BmProposalPool),
                           hipMemcpyHostToDevice, a->stream));
    AHIP(a, hipStreamSynchronize(a->stream));
    return BM_OK;
}

extern "C" int bm_anneal_set_pool(bm_anneal* a, int32_t inst, const int32_t* uv, int64_t k) {
    if (!a) return afail(nullptr, BM_ERR_ARG, "bm_anneal_set_pool: NULL object");
    if (inst < 0 || inst >= a->n_inst || k < 0 || k > (1LL << 31) || (k > 0 && !uv))
        return afail(a, BM_ERR_ARG, "bm_anneal_set_pool: instance out of range or k outside 0..2^31");
    const int32_t n = a->graphs[(size_t)inst].n;
    for (int64_t i = 0; i < k; i++)
        if (uv[2 * i] < 1 || uv[2 * i] > n || uv[2 * i + 1] < 1 || uv[2 * i + 1] > n || uv[2 * i] == uv[2 * i + 1])
            return afail(a, BM_ERR_ARG, "bm_anneal_set_pool: pair %lld is not two distinct vertices in 1..n", (long long)i);
    AHIP(a, hipSetDevice(a->device));
    int rc = pool_reserve(a, inst, k);
    if (rc != BM_OK) return rc;
    if (k) AHIP(a, hipMemcpyAsync(a->prop_buf[(size_t)inst], uv, sizeof(int2) * (size_t)k, hipMemcpyHostToDevice,
                                 a->stream));
    return pool_publish(a, inst, k);
}

extern "C" int bm_anneal_pool_learned(bm_anneal* a, int32_t inst, bm_mlp* m, uint64_t seed, int64_t pool, int64_t top) {
    if (!a || !m) return afail(a, BM_ERR_ARG, "bm_anneal_pool_learned: null argument");
    if (inst < 0 || inst >= a->n_inst || top < 1 || top > pool || top > (1LL << 31))
        return afail(a, BM_ERR_ARG, "bm_anneal_pool_learned: instance out of range or not 1 <= top <= pool");
    if (bm_mlp_context(m) != a->ctxs[(size_t)inst])
        return afail(a, BM_ERR_ARG, "bm_anneal_pool_learned: the network belongs to another instance's context");
    AHIP(a, hipSetDevice(a->device));
    int rc = pool_reserve(a, inst, top);
    if (rc != BM_OK) return rc;
    rc = bm_mlp_rank_into(m, seed, pool, top, a->prop_buf[(size_t)inst]);
    if (rc != BM_OK) return afail(a, rc, "bm_anneal_pool_learned: %s", bm_last_error(a->ctxs[(size_t)inst]));
    return pool_publish(a, inst, top);
}

extern "C" int bm_anneal_get_pool(bm_anneal* a, int32_t inst, int32_t* uv, int64_t* k) {
    if (!a || !k) return afail(a, BM_ERR_ARG, "bm_anneal_get_pool: null argument");
    if (inst < 0 || inst >= a->n_inst) return afail(a, BM_ERR_ARG, "bm_anneal_get_pool: instance out of range");
    const int64_t cnt = a->h_props[(size_t)inst].count;
    if (uv && cnt) {
        AHIP(a, hipSetDevice(a->device));
        AHIP(a, hipMemcpyAsync(uv, a->prop_buf[(size_t)inst], sizeof(int2) * (size_t)cnt, hipMemcpyDeviceToHost,
                               a->stream));
        AHIP(a, hipStreamSynchronize(a->stream));
    }
    *k = cnt;
    return BM_OK;
}

extern "C" int bm_anneal_schedule(const bm_anneal* a, double* beta0, double* lam, uint64_t* p_rot3) {
    if (!a) return BM_ERR_ARG;
    if (beta0) *beta0 = a->sc.beta0;
    if (lam) *lam = a->sc.lam;
    if (p_rot3) *p_rot3 = a->sc.p_rot3;
    return BM_OK;
}

extern "C" uint64_t bm_anneal_p_pool(const bm_anneal* a) { return a ? a->sc.p_pool : 0; }

/* Chains back to their creation state (matchings, scores, counters). */
static int reset_chains(bm_anneal* a) {
    AHIP(a, hipSetDevice(a->device));
    AHIP(a, hipMemcpyAsync(a->d_chains, a->h_init.data(), sizeof(BmChain) * (size_t)a->n_chains, hipMemcpyHostToDevice,
                           a->stream));
    for (int64_t c = 0; c < a->n_chains; c++) {
        const int32_t i = a->h_init[c].inst;
        const size_t len = (size_t)a->graphs[i].n + 1;
        AHIP(a, hipMemcpyAsync(a->d_pool + a->pi_off[c], a->d_init + a->init_off[i], len * sizeof(int32_t),
                               hipMemcpyDeviceToDevice, a->stream));
        AHIP(a, hipMemcpyAsync(a->d_pool + a->pi_off[c] + len, a->d_init + a->init_off[i], len * sizeof(int32_t),
                               hipMemcpyDeviceToDevice, a->stream));
    }
    AHIP(a, hipStreamSynchronize(a->stream));
    a->launches = 0;
    a->total_ms = 0;
    return BM_OK;
}

extern "C" int bm_anneal_reset(bm_anneal* a) {
    if (!a) return afail(nullptr, BM_ERR_ARG, "bm_anneal_reset: NULL object");
    return reset_chains(a);
}

extern "C" int bm_anneal_get_tuning(const bm_anneal* a, bm_tuning* t) {
    if (!a || !t) return BM_ERR_ARG;
    t->waves_per_block = a->L.waves_per_block;
    t->lds_slots = a->L.lds_slots;
    t->blocks_per_cu = a->L.blocks_per_cu;
    t->group_lanes = 64;
    t->unroll = a->L.unroll == 2 ? 2 : 4;
    return BM_OK;
}

extern "C" int bm_anneal_set_tuning(bm_anneal* a, const bm_tuning* t) {
    if (!a || !t) return afail(a, BM_ERR_ARG, "bm_anneal_set_tuning: null argument");
    const int w = t->waves_per_block;
    const size_t sb = a->compact ? 4 : 8;
    if ((w != 1 && w != 2 && w != 4 && w != 8) || t->lds_slots < 4 || (t->lds_slots & 3) ||
        (size_t)w * (size_t)t->lds_slots * sb > 160 * 1024 || t->blocks_per_cu < 1 || t->blocks_per_cu > 64 ||
        (t->unroll != 2 && t->unroll != 4))
        return afail(a, BM_ERR_ARG, "bm_anneal_set_tuning: waves 1/2/4/8, lds_slots a multiple of 4 within 160 KiB, "
                                    "blocks_per_cu 1..64, unroll 2 or 4");
    BmLaunch L = a->L;
    L.waves_per_block = w;
    L.lds_slots = t->lds_slots;
    L.blocks_per_cu = t->blocks_per_cu;
    L.unroll = t->unroll;
    dim3 grid, block;
    size_t lds;
    AHIP(a, hipSetDevice(a->device));
    AHIP(a, bm_anneal_grid(a->compact, a->sc.p_rot3 != 0, L, a->n_chains, &grid, &block, &lds));
    a->L = L;
    a->grid = grid;
    a->block = block;
    a->lds = lds;
    if (a->exec) (void)hipGraphExecDestroy(a->exec);
    if (a->graph) (void)hipGraphDestroy(a->graph);
    a->exec = nullptr;
    a->graph = nullptr;
    return BM_OK;
}

/* Times one launch of `steps` steps for every launch shape in a small grid
 * (list entries per lane x waves per block x table slots per wave x blocks
 * per CU), keeps the fastest, and resets the chains.  trials (may be NULL)
 * gets up to max_trials records {waves, slots, blocks_per_cu, unroll, ms}. */
extern "C" int bm_anneal_autotune(bm_anneal* a, int64_t steps, bm_tuning* best, double* trials, int32_t max_trials,
                                  int32_t* n_trials) {
    if (!a || steps < 1) return afail(a, BM_ERR_ARG, "bm_anneal_autotune: invalid argument");
    const int waves[4] = {1, 2, 4, 8};
    const int slots_c[3] = {640, 1280, 2560}, slots_w[3] = {320, 640, 1280};
    const int bpc[2] = {4, 16};
    const int unrolls[2] = {2, 4};
    bm_tuning bt;
    memset(&bt, 0, sizeof bt);
    double best_ms = 1e300;
    int32_t nt = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    AHIP(a, hipSetDevice(a->device));
    AHIP(a, hipEventCreate(&e0));
    AHIP(a, hipEventCreate(&e1));
    int rc = BM_OK;
    for (int ui = 0; ui < 2 && rc == BM_OK; ui++)
    for (int wi = 0; wi < 4 && rc == BM_OK; wi++)
        for (int si = 0; si < 3 && rc == BM_OK; si++)
            for (int bi = 0; bi < 2 && rc == BM_OK; bi++) {
                bm_tuning t;
                t.waves_per_block = waves[wi];
                t.lds_slots = a->compact ? slots_c[si] : slots_w[si];
                t.blocks_per_cu = bpc[bi];
                t.group_lanes = 64;
                t.unroll = unrolls[ui];
                if ((size_t)t.waves_per_block * t.lds_slots * (a->compact ? 4 : 8) > 160 * 1024) continue;
                if ((rc = bm_anneal_set_tuning(a, &t)) != BM_OK) break;
                if ((rc = reset_chains(a)) != BM_OK) break;
                hipError_t e = hipEventRecord(e0, a->stream);
                if (e == hipSuccess)
                    e = bm_launch_anneal(a->d_insts, a->compact, a->d_chains, a->n_chains, a->d_pool, steps, a->sc,
                                         nullptr, a->d_props, a->L, a->grid, a->block, a->lds, a->stream);
                if (e == hipSuccess) e = hipEventRecord(e1, a->stream);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e != hipSuccess) {
                    rc = afail(a, BM_ERR_HIP, "bm_anneal_autotune: trial launch failed: %s", hipGetErrorString(e));
                    break;
                }
                if (trials && nt < max_trials) {
                    trials[5 * nt] = t.waves_per_block;
                    trials[5 * nt + 1] = t.lds_slots;
                    trials[5 * nt + 2] = t.blocks_per_cu;
                    trials[5 * nt + 3] = t.unroll;
                    trials[5 * nt + 4] = ms;
                }
                nt++;
                if (ms < best_ms) {
                    best_ms = ms;
                    bt = t;
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != BM_OK) return rc;
    if (n_trials) *n_trials = nt;
    if ((rc = bm_anneal_set_tuning(a, &bt)) != BM_OK) return rc;
    if ((rc = reset_chains(a)) != BM_OK) return rc;
    if (best) *best = bt;
    return BM_OK;
}
