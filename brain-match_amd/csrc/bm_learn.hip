/*
 * Learned proposal (include/brainmatch.h, bm_mlp_* and
 * bm_batch_gen_learned): host side of the perceptron in bm_mlp.hip.
 *
 * Device memory per network: fp32 parameters, gradient and Adam moments
 * (18,817 floats each), the bf16 weights in MFMA fragment order (68 KB), the
 * per-vertex features of the context's matching ((n+1) x 8 floats), and
 * scratch for training (activations of one batch, targets), explicit pair
 * scoring and pool ranking, grown on demand.
 */
#include "bm_device.h"
#include "bm_host.h"
#include "bm_rng.h"

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

/* Device scratch of bm_batch_gen_learned_noalloc: the pool's scores, the
 * select state and the selected keys for pools of up to BM_NOALLOC_POOL pairs
 * (bm_mlp_select_bytes(BM_NOALLOC_POOL) = 4,325,928 bytes fit).  Each device
 * holds its own copy for the life of the process; host code uses the address
 * hipGetSymbolAddress returns for the device (g_scratch_dev), never the
 * array's name.  Every call on a device shares that device's copy. */
#define BM_NOALLOC_SCRATCH_BYTES ((4u << 20) + (256u << 10))
static __device__ __align__(256) unsigned char scratch[BM_NOALLOC_SCRATCH_BYTES];

static constexpr int kMaxDevices = 64;
static void* g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

/* Device address of the scratch array on device dev (current), taken once
 * per device. */
static hipError_t scratch_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[dev]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[dev] = ptr;
    return hipSuccess;
}

struct bm_mlp {
    bm_ctx* ctx;
    int device;
    int num_cus;
    int32_t n;
    hipStream_t stream;
    BmMlpDev d;
    float* d_vf;
    int4* d_chunks;         /* male edge chunks of the feature pass */
    int64_t nchunks;
    void* d_feat_acc;       /* (n+1) x 32 bytes of feature accumulators, zero between passes */
    int features_ready;
    float* d_loss;
    float* d_target;
    int64_t target_cap;
    void* d_act;
    size_t act_bytes;
    int2* d_pairs;
    int64_t pairs_cap;
    float* d_score;
    int64_t score_cap;
    void* d_rank;
    size_t rank_bytes;
    bm_batch* train_batch;
    int64_t train_cap;
    int64_t adam_step;
};

static int mfail(bm_mlp* m, int code, const char* msg) { return bm_ctx_set_error(m ? m->ctx : nullptr, code, msg); }

#define MHIP(m, expr)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            char b_[256];                                                                         \
            snprintf(b_, sizeof b_, "%s failed: %s", #expr, hipGetErrorString(e_));               \
            return mfail((m), e_ == hipErrorOutOfMemory ? BM_ERR_NOMEM : BM_ERR_HIP, b_);         \
        }                                                                                         \
    } while (0)

template <typename T>
static int grow(bm_mlp* m, T** p, int64_t* cap, int64_t need) {
    if (*cap >= need) return BM_OK;
    MHIP(m, hipStreamSynchronize(m->stream));
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    MHIP(m, hipMalloc((void**)p, (size_t)(need > 0 ? need : 1) * sizeof(T)));
    *cap = need;
    return BM_OK;
}

static int grow_bytes(bm_mlp* m, void** p, size_t* cap, size_t need) {
    if (*cap >= need) return BM_OK;
    MHIP(m, hipStreamSynchronize(m->stream));
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    MHIP(m, hipMalloc(p, need ? need : 1));
    *cap = need;
    return BM_OK;
}

extern "C" void bm_mlp_destroy(bm_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    (void)hipStreamSynchronize(m->stream);
    (void)hipFree(m->d.params);
    (void)hipFree(m->d.grad);
    (void)hipFree(m->d.adam_m);
    (void)hipFree(m->d.adam_v);
    (void)hipFree(m->d.W1f);
    (void)hipFree(m->d.W2f);
    (void)hipFree(m->d.W2Tf);
    (void)hipFree(m->d_vf);
    (void)hipFree(m->d_chunks);
    (void)hipFree(m->d_feat_acc);
    (void)hipFree(m->d_loss);
    (void)hipFree(m->d_target);
    (void)hipFree(m->d_act);
    (void)hipFree(m->d_pairs);
    (void)hipFree(m->d_score);
    (void)hipFree(m->d_rank);
    if (m->train_batch) bm_batch_destroy(m->train_batch);
    delete m;
}

extern "C" int32_t bm_mlp_num_params(void) { return bm_mlp_param_count(); }

/* He-normal weights (Box-Muller over xoshiro256**), zero biases. */
static void init_params(std::vector<float>& p, uint64_t seed) {
    bm_xoshiro rng;
    bm_xoshiro_seed(&rng, seed);
    auto normal = [&rng]() {
        double u1 = bm_xoshiro_unit(&rng), u2 = bm_xoshiro_unit(&rng);
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    };
    const int F = 16, H = 128;
    size_t o = 0;
    for (int i = 0; i < F * H; i++) p[o++] = (float)(normal() * sqrt(2.0 / F));
    for (int i = 0; i < H; i++) p[o++] = 0.f;
    for (int i = 0; i < H * H; i++) p[o++] = (float)(normal() * sqrt(2.0 / H));
    for (int i = 0; i < H; i++) p[o++] = 0.f;
    for (int i = 0; i < H; i++) p[o++] = (float)(normal() * sqrt(1.0 / H));
    p[o++] = 0.f;
}

extern "C" int bm_mlp_set_params(bm_mlp* m, const float* p) {
    if (!m || !p) return mfail(m, BM_ERR_ARG, "bm_mlp_set_params: null argument");
    MHIP(m, hipSetDevice(m->device));
    const size_t np = (size_t)bm_mlp_param_count();
    MHIP(m, hipMemcpyAsync(m->d.params, p, np * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipMemsetAsync(m->d.grad, 0, np * sizeof(float), m->stream));
    MHIP(m, hipMemsetAsync(m->d.adam_m, 0, np * sizeof(float), m->stream));
    MHIP(m, hipMemsetAsync(m->d.adam_v, 0, np * sizeof(float), m->stream));
    MHIP(m, bm_launch_mlp_pack(m->d, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    m->adam_step = 0;
    return BM_OK;
}

extern "C" int bm_mlp_get_params(bm_mlp* m, float* p) {
    if (!m || !p) return mfail(m, BM_ERR_ARG, "bm_mlp_get_params: null argument");
    MHIP(m, hipSetDevice(m->device));
    MHIP(m, hipMemcpyAsync(p, m->d.params, (size_t)bm_mlp_param_count() * sizeof(float), hipMemcpyDeviceToHost,
                           m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return BM_OK;
}

extern "C" int bm_mlp_create(bm_ctx* ctx, uint64_t seed, bm_mlp** out) {
    if (!out || !ctx) return bm_ctx_set_error(ctx, BM_ERR_ARG, "bm_mlp_create: null argument");
    *out = nullptr;
    bm_mlp* m = new bm_mlp();
    memset(m, 0, sizeof *m);
    m->ctx = ctx;
    m->device = bm_ctx_device(ctx);
    m->num_cus = bm_ctx_num_cus(ctx);
    m->n = bm_num_vertices(ctx);
    m->stream = bm_ctx_stream(ctx);
    const size_t np = (size_t)bm_mlp_param_count();
    int rc = BM_OK;
    auto alloc = [&](void** p, size_t bytes) {
        if (rc == BM_OK && hipMalloc(p, bytes) != hipSuccess)
            rc = bm_ctx_set_error(ctx, BM_ERR_NOMEM, "bm_mlp_create: device allocation failed");
    };
    if (hipSetDevice(m->device) != hipSuccess) rc = bm_ctx_set_error(ctx, BM_ERR_HIP, "bm_mlp_create: hipSetDevice");
    if (rc == BM_OK && scratch_setup(m->device) != hipSuccess)
        rc = bm_ctx_set_error(ctx, BM_ERR_HIP, "bm_mlp_create: hipGetSymbolAddress of the ranking scratch failed");
    alloc((void**)&m->d.params, np * sizeof(float));
    alloc((void**)&m->d.grad, np * sizeof(float));
    alloc((void**)&m->d.adam_m, np * sizeof(float));
    alloc((void**)&m->d.adam_v, np * sizeof(float));
    alloc(&m->d.W1f, 16 * 128 * 2);
    alloc(&m->d.W2f, 128 * 128 * 2);
    alloc(&m->d.W2Tf, 128 * 128 * 2);
    alloc((void**)&m->d_vf, ((size_t)m->n + 1) * 8 * sizeof(float));
    alloc((void**)&m->d_loss, sizeof(float));
    alloc(&m->d_feat_acc, ((size_t)m->n + 1) * 32);
    if (rc == BM_OK && hipMemset(m->d_feat_acc, 0, ((size_t)m->n + 1) * 32) != hipSuccess)
        rc = bm_ctx_set_error(ctx, BM_ERR_HIP, "bm_mlp_create: hipMemset failed");
    if (rc == BM_OK) {
        /* chunks of at most bm_feature_chunk_edges() edges of every male list */
        const bm_graph_host* g = bm_ctx_male(ctx);
        const int64_t cap = bm_feature_chunk_edges();
        std::vector<int4> ch;
        for (int32_t x = 1; x <= m->n; x++)
            for (int dir = 0; dir < 2; dir++) {
                const bm_csr_host& L = dir ? g->out : g->in;
                const int64_t deg = L.ptr[x + 1] - L.ptr[x];
                for (int64_t e0 = 0; e0 < deg; e0 += cap)
                    ch.push_back(make_int4(x, dir, (int32_t)e0, (int32_t)(deg - e0 < cap ? deg - e0 : cap)));
            }
        m->nchunks = (int64_t)ch.size();
        alloc((void**)&m->d_chunks, ch.size() * sizeof(int4) + sizeof(int4));
        if (rc == BM_OK && !ch.empty() &&
            hipMemcpy(m->d_chunks, ch.data(), ch.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess)
            rc = bm_ctx_set_error(ctx, BM_ERR_HIP, "bm_mlp_create: chunk upload failed");
    }
    if (rc == BM_OK) {
        std::vector<float> p(np);
        init_params(p, seed);
        rc = bm_mlp_set_params(m, p.data());
    }
    if (rc != BM_OK) {
        bm_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return BM_OK;
}

extern "C" int bm_mlp_refresh_features(bm_mlp* m) {
    if (!m) return BM_ERR_ARG;
    MHIP(m, hipSetDevice(m->device));
    MHIP(m, bm_launch_vertex_features(bm_ctx_devgraph(m->ctx), m->d_chunks, m->nchunks, m->d_feat_acc, m->d_vf,
                                      m->num_cus, m->stream));
    m->features_ready = 1;
    return BM_OK;
}

extern "C" int bm_mlp_vertex_features(bm_mlp* m, float* out) {
    if (!m || !out) return mfail(m, BM_ERR_ARG, "bm_mlp_vertex_features: null argument");
    if (!m->features_ready) return mfail(m, BM_ERR_STATE, "bm_mlp_vertex_features: call bm_mlp_refresh_features first");
    MHIP(m, hipSetDevice(m->device));
    MHIP(m, hipMemcpyAsync(out, m->d_vf, ((size_t)m->n + 1) * 8 * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return BM_OK;
}

static int upload_pairs(bm_mlp* m, const int32_t* uv, int64_t k) {
    for (int64_t i = 0; i < 2 * k; i++)
        if (uv[i] < 1 || uv[i] > m->n) return mfail(m, BM_ERR_ARG, "pair vertex outside 1..n");
    int rc = grow(m, &m->d_pairs, &m->pairs_cap, k);
    if (rc != BM_OK) return rc;
    MHIP(m, hipMemcpyAsync(m->d_pairs, uv, (size_t)k * sizeof(int2), hipMemcpyHostToDevice, m->stream));
    return BM_OK;
}

extern "C" int bm_mlp_score_pairs(bm_mlp* m, const int32_t* uv, int64_t k, float* out) {
    if (!m || k < 0 || (k > 0 && (!uv || !out))) return mfail(m, BM_ERR_ARG, "bm_mlp_score_pairs: invalid argument");
    if (!m->features_ready) return mfail(m, BM_ERR_STATE, "bm_mlp_score_pairs: call bm_mlp_refresh_features first");
    if (k == 0) return BM_OK;
    MHIP(m, hipSetDevice(m->device));
    int rc = upload_pairs(m, uv, k);
    if (rc == BM_OK) rc = grow(m, &m->d_score, &m->score_cap, k);
    if (rc != BM_OK) return rc;
    MHIP(m, bm_launch_mlp_forward(m->d, m->d_vf, m->d_pairs, 0, m->n, k, m->d_score, m->num_cus, m->stream));
    MHIP(m, hipMemcpyAsync(out, m->d_score, (size_t)k * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return BM_OK;
}

/* Gradient of one batch into m->d.grad (which the caller has cleared); the
 * squared error into *d_loss (cleared here). */
static int batch_gradient(bm_mlp* m, const int2* list, uint64_t seed, int64_t k, const float* d_target) {
    int rc = grow_bytes(m, &m->d_act, &m->act_bytes, bm_mlp_train_bytes(k));
    if (rc != BM_OK) return rc;
    MHIP(m, hipMemsetAsync(m->d_loss, 0, sizeof(float), m->stream));
    MHIP(m, bm_launch_mlp_gradient(m->d, m->d_vf, list, seed, m->n, k, d_target, m->d_act, m->d_loss, m->num_cus,
                                   m->stream));
    return BM_OK;
}

extern "C" int bm_mlp_gradient(bm_mlp* m, const int32_t* uv, const float* target, int64_t k, float* grad,
                               float* mse) {
    if (!m || k < 1 || !uv || !target || !grad) return mfail(m, BM_ERR_ARG, "bm_mlp_gradient: invalid argument");
    if (!m->features_ready) return mfail(m, BM_ERR_STATE, "bm_mlp_gradient: call bm_mlp_refresh_features first");
    MHIP(m, hipSetDevice(m->device));
    int rc = upload_pairs(m, uv, k);
    if (rc == BM_OK) rc = grow(m, &m->d_target, &m->target_cap, k);
    if (rc != BM_OK) return rc;
    MHIP(m, hipMemcpyAsync(m->d_target, target, (size_t)k * sizeof(float), hipMemcpyHostToDevice, m->stream));
    const size_t np = (size_t)bm_mlp_param_count();
    MHIP(m, hipMemsetAsync(m->d.grad, 0, np * sizeof(float), m->stream));
    rc = batch_gradient(m, m->d_pairs, 0, k, m->d_target);
    if (rc != BM_OK) return rc;
    float se = 0.f;
    MHIP(m, hipMemcpyAsync(grad, m->d.grad, np * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipMemcpyAsync(&se, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipMemsetAsync(m->d.grad, 0, np * sizeof(float), m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    if (mse) *mse = se / (float)k;
    return BM_OK;
}

extern "C" int bm_mlp_train(bm_mlp* m, uint64_t seed, int64_t steps, int32_t batch, float lr, float* mse_trace) {
    if (!m || steps < 0 || batch < 16 || !(lr > 0.f)) return mfail(m, BM_ERR_ARG, "bm_mlp_train: invalid argument");
    MHIP(m, hipSetDevice(m->device));
    if (!m->features_ready) {
        int rc = bm_mlp_refresh_features(m);
        if (rc != BM_OK) return rc;
    }
    if (!m->train_batch || m->train_cap < batch) {
        if (m->train_batch) bm_batch_destroy(m->train_batch);
        m->train_batch = nullptr;
        int rc = bm_batch_create(m->ctx, batch, &m->train_batch);
        if (rc != BM_OK) return rc;
        m->train_cap = batch;
    }
    int rc = grow(m, &m->d_target, &m->target_cap, batch);
    if (rc != BM_OK) return rc;
    for (int64_t s = 0; s < steps; s++) {
        const uint64_t sd = bm_rand_at(seed, (uint64_t)s);
        /* exact deltas of the batch from the context's own evaluator */
        rc = bm_batch_eval_swaps_uniform(m->train_batch, sd, batch);
        if (rc != BM_OK) return rc;
        MHIP(m, bm_launch_mlp_targets(bm_batch_deltas_dev(m->train_batch), batch, m->d_target, m->stream));
        rc = batch_gradient(m, nullptr, sd, batch, m->d_target);
        if (rc != BM_OK) return rc;
        m->adam_step++;
        MHIP(m, bm_launch_mlp_adam(m->d, lr, m->adam_step, m->stream));
        if (mse_trace) {
            float se = 0.f;
            MHIP(m, hipMemcpyAsync(&se, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
            MHIP(m, hipStreamSynchronize(m->stream));
            mse_trace[s] = se / (float)batch;
        }
    }
    MHIP(m, hipStreamSynchronize(m->stream));
    return BM_OK;
}

bm_ctx* bm_mlp_context(const bm_mlp* m) { return m ? m->ctx : nullptr; }

int bm_mlp_rank_into(bm_mlp* m, uint64_t seed, int64_t pool, int64_t top, int2* out) {
    if (pool < 1 || top < 1 || top > pool || pool > (1LL << 30))
        return mfail(m, BM_ERR_ARG, "learned ranking: need 1 <= top <= pool <= 2^30");
    if (!m->features_ready) return mfail(m, BM_ERR_STATE, "learned ranking: call bm_mlp_refresh_features first");
    MHIP(m, hipSetDevice(m->device));
    int rc = grow_bytes(m, &m->d_rank, &m->rank_bytes, bm_mlp_rank_scratch(pool));
    if (rc != BM_OK) return rc;
    MHIP(m, bm_launch_mlp_rank(m->d, m->d_vf, seed, m->n, pool, top, m->d_rank, m->rank_bytes, out, m->num_cus,
                               m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return BM_OK;
}

extern "C" int bm_batch_gen_learned(bm_batch* b, bm_mlp* m, uint64_t seed, int64_t pool, int64_t top) {
    if (!b || !m) return BM_ERR_ARG;
    if (pool < 1 || top < 1 || top > pool || pool > (1LL << 30) || top > bm_batch_capacity(b))
        return mfail(m, BM_ERR_ARG, "bm_batch_gen_learned: need 1 <= top <= pool <= 2^30 and top <= batch capacity");
    if (!m->features_ready) return mfail(m, BM_ERR_STATE, "bm_batch_gen_learned: call bm_mlp_refresh_features first");
    MHIP(m, hipSetDevice(m->device));
    int rc = grow_bytes(m, &m->d_rank, &m->rank_bytes, bm_mlp_rank_scratch(pool));
    if (rc != BM_OK) return rc;
    MHIP(m, bm_launch_mlp_rank(m->d, m->d_vf, seed, m->n, pool, top, m->d_rank, m->rank_bytes, bm_batch_pairs_dev(b),
                               m->num_cus, m->stream));
    bm_batch_mark_pairs(b, top);
    return BM_OK;
}

extern "C" int bm_mlp_train_batch(bm_mlp* m, bm_batch* b, float lr, float* mse) {
    if (!m || !b || !(lr > 0.f)) return mfail(m, BM_ERR_ARG, "bm_mlp_train_batch: invalid argument");
    if (!bm_batch_has_pair_deltas(b))
        return mfail(m, BM_ERR_STATE, "bm_mlp_train_batch: the batch holds no evaluated pairs");
    if (!m->features_ready) return mfail(m, BM_ERR_STATE, "bm_mlp_train_batch: call bm_mlp_refresh_features first");
    const int64_t k = bm_batch_size(b);
    if (k < 1) return BM_OK;
    MHIP(m, hipSetDevice(m->device));
    int rc = grow(m, &m->d_target, &m->target_cap, k);
    if (rc != BM_OK) return rc;
    MHIP(m, bm_launch_mlp_targets(bm_batch_deltas_dev(b), k, m->d_target, m->stream));
    rc = batch_gradient(m, bm_batch_pairs_dev(b), 0, k, m->d_target);
    if (rc != BM_OK) return rc;
    m->adam_step++;
    MHIP(m, bm_launch_mlp_adam(m->d, lr, m->adam_step, m->stream));
    if (mse) {
        float se = 0.f;
        MHIP(m, hipMemcpyAsync(&se, m->d_loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MHIP(m, hipStreamSynchronize(m->stream));
        *mse = se / (float)k;
    }
    return BM_OK;
}

extern "C" int bm_batch_gen_learned_noalloc(bm_batch* b, bm_mlp* m, uint64_t seed, int64_t pool, int64_t top) {
    if (!b || !m) return BM_ERR_ARG;
    if (pool < 1 || top < 1 || top > pool || pool > BM_NOALLOC_POOL || top > bm_mlp_select_max_top() ||
        top > bm_batch_capacity(b))
        return mfail(m, BM_ERR_ARG,
                     "bm_batch_gen_learned_noalloc: need 1 <= top <= pool, pool <= 2^20, top <= 8192 and top <= "
                     "batch capacity");
    if (!m->features_ready)
        return mfail(m, BM_ERR_STATE, "bm_batch_gen_learned_noalloc: call bm_mlp_refresh_features first");
    if (bm_mlp_select_bytes(pool) > (size_t)(BM_NOALLOC_SCRATCH_BYTES))
        return mfail(m, BM_ERR_ARG, "bm_batch_gen_learned_noalloc: pool exceeds the scratch");
    void* ptr = (m->device >= 0 && m->device < kMaxDevices) ? g_scratch_dev[m->device] : nullptr;
    if (!ptr) return mfail(m, BM_ERR_STATE, "bm_batch_gen_learned_noalloc: no scratch address for this device");
    MHIP(m, hipSetDevice(m->device));
    MHIP(m, bm_launch_mlp_rank_select(m->d, m->d_vf, seed, m->n, pool, top, ptr, (size_t)(BM_NOALLOC_SCRATCH_BYTES),
                                      bm_batch_pairs_dev(b), m->num_cus, m->stream));
    bm_batch_mark_pairs(b, top);
    return BM_OK;
}
