/*
 * Seeded synthetic graph pairs for the benchmark workloads.
 *
 * Male graph: out-degrees from a discrete power law (exponent alpha, capped
 * at n-1, the scale chosen by bisection so the realised edge count is
 * n*mean_out_degree), targets uniform without replacement, no self-loops, no
 * duplicate (src,dst), weights >= 1.
 * Female graph: the male graph relabelled by a hidden permutation P, then
 * modify_frac of its edges changed: the first half of the chosen edges get a
 * fresh weight from the same law, the second half a fresh uniform target
 * (still no self-loop and no duplicate).
 * pi0: P with shuffle_frac of its positions permuted among themselves.
 *
 * The reference has no generator (its inputs are the Flywire VNC CSVs that
 * are not shipped, .MISSING_LARGE_BLOBS:1-2); the rules avoid the three
 * reference quirks (self-loops, duplicate lines, weights above int16) so the
 * reference kernels, the exact delta and score-matching all agree on them.
 */
#include "bm_host.h"
#include "bm_rng.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int32_t draw_weight(bm_xoshiro* r, int32_t law) {
    if (law == BM_WEIGHTS_GEOMETRIC) {
        /* failures before the first success of Bernoulli(0.3), plus one */
        double u = 1.0 - bm_xoshiro_unit(r); /* (0,1] */
        double k = floor(log(u) / log(1.0 - 0.3));
        if (k > 2.0e9) k = 2.0e9;
        return (int32_t)k + 1;
    }
    /* lognormal(mu=0.5, sigma=1.2), rounded, at least 1 (Box-Muller) */
    double u1 = 1.0 - bm_xoshiro_unit(r);
    double u2 = bm_xoshiro_unit(r);
    double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    double x = exp(0.5 + 1.2 * z);
    double w = floor(x + 0.5);
    if (w < 1.0) w = 1.0;
    if (w > 2147483647.0) w = 2147483647.0;
    return (int32_t)w;
}

static void shuffle_i32(bm_xoshiro* r, int32_t* a, int64_t k) {
    for (int64_t i = k - 1; i > 0; i--) {
        int64_t j = (int64_t)bm_xoshiro_below(r, (uint32_t)(i + 1));
        int32_t t = a[i];
        a[i] = a[j];
        a[j] = t;
    }
}

/* Open-addressing set of (src,dst) keys for the duplicate rule. */
typedef struct { uint64_t* k; uint64_t mask; } keyset;
#define KS_EMPTY 0ull
#define KS_DEAD 1ull
static uint64_t ks_key(int32_t s, int32_t d) { return ((uint64_t)(uint32_t)s << 32) | (uint32_t)d; }
static uint64_t ks_h(uint64_t k, uint64_t mask) { return bm_mix64(k) & mask; }
static int ks_has(const keyset* ks, uint64_t key) {
    for (uint64_t i = ks_h(key, ks->mask);; i = (i + 1) & ks->mask) {
        if (ks->k[i] == KS_EMPTY) return 0;
        if (ks->k[i] == key) return 1;
    }
}
static void ks_put(keyset* ks, uint64_t key) {
    for (uint64_t i = ks_h(key, ks->mask);; i = (i + 1) & ks->mask) {
        if (ks->k[i] == KS_EMPTY || ks->k[i] == KS_DEAD) { ks->k[i] = key; return; }
    }
}
static void ks_del(keyset* ks, uint64_t key) {
    for (uint64_t i = ks_h(key, ks->mask);; i = (i + 1) & ks->mask) {
        if (ks->k[i] == KS_EMPTY) return;
        if (ks->k[i] == key) { ks->k[i] = KS_DEAD; return; }
    }
}

int bm_synth_pair(const bm_synth_params* p, bm_edge_list* male, bm_edge_list* female,
                  int32_t* perm_true, int32_t* pi0) {
    memset(male, 0, sizeof *male);
    memset(female, 0, sizeof *female);
    if (!p || p->n < 3 || p->mean_out_degree <= 0 || p->alpha <= 1.0 ||
        (p->weights != BM_WEIGHTS_GEOMETRIC && p->weights != BM_WEIGHTS_LOGNORMAL) ||
        p->modify_frac < 0 || p->modify_frac > 1 || p->shuffle_frac < 0 || p->shuffle_frac > 1) {
        bm_set_global_error("bm_synth_pair: invalid parameters");
        return BM_ERR_ARG;
    }
    const int32_t n = p->n;
    bm_xoshiro rng;
    bm_xoshiro_seed(&rng, p->seed);

    /* ---- out-degrees: d_i(x) = min(n-1, floor(x * U_i^(-1/(alpha-1)))) ---- */
    double* base = (double*)malloc(sizeof(double) * (size_t)n);
    int32_t* deg = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    if (!base || !deg) { free(base); free(deg); return BM_ERR_NOMEM; }
    for (int32_t i = 0; i < n; i++) {
        double u = 1.0 - bm_xoshiro_unit(&rng);
        base[i] = pow(u, -1.0 / (p->alpha - 1.0));
    }
    const double target = (double)n * p->mean_out_degree;
    double lo = 0.0, hi = (double)n;
    for (int it = 0; it < 100; it++) {
        double x = 0.5 * (lo + hi);
        double tot = 0;
        for (int32_t i = 0; i < n; i++) {
            double d = floor(x * base[i]);
            if (d > n - 1) d = n - 1;
            tot += d;
        }
        if (tot < target) lo = x; else hi = x;
    }
    int64_t m = 0;
    deg[0] = 0;
    for (int32_t i = 0; i < n; i++) {
        double d = floor(hi * base[i]);
        if (d > n - 1) d = n - 1;
        if (d < 1) d = 1;
        deg[i + 1] = (int32_t)d;
        m += (int64_t)d;
    }
    free(base);

    male->src = (int32_t*)malloc(sizeof(int32_t) * (size_t)m);
    male->dst = (int32_t*)malloc(sizeof(int32_t) * (size_t)m);
    male->w = (int32_t*)malloc(sizeof(int32_t) * (size_t)m);
    int32_t* stamp = (int32_t*)calloc((size_t)(n + 1), sizeof(int32_t));
    if (!male->src || !male->dst || !male->w || !stamp) {
        free(deg); free(stamp); bm_edge_list_free(male);
        return BM_ERR_NOMEM;
    }
    male->m = m;
    male->n = n;
    int64_t e = 0;
    for (int32_t u = 1; u <= n; u++) {
        int32_t d = deg[u];
        if (2 * (int64_t)d <= n - 1) {
            /* rejection with per-source stamps */
            int32_t got = 0;
            stamp[u] = u;
            while (got < d) {
                int32_t t = (int32_t)bm_xoshiro_below(&rng, (uint32_t)n) + 1;
                if (stamp[t] == u) continue;
                stamp[t] = u;
                male->src[e] = u;
                male->dst[e] = t;
                male->w[e] = draw_weight(&rng, p->weights);
                e++;
                got++;
            }
        } else {
            /* choose the n-1-d excluded targets, keep the rest in id order */
            int32_t excl = (n - 1) - d, got = 0;
            stamp[u] = u;
            while (got < excl) {
                int32_t t = (int32_t)bm_xoshiro_below(&rng, (uint32_t)n) + 1;
                if (stamp[t] == u) continue;
                stamp[t] = u;
                got++;
            }
            for (int32_t t = 1; t <= n; t++) {
                if (stamp[t] == u) continue;
                male->src[e] = u;
                male->dst[e] = t;
                male->w[e] = draw_weight(&rng, p->weights);
                e++;
            }
        }
    }
    free(deg);
    free(stamp);

    /* ---- hidden permutation P and the relabelled female graph ---- */
    int32_t* P = perm_true;
    P[0] = 0;
    for (int32_t i = 1; i <= n; i++) P[i] = i;
    shuffle_i32(&rng, P + 1, n);

    female->m = m;
    female->n = n;
    female->src = (int32_t*)malloc(sizeof(int32_t) * (size_t)m);
    female->dst = (int32_t*)malloc(sizeof(int32_t) * (size_t)m);
    female->w = (int32_t*)malloc(sizeof(int32_t) * (size_t)m);
    keyset ks = {0, 0};
    uint64_t cap = 2;
    while (cap < 2 * (uint64_t)m + 2) cap <<= 1;
    ks.k = (uint64_t*)calloc(cap, sizeof(uint64_t));
    ks.mask = cap - 1;
    int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (size_t)m);
    if (!female->src || !female->dst || !female->w || !ks.k || !order) {
        free(ks.k); free(order); bm_edge_list_free(male); bm_edge_list_free(female);
        return BM_ERR_NOMEM;
    }
    for (int64_t i = 0; i < m; i++) {
        female->src[i] = P[male->src[i]];
        female->dst[i] = P[male->dst[i]];
        female->w[i] = male->w[i];
        ks_put(&ks, ks_key(female->src[i], female->dst[i]));
        order[i] = i;
    }
    int64_t nmod = (int64_t)floor(p->modify_frac * (double)m);
    /* partial Fisher-Yates: the first nmod entries of order are the chosen edges */
    for (int64_t i = 0; i < nmod; i++) {
        uint64_t span = (uint64_t)(m - i);
        int64_t j = i + (int64_t)(span <= 0xFFFFFFFFull ? bm_xoshiro_below(&rng, (uint32_t)span)
                                                       : (bm_xoshiro_next(&rng) % span));
        int64_t t = order[i];
        order[i] = order[j];
        order[j] = t;
    }
    int64_t nre = nmod / 2;
    for (int64_t i = 0; i < nmod; i++) {
        int64_t ei = order[i];
        if (i < nre) {
            female->w[ei] = draw_weight(&rng, p->weights);
        } else {
            int32_t s = female->src[ei];
            /* Up to 64 draws for a target that is neither s nor already an
             * out-neighbour of s; a hub source that finds none in 64 draws
             * (it is adjacent to almost every vertex) gets a new weight
             * instead. */
            int done = 0;
            for (int tries = 0; tries < 64 && !done; tries++) {
                int32_t t = (int32_t)bm_xoshiro_below(&rng, (uint32_t)n) + 1;
                if (t == s || ks_has(&ks, ks_key(s, t))) continue;
                ks_del(&ks, ks_key(s, female->dst[ei]));
                ks_put(&ks, ks_key(s, t));
                female->dst[ei] = t;
                done = 1;
            }
            if (!done) female->w[ei] = draw_weight(&rng, p->weights);
        }
    }
    free(ks.k);
    free(order);

    /* ---- pi0: P with shuffle_frac of its positions permuted ---- */
    if (pi0) {
        memcpy(pi0, P, sizeof(int32_t) * (size_t)(n + 1));
        int64_t k = (int64_t)floor(p->shuffle_frac * (double)n);
        int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
        int32_t* val = (int32_t*)malloc(sizeof(int32_t) * (size_t)(k ? k : 1));
        if (!pos || !val) { free(pos); free(val); bm_edge_list_free(male); bm_edge_list_free(female); return BM_ERR_NOMEM; }
        for (int32_t i = 0; i < n; i++) pos[i] = i + 1;
        for (int64_t i = 0; i < k; i++) {
            int64_t j = i + (int64_t)bm_xoshiro_below(&rng, (uint32_t)(n - i));
            int32_t t = pos[i];
            pos[i] = pos[j];
            pos[j] = t;
        }
        for (int64_t i = 0; i < k; i++) val[i] = P[pos[i]];
        shuffle_i32(&rng, val, k);
        for (int64_t i = 0; i < k; i++) pi0[pos[i]] = val[i];
        free(pos);
        free(val);
    }
    return BM_OK;
}
