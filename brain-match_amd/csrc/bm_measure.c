/*
 * Host-side byte model for the roofline of the swap-delta kernel.
 *
 * Algorithmic bytes of one swap candidate (u,v), independent of the lookup
 * structure and of cache hits: 32 bytes of list bounds (begin/end of the four
 * male lists), 8 of pi(u), pi(v), 8 of int64 delta out, and per incident male
 * edge 28 bytes = 4 neighbour id + 4 weight + 4 pi(neighbour) + 2 x 8 for the
 * old and the new female probe:  B = 48 + 28 * D(u,v),
 * D(u,v) = d_out(u) + d_in(u) + d_out(v) + d_in(v) in the male graph.
 * Candidates generated inside the kernel read nothing for the pair itself.
 */
#include "bm_host.h"
#include "bm_rng.h"

#include <stdlib.h>

int bm_uniform_algo_bytes(bm_ctx* ctx, uint64_t seed, int64_t k, int64_t* bytes, int64_t* sum_degree) {
    if (!ctx || k < 0) return BM_ERR_ARG;
    const int32_t n = bm_num_vertices(ctx);
    int32_t* od = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t* id = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    if (!od || !id) {
        free(od);
        free(id);
        return BM_ERR_NOMEM;
    }
    int rc = bm_degrees(ctx, 0, od, id);
    if (rc != BM_OK) {
        free(od);
        free(id);
        return rc;
    }
    for (int32_t x = 0; x <= n; x++) od[x] += id[x];
    int64_t D = 0;
#pragma omp parallel for reduction(+ : D) schedule(static)
    for (int64_t i = 0; i < k; i++) {
        int32_t u, v;
        bm_uniform_pair(seed, (uint64_t)i, n, &u, &v);
        D += (int64_t)od[u] + od[v];
    }
    free(od);
    free(id);
    if (bytes) *bytes = 48 * k + 28 * D;
    if (sum_degree) *sum_degree = D;
    return BM_OK;
}

/* Bytes of the annealing steps taken so far: every chain's proposals are
 * replayed from its stream (the same draws as k_anneal) and priced with the
 * swap model above or the triple model 69 + 36 * D3 (6 list bounds, 3 pi, the
 * delta and perm, and per incident edge 12 bytes of male data plus three
 * female probes).  graphs[inst[c]] is chain c's male graph. */
int64_t bm_chain_algo_bytes(const bm_graph_host* const* graphs, const int32_t* inst, const uint64_t* seeds,
                            const int64_t* steps, int64_t n_chains, uint64_t p_rot3) {
    int64_t B = 0;
#pragma omp parallel for reduction(+ : B) schedule(dynamic, 1)
    for (int64_t c = 0; c < n_chains; c++) {
        const bm_graph_host* G = graphs[inst[c]];
        const int32_t n = G->n;
        const uint64_t q0 = bm_rand_at(seeds[c], 0), q2 = bm_rand_at(seeds[c], 2), q3 = bm_rand_at(seeds[c], 3);
#define BM_DEG(x) ((G->out.ptr[(x) + 1] - G->out.ptr[(x)]) + (G->in.ptr[(x) + 1] - G->in.ptr[(x)]))
        for (int64_t s = 0; s < steps[c]; s++) {
            const uint64_t rk = bm_rand_at(q2, (uint64_t)s);
            if (n >= 3 && (uint64_t)(uint32_t)rk < p_rot3) {
                int32_t x, y, z;
                bm_uniform_triple(q3, (uint64_t)s, n, &x, &y, &z);
                B += 69 + 36 * (BM_DEG(x) + BM_DEG(y) + BM_DEG(z));
            } else {
                int32_t u, v;
                bm_uniform_pair(q0, (uint64_t)s, n, &u, &v);
                B += 48 + 28 * (BM_DEG(u) + BM_DEG(v));
            }
        }
#undef BM_DEG
    }
    return B;
}

/* Swap byte model (48 + 28 * D per pair) summed over the u-major stream
 * (seed, k): the first-vertex counts are recounted from the draws and each
 * partner is regenerated. */
int bm_umajor_algo_bytes(bm_ctx* ctx, uint64_t seed, int64_t k, int64_t* bytes, int64_t* sum_degree) {
    if (!ctx || k < 0) return BM_ERR_ARG;
    const int32_t n = bm_num_vertices(ctx);
    int32_t* od = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t* id = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int64_t* cnt = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
    if (!od || !id || !cnt) {
        free(od);
        free(id);
        free(cnt);
        return BM_ERR_NOMEM;
    }
    int rc = bm_degrees(ctx, 0, od, id);
    if (rc == BM_OK) {
        for (int32_t x = 0; x <= n; x++) od[x] += id[x];
#pragma omp parallel
        {
            int64_t* mine = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
#pragma omp for schedule(static)
            for (int64_t c = 0; c < k; c++) mine[bm_umajor_first(seed, (uint64_t)c, n)]++;
#pragma omp critical
            for (int32_t x = 1; x <= n; x++) cnt[x] += mine[x];
            free(mine);
        }
        int64_t D = 0;
        for (int32_t u = 1; u <= n; u++) D += cnt[u] * od[u];
        int64_t* off = cnt;   /* in place: cnt[u] becomes the start of u's range */
        int64_t run = 0;
        for (int32_t u = 1; u <= n; u++) {
            const int64_t c = off[u];
            off[u] = run;
            run += c;
        }
        off[n + 1] = run;
        int64_t Dv = 0;
#pragma omp parallel for reduction(+ : Dv) schedule(dynamic, 64)
        for (int32_t u = 1; u <= n; u++)
            for (int64_t i = off[u]; i < off[u + 1]; i++) Dv += od[bm_umajor_partner(seed, (uint64_t)i, u, n)];
        D += Dv;
        if (bytes) *bytes = 48 * k + 28 * D;
        if (sum_degree) *sum_degree = D;
    }
    free(od);
    free(id);
    free(cnt);
    return rc;
}
