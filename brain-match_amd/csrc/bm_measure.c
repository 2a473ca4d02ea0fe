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
