/*
 * Greedy initial matching in the manner of greedy-match
 * (src/greedy-match.c:213-530).
 *
 * Per vertex: out-degree, in-degree, the number of distinct vertices reachable
 * by two out-edges, and the number of distinct vertices that reach it by two
 * in-edges (compute_2hop_metrics, :213-252; the start vertex itself counts if
 * it is reached).  total = sum of the four.  Both graphs' vertex lists are
 * sorted by total descending (:489-490); each male vertex in that order takes
 * the first unmatched female vertex, in female order, with the smallest L1
 * distance of the four features (:495-522).
 *
 * Differences, by design: ties of the sort are broken by the smaller vertex id
 * (the reference's qsort leaves them to libc), and duplicate CSV lines count
 * once (the reference's adjacency lists keep every line).  The vertex set of
 * each graph is the set of ids with at least one incident edge, as in the
 * reference's loader (:302-399); vertices outside it stay unmatched (0).
 */
#include "bm_host.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t id;
    int32_t in_deg, out_deg, reach2, coreach2;
    int64_t total;
} feat_t;

static void features(const bm_graph_host* g, feat_t* f, int32_t* count) {
    const int32_t n = g->n;
    int32_t k = 0;
    int32_t* present = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    for (int32_t v = 1; v <= n; v++) {
        int64_t od = g->out.ptr[v + 1] - g->out.ptr[v], id = g->in.ptr[v + 1] - g->in.ptr[v];
        if (od + id > 0) present[v] = 1;
    }
    int32_t* list = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    for (int32_t v = 1; v <= n; v++)
        if (present[v]) list[k++] = v;
    free(present);
#pragma omp parallel
    {
        int32_t* stamp = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
#pragma omp for schedule(dynamic, 64)
        for (int32_t i = 0; i < k; i++) {
            const int32_t v = list[i];
            feat_t* q = &f[i];
            q->id = v;
            q->out_deg = (int32_t)(g->out.ptr[v + 1] - g->out.ptr[v]);
            q->in_deg = (int32_t)(g->in.ptr[v + 1] - g->in.ptr[v]);
            int32_t r = 0, s = 0;
            const int32_t tag1 = 2 * v, tag2 = 2 * v + 1; /* stamps unique per (vertex, pass) */
            for (int64_t a = g->out.ptr[v]; a < g->out.ptr[v + 1]; a++) {
                const int32_t x = g->out.adj[a].x;
                for (int64_t b = g->out.ptr[x]; b < g->out.ptr[x + 1]; b++) {
                    const int32_t y = g->out.adj[b].x;
                    if (stamp[y] != tag1) { stamp[y] = tag1; r++; }
                }
            }
            for (int64_t a = g->in.ptr[v]; a < g->in.ptr[v + 1]; a++) {
                const int32_t x = g->in.adj[a].x;
                for (int64_t b = g->in.ptr[x]; b < g->in.ptr[x + 1]; b++) {
                    const int32_t y = g->in.adj[b].x;
                    if (stamp[y] != tag2) { stamp[y] = tag2; s++; }
                }
            }
            q->reach2 = r;
            q->coreach2 = s;
            q->total = (int64_t)q->in_deg + q->out_deg + r + s;
        }
        free(stamp);
    }
    free(list);
    *count = k;
}

static int by_total_desc(const void* a, const void* b) {
    const feat_t* x = (const feat_t*)a;
    const feat_t* y = (const feat_t*)b;
    if (x->total != y->total) return x->total > y->total ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id);
}

static inline int64_t fdist(const feat_t* a, const feat_t* b) {
    return llabs((long long)a->in_deg - b->in_deg) + llabs((long long)a->out_deg - b->out_deg) +
           llabs((long long)a->reach2 - b->reach2) + llabs((long long)a->coreach2 - b->coreach2);
}

int bm_greedy_match(int32_t n, const bm_edges* male, const bm_edges* female, int32_t* pi) {
    bm_graph_host gm, gf;
    const char* err = NULL;
    int rc = bm_graph_build(n, male, &gm, &err);
    if (rc != BM_OK) { bm_set_global_error("bm_greedy_match: male graph: %s", err); return rc; }
    rc = bm_graph_build(n, female, &gf, &err);
    if (rc != BM_OK) { bm_graph_free(&gm); bm_set_global_error("bm_greedy_match: female graph: %s", err); return rc; }
    feat_t* f1 = (feat_t*)malloc(sizeof(feat_t) * (size_t)(n + 1));
    feat_t* f2 = (feat_t*)malloc(sizeof(feat_t) * (size_t)(n + 1));
    uint8_t* used = (uint8_t*)calloc((size_t)n + 1, 1);
    if (!f1 || !f2 || !used) {
        free(f1); free(f2); free(used); bm_graph_free(&gm); bm_graph_free(&gf);
        return BM_ERR_NOMEM;
    }
    int32_t k1 = 0, k2 = 0;
    features(&gm, f1, &k1);
    features(&gf, f2, &k2);
    bm_graph_free(&gm);
    bm_graph_free(&gf);
    if (k1 != k2) {
        free(f1); free(f2); free(used);
        bm_set_global_error("bm_greedy_match: graphs have different vertex counts (%d vs %d)", k1, k2);
        return BM_ERR_ARG;
    }
    qsort(f1, (size_t)k1, sizeof(feat_t), by_total_desc);
    qsort(f2, (size_t)k2, sizeof(feat_t), by_total_desc);
    memset(pi, 0, sizeof(int32_t) * (size_t)(n + 1));
    /* Alive list of unmatched female positions, kept in sorted order. */
    int32_t* alive = (int32_t*)malloc(sizeof(int32_t) * (size_t)(k2 ? k2 : 1));
    int32_t nalive = k2;
    for (int32_t j = 0; j < k2; j++) alive[j] = j;
    for (int32_t i = 0; i < k1; i++) {
        int64_t best = INT64_MAX;
        int32_t bj = -1;
        for (int32_t q = 0; q < nalive; q++) {
            const int64_t d = fdist(&f1[i], &f2[alive[q]]);
            if (d < best) { best = d; bj = q; if (d == 0) break; }
        }
        if (bj < 0) break;
        pi[f1[i].id] = f2[alive[bj]].id;
        memmove(&alive[bj], &alive[bj + 1], sizeof(int32_t) * (size_t)(nalive - bj - 1));
        nalive--;
    }
    free(alive);
    free(f1);
    free(f2);
    free(used);
    return BM_OK;
}
