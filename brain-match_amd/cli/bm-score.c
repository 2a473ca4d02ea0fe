/* Scores a matching on the device: the same integer as score-matching
 * (src/score-matching.c:599-704), printed alone on stdout.
 * Usage: bm-score <g1.csv> <g2.csv> <matching.csv> */
#include <stdio.h>
#include <stdlib.h>

#include "brainmatch.h"

int main(int argc, char** argv) {
    if (argc != 4) {
        fprintf(stderr, "Usage: %s <g1.csv> <g2.csv> <matching.csv>\n", argv[0]);
        return 1;
    }
    bm_edge_list a, b;
    if (bm_load_graph_csv(argv[1], &a) != BM_OK || bm_load_graph_csv(argv[2], &b) != BM_OK) {
        fprintf(stderr, "%s\n", bm_last_error(NULL));
        return 1;
    }
    int32_t n = a.n > b.n ? a.n : b.n;
    if (n < 2) n = 2;
    int32_t* pi = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    int64_t lines = 0;
    if (bm_load_mapping_csv(argv[3], n, pi, &lines) != BM_OK) {
        fprintf(stderr, "%s\n", bm_last_error(NULL));
        return 1;
    }
    bm_edges me = {a.m, a.src, a.dst, a.w}, fe = {b.m, b.src, b.dst, b.w};
    bm_ctx* ctx = NULL;
    if (bm_create(&ctx, n, &me, &fe, 0) != BM_OK) {
        fprintf(stderr, "%s\n", bm_last_error(NULL));
        return 1;
    }
    int64_t s = 0;
    if (bm_set_mapping(ctx, pi) != BM_OK || bm_score(ctx, &s) != BM_OK) {
        fprintf(stderr, "%s\n", bm_last_error(ctx));
        return 1;
    }
    printf("%lld\n", (long long)s);
    bm_destroy(ctx);
    bm_edge_list_free(&a);
    bm_edge_list_free(&b);
    free(pi);
    return 0;
}
