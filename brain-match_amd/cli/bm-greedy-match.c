/* Greedy initial matching (as greedy-match, src/greedy-match.c:532-629):
 * writes "Node G1,Node G2" then "<male>,<female>" lines.
 * Usage: bm-greedy-match <g1.csv> <g2.csv> <output_matching.csv> */
#include <stdio.h>
#include <stdlib.h>

#include "brainmatch.h"

int main(int argc, char** argv) {
    if (argc != 4) {
        fprintf(stderr, "Usage: %s <g1.csv> <g2.csv> <output_matching.csv>\n", argv[0]);
        return 1;
    }
    bm_edge_list a, b;
    if (bm_load_graph_csv(argv[1], &a) != BM_OK || bm_load_graph_csv(argv[2], &b) != BM_OK) {
        fprintf(stderr, "%s\n", bm_last_error(NULL));
        return 1;
    }
    int32_t n = a.n > b.n ? a.n : b.n;
    int32_t* pi = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    bm_edges me = {a.m, a.src, a.dst, a.w}, fe = {b.m, b.src, b.dst, b.w};
    if (!pi || bm_greedy_match(n, &me, &fe, pi) != BM_OK) {
        fprintf(stderr, "%s\n", bm_last_error(NULL));
        return 1;
    }
    FILE* f = fopen(argv[3], "w");
    if (!f) {
        fprintf(stderr, "cannot write %s\n", argv[3]);
        return 1;
    }
    fprintf(f, "Node G1,Node G2\n");
    for (int32_t i = 1; i <= n; i++)
        if (pi[i]) fprintf(f, "%d,%d\n", i, pi[i]);
    fclose(f);
    bm_edge_list_free(&a);
    bm_edge_list_free(&b);
    free(pi);
    return 0;
}
