/*
 * Harness around the reference's own CPU delta code, for the oracle/_ref
 * build only (never shipped in the library).  It is linked with
 * src/graph-alignment-optimized.c compiled unmodified from /root/reference
 * (its main() renamed at compile time) and calls that file's new_graph,
 * add_edge, calculate_alignment_score and calculate_swap_delta
 * (src/graph-alignment-optimized.c:125-147, :183-204, :292-305, :321-408).
 * The reference sizes everything by NUM_NODES = 18524 (:15), so vertex ids
 * must stay within 1..18524 and the matching passed in is padded to 18525
 * entries here.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct Graph Graph;
Graph* new_graph(void);
void add_edge(Graph* g, int from, int to, int weight);
int calculate_alignment_score(Graph* gm, Graph* gf, int* mapping);
int calculate_swap_delta(Graph* gm, Graph* gf, int* mapping, int node_m1, int node_m2);
void free_graph(Graph* g);

#define REF_NUM_NODES 18524

int refh_num_nodes(void) { return REF_NUM_NODES; }

void* refh_build(int64_t m, const int32_t* s, const int32_t* d, const int32_t* w) {
    Graph* g = new_graph();
    for (int64_t i = 0; i < m; i++) add_edge(g, s[i], d[i], w[i]);
    return g;
}

void refh_free(void* g) { free_graph((Graph*)g); }

static int* padded(const int32_t* map, int32_t n) {
    int* p = (int*)calloc(REF_NUM_NODES + 1, sizeof(int));
    memcpy(p, map, sizeof(int32_t) * (size_t)(n + 1));
    return p;
}

int32_t refh_score(void* gm, void* gf, const int32_t* map, int32_t n) {
    int* p = padded(map, n);
    int s = calculate_alignment_score((Graph*)gm, (Graph*)gf, p);
    free(p);
    return s;
}

/* Deltas for k pairs, OpenMP dynamic schedule; threads <= 0 keeps the default. */
void refh_deltas(void* gm, void* gf, const int32_t* map, int32_t n, const int32_t* uv, int64_t k, int32_t* out,
                 int threads) {
    int* p = padded(map, n);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < k; i++) out[i] = calculate_swap_delta((Graph*)gm, (Graph*)gf, p, uv[2 * i], uv[2 * i + 1]);
    free(p);
}
