/* Writes a synthetic benchmark pair as CSV files.
 * Usage: bm-synth <n> <mean out-degree> <weights: geometric|lognormal> <shuffle frac>
 *                 <seed> <male.csv> <female.csv> <pi0.csv> [truth.csv] */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "brainmatch.h"

int main(int argc, char** argv) {
    if (argc < 9) {
        fprintf(stderr, "Usage: %s <n> <mean out-degree> <geometric|lognormal> <shuffle frac> <seed> <male.csv> <female.csv> <pi0.csv> [truth.csv]\n", argv[0]);
        return 1;
    }
    bm_synth_params p;
    memset(&p, 0, sizeof p);
    p.n = atoi(argv[1]);
    p.mean_out_degree = atof(argv[2]);
    p.alpha = 2.1;
    p.weights = strcmp(argv[3], "geometric") == 0 ? BM_WEIGHTS_GEOMETRIC : BM_WEIGHTS_LOGNORMAL;
    p.modify_frac = 0.2;
    p.shuffle_frac = atof(argv[4]);
    p.seed = strtoull(argv[5], NULL, 10);
    int32_t* truth = (int32_t*)calloc((size_t)p.n + 1, sizeof(int32_t));
    int32_t* pi0 = (int32_t*)calloc((size_t)p.n + 1, sizeof(int32_t));
    bm_edge_list m, f;
    if (!truth || !pi0 || bm_synth_pair(&p, &m, &f, truth, pi0) != BM_OK) {
        fprintf(stderr, "%s\n", bm_last_error(NULL));
        return 1;
    }
    if (bm_write_graph_csv(argv[6], &m) != BM_OK || bm_write_graph_csv(argv[7], &f) != BM_OK ||
        bm_write_mapping_csv(argv[8], p.n, pi0) != BM_OK || (argc > 9 && bm_write_mapping_csv(argv[9], p.n, truth) != BM_OK)) {
        fprintf(stderr, "%s\n", bm_last_error(NULL));
        return 1;
    }
    printf("n=%d male_edges=%lld female_edges=%lld\n", p.n, (long long)m.m, (long long)f.m);
    bm_edge_list_free(&m);
    bm_edge_list_free(&f);
    free(truth);
    free(pi0);
    return 0;
}
