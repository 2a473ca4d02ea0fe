/*
 * Shared main() of the drop-in re-alignment tools.
 *
 * Command line as the reference tools (src/graph-re-alignment-gpu-swap.cu:1469-1474):
 *     <male graph.csv> <female graph.csv> <in mapping.csv> <out mapping.csv> [seconds]
 * The best matching is written to "<out mapping minus its last 4 chars>-rank0.csv"
 * on every new best (:1519), in the "Male Node ID,Female Node ID" / "m<u>,f<v>"
 * format (:681-696).  Every 1800 s "../data/best.csv" is adopted when it scores
 * higher (:1114-1132).  The reference loops until killed; here the optional
 * fifth argument (or BM_TIME_LIMIT) bounds the run, and SIGINT/SIGTERM end it
 * after saving.  BM_SEED fixes the random streams (default: the clock).
 * Exit status 1 on usage, load or device errors, as the reference (:1469-1509).
 */
#ifndef BM_CLI_H
#define BM_CLI_H

#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "brainmatch.h"

static volatile int g_stop = 0;
static void on_signal(int sig) {
    (void)sig;
    g_stop = 1;
}

static int bm_tool_main(int argc, char** argv, int mode, const char* title) {
    if (argc < 5) {
        fprintf(stderr, "[ERROR] Usage: %s <male graph> <female graph> <in mapping> <out mapping> [seconds]\n", argv[0]);
        return 1;
    }
    printf("[INFO] %s\n", title);
    printf("[INFO]   - Male graph: %s\n[INFO]   - Female graph: %s\n[INFO]   - Output mapping: %s\n", argv[1], argv[2],
           argv[4]);
    bm_edge_list gm, gf;
    if (bm_load_graph_csv(argv[1], &gm) != BM_OK) {
        fprintf(stderr, "[ERROR] Failed to load male graph: %s\n", bm_last_error(NULL));
        return 1;
    }
    if (bm_load_graph_csv(argv[2], &gf) != BM_OK) {
        fprintf(stderr, "[ERROR] Failed to load female graph: %s\n", bm_last_error(NULL));
        bm_edge_list_free(&gm);
        return 1;
    }
    int32_t n = gm.n > gf.n ? gm.n : gf.n;
    printf("[INFO] Graphs loaded: %d vertices, %lld male lines, %lld female lines\n", n, (long long)gm.m,
           (long long)gf.m);
    bm_edges me = {gm.m, gm.src, gm.dst, gm.w}, fe = {gf.m, gf.src, gf.dst, gf.w};
    bm_ctx* ctx = NULL;
    if (bm_create(&ctx, n, &me, &fe, 0) != BM_OK) {
        fprintf(stderr, "[ERROR] %s\n", bm_last_error(NULL));
        bm_edge_list_free(&gm);
        bm_edge_list_free(&gf);
        return 1;
    }
    bm_edge_list_free(&gm);
    bm_edge_list_free(&gf);
    char info[128];
    bm_device_info(ctx, info, sizeof info);
    printf("[INFO] Using device: %s\n", info);
    int32_t* pi = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    int64_t lines = 0;
    if (!pi || bm_load_mapping_csv(argv[3], n, pi, &lines) != BM_OK) {
        fprintf(stderr, "[ERROR] Failed to load benchmark mapping: %s\n", bm_last_error(NULL));
        bm_destroy(ctx);
        free(pi);
        return 1;
    }
    if (bm_set_mapping(ctx, pi) != BM_OK) {
        fprintf(stderr, "[ERROR] %s\n", bm_last_error(ctx));
        bm_destroy(ctx);
        free(pi);
        return 1;
    }
    int64_t s0 = 0;
    if (bm_score(ctx, &s0) != BM_OK) {
        fprintf(stderr, "[ERROR] %s\n", bm_last_error(ctx));
        bm_destroy(ctx);
        free(pi);
        return 1;
    }
    printf("[INFO] Initial alignment score: %lld\n", (long long)s0);
    const size_t L = strlen(argv[4]);
    char* outname = (char*)malloc(L + 16);
    snprintf(outname, L + 16, "%.*s-rank0.csv", (int)(L >= 4 ? L - 4 : 0), argv[4]);
    printf("[INFO] output file name is %s\n", outname);

    signal(SIGINT, on_signal);
    signal(SIGTERM, on_signal);
    bm_search_opts o;
    memset(&o, 0, sizeof o);
    o.mode = mode;
    const char* seed_env = getenv("BM_SEED");
    o.seed = seed_env ? strtoull(seed_env, NULL, 10) : (uint64_t)time(NULL);
    o.time_limit_s = argc > 5 ? atof(argv[5]) : (getenv("BM_TIME_LIMIT") ? atof(getenv("BM_TIME_LIMIT")) : 0);
    o.out_path = outname;
    o.sync_path = getenv("BM_SYNC_PATH") ? getenv("BM_SYNC_PATH") : "../data/best.csv";
    o.sync_interval_s = getenv("BM_SYNC_INTERVAL") ? atof(getenv("BM_SYNC_INTERVAL")) : 1800.0;
    o.verify_every = 256;
    o.verbose = 1;
    o.stop_flag = &g_stop;
    o.restart_gap = 0.002;
    bm_search_result r;
    int rc = bm_search(ctx, &o, pi, &r);
    if (rc != BM_OK) {
        fprintf(stderr, "[ERROR] search failed: %s\n", bm_last_error(ctx));
        bm_destroy(ctx);
        free(pi);
        free(outname);
        return 1;
    }
    if (bm_write_mapping_csv(outname, n, pi) != BM_OK)
        fprintf(stderr, "[ERROR] could not write %s\n", outname);
    printf("[INFO] Process completed:\n");
    printf("[INFO]   - Initial score: %lld\n", (long long)r.initial_score);
    printf("[INFO]   - Final score: %lld\n", (long long)r.best_score);
    printf("[INFO]   - Improvement: %.2f%%\n",
           r.initial_score ? 100.0 * (double)(r.best_score - r.initial_score) / (double)r.initial_score : 0.0);
    printf("[INFO]   - Kernel launches: %lld, candidates evaluated: %lld, moves: %lld\n", (long long)r.iterations,
           (long long)r.evaluations, (long long)r.accepted);
    printf("[INFO]   - Total time: %.1f minutes\n", r.seconds / 60.0);
    if (r.verify_failures) fprintf(stderr, "[ERROR] %lld score verification failures\n", (long long)r.verify_failures);
    bm_destroy(ctx);
    free(pi);
    free(outname);
    return r.verify_failures ? 1 : 0;
}

#endif
