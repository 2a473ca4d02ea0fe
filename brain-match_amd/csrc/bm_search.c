/*
 * Host search drivers over the device calls.
 *
 * bm_search keeps the outer shape of the reference's optimize_mapping
 * (src/graph-re-alignment-gpu-swap.cu:1002-1350 and its 3-swap / uniform
 * variants): perturb the current (or best) matching, evaluate every candidate
 * that touches a changed vertex in the changed-first order, then walk:
 * while a positive move exists take a random positive one or the argmax with
 * equal odds, otherwise a random zero move (at most 1000 zero moves in a
 * row), re-evaluating the candidates of the vertices just moved after every
 * move (:1220-1325).  What moved to the device: the evaluation, the argmax
 * and the random picks (bm_batch_stats / bm_batch_pick) -- the reference
 * copies P*12 bytes back and scans them on the host (:1194-1218).  The
 * running score is carried by the exact deltas and checked against a device
 * recount, where the reference recounts on the host after every move (:1249).
 * The random streams are seeded xoshiro256**, not time-seeded rand().
 *
 * bm_sweep is the OpenMP tool's loop (src/re-alignment-parallel.c:556-676):
 * passes over the vertices, each swapped with its best partner when that
 * delta is positive, restarts from a perturbed best matching after a pass
 * without a swap; the best partners of a run of consecutive vertices come
 * from one device launch, and the run is cut at the first swap applied.
 *
 * bm_hillclimb is the greedy disjoint-swap climb: each round evaluates a
 * batch of uniform pairs and applies a disjoint set of positive swaps.
 * bm_hillclimb_twohop does the same with 2-hop proposals, the device-wide
 * form of refine-twohop-random's loop (src/refine-twohop-random.c:676-760:
 * random anchors, up to 1000 targets each from the 2-hop set, best delta
 * applied), which applies a disjoint set of positive swaps per round instead
 * of the single best one.
 */
#include "bm_host.h"
#include "bm_rng.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* 0.7 * out-weight + 0.3 * in-weight over positive male weights
 * (calculate_vertex_scores, src/graph-re-alignment-gpu-swap.cu:767-791). */
static void vertex_scores(const bm_graph_host* g, double* s) {
    for (int32_t v = 1; v <= g->n; v++) {
        double o = 0, i = 0;
        for (int64_t e = g->out.ptr[v]; e < g->out.ptr[v + 1]; e++)
            if (g->out.adj[e].y > 0) o += g->out.adj[e].y;
        for (int64_t e = g->in.ptr[v]; e < g->in.ptr[v + 1]; e++)
            if (g->in.adj[e].y > 0) i += g->in.adj[e].y;
        s[v] = 0.7 * o + 0.3 * i;
    }
}

/* Perturbation: choose 2k distinct vertices -- the first k by score
 * (select_vertex_by_score, :794-814) when `scores` is given, otherwise all
 * uniformly -- and swap selected[i] with selected[i+k] (:816-887). */
static int perturb(bm_ctx* ctx, bm_xoshiro* rng, int32_t n, int32_t k, const double* scores, double total,
                   int32_t* pi, int32_t* changed, uint8_t* used) {
    memset(used, 0, (size_t)n + 1);
    int32_t cnt = 0;
    if (scores) {
        double remaining = total;
        while (cnt < k) {
            double r = bm_xoshiro_unit(rng) * remaining;
            double cum = 0;
            int32_t pick = 0;
            for (int32_t i = 1; i <= n; i++) {
                if (used[i]) continue;
                cum += scores[i];
                if (cum >= r) { pick = i; break; }
            }
            if (!pick)
                for (int32_t i = 1; i <= n && !pick; i++)
                    if (!used[i]) pick = i;
            used[pick] = 1;
            remaining -= scores[pick];
            changed[cnt++] = pick;
        }
    }
    while (cnt < 2 * k) {
        int32_t x = (int32_t)bm_xoshiro_below(rng, (uint32_t)n) + 1;
        if (used[x]) continue;
        used[x] = 1;
        changed[cnt++] = x;
    }
    for (int32_t i = 0; i < k; i++) {
        const int32_t a = changed[i], b = changed[i + k];
        const int32_t t = pi[a];
        pi[a] = pi[b];
        pi[b] = t;
        int rc = bm_apply_swap(ctx, a, b);
        if (rc != BM_OK) return rc;
    }
    return BM_OK;
}

static void log_info(const bm_search_opts* o, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static void log_info(const bm_search_opts* o, const char* fmt, ...) {
    if (!o->verbose) return;
    va_list ap;
    va_start(ap, fmt);
    printf("[INFO] ");
    vprintf(fmt, ap);
    printf("\n");
    fflush(stdout);
    va_end(ap);
}

/* Write the best matching, re-read it and rescore it on the device
 * (save_intermediate_mapping, :749-763).  The context matching is restored. */
static int save_best(bm_ctx* ctx, const bm_search_opts* o, const int32_t* best, const int32_t* cur, int32_t n,
                     int64_t best_score, int32_t* scratch) {
    if (!o->out_path) return BM_OK;
    int rc = bm_write_mapping_csv(o->out_path, n, best);
    if (rc != BM_OK) return rc;
    int64_t lines = 0;
    rc = bm_load_mapping_csv(o->out_path, n, scratch, &lines);
    if (rc != BM_OK) return rc;
    rc = bm_set_mapping(ctx, scratch);
    if (rc != BM_OK) return rc;
    int64_t s = 0;
    rc = bm_score(ctx, &s);
    if (rc != BM_OK) return rc;
    if (s != best_score) {
        fprintf(stderr, "[ERROR] save_best: score mismatch - internal: %lld, written: %lld\n", (long long)best_score,
                (long long)s);
    }
    return bm_set_mapping(ctx, cur);
}

static int try_sync(bm_ctx* ctx, const bm_search_opts* o, int32_t n, int32_t* best, int64_t* best_score,
                    const int32_t* cur, int32_t* scratch) {
    FILE* f = fopen(o->sync_path, "r");
    if (!f) return BM_OK;
    fclose(f);
    int64_t lines = 0;
    if (bm_load_mapping_csv(o->sync_path, n, scratch, &lines) != BM_OK) return BM_OK;
    int rc = bm_set_mapping(ctx, scratch);
    if (rc != BM_OK) return rc;
    int64_t s = 0;
    rc = bm_score(ctx, &s);
    if (rc != BM_OK) return rc;
    if (s > *best_score) {
        *best_score = s;
        memcpy(best, scratch, sizeof(int32_t) * (size_t)(n + 1));
        log_info(o, "adopted %s with score %lld", o->sync_path, (long long)s);
    }
    return bm_set_mapping(ctx, cur);
}

int bm_search(bm_ctx* ctx, const bm_search_opts* o, int32_t* best_pi, bm_search_result* res) {
    if (!ctx || !o || !res) return BM_ERR_ARG;
    if (o->mode < BM_SEARCH_SWAP_WEIGHTED || o->mode > BM_SEARCH_ROT3)
        return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_search: unknown mode");
    memset(res, 0, sizeof *res);
    const int32_t n = bm_num_vertices(ctx);
    const int rot3 = o->mode == BM_SEARCH_ROT3;
    if (n < 8) return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_search: n must be >= 8");
    const int32_t kmax = o->mode == BM_SEARCH_SWAP_WEIGHTED ? 101 : 5;
    if (2 * kmax > n) return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_search: graph too small for the perturbation size");
    const int32_t cmax = 2 * kmax;
    int64_t cap;
    if (rot3)
        cap = (int64_t)cmax * (cmax - 1) * (cmax - 2) / 6 + (int64_t)cmax * (cmax - 1) / 2 * (n - cmax) +
              (int64_t)cmax * (n - cmax - 1) + 16;
    else
        cap = (int64_t)cmax * (cmax - 1) / 2 + (int64_t)cmax * (n - cmax) + 16;
    bm_batch* b = NULL;
    int rc = bm_batch_create(ctx, cap, &b);
    if (rc != BM_OK) return rc;
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t* best = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t* scratch = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t* changed = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * kmax + 3));
    uint8_t* used = (uint8_t*)malloc((size_t)n + 1);
    double* scores = NULL;
    double total_score = 0;
    if (!cur || !best || !scratch || !changed || !used) { rc = BM_ERR_NOMEM; goto done; }
    if (o->mode == BM_SEARCH_SWAP_WEIGHTED) {
        scores = (double*)calloc((size_t)n + 1, sizeof(double));
        if (!scores) { rc = BM_ERR_NOMEM; goto done; }
        vertex_scores(bm_ctx_male(ctx), scores);
        for (int32_t i = 1; i <= n; i++) total_score += scores[i];
    }
    bm_xoshiro rng;
    bm_xoshiro_seed(&rng, o->seed);
    if ((rc = bm_get_mapping(ctx, cur)) != BM_OK) goto done;
    memcpy(best, cur, sizeof(int32_t) * (size_t)(n + 1));
    int64_t score = 0;
    if ((rc = bm_score(ctx, &score)) != BM_OK) goto done;
    int64_t best_score = score;
    res->initial_score = score;
    const double t0 = now_s();
    double last_sync = t0;
    int64_t iter = 0, last = 0, since_verify = 0;
    const int64_t restart_window = rot3 ? 1500 : 500;
    const double gap = o->restart_gap > 0 ? o->restart_gap : 0.002;
    for (;;) {
        if (o->stop_flag && *o->stop_flag) break;
        if (o->max_iters > 0 && iter >= o->max_iters) break;
        if (o->time_limit_s > 0 && now_s() - t0 >= o->time_limit_s) break;
        if (o->sync_path && o->sync_interval_s > 0 && now_s() - last_sync >= o->sync_interval_s) {
            if ((rc = try_sync(ctx, o, n, best, &best_score, cur, scratch)) != BM_OK) goto done;
            last_sync = now_s();
        }
        /* ---- perturb ---- */
        int32_t k;
        if (o->mode == BM_SEARCH_SWAP_WEIGHTED)
            k = 100 + (int32_t)bm_xoshiro_below(&rng, 2);
        else
            k = 3 + (int32_t)bm_xoshiro_below(&rng, 3);
        if (iter - last < restart_window || (best_score > 0 && (double)(best_score - score) / (double)best_score > gap)) {
            memcpy(cur, best, sizeof(int32_t) * (size_t)(n + 1));
            if ((rc = bm_set_mapping(ctx, cur)) != BM_OK) goto done;
        }
        rc = perturb(ctx, &rng, n, k, o->mode == BM_SEARCH_SWAP_WEIGHTED ? scores : NULL, total_score, cur, changed,
                     used);
        if (rc != BM_OK) goto done;
        if ((rc = bm_score(ctx, &score)) != BM_OK) goto done;
        int32_t cnum = 2 * k;
        const int64_t iter_at_perturb = iter;
        /* ---- evaluate + walk ---- */
        int64_t zerocheck = 0;
        for (;;) {
            if (rot3) {
                if ((rc = bm_batch_gen_changed_triples(b, changed, cnum)) != BM_OK) goto done;
                if ((rc = bm_batch_eval_rot3(b)) != BM_OK) goto done;
            } else {
                if ((rc = bm_batch_gen_changed_pairs(b, changed, cnum)) != BM_OK) goto done;
                if ((rc = bm_batch_eval_swaps(b)) != BM_OK) goto done;
            }
            iter++;
            res->evaluations += bm_batch_size(b);
            bm_stats st;
            if ((rc = bm_batch_stats(b, &st)) != BM_OK) goto done;
            if (!(st.max > 0 || (st.n_zero > 0 && zerocheck < 1000))) break;
            if (o->stop_flag && *o->stop_flag) break;
            if (o->max_iters > 0 && iter >= o->max_iters) break;
            if (o->time_limit_s > 0 && now_s() - t0 >= o->time_limit_s) break;
            int64_t pick = -1;
            if (st.max > 0) {
                if (bm_xoshiro_below(&rng, 2) == 0) {
                    int64_t r = (int64_t)(bm_xoshiro_next(&rng) % (uint64_t)st.n_pos);
                    if ((rc = bm_batch_pick(b, 1, r, &pick)) != BM_OK) goto done;
                } else {
                    pick = st.argmax;
                }
                zerocheck = 0;
            } else {
                int64_t r = (int64_t)(bm_xoshiro_next(&rng) % (uint64_t)st.n_zero);
                if ((rc = bm_batch_pick(b, 0, r, &pick)) != BM_OK) goto done;
                zerocheck++;
            }
            if (pick < 0) { rc = bm_ctx_fail(ctx, BM_ERR_STATE, "bm_search: pick failed"); goto done; }
            int64_t d = 0;
            int32_t q[4] = {0, 0, 0, 0};
            uint8_t p8 = 0;
            if ((rc = bm_batch_get(b, pick, 1, &d, rot3 ? &p8 : NULL, q)) != BM_OK) goto done;
            if (rot3) {
                const int32_t f[3] = {cur[q[0]], cur[q[1]], cur[q[2]]};
                static const int8_t P[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
                for (int i = 0; i < 3; i++) cur[q[i]] = f[P[p8][i]];
                if ((rc = bm_apply_perm3(ctx, q[0], q[1], q[2], p8)) != BM_OK) goto done;
                changed[0] = q[0];
                changed[1] = q[1];
                changed[2] = q[2];
                cnum = 3;
            } else {
                const int32_t t = cur[q[0]];
                cur[q[0]] = cur[q[1]];
                cur[q[1]] = t;
                if ((rc = bm_apply_swap(ctx, q[0], q[1])) != BM_OK) goto done;
                changed[0] = q[0];
                changed[1] = q[1];
                cnum = 2;
            }
            score += d;
            res->accepted++;
            if (o->verify_every > 0 && ++since_verify >= o->verify_every) {
                since_verify = 0;
                int64_t s = 0;
                if ((rc = bm_score(ctx, &s)) != BM_OK) goto done;
                if (s != score) {
                    res->verify_failures++;
                    fprintf(stderr, "[ERROR] running score %lld != recount %lld\n", (long long)score, (long long)s);
                    score = s;
                }
            }
            if (score > best_score) {
                best_score = score;
                memcpy(best, cur, sizeof(int32_t) * (size_t)(n + 1));
                last = iter;
                log_info(o, "new best score: %lld in iter %lld", (long long)best_score, (long long)iter);
                if ((rc = save_best(ctx, o, best, cur, n, best_score, scratch)) != BM_OK) goto done;
            }
        }
        if (o->verbose && (iter / 2400) != (iter_at_perturb / 2400)) {
            log_info(o, "iteration %lld: current score %lld, best score %lld, %.1f s", (long long)iter,
                     (long long)score, (long long)best_score, now_s() - t0);
        }
    }
    res->seconds = now_s() - t0;
    res->iterations = iter;
    res->best_score = best_score;
    if ((rc = bm_score(ctx, &res->final_score)) != BM_OK) goto done;
    if (res->final_score != score) res->verify_failures++;
    if (best_pi) memcpy(best_pi, best, sizeof(int32_t) * (size_t)(n + 1));
    rc = BM_OK;
done:
    bm_batch_destroy(b);
    free(cur);
    free(best);
    free(scratch);
    free(changed);
    free(used);
    free(scores);
    return rc;
}

/* Proposal source of a climb round: uniform pairs (batch of them) or 2-hop
 * pairs (n_anchors anchors x per_anchor targets). */
typedef struct climb_src {
    int twohop;
    int32_t n_anchors, per_anchor;
} climb_src;

static int climb(bm_ctx* ctx, uint64_t seed, int64_t rounds, int64_t batch, climb_src src, int32_t verify,
                 int64_t* score_trace, bm_search_result* res) {
    if (!ctx || rounds < 0 || batch < 1 || !res) return BM_ERR_ARG;
    memset(res, 0, sizeof *res);
    bm_batch* b = NULL;
    int rc = bm_batch_create(ctx, batch, &b);
    if (rc != BM_OK) return rc;
    int64_t score = 0;
    if ((rc = bm_score(ctx, &score)) != BM_OK) goto done;
    res->initial_score = score;
    if (score_trace) score_trace[0] = score;
    const double t0 = now_s();
    for (int64_t r = 0; r < rounds; r++) {
        if (src.twohop)
            rc = bm_batch_gen_twohop(b, NULL, src.n_anchors, src.per_anchor, seed + (uint64_t)r);
        else
            rc = bm_batch_gen_pairs(b, seed + (uint64_t)r, batch);
        if (rc != BM_OK) goto done;
        if ((rc = bm_batch_eval_swaps(b)) != BM_OK) goto done;
        int64_t nacc = 0, gain = 0;
        if ((rc = bm_batch_accept_disjoint(b, 0, &nacc, &gain)) != BM_OK) goto done;
        score += gain;
        res->accepted += nacc;
        res->evaluations += batch;
        res->iterations++;
        if (verify) {
            int64_t s = 0;
            if ((rc = bm_score(ctx, &s)) != BM_OK) goto done;
            if (s != score) {
                res->verify_failures++;
                score = s;
            }
        }
        if (score_trace) score_trace[r + 1] = score;
    }
    res->seconds = now_s() - t0;
    res->best_score = score;
    if ((rc = bm_score(ctx, &res->final_score)) != BM_OK) goto done;
    if (res->final_score != score) res->verify_failures++;
    rc = BM_OK;
done:
    bm_batch_destroy(b);
    return rc;
}

int bm_hillclimb(bm_ctx* ctx, uint64_t seed, int64_t rounds, int64_t batch, int32_t verify, int64_t* score_trace,
                 bm_search_result* res) {
    climb_src src = {0, 0, 0};
    return climb(ctx, seed, rounds, batch, src, verify, score_trace, res);
}

/* Steepest ascent (src/best-swap/score-lookup.c:519-716): the best pair
 * over all pairs, applied while its delta is positive. */
int bm_best_swap_climb(bm_ctx* ctx, int64_t max_iters, int64_t verify_every, int64_t* score_trace,
                       bm_search_result* res) {
    if (!ctx || !res) return BM_ERR_ARG;
    memset(res, 0, sizeof *res);
    if (max_iters < 0 || verify_every < 0)
        return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_best_swap_climb: max_iters and verify_every must be >= 0");
    if (!bm_gain_enabled(ctx))
        return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_best_swap_climb: needs the placement-gain table (bm_gain_enable)");
    const double t0 = now_s();
    int64_t score = 0;
    int rc = bm_score(ctx, &score);
    if (rc != BM_OK) return rc;
    res->initial_score = score;
    if (score_trace) score_trace[0] = score;
    for (int64_t it = 0; it < max_iters; it++) {
        int64_t d = 0;
        int32_t u = 0, v = 0;
        if ((rc = bm_best_swap(ctx, &d, &u, &v)) != BM_OK) return rc;
        res->evaluations += (int64_t)bm_num_vertices(ctx) * (bm_num_vertices(ctx) - 1) / 2;
        if (d <= 0) break;
        if ((rc = bm_apply_swap(ctx, u, v)) != BM_OK) return rc;
        score += d;
        res->iterations++;
        res->accepted++;
        if (score_trace) score_trace[it + 1] = score;
        if (verify_every > 0 && res->iterations % verify_every == 0) {
            int64_t s = 0;
            if ((rc = bm_score(ctx, &s)) != BM_OK) return rc;
            if (s != score) {
                res->verify_failures++;
                score = s;
            }
        }
    }
    if (verify_every > 0) {
        int64_t s = 0;
        if ((rc = bm_score(ctx, &s)) != BM_OK) return rc;
        if (s != score) {
            res->verify_failures++;
            score = s;
        }
    }
    res->best_score = res->final_score = score;
    res->seconds = now_s() - t0;
    return BM_OK;
}

/* Random disjoint swaps of perturbation p: 2k distinct vertices drawn from
 * the stream bm_rand_at(bm_rand_at(seed, p), j), the i-th swapped with the
 * (i + k)-th (random_swap_k_vertices, src/re-alignment-parallel.c:495-537,
 * draws them with rand() % n). */
static int sweep_perturb(int32_t* pi, int32_t n, uint64_t seed, uint64_t p, int32_t k, uint8_t* used, int32_t* sel) {
    if (k <= 0) return 0;
    if ((int64_t)2 * k > n) return -1;
    memset(used, 0, (size_t)n + 1);
    const uint64_t st = bm_rand_at(seed, p);
    int32_t cnt = 0;
    for (uint64_t j = 0; cnt < 2 * k; j++) {
        const int32_t x = 1 + (int32_t)bm_range32((uint32_t)(bm_rand_at(st, j) >> 32), (uint32_t)n);
        if (!used[x]) {
            used[x] = 1;
            sel[cnt++] = x;
        }
    }
    for (int32_t i = 0; i < k; i++) {
        const int32_t t = pi[sel[i]];
        pi[sel[i]] = pi[sel[i + k]];
        pi[sel[i + k]] = t;
    }
    return 0;
}

int bm_sweep(bm_ctx* ctx, const bm_sweep_opts* o, int32_t* best_pi, int64_t* pass_trace, bm_search_result* res) {
    if (!ctx || !o || !res) return BM_ERR_ARG;
    memset(res, 0, sizeof *res);
    if (!bm_gain_enabled(ctx))
        return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_sweep: needs the placement-gain table (bm_gain_enable)");
    if (o->max_passes < 0 || o->time_limit_s < 0 || (o->max_passes == 0 && o->time_limit_s <= 0))
        return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_sweep: set max_passes > 0 or time_limit_s > 0 (both >= 0)");
    if (o->initial_swaps < 0 || o->restart_swaps < 0 || o->batch < 0)
        return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_sweep: initial_swaps, restart_swaps and batch must be >= 0");
    const int32_t n = bm_num_vertices(ctx);
    if ((int64_t)2 * o->initial_swaps > n || (int64_t)2 * o->restart_swaps > n)
        return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_sweep: a perturbation of k swaps needs 2k <= n vertices");
    const int32_t batch = o->batch > 0 ? o->batch : 512;
    const double t0 = now_s();
    int32_t* pi = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t* best = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t* sel = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    uint8_t* used = (uint8_t*)malloc((size_t)n + 1);
    int64_t* dl = (int64_t*)malloc(sizeof(int64_t) * (size_t)batch);
    int32_t* vl = (int32_t*)malloc(sizeof(int32_t) * (size_t)batch);
    int rc = BM_OK;
    if (!pi || !best || !sel || !used || !dl || !vl) {
        rc = bm_ctx_fail(ctx, BM_ERR_NOMEM, "bm_sweep: out of host memory");
        goto done;
    }
    int64_t score = 0;
    if ((rc = bm_score(ctx, &score)) != BM_OK) goto done;
    res->initial_score = score;
    if ((rc = bm_get_mapping(ctx, pi)) != BM_OK) goto done;
    uint64_t nperturb = 0;
    sweep_perturb(pi, n, o->seed, nperturb++, o->initial_swaps, used, sel);
    if (o->initial_swaps > 0) {
        if ((rc = bm_set_mapping(ctx, pi)) != BM_OK || (rc = bm_score(ctx, &score)) != BM_OK) goto done;
    }
    int64_t best_score = score;
    memcpy(best, pi, sizeof(int32_t) * ((size_t)n + 1));
    int64_t pass = 0, improvements = 0, last = 0, passes = 0;
    int out_of_time = 0;
    while (!out_of_time && (o->max_passes == 0 || passes < o->max_passes)) {
        for (int32_t u = 1; u <= n;) {
            if (o->time_limit_s > 0 && now_s() - t0 >= o->time_limit_s) {
                out_of_time = 1;
                break;
            }
            const int32_t cnt = n - u + 1 < batch ? n - u + 1 : batch;
            if ((rc = bm_best_partners(ctx, u, cnt, dl, vl)) != BM_OK) goto done;
            int32_t i = 0;
            while (i < cnt && !(dl[i] > 0)) {
                res->evaluations += n - (u + i);
                i++;
            }
            if (i == cnt) {
                u += cnt;
                continue;
            }
            res->evaluations += n - (u + i);
            const int32_t a = u + i, b = vl[i];
            if ((rc = bm_apply_swap(ctx, a, b)) != BM_OK) goto done;
            const int32_t t = pi[a];
            pi[a] = pi[b];
            pi[b] = t;
            score += dl[i];
            improvements++;
            res->accepted++;
            if (score > best_score) {
                best_score = score;
                memcpy(best, pi, sizeof(int32_t) * ((size_t)n + 1));
            }
            if (o->verify) {
                int64_t s2 = 0;
                if ((rc = bm_score(ctx, &s2)) != BM_OK) goto done;
                if (s2 != score) {
                    res->verify_failures++;
                    score = s2;
                }
            }
            u = a + 1;
        }
        if (out_of_time) break;
        pass++;
        passes++;
        if (pass_trace && o->max_passes > 0) pass_trace[passes - 1] = best_score;
        if (pass > 2 && improvements == last) {   /* no swap in this pass: restart from the best, perturbed */
            memcpy(pi, best, sizeof(int32_t) * ((size_t)n + 1));
            sweep_perturb(pi, n, o->seed, nperturb++, o->restart_swaps, used, sel);
            if ((rc = bm_set_mapping(ctx, pi)) != BM_OK || (rc = bm_score(ctx, &score)) != BM_OK) goto done;
            improvements = last = pass = 0;
            res->restarts++;
        }
        if (last < improvements) last = improvements;
    }
    res->iterations = passes;
    if ((rc = bm_set_mapping(ctx, best)) != BM_OK) goto done;
    res->best_score = res->final_score = best_score;
    if (o->verify) {
        int64_t s2 = 0;
        if ((rc = bm_score(ctx, &s2)) != BM_OK) goto done;
        if (s2 != best_score) res->verify_failures++;
    }
    if (best_pi) memcpy(best_pi, best, sizeof(int32_t) * ((size_t)n + 1));
    res->seconds = now_s() - t0;
done:
    free(pi);
    free(best);
    free(sel);
    free(used);
    free(dl);
    free(vl);
    return rc;
}

int bm_hillclimb_twohop(bm_ctx* ctx, uint64_t seed, int64_t rounds, int32_t anchors, int32_t per_anchor,
                        int32_t verify, int64_t* score_trace, bm_search_result* res) {
    if (anchors < 1 || per_anchor < 1) return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_hillclimb_twohop: anchors and per_anchor must be >= 1");
    climb_src src = {1, anchors, per_anchor};
    return climb(ctx, seed, rounds, (int64_t)anchors * per_anchor, src, verify, score_trace, res);
}

/* Disjoint-swap climb with a chosen proposal source.  Round r: candidates
 * from stream bm_rand_at(seed, r) -- per_round uniform pairs, or 2-hop pairs
 * (per_round / per_anchor anchors), or the per_round best-scored of `pool`
 * uniform pairs under the learned proposal -- are evaluated exactly and a
 * disjoint set of positive swaps is applied.  A learned climb may retrain the
 * network on the candidates it has just evaluated (their exact deltas are
 * the targets) and, with fallback, take uniform rounds while those gain more
 * per second (the uniform and 2-hop proposals stay the fallback of the
 * learned one).  trace_score / trace_time (trace_cap entries each, may be
 * NULL) get the score and elapsed seconds after each round. */
int bm_climb(bm_ctx* ctx, const bm_climb_opts* o, int64_t* trace_score, double* trace_time, int64_t trace_cap,
             bm_search_result* res) {
    if (!ctx || !o || !res) return BM_ERR_ARG;
    memset(res, 0, sizeof *res);
    if (o->per_round < 1 || (o->time_limit_s <= 0 && o->max_rounds <= 0))
        return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_climb: per_round >= 1 and a time or round limit are required");
    if (o->proposal == BM_PROPOSAL_LEARNED && (!o->mlp || o->pool < o->per_round || o->refresh_every < 1))
        return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_climb: the learned proposal needs mlp, pool >= per_round, refresh_every >= 1");
    if (o->proposal == BM_PROPOSAL_LEARNED && (o->retrain_every < 0 || (o->retrain_every > 0 && !(o->retrain_lr > 0.f))))
        return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_climb: retrain_every >= 0, and retrain_lr > 0 when retraining");
    if (o->proposal == BM_PROPOSAL_TWOHOP && (o->per_anchor < 1 || o->per_round < o->per_anchor))
        return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_climb: the 2-hop proposal needs 1 <= per_anchor <= per_round");
    if (o->proposal < BM_PROPOSAL_UNIFORM || o->proposal > BM_PROPOSAL_LEARNED)
        return bm_ctx_fail(ctx, BM_ERR_ARG, "bm_climb: unknown proposal");
    const int learned = o->proposal == BM_PROPOSAL_LEARNED;
    const int mixed = learned && o->fallback;
    bm_batch* b = NULL;
    int rc = bm_batch_create(ctx, o->per_round, &b);
    if (rc != BM_OK) return rc;
    int64_t score = 0;
    if ((rc = bm_score(ctx, &score)) != BM_OK) goto done;
    res->initial_score = score;
    const double t0 = now_s();
    /* moving averages of exact gain per second: [0] learned rounds, [1] uniform rounds */
    double rate[2] = {0.0, 0.0};
    int64_t learned_rounds = 0;
    for (int64_t r = 0;; r++) {
        if (o->max_rounds > 0 && r >= o->max_rounds) break;
        const double tr = now_s();
        if (o->time_limit_s > 0 && tr - t0 >= o->time_limit_s) break;
        const uint64_t sd = bm_rand_at(o->seed, (uint64_t)r);
        int use_learned = learned;
        if (mixed) {
            if (r < 2)
                use_learned = r == 0;
            else {
                use_learned = rate[0] >= rate[1];
                if (r % 16 == 15) use_learned = !use_learned;
            }
        }
        if (o->proposal == BM_PROPOSAL_TWOHOP) {
            rc = bm_batch_gen_twohop(b, NULL, (int32_t)(o->per_round / o->per_anchor), o->per_anchor, sd);
        } else if (use_learned) {
            if (learned_rounds % o->refresh_every == 0 && (rc = bm_mlp_refresh_features(o->mlp)) != BM_OK) goto done;
            rc = bm_batch_gen_learned(b, o->mlp, sd, o->pool, o->per_round);
        } else {
            rc = bm_batch_gen_pairs(b, sd, o->per_round);
        }
        if (rc != BM_OK) goto done;
        if ((rc = bm_batch_eval_swaps(b)) != BM_OK) goto done;
        if (use_learned) {
            /* the features still describe the matching the deltas were taken on */
            if (o->retrain_every > 0 && learned_rounds % o->retrain_every == o->retrain_every - 1 &&
                (rc = bm_mlp_train_batch(o->mlp, b, o->retrain_lr, NULL)) != BM_OK)
                goto done;
            learned_rounds++;
        } else if (learned) {
            res->fallback_rounds++;
        }
        int64_t nacc = 0, gain = 0;
        if ((rc = bm_batch_accept_disjoint(b, 0, &nacc, &gain)) != BM_OK) goto done;
        score += gain;
        res->accepted += nacc;
        res->evaluations += bm_batch_size(b);
        res->iterations++;
        if (mixed) {
            const double dt = now_s() - tr;
            const int src = use_learned ? 0 : 1;
            const double g = dt > 0 ? (double)gain / dt : 0.0;
            rate[src] = r < 2 ? g : rate[src] + (g - rate[src]) * 0.125;
        }
        if (o->verify) {
            int64_t s = 0;
            if ((rc = bm_score(ctx, &s)) != BM_OK) goto done;
            if (s != score) {
                res->verify_failures++;
                score = s;
            }
        }
        if (r < trace_cap) {
            if (trace_score) trace_score[r] = score;
            if (trace_time) trace_time[r] = now_s() - t0;
        }
    }
    res->seconds = now_s() - t0;
    res->best_score = score;
    if ((rc = bm_score(ctx, &res->final_score)) != BM_OK) goto done;
    if (res->final_score != score) res->verify_failures++;
    rc = BM_OK;
done:
    bm_batch_destroy(b);
    return rc;
}
