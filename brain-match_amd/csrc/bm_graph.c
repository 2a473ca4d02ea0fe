/*
 * Edge-list -> CSR/CSC build with score-matching's duplicate rule, and the
 * per-row female hash tables.
 *
 * Duplicates: score-matching's add_edge keeps one edge per (src,dst) and
 * overwrites its weight, so the last line wins (src/score-matching.c:339-359).
 * The reference swap tools instead append every line to their edge lists
 * while the dense matrix keeps the last weight
 * (src/graph-re-alignment-gpu-swap.cu:454-475); the oracle's rule is kept.
 */
#include "bm_host.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static _Thread_local char g_err[512];

void bm_set_global_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

const char* bm_global_error(void) { return g_err; }

void bm_graph_free(bm_graph_host* g) {
    if (!g) return;
    free(g->out.ptr);
    free(g->out.adj);
    free(g->in.ptr);
    free(g->in.adj);
    memset(g, 0, sizeof *g);
}

/* Stable counting sort of edge indices `idx` by key[idx[i]] in 0..n. */
static int counting_sort(const int32_t* key, int32_t n, int64_t m, const int64_t* idx_in,
                         int64_t* idx_out, int64_t* count /* n+2 */) {
    memset(count, 0, sizeof(int64_t) * (size_t)(n + 2));
    for (int64_t i = 0; i < m; i++) count[key[idx_in[i]] + 1]++;
    for (int32_t r = 0; r <= n; r++) count[r + 1] += count[r];
    /* count[r] is now the start of bucket r; use a running copy. */
    int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    if (!pos) return BM_ERR_NOMEM;
    memcpy(pos, count, sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t i = 0; i < m; i++) {
        int64_t e = idx_in[i];
        idx_out[pos[key[e]]++] = e;
    }
    free(pos);
    return BM_OK;
}

int bm_graph_build(int32_t n, const bm_edges* e, bm_graph_host* g, const char** err) {
    memset(g, 0, sizeof *g);
    if (n < 1) { *err = "n must be >= 1"; return BM_ERR_ARG; }
    if (!e || e->m < 0 || (e->m > 0 && (!e->src || !e->dst || !e->w))) {
        *err = "invalid edge list";
        return BM_ERR_ARG;
    }
    const int64_t m = e->m;
    for (int64_t i = 0; i < m; i++) {
        if (e->src[i] < 1 || e->src[i] > n || e->dst[i] < 1 || e->dst[i] > n) {
            *err = "edge endpoint outside 1..n";
            return BM_ERR_ARG;
        }
        if (e->w[i] == INT32_MIN) { *err = "edge weight INT32_MIN is not supported"; return BM_ERR_ARG; }
    }
    int64_t* a = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m ? m : 1));
    int64_t* b = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m ? m : 1));
    int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 2));
    if (!a || !b || !cnt) { free(a); free(b); free(cnt); *err = "out of host memory"; return BM_ERR_NOMEM; }
    for (int64_t i = 0; i < m; i++) a[i] = i;
    /* LSD: stable by dst, then stable by src -> sorted by (src, dst), and
     * lines with equal (src, dst) stay in file order. */
    int rc = counting_sort(e->dst, n, m, a, b, cnt);
    if (rc == BM_OK) rc = counting_sort(e->src, n, m, b, a, cnt);
    if (rc != BM_OK) { free(a); free(b); free(cnt); *err = "out of host memory"; return rc; }
    /* Deduplicate: keep the last line of each (src, dst) run. */
    int64_t mu = 0;
    for (int64_t i = 0; i < m; i++) {
        int64_t ei = a[i];
        if (i + 1 < m) {
            int64_t ej = a[i + 1];
            if (e->src[ej] == e->src[ei] && e->dst[ej] == e->dst[ei]) continue;
        }
        b[mu++] = ei;
    }
    g->n = n;
    g->m = mu;
    g->out.n = g->in.n = n;
    g->out.m = g->in.m = mu;
    g->out.ptr = (int64_t*)calloc((size_t)(n + 2), sizeof(int64_t));
    g->in.ptr = (int64_t*)calloc((size_t)(n + 2), sizeof(int64_t));
    g->out.adj = (bm_pair32*)malloc(sizeof(bm_pair32) * (size_t)(mu ? mu : 1));
    g->in.adj = (bm_pair32*)malloc(sizeof(bm_pair32) * (size_t)(mu ? mu : 1));
    if (!g->out.ptr || !g->in.ptr || !g->out.adj || !g->in.adj) {
        free(a); free(b); free(cnt); bm_graph_free(g);
        *err = "out of host memory";
        return BM_ERR_NOMEM;
    }
    for (int64_t i = 0; i < mu; i++) {
        int64_t ei = b[i];
        g->out.ptr[e->src[ei] + 1]++;
        g->in.ptr[e->dst[ei] + 1]++;
    }
    for (int32_t r = 0; r <= n; r++) {
        g->out.ptr[r + 1] += g->out.ptr[r];
        g->in.ptr[r + 1] += g->in.ptr[r];
    }
    /* Out rows: b is already sorted by (src, dst). */
    for (int64_t i = 0; i < mu; i++) {
        int64_t ei = b[i];
        g->out.adj[i].x = e->dst[ei];
        g->out.adj[i].y = e->w[ei];
    }
    /* In rows: scatter in (src, dst) order -> sources ascending per column. */
    for (int32_t r = 0; r <= n; r++) cnt[r] = g->in.ptr[r];
    for (int64_t i = 0; i < mu; i++) {
        int64_t ei = b[i];
        int64_t p = cnt[e->dst[ei]]++;
        g->in.adj[p].x = e->src[ei];
        g->in.adj[p].y = e->w[ei];
    }
    free(a);
    free(b);
    free(cnt);
    return BM_OK;
}

void bm_rowhash_free(bm_rowhash_host* h) {
    if (!h) return;
    free(h->desc);
    free(h->tab);
    memset(h, 0, sizeof *h);
}

static uint32_t table_bits_for(int64_t deg) {
    if (deg <= 0) return 0;
    uint32_t bits = 1;
    while (((int64_t)1 << bits) < 2 * deg) bits++;
    return bits;
}

int bm_rowhash_build(const bm_csr_host* rows, bm_rowhash_host* h) {
    memset(h, 0, sizeof *h);
    const int32_t n = rows->n;
    h->n = n;
    h->desc = (uint64_t*)calloc((size_t)(n + 1), sizeof(uint64_t));
    if (!h->desc) return BM_ERR_NOMEM;
    int64_t total = 0;
    for (int32_t r = 0; r <= n; r++) {
        int64_t d = rows->ptr[r + 1] - rows->ptr[r];
        uint32_t bits = table_bits_for(d);
        h->desc[r] = ((uint64_t)total << 6) | bits;
        if (bits) total += (int64_t)1 << bits;
    }
    h->slots = total;
    h->tab = (bm_pair32*)calloc((size_t)(total ? total : 2), sizeof(bm_pair32));
    if (!h->tab) { bm_rowhash_free(h); return BM_ERR_NOMEM; }
    for (int32_t r = 0; r <= n; r++) {
        uint32_t bits = (uint32_t)(h->desc[r] & 63u);
        if (!bits) continue;
        bm_pair32* t = h->tab + (h->desc[r] >> 6);
        uint32_t mask = (1u << bits) - 1u;
        for (int64_t p = rows->ptr[r]; p < rows->ptr[r + 1]; p++) {
            uint32_t s = bm_slot_hash(rows->adj[p].x, bits);
            while (t[s].x != 0) s = (s + 1u) & mask;
            t[s] = rows->adj[p];
        }
    }
    return BM_OK;
}

int bm_rowhash_get(const bm_rowhash_host* h, int32_t row, int32_t key, int32_t* w) {
    if (row < 0 || row > h->n) return 0;
    uint32_t bits = (uint32_t)(h->desc[row] & 63u);
    if (!bits) return 0;
    const bm_pair32* t = h->tab + (h->desc[row] >> 6);
    uint32_t mask = (1u << bits) - 1u;
    uint32_t s = bm_slot_hash(key, bits);
    for (;;) {
        if (t[s].x == 0) return 0;
        if (t[s].x == key) { *w = t[s].y; return 1; }
        s = (s + 1u) & mask;
    }
}
