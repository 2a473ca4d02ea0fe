/*
 * CSV front end: graph and matching files in the reference's formats.
 *
 * Graph lines are parsed with score-matching's rules
 * (src/score-matching.c:361-398 parse_vertex_id, :400-488 load_graph_sparse):
 * the first line is a header and is skipped, a line must have three
 * comma-separated fields, an id may carry one 'm' or 'f' prefix and any other
 * non-digit characters before its digits, the weight is a decimal int, and
 * lines that fail are skipped.  Matching lines follow load_matching
 * (:490-556) and load_benchmark_mapping (src/graph-re-alignment-gpu-swap.cu:699-746):
 * "male,female", generic or m/f-prefixed.
 */
#include "bm_host.h"

#include <ctype.h>
#include <errno.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void bm_edge_list_free(bm_edge_list* g) {
    if (!g) return;
    free(g->src);
    free(g->dst);
    free(g->w);
    memset(g, 0, sizeof *g);
}

static char* slurp(const char* path, size_t* len) {
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    size_t cap = 1 << 20, n = 0;
    char* buf = (char*)malloc(cap + 1);
    if (!buf) { fclose(f); return NULL; }
    for (;;) {
        if (n == cap) {
            cap *= 2;
            char* nb = (char*)realloc(buf, cap + 1);
            if (!nb) { free(buf); fclose(f); return NULL; }
            buf = nb;
        }
        size_t r = fread(buf + n, 1, cap - n, f);
        n += r;
        if (r == 0) break;
    }
    fclose(f);
    buf[n] = '\0';
    *len = n;
    return buf;
}

/* parse_vertex_id semantics on the field [s, e): -1 when no digits. */
static int64_t parse_id(const char* s, const char* e) {
    while (s < e && isspace((unsigned char)*s)) s++;
    if (s < e && (*s == 'm' || *s == 'f')) s++;
    while (s < e && !isdigit((unsigned char)*s)) s++;
    if (s >= e) return -1;
    int64_t v = 0;
    while (s < e && isdigit((unsigned char)*s)) {
        v = v * 10 + (*s - '0');
        if (v > INT_MAX) return -1;
        s++;
    }
    return v;
}

/* sscanf "%d" semantics on [s, e): leading space, optional sign, digits. */
static int parse_int(const char* s, const char* e, int32_t* out) {
    while (s < e && isspace((unsigned char)*s)) s++;
    int neg = 0;
    if (s < e && (*s == '+' || *s == '-')) { neg = (*s == '-'); s++; }
    if (s >= e || !isdigit((unsigned char)*s)) return 0;
    int64_t v = 0;
    while (s < e && isdigit((unsigned char)*s)) {
        v = v * 10 + (*s - '0');
        if (v > (int64_t)INT_MAX + 1) v = (int64_t)INT_MAX + 1;
        s++;
    }
    if (neg) v = -v;
    if (v > INT_MAX) v = INT_MAX;
    if (v < INT_MIN) v = INT_MIN;
    *out = (int32_t)v;
    return 1;
}

int bm_load_graph_csv(const char* path, bm_edge_list* out) {
    memset(out, 0, sizeof *out);
    size_t len = 0;
    char* buf = slurp(path, &len);
    if (!buf) { bm_set_global_error("cannot read %s: %s", path, strerror(errno)); return BM_ERR_IO; }
    int64_t lines = 0;
    for (size_t i = 0; i < len; i++) lines += (buf[i] == '\n');
    int64_t cap = lines + 1;
    out->src = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
    out->dst = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
    out->w = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
    if (!out->src || !out->dst || !out->w) { free(buf); bm_edge_list_free(out); return BM_ERR_NOMEM; }
    const char* p = buf;
    const char* end = buf + len;
    /* skip the header line */
    while (p < end && *p != '\n') p++;
    if (p < end) p++;
    int64_t m = 0;
    int32_t nmax = 0;
    while (p < end) {
        const char* ls = p;
        while (p < end && *p != '\n') p++;
        const char* le = p;
        if (p < end) p++;
        const char* c1 = memchr(ls, ',', (size_t)(le - ls));
        if (!c1 || c1 == ls) continue;
        const char* c2 = memchr(c1 + 1, ',', (size_t)(le - c1 - 1));
        if (!c2 || c2 == c1 + 1) continue;
        int32_t w;
        if (!parse_int(c2 + 1, le, &w)) continue;
        int64_t s = parse_id(ls, c1), t = parse_id(c1 + 1, c2);
        if (s < 0 || t < 0) continue;
        if (s == 0 || t == 0) {
            free(buf);
            bm_edge_list_free(out);
            bm_set_global_error("%s: vertex id 0 is not supported (ids start at 1)", path);
            return BM_ERR_IO;
        }
        out->src[m] = (int32_t)s;
        out->dst[m] = (int32_t)t;
        out->w[m] = w;
        if (s > nmax) nmax = (int32_t)s;
        if (t > nmax) nmax = (int32_t)t;
        m++;
    }
    free(buf);
    out->m = m;
    out->n = nmax;
    return BM_OK;
}

int bm_write_graph_csv(const char* path, const bm_edge_list* g) {
    FILE* f = fopen(path, "w");
    if (!f) { bm_set_global_error("cannot write %s: %s", path, strerror(errno)); return BM_ERR_IO; }
    fprintf(f, "from,to,weight\n");
    for (int64_t i = 0; i < g->m; i++) fprintf(f, "%d,%d,%d\n", g->src[i], g->dst[i], g->w[i]);
    if (fclose(f) != 0) return BM_ERR_IO;
    return BM_OK;
}

int bm_load_mapping_csv(const char* path, int32_t n, int32_t* pi, int64_t* n_lines) {
    size_t len = 0;
    char* buf = slurp(path, &len);
    if (!buf) { bm_set_global_error("cannot read %s: %s", path, strerror(errno)); return BM_ERR_IO; }
    memset(pi, 0, sizeof(int32_t) * (size_t)(n + 1));
    const char* p = buf;
    const char* end = buf + len;
    while (p < end && *p != '\n') p++;
    if (p < end) p++;
    int64_t cnt = 0;
    while (p < end) {
        const char* ls = p;
        while (p < end && *p != '\n') p++;
        const char* le = p;
        if (p < end) p++;
        const char* c1 = memchr(ls, ',', (size_t)(le - ls));
        if (!c1 || c1 == ls || c1 + 1 >= le) continue;
        int64_t a = parse_id(ls, c1), b = parse_id(c1 + 1, le);
        if (a < 0 || b < 0) continue;
        if (a < 1 || a > n || b > n) {
            free(buf);
            bm_set_global_error("%s: matching line %lld maps %lld -> %lld outside 1..%d", path,
                                (long long)(cnt + 2), (long long)a, (long long)b, n);
            return BM_ERR_ARG;
        }
        pi[a] = (int32_t)b;
        cnt++;
    }
    free(buf);
    if (n_lines) *n_lines = cnt;
    return BM_OK;
}

int bm_write_mapping_csv(const char* path, int32_t n, const int32_t* pi) {
    FILE* f = fopen(path, "w");
    if (!f) { bm_set_global_error("cannot write %s: %s", path, strerror(errno)); return BM_ERR_IO; }
    fprintf(f, "Male Node ID,Female Node ID\n");
    for (int32_t i = 1; i <= n; i++)
        if (pi[i] != 0) fprintf(f, "m%d,f%d\n", i, pi[i]);
    if (fclose(f) != 0) { bm_set_global_error("write of %s failed", path); return BM_ERR_IO; }
    return BM_OK;
}

/* A field of a matching line, as the reference converters scan it with
 * sscanf's %[...]: the longest non-empty run of characters from s that are
 * not in stop.  Returns its end, or NULL when the run is empty. */
static const char* scan_field(const char* s, const char* stop) {
    const char* p = s;
    while (*p && !strchr(stop, *p)) p++;
    return p > s ? p : NULL;
}

static int all_digits(const char* s, const char* e) {
    for (; s < e; s++)
        if (!isdigit((unsigned char)*s)) return 0;
    return 1;
}

int bm_convert_mapping_csv(const char* in_path, const char* out_path, int to_prefixed, int64_t* n_lines) {
    FILE* in = fopen(in_path, "r");
    if (!in) { bm_set_global_error("Error: Cannot open input file '%s'", in_path); return BM_ERR_IO; }
    FILE* out = fopen(out_path, "w");
    if (!out) {
        fclose(in);
        bm_set_global_error("Error: Cannot create output file '%s'", out_path);
        return BM_ERR_IO;
    }
    fputs(to_prefixed ? "Male Node ID,Female Node ID\n" : "Male ID,Female ID\n", out);
    /* lines are read as the reference reads them: fgets into 1,024 bytes,
     * the first one (the header) skipped */
    char line[1024];
    int64_t line_number = 0;
    int rc = BM_OK;
    if (fgets(line, sizeof line, in)) line_number++;
    while (fgets(line, sizeof line, in)) {
        line_number++;
        const char* m = line;
        if (!to_prefixed) {   /* "m<male>,f<female>" (src/convert-VNC-matching.c:64-69) */
            if (*m != 'm') { rc = BM_ERR_ARG; break; }
            m++;
        }
        const char* me = scan_field(m, ",");
        if (!me || *me != ',') { rc = BM_ERR_ARG; break; }
        const char* f = me + 1;
        if (!to_prefixed) {
            if (*f != 'f') { rc = BM_ERR_ARG; break; }
            f++;
        }
        const char* fe = scan_field(f, "\n\r");
        if (!fe) { rc = BM_ERR_ARG; break; }
        if (to_prefixed) {   /* "<male>,<female>", digits only (src/transform-solution.c:70-92) */
            if (!all_digits(m, me) || !all_digits(f, fe)) {
                bm_set_global_error("Error on line %lld: IDs must contain only numbers", (long long)line_number);
                rc = BM_ERR_IO;
                break;
            }
            fprintf(out, "m%.*s,f%.*s\n", (int)(me - m), m, (int)(fe - f), f);
        } else {
            fprintf(out, "%.*s,%.*s\n", (int)(me - m), m, (int)(fe - f), f);
        }
    }
    if (rc == BM_ERR_ARG) {   /* a line the reference's sscanf would not match */
        bm_set_global_error("Error on line %lld: Malformed input line", (long long)line_number);
        rc = BM_ERR_IO;
    }
    fclose(in);
    if (fclose(out) != 0 && rc == BM_OK) {
        bm_set_global_error("write of %s failed", out_path);
        rc = BM_ERR_IO;
    }
    if (n_lines) *n_lines = line_number;
    return rc;
}
