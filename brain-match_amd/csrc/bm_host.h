/*
 * Host-side graph layout shared by the C ABI implementation, the CSV/synthetic
 * front ends and the search drivers.  Not part of the public ABI.
 *
 * Layout (identical on host and device):
 *   - male out-lists (CSR) and in-lists (CSC): ptr[n+2] int64 row offsets for
 *     rows 0..n (row 0 is always empty) and adj[m] of {neighbour, weight}
 *     int32 pairs, neighbours ascending within a row;
 *   - female out-rows and in-rows: the same CSR/CSC, plus one open-addressing
 *     hash table per row (linear probing, power-of-two size >= 2*deg, slot =
 *     {key, weight}, key 0 = empty since vertex ids start at 1).  A row's
 *     descriptor is (slot_offset << 6) | log2(size); log2(size) = 0 marks an
 *     empty row.  The tables replace the reference's dense (n+1)^2 int16
 *     female matrix (src/graph-re-alignment-gpu-swap.cu:396-418).
 */
#ifndef BM_HOST_H
#define BM_HOST_H

#include <stdint.h>
#include <stddef.h>
#include "../../include/brainmatch.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bm_pair32 { int32_t x; int32_t y; } bm_pair32;  /* {neighbour|key, weight} */

typedef struct bm_csr_host {
    int32_t n;
    int64_t m;
    int64_t* ptr;      /* n+2 */
    bm_pair32* adj;    /* m */
} bm_csr_host;

typedef struct bm_rowhash_host {
    int32_t n;
    int64_t slots;     /* total slots over all rows */
    uint64_t* desc;    /* n+1 descriptors */
    bm_pair32* tab;    /* slots */
} bm_rowhash_host;

typedef struct bm_graph_host {
    int32_t n;
    int64_t m;
    bm_csr_host out;   /* CSR by source */
    bm_csr_host in;    /* CSC by destination (neighbour = source) */
} bm_graph_host;

/* Hash of a vertex id into a table of 2^bits slots (bits >= 1). */
static inline uint32_t bm_slot_hash(int32_t key, uint32_t bits) {
    return ((uint32_t)key * 0x9E3779B1u) >> (32u - bits);
}

/* Sets *err to a static message on failure and returns a BM_ERR_* code. */
int bm_graph_build(int32_t n, const bm_edges* e, bm_graph_host* g, const char** err);
void bm_graph_free(bm_graph_host* g);
int bm_rowhash_build(const bm_csr_host* rows, bm_rowhash_host* h);
void bm_rowhash_free(bm_rowhash_host* h);
/* Lookup in a host row table: returns 1 and the weight if present. */
int bm_rowhash_get(const bm_rowhash_host* h, int32_t row, int32_t key, int32_t* w);

/* Thread-local error string for calls without a context. */
void bm_set_global_error(const char* fmt, ...);

/* Host copy of the context's male graph (read-only). */
const bm_graph_host* bm_ctx_male(const bm_ctx* ctx);
/* Sets the context error message (for host drivers). */
int bm_ctx_fail(bm_ctx* ctx, int code, const char* msg);

#ifdef __cplusplus
}
#endif

#endif
