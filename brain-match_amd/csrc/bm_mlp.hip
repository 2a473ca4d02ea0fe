/*
 * Learned swap proposal (gfx950): a 16 -> 128 -> 128 -> 1 perceptron in BF16
 * with FP32 accumulation on the matrix cores, with hand-written forward and
 * backward kernels.  It scores candidate pairs from per-vertex features of
 * the current matching; a search sends only the top candidates of a large
 * pool to the exact delta kernels, so the score stays exact whatever the
 * network predicts.  There is no counterpart in the reference (its proposals
 * are uniform, score-weighted or 2-hop: src/graph-re-alignment-gpu-swap.cu:816-887,
 * src/refine-twohop-random.c:676-693).
 *
 * Model (x = the 16 pair features, bf() = rounding to bfloat16, nearest even):
 *   z1 = bf(x) @ bf(W1) + b1          W1: 16 x 128
 *   h1 = bf(relu(z1))
 *   z2 = h1 @ bf(W2) + b2             W2: 128 x 128
 *   h2 = relu(z2)                     (fp32)
 *   y  = h2 . w3 + b3                 (fp32)
 * Loss: mean (y - t)^2 with t = sign(delta) log(1 + |delta|).  Backward:
 *   dy = 2 (y - t) / B;  dw3 = sum h2 dy;  db3 = sum dy
 *   g2 = bf(dy w3 * [z2 > 0]);  dW2 = h1^T g2;  db2 = sum g2 (unrounded)
 *   g1 = bf((g2 @ bf(W2)^T) * [z1 > 0]);  dW1 = bf(x)^T g1;  db1 = sum g1 (unrounded)
 * Products with K = 16 (layer 1) run on v_mfma_f32_16x16x16_bf16, those with
 * K = 128 or K = the batch (layer 2, the backward through W2, the weight
 * gradients) on gfx950's v_mfma_f32_16x16x32_bf16.  The scoring kernel
 * computes the transposed products (z^T = W^T x^T: the weights are the A
 * operand), so each lane ends a layer holding four consecutive hidden units of
 * one pair: h1 goes to LDS in 8-byte stores and the output layer's sum over
 * hidden units is one in-lane sum plus two shuffles.  The weight-gradient
 * products use activations stored transposed (feature-major) so that every
 * operand fragment is one 16-byte load.
 *
 * Fragment layouts (64 lanes, l = lane, q = l / 16):
 *   16x16x16: A lane l holds A[l % 16][4q + i], B holds B[4q + i][l % 16], i < 4;
 *   16x16x32: A lane l holds A[l % 16][8q + i], B holds B[8q + i][l % 16], i < 8;
 *   both:     D lane l holds D[4q + j][l % 16], j < 4.
 * Weights are kept in "fragment order": for a K x N matrix W and k-steps of
 * KB = 16 or 32 rows, fragment (s, t) (rows KB s.., columns 16 t..) stores
 * lane l's KB / 4 values W[KB s + (KB / 4) q + i][16 t + l % 16] contiguously.
 * The same array serves as the B operand of W and as the A operand of W^T.
 *
 * Per-vertex features (8 per vertex; a pair's 16 are u's then v's):
 *   log1p d_out, log1p d_in, log1p (weight sum), L / (weight sum + 1),
 *   unmatched-edge fraction, log1p d_out_F(pi), log1p d_in_F(pi), log1p L,
 * with L = the score carried by the vertex's incident edges under pi and an
 * edge "unmatched" when its contribution is 0.
 */
#include "bm_device.h"
#include "bm_devutil.h"
#include "bm_rng.h"

#include <hipcub/hipcub.hpp>

using namespace bmd;

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kF = 16;     /* features per pair */
constexpr int kH = 128;    /* hidden width */
constexpr int kT = kH / 16;   /* 16-wide tiles per hidden layer */
constexpr int kS = kH / 32;   /* 32-deep k-steps over a hidden layer */

/* bfloat16 bits of x, rounded to nearest even (v_cvt_pk_bf16_f32). */
__device__ __forceinline__ short bf16_of(float x) { return __builtin_bit_cast(short, (__bf16)x); }

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4f mfma32(v8s a, v8s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0,
                                                    0);
}

/* w_F(a, b) through a row hash table (descriptor (offset << 6) | bits). */
__device__ __forceinline__ int32_t row_w(const uint64_t* __restrict__ desc, const int2* __restrict__ tab, int32_t row,
                                         int32_t key) {
    if (row == 0 || key == 0) return BM_ABSENT;
    const uint64_t d = desc[row];
    const uint32_t bits = (uint32_t)(d & 63u);
    if (!bits) return BM_ABSENT;
    const int2* T = tab + (d >> 6);
    const uint32_t mask = (1u << bits) - 1u;
    uint32_t s = slot_hash(key, bits);
    for (;;) {
        const int2 e = T[s];
        if (e.x == 0) return BM_ABSENT;
        if (e.x == key) return e.y;
        s = (s + 1u) & mask;
    }
}

/* Per-vertex features in two passes.  k_feature_chunks: one wave per chunk
 * {x, direction, first, count} of at most kFeatChunk edges of one of x's male
 * lists (the chunks depend only on the male graph), so a hub's lists spread
 * over many waves; each edge's contribution under pi (a probe of the female
 * row of pi(x)) is summed with its weight and whether it is 0, and the wave
 * adds the sums to x's accumulators {L, weight sum, zero count} (a self-loop
 * counts once, with the out-edges).  k_feature_finish: one thread per vertex
 * turns the exact integer sums into the 8 floats and clears the
 * accumulators. */
__global__ void __launch_bounds__(256) k_feature_chunks(BmDevGraph g, const int4* __restrict__ chunks, int64_t nchunks,
                                                        longlong4* __restrict__ acc) {
    const int lane = threadIdx.x & (BM_WAVE - 1);
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / BM_WAVE);
    for (int64_t ci = (int64_t)blockIdx.x * (blockDim.x / BM_WAVE) + threadIdx.x / BM_WAVE; ci < nchunks; ci += nw) {
        const int4 ch = chunks[ci];
        const int32_t x = uni(ch.x);
        const int out = uni(ch.y);
        const int32_t fx = uni(g.pi[x]);
        const int64_t base = (out ? g.m_out_ptr[x] : g.m_in_ptr[x]) + uni(ch.z);
        const int2* adj = out ? g.m_out : g.m_in;
        const uint64_t* desc = out ? g.f_out_desc : g.f_in_desc;
        const int2* tab = out ? g.f_out_tab : g.f_in_tab;
        int64_t L = 0, ws = 0, miss = 0;
        for (int32_t i = lane; i < ch.w; i += BM_WAVE) {
            const int2 ed = adj[base + i];
            if (!out && ed.x == x) continue;   /* a self-loop counted once, with the out-edges */
            const int32_t fy = ed.x == x ? fx : g.pi[ed.x];
            const int32_t c = contrib(ed.y, row_w(desc, tab, fx, fy));
            L += c;
            ws += ed.y;
            miss += c == 0;
        }
        L = wave_sum64(L);
        ws = wave_sum64(ws);
        miss = wave_sum64(miss);
        if (lane == 0) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&acc[x].x), (unsigned long long)L);
            atomicAdd(reinterpret_cast<unsigned long long*>(&acc[x].y), (unsigned long long)ws);
            atomicAdd(reinterpret_cast<unsigned long long*>(&acc[x].z), (unsigned long long)miss);
        }
    }
}

__global__ void k_feature_finish(BmDevGraph g, longlong4* __restrict__ acc, float4* __restrict__ vf) {
    for (int32_t x = blockIdx.x * blockDim.x + threadIdx.x; x <= g.n; x += gridDim.x * blockDim.x) {
        if (x == 0) {
            vf[0] = make_float4(0.f, 0.f, 0.f, 0.f);
            vf[1] = make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        const longlong4 a = acc[x];
        acc[x] = make_longlong4(0, 0, 0, 0);
        const int64_t L = a.x, wsum = a.y, miss = a.z;
        const int32_t fx = g.pi[x];
        const float dout = (float)(g.m_out_ptr[x + 1] - g.m_out_ptr[x]), din = (float)(g.m_in_ptr[x + 1] - g.m_in_ptr[x]);
        const float fo = fx ? (float)(g.f_out_ptr[fx + 1] - g.f_out_ptr[fx]) : 0.f;
        const float fi = fx ? (float)(g.f_in_ptr[fx + 1] - g.f_in_ptr[fx]) : 0.f;
        const float ws = (float)(wsum > 0 ? wsum : 0);
        const float Lf = (float)(L > 0 ? L : 0);
        vf[2 * (int64_t)x] = make_float4(log1pf(dout), log1pf(din), log1pf(ws), Lf / (ws + 1.f));
        vf[2 * (int64_t)x + 1] = make_float4((float)miss / (dout + din + 1.f), log1pf(fo), log1pf(fi), log1pf(Lf));
    }
}

/* Weights in fragment order with k-steps of KB rows (16 or 32):
 * frag[((s * (NN / 16) + t) * 64 + l) * (KB / 4) + i] = M[KB s + (KB / 4) (l / 16) + i][16 t + l % 16]
 * for the KK x NN matrix M = W (K x N, row-major) or, with transpose, W^T. */
__global__ void k_mlp_pack(const float* __restrict__ W, int K, int N, int transpose, int KB, short* __restrict__ frag) {
    const int KK = transpose ? N : K, NN = transpose ? K : N;   /* shape of the packed matrix */
    const int per = KB / 4;
    const int total = (KK / KB) * (NN / 16) * 64;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
        const int l = q & 63, f = q >> 6;
        const int s = f / (NN / 16), t = f % (NN / 16);
        for (int i = 0; i < per; i++) {
            const int r = KB * s + per * (l / 16) + i, c = 16 * t + (l % 16);
            frag[(size_t)q * per + i] = bf16_of(transpose ? W[c * N + r] : W[r * N + c]);
        }
    }
}

struct MlpDev {
    const v4s* W1f;     /* 1 x 8 fragments (k-step 16) of W1 (16 x 128) */
    const v8s* W2f;     /* 4 x 8 fragments (k-step 32) of W2 (128 x 128) */
    const v8s* W2Tf;    /* 4 x 8 fragments (k-step 32) of W2^T */
    const float* b1;
    const float* b2;
    const float* w3;
    const float* b3;
};

/* Candidate pair p of a proposal source. */
struct PairSrc {
    const int2* list;   /* explicit pairs, or NULL for the uniform stream */
    uint64_t seed;
    int32_t n;
};

__device__ __forceinline__ int2 pair_at(const PairSrc& ps, int64_t p) {
    if (ps.list) return ps.list[p];
    int32_t u, v;
    bm_uniform_pair(ps.seed, (uint64_t)p, ps.n, &u, &v);
    return make_int2(u, v);
}

constexpr int kFwdThreads = 256;
constexpr int kFwdWaves = kFwdThreads / BM_WAVE;
constexpr int kHPad = kH + 8;   /* padded bf16 row of a staged 16 x 128 tile (272 bytes: 16-byte reads hit distinct banks) */

struct FwdShared {
    v4s W1[kT * 64];
    v8s W2[kS * kT * 64];
    float b1[kH], b2[kH], w3[kH];
    short Ht[kFwdWaves][16][kHPad];
};

__device__ __forceinline__ void stage_weights(FwdShared& S, const MlpDev& M) {
    for (int q = threadIdx.x; q < kT * 64; q += blockDim.x) S.W1[q] = M.W1f[q];
    for (int q = threadIdx.x; q < kS * kT * 64; q += blockDim.x) S.W2[q] = M.W2f[q];
    for (int q = threadIdx.x; q < kH; q += blockDim.x) {
        S.b1[q] = M.b1[q];
        S.b2[q] = M.b2[q];
        S.w3[q] = M.w3[q];
    }
}

/* Input fragment of a 16-pair tile: lane l holds features 4 (l / 16) .. +3
 * of pair l % 16 (u's features 0..7, v's 8..15) -- the A operand of x W1
 * and, equally, the B operand of W1^T x^T. */
__device__ __forceinline__ v4s input_frag(const float4* __restrict__ vf, int2 uv, bool valid, int lane) {
    const int q = lane / 16;
    float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid) f = vf[2 * (int64_t)(q < 2 ? uv.x : uv.y) + (q & 1)];
    v4s a;
    a[0] = bf16_of(f.x);
    a[1] = bf16_of(f.y);
    a[2] = bf16_of(f.z);
    a[3] = bf16_of(f.w);
    return a;
}

/* Layer 1 for one tile in the pair-major form (training): z1 (D layout:
 * lane l has pairs 4 (l / 16) + j of hidden unit 16 t + l % 16, kept for the
 * backward mask) and h1 into the wave's LDS tile. */
__device__ __forceinline__ void layer1(const FwdShared& S, short (*Ht)[kHPad], v4s xa, v4f* z1, int lane) {
#pragma unroll
    for (int t = 0; t < kT; t++) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(xa, S.W1[t * 64 + lane], acc);
        const int col = 16 * t + (lane % 16);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            acc[j] += S.b1[col];
            Ht[4 * (lane / 16) + j][col] = bf16_of(acc[j] > 0.f ? acc[j] : 0.f);
        }
        z1[t] = acc;
    }
}

/* 32-deep fragments (kS k-steps) of the wave's staged 16 x 128 bf16 tile:
 * lane l gets row l % 16, columns 32 s + 8 (l / 16) .. +7 (one 16-byte read). */
__device__ __forceinline__ void load_tile32(const short (*Ht)[kHPad], v8s* a, int lane) {
#pragma unroll
    for (int s = 0; s < kS; s++) a[s] = *reinterpret_cast<const v8s*>(&Ht[lane % 16][32 * s + 8 * (lane / 16)]);
}

/* Forward: scores of k pairs (one wave per 16-pair tile), in the transposed
 * form: z1^T = W1^T x^T (A = W1^T fragments, B = the input fragment) leaves
 * lane l with hidden units 16 t + 4 (l / 16) .. +3 of pair l % 16, stored to
 * the pair's LDS row in one 8-byte write per t; z2^T = W2^T h1^T takes h1^T
 * fragments as 16-byte reads of those rows; the output layer sums the lane's
 * 32 hidden units, then over the four lanes of the pair. */
__global__ void __launch_bounds__(kFwdThreads) k_mlp_forward(const float4* __restrict__ vf, PairSrc ps, int64_t k,
                                                             MlpDev M, float* __restrict__ score) {
    extern __shared__ uint4 smem[];
    FwdShared& S = *reinterpret_cast<FwdShared*>(smem);
    stage_weights(S, M);
    __syncthreads();
    const int lane = threadIdx.x & (BM_WAVE - 1), wave = threadIdx.x / BM_WAVE;
    const int q = lane / 16;
    short(*Ht)[kHPad] = S.Ht[wave];
    const float b3 = M.b3[0];
    const int64_t tiles = (k + 15) / 16;
    for (int64_t tile = (int64_t)blockIdx.x * kFwdWaves + wave; tile < tiles; tile += (int64_t)gridDim.x * kFwdWaves) {
        const int64_t p = tile * 16 + (lane % 16);
        const bool valid = p < k;
        const int2 uv = valid ? pair_at(ps, p) : make_int2(0, 0);
        const v4s xb = input_frag(vf, uv, valid, lane);
#pragma unroll
        for (int t = 0; t < kT; t++) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(S.W1[t * 64 + lane], xb, acc);
            v4s h;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float z = acc[j] + S.b1[16 * t + 4 * q + j];
                h[j] = bf16_of(z > 0.f ? z : 0.f);
            }
            *reinterpret_cast<v4s*>(&Ht[lane % 16][16 * t + 4 * q]) = h;
        }
        wave_sync();
        v8s hb[kS];
        load_tile32(Ht, hb, lane);
        float part = 0.f;
#pragma unroll
        for (int t = 0; t < kT; t++) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s2 = 0; s2 < kS; s2++) acc = mfma32(S.W2[(s2 * kT + t) * 64 + lane], hb[s2], acc);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int u = 16 * t + 4 * q + j;
                const float h2 = acc[j] + S.b2[u];
                part += (h2 > 0.f ? h2 : 0.f) * S.w3[u];
            }
        }
        part += __shfl_xor(part, 16, BM_WAVE);
        part += __shfl_xor(part, 32, BM_WAVE);
        if (lane < 16 && valid) score[p] = part + b3;
        wave_sync();
    }
}

/* Training pass over a batch of B pairs (B a multiple of 32): forward, loss,
 * and the backward through the activations.  Stores bf(x)^T (16 x B), h1^T,
 * g1^T and g2^T (128 x B) for the weight-gradient products and accumulates
 * the bias gradients, dw3 and the summed squared error with atomics. */
struct TrainBuf {
    short* XT;      /* 16 x B */
    short* H1T;     /* 128 x B */
    short* G1T;     /* 128 x B */
    short* G2T;     /* 128 x B */
    float* grad;    /* gradient in the parameter layout */
    float* loss;    /* summed squared error */
    int64_t B;
};

constexpr int kOffW1 = 0, kOffB1 = kF * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH,
              kOffB3 = kOffW3 + kH, kParams = kOffB3 + 1;

struct TrainShared {
    FwdShared f;
    v8s W2T[kS * kT * 64];
};

__global__ void __launch_bounds__(kFwdThreads) k_mlp_train_act(const float4* __restrict__ vf, PairSrc ps, int64_t k,
                                                               const float* __restrict__ target, MlpDev M, TrainBuf T) {
    extern __shared__ uint4 smem[];
    TrainShared& S = *reinterpret_cast<TrainShared*>(smem);
    stage_weights(S.f, M);
    for (int q = threadIdx.x; q < kS * kT * 64; q += blockDim.x) S.W2T[q] = M.W2Tf[q];
    __syncthreads();
    const int lane = threadIdx.x & (BM_WAVE - 1), wave = threadIdx.x / BM_WAVE;
    short(*Ht)[kHPad] = S.f.Ht[wave];
    const float b3 = M.b3[0];
    const float inv = 2.f / (float)k;
    const int64_t B = T.B;
    const int64_t tiles = B / 16;
    for (int64_t tile = (int64_t)blockIdx.x * kFwdWaves + wave; tile < tiles; tile += (int64_t)gridDim.x * kFwdWaves) {
        const int64_t base = tile * 16;
        const int64_t p = base + (lane % 16);
        const bool valid = p < k;
        const int2 uv = valid ? pair_at(ps, p) : make_int2(0, 0);
        const v4s xa = input_frag(vf, uv, valid, lane);
#pragma unroll
        for (int i = 0; i < 4; i++) T.XT[(int64_t)(4 * (lane / 16) + i) * B + p] = xa[i];
        v4f z1[kT];
        layer1(S.f, Ht, xa, z1, lane);
        wave_sync();
        v8s a[kS];
        load_tile32(Ht, a, lane);
        /* h1^T: rows 4 (l/16) + j of column c are 4 consecutive pairs */
#pragma unroll
        for (int t = 0; t < kT; t++) {
            const int col = 16 * t + (lane % 16);
            v4s h;
#pragma unroll
            for (int j = 0; j < 4; j++) h[j] = Ht[4 * (lane / 16) + j][col];
            *reinterpret_cast<v4s*>(&T.H1T[(int64_t)col * B + base + 4 * (lane / 16)]) = h;
        }
        v4f z2[kT];
        float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < kT; t++) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < kS; s++) acc = mfma32(a[s], S.f.W2[(s * kT + t) * 64 + lane], acc);
            const int col = 16 * t + (lane % 16);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                acc[j] += S.f.b2[col];
                part[j] += (acc[j] > 0.f ? acc[j] : 0.f) * S.f.w3[col];
            }
            z2[t] = acc;
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) part[j] += __shfl_xor(part[j], o, BM_WAVE);
        /* dy for this lane's four rows (every lane of a 16-lane group has the sums) */
        float dy[4], se = 0.f;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t q = base + 4 * (lane / 16) + j;
            const bool ok = q < k;
            const float y = part[j] + b3;
            const float r = ok ? y - target[q] : 0.f;
            dy[j] = r * inv;
            se += r * r;
        }
        if ((lane % 16) == 0) atomicAdd(T.loss, se);
        /* g2 = bf(dy w3 [z2 > 0]); dw3 += h2 dy; db2 += dy w3 [z2 > 0] */
        wave_sync();
#pragma unroll
        for (int t = 0; t < kT; t++) {
            const int col = 16 * t + (lane % 16);
            float gsum = 0.f, wsum = 0.f;
            v4s gv;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float h = z2[t][j] > 0.f ? z2[t][j] : 0.f;
                const float g = z2[t][j] > 0.f ? dy[j] * S.f.w3[col] : 0.f;
                gsum += g;
                wsum += h * dy[j];
                gv[j] = bf16_of(g);
                Ht[4 * (lane / 16) + j][col] = gv[j];
            }
            *reinterpret_cast<v4s*>(&T.G2T[(int64_t)col * B + base + 4 * (lane / 16)]) = gv;
            gsum += __shfl_xor(gsum, 16, BM_WAVE);
            gsum += __shfl_xor(gsum, 32, BM_WAVE);
            wsum += __shfl_xor(wsum, 16, BM_WAVE);
            wsum += __shfl_xor(wsum, 32, BM_WAVE);
            if (lane < 16) {
                atomicAdd(&T.grad[kOffB2 + col], gsum);
                atomicAdd(&T.grad[kOffW3 + col], wsum);
            }
        }
        {
            float d = dy[0] + dy[1] + dy[2] + dy[3];
            d += __shfl_xor(d, 16, BM_WAVE);
            d += __shfl_xor(d, 32, BM_WAVE);
            if (lane == 0) atomicAdd(&T.grad[kOffB3], d);
        }
        wave_sync();
        /* g1 = bf((g2 @ W2^T) [z1 > 0]) */
        v8s ga[kS];
        load_tile32(Ht, ga, lane);
#pragma unroll
        for (int t = 0; t < kT; t++) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < kS; s++) acc = mfma32(ga[s], S.W2T[(s * kT + t) * 64 + lane], acc);
            const int col = 16 * t + (lane % 16);
            float gsum = 0.f;
            v4s gv;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float g = z1[t][j] > 0.f ? acc[j] : 0.f;
                gsum += g;
                gv[j] = bf16_of(g);
            }
            *reinterpret_cast<v4s*>(&T.G1T[(int64_t)col * B + base + 4 * (lane / 16)]) = gv;
            gsum += __shfl_xor(gsum, 16, BM_WAVE);
            gsum += __shfl_xor(gsum, 32, BM_WAVE);
            if (lane < 16) atomicAdd(&T.grad[kOffB1 + col], gsum);
        }
        wave_sync();
    }
}

/* Weight gradients: dW1 = bf(x)^T g1 (16 x 128), dW2 = h1^T g2 (128 x 128),
 * 32 batch entries per MFMA.  Work item = (output tile, slice of the batch);
 * one wave each, summed into the gradient with atomics. */
constexpr int kWgTiles = 1 * kT + kT * kT;

__global__ void __launch_bounds__(256) k_mlp_wgrad(TrainBuf T, int32_t slices) {
    const int lane = threadIdx.x & (BM_WAVE - 1);
    const int64_t items = (int64_t)kWgTiles * slices;
    const int64_t B = T.B;
    const int64_t steps = B / 32;
    for (int64_t it = (int64_t)blockIdx.x * (blockDim.x / BM_WAVE) + threadIdx.x / BM_WAVE; it < items;
         it += (int64_t)gridDim.x * (blockDim.x / BM_WAVE)) {
        const int tile = (int)(it % kWgTiles), slice = (int)(it / kWgTiles);
        const bool first = tile < kT;                     /* dW1 tiles, else dW2 */
        const int mt = first ? 0 : (tile - kT) / kT, nt = first ? tile : (tile - kT) % kT;
        const short* AT = first ? T.XT : T.H1T;           /* rows = input features / neurons */
        const short* BT = first ? T.G1T : T.G2T;          /* rows = output neurons */
        const int64_t s0 = steps * slice / slices, s1 = steps * (slice + 1) / slices;
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        const int64_t arow = (int64_t)(16 * mt + lane % 16) * B, brow = (int64_t)(16 * nt + lane % 16) * B;
        for (int64_t s = s0; s < s1; s++) {
            const int64_t kk = 32 * s + 8 * (lane / 16);
            const v8s a = *reinterpret_cast<const v8s*>(&AT[arow + kk]);
            const v8s b = *reinterpret_cast<const v8s*>(&BT[brow + kk]);
            acc = mfma32(a, b, acc);
        }
        const int off = first ? kOffW1 : kOffW2;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = 16 * mt + 4 * (lane / 16) + j, c = 16 * nt + lane % 16;
            atomicAdd(&T.grad[off + r * kH + c], acc[j]);
        }
    }
}

/* Adam on the fp32 master parameters; clears the gradient. */
__global__ void k_mlp_adam(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                           int32_t count, float lr, float b1c, float b2c) {
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = beta1 * m[i] + (1.f - beta1) * gi;
        const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / b1c) / (sqrtf(vi / b2c) + eps);
        g[i] = 0.f;
    }
}

/* Regression target t = sign(d) log(1 + |d|) of exact int64 deltas. */
__global__ void k_mlp_targets(const int64_t* __restrict__ d, int64_t k, float* __restrict__ t) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (int64_t)gridDim.x * blockDim.x) {
        const double x = (double)d[i];
        t[i] = (float)(x < 0 ? -log1p(-x) : log1p(x));
    }
}

__global__ void k_iota(int32_t* __restrict__ p, int64_t k) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (int32_t)i;
}

/* Pairs of the top-ranked stream indices into a pair list. */
__global__ void k_gather_stream_pairs(uint64_t seed, int32_t n, const int32_t* __restrict__ idx, int64_t top,
                                      int2* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < top; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t u, v;
        bm_uniform_pair(seed, (uint64_t)idx[i], n, &u, &v);
        out[i] = make_int2(u, v);
    }
}

/* ---- top of the ranked pool: radix select over distinct 64-bit keys ---- */

/* key(p) = order-preserving bits of score p (the float order, -0 below +0)
 * in the high half, 2^32 - 1 - p in the low half: a larger key is a larger
 * score, ties to the smaller pool index, and no two keys are equal -- the
 * order of a stable descending sort of the scores. */
__device__ __forceinline__ uint64_t rank_key(float x, int64_t p) {
    uint32_t u = __float_as_uint(x);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ((uint64_t)u << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)p);
}

constexpr int kSelBits = 11, kSelBins = 1 << kSelBits, kSelThreads = 256, kSelPasses = 6;
constexpr int kSelPer = 16;          /* consecutive keys per lane, loaded before any is used */
constexpr int kSelCopies = 8;        /* copies of the global histogram (block b adds to copy b % 8) */
constexpr int kSortMax = 8192;       /* largest top sorted in one block */

struct SelState {
    uint64_t prefix;      /* key bits fixed so far (the bits of the current digit and below are 0) */
    uint64_t fixed;       /* mask of the fixed bits */
    int64_t remaining;    /* rank (1-based) of the wanted key within the keys sharing the prefix */
    uint32_t done;        /* blocks of the current pass that have merged their histogram */
    uint32_t count;       /* keys written by the compaction */
    uint32_t settled;     /* 1: the keys >= prefix under the mask are exactly the top */
    uint32_t pad;
    uint32_t hist[kSelCopies][kSelBins];
};

__device__ __forceinline__ void sel_digit(int pass, int* shift, int* width) {
    *width = pass < kSelPasses - 1 ? kSelBits : 64 - kSelBits * (kSelPasses - 1);
    *shift = 64 - kSelBits * (pass + 1);
    if (*shift < 0) *shift = 0;
}

__global__ void k_select_init(SelState* st, int64_t top) {
    if (threadIdx.x == 0) {
        st->prefix = 0;
        st->fixed = 0;
        st->remaining = top;
        st->done = 0;
        st->count = 0;
        st->settled = 0;
    }
    for (int i = threadIdx.x; i < kSelCopies * kSelBins; i += blockDim.x) (&st->hist[0][0])[i] = 0;
}

/* Adds 1 to bin b of the LDS histogram for every lane with `valid`: lanes
 * sharing the first valid lane's bin (the common case: scores crowd into a
 * few bins) add together in one atomic. */
__device__ __forceinline__ void wave_hist_add(uint32_t* h, bool valid, uint32_t b, int lane) {
    const uint64_t act = __ballot(valid);
    if (!act) return;
    const int lead = __ffsll((unsigned long long)act) - 1;
    const uint32_t lb = (uint32_t)__shfl((int)b, lead, BM_WAVE);
    const uint64_t same = __ballot(valid && b == lb);
    if (lane == lead) atomicAdd(&h[lb], (uint32_t)__popcll(same));
    if (valid && b != lb) atomicAdd(&h[b], 1u);
}

/* One digit of the select: every block histograms the digit of the keys that
 * share the prefix (LDS counters, merged into one of kSelCopies global
 * copies with atomics); the last block to finish sums the copies, finds the
 * digit holding the wanted rank (suffix sums from the top bin), extends the
 * prefix and clears the histograms.  When the wanted rank is the last key of
 * its digit's bin, every key of that bin and above belongs to the top: the
 * select is settled and the remaining passes do nothing. */
__global__ void __launch_bounds__(kSelThreads) k_select_pass(const float* __restrict__ sc, int64_t P,
                                                             SelState* st, int pass) {
    __shared__ uint32_t h[kSelBins];
    __shared__ uint32_t part[kSelThreads];
    __shared__ int last;
    if (st->settled) return;
    int shift, width;
    sel_digit(pass, &shift, &width);
    const int bins = 1 << width;
    const int top = shift + width;
    const uint64_t fixed = top >= 64 ? 0ull : (~0ull << top);
    const uint64_t prefix = st->prefix;
    const int lane = threadIdx.x & (BM_WAVE - 1);
    for (int i = threadIdx.x; i < bins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    /* a wave takes 64 kSelPer consecutive keys, kSelPer per lane, loaded
     * before any is used; every lane of a wave runs the same iterations (the
     * ballots need them all) */
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / BM_WAVE) * BM_WAVE * kSelPer;
    for (int64_t w0 = ((int64_t)blockIdx.x * (blockDim.x / BM_WAVE) + threadIdx.x / BM_WAVE) * BM_WAVE * kSelPer;
         w0 < P; w0 += wstride) {
        const int64_t p0 = w0 + (int64_t)lane * kSelPer;
        float v[kSelPer];
        if (p0 + kSelPer <= P) {
#pragma unroll
            for (int j = 0; j < kSelPer; j += 4) {
                const float4 f = *reinterpret_cast<const float4*>(&sc[p0 + j]);
                v[j] = f.x;
                v[j + 1] = f.y;
                v[j + 2] = f.z;
                v[j + 3] = f.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kSelPer; j++) v[j] = p0 + j < P ? sc[p0 + j] : 0.f;
        }
#pragma unroll
        for (int j = 0; j < kSelPer; j++) {
            const uint64_t k = rank_key(v[j], p0 + j);
            const bool valid = p0 + j < P && (k & fixed) == prefix;
            wave_hist_add(h, valid, (uint32_t)(k >> shift) & (uint32_t)(bins - 1), lane);
        }
    }
    __syncthreads();
    uint32_t* gh = st->hist[blockIdx.x % kSelCopies];
    for (int i = threadIdx.x; i < bins; i += blockDim.x)
        if (h[i]) atomicAdd(&gh[i], h[i]);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(&st->done, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    /* every thread reads the rank before the barriers below; one thread rewrites it */
    const int64_t rem = st->remaining;
    /* thread t owns bins [t * per, (t + 1) * per) (per = 8 or 2); the
     * device-scope fence above makes the other blocks' counts visible to
     * plain loads, all issued before any is summed */
    const int per = bins / kSelThreads;
    const int t = threadIdx.x;
    uint32_t mine = 0;
    if (per == 8) {
        uint4 a[kSelCopies][2];
#pragma unroll
        for (int cp = 0; cp < kSelCopies; cp++) {
            a[cp][0] = *reinterpret_cast<const uint4*>(&st->hist[cp][t * 8]);
            a[cp][1] = *reinterpret_cast<const uint4*>(&st->hist[cp][t * 8 + 4]);
        }
        uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int cp = 0; cp < kSelCopies; cp++) {
            c[0] += a[cp][0].x;
            c[1] += a[cp][0].y;
            c[2] += a[cp][0].z;
            c[3] += a[cp][0].w;
            c[4] += a[cp][1].x;
            c[5] += a[cp][1].y;
            c[6] += a[cp][1].z;
            c[7] += a[cp][1].w;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            h[t * 8 + j] = c[j];
            mine += c[j];
        }
    } else {
        for (int j = 0; j < per; j++) {
            uint32_t c = 0;
            for (int cp = 0; cp < kSelCopies; cp++) c += st->hist[cp][t * per + j];
            h[t * per + j] = c;
            mine += c;
        }
    }
    part[t] = mine;
    __syncthreads();
    for (int o = 1; o < kSelThreads; o <<= 1) {
        const uint32_t add = t + o < kSelThreads ? part[t + o] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    /* part[t] = keys in bins >= t * per; the digit lies with the thread where
     * part[t + 1] < remaining <= part[t] */
    const uint32_t above = t + 1 < kSelThreads ? part[t + 1] : 0u;
    if ((int64_t)above < rem && rem <= (int64_t)part[t]) {
        int64_t r = rem - above;
        int d = t * per + per - 1;
        for (; d > t * per; d--) {
            if (r <= (int64_t)h[d]) break;
            r -= h[d];
        }
        st->prefix = prefix | ((uint64_t)d << shift);
        st->fixed = top - width >= 64 ? ~0ull : (~0ull << (top - width));
        st->remaining = r;
        st->settled = r == (int64_t)h[d];
    }
    if (t == 0) st->done = 0;
    for (int cp = 0; cp < kSelCopies; cp++)
        for (int i = 4 * t; i < bins; i += 4 * blockDim.x) *reinterpret_cast<uint4*>(&st->hist[cp][i]) = make_uint4(0, 0, 0, 0);
}

/* Keys at or above the selected prefix under its mask (exactly `top` of
 * them; the bound only guards the buffer).  Each wave ballots its lanes'
 * keys, the block scans its waves' counts in LDS, and one global atomic per
 * block reserves the block's slots. */
__global__ void __launch_bounds__(kSelThreads) k_select_compact(const float* __restrict__ sc, int64_t P, SelState* st,
                                                                uint64_t* __restrict__ out, uint32_t top) {
    __shared__ uint32_t wcount[kSelThreads / BM_WAVE];
    __shared__ uint32_t base;
    const uint64_t kstar = st->prefix, fixed = st->fixed;
    const int lane = threadIdx.x & (BM_WAVE - 1), wave = threadIdx.x / BM_WAVE;
    const int64_t bstride = (int64_t)gridDim.x * blockDim.x * kSelPer;
    for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x * kSelPer; b0 < P; b0 += bstride) {
        const int64_t p0 = b0 + ((int64_t)wave * BM_WAVE + lane) * kSelPer;
        float v[kSelPer];
        if (p0 + kSelPer <= P) {
#pragma unroll
            for (int j = 0; j < kSelPer; j += 4) {
                const float4 f = *reinterpret_cast<const float4*>(&sc[p0 + j]);
                v[j] = f.x;
                v[j + 1] = f.y;
                v[j + 2] = f.z;
                v[j + 3] = f.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kSelPer; j++) v[j] = p0 + j < P ? sc[p0 + j] : 0.f;
        }
        uint32_t inmask = 0;   /* bit j: this lane's key j is in */
#pragma unroll
        for (int j = 0; j < kSelPer; j++) {
            const uint64_t k = rank_key(v[j], p0 + j);
            if (p0 + j < P && (k & fixed) >= kstar) inmask |= 1u << j;
        }
        uint32_t mine = (uint32_t)__popc(inmask);
        /* exclusive prefix of the lanes' counts within the wave */
        uint32_t incl = mine;
#pragma unroll
        for (int o = 1; o < BM_WAVE; o <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, o, BM_WAVE);
            if (lane >= o) incl += y;
        }
        if (lane == BM_WAVE - 1) wcount[wave] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (int w = 0; w < kSelThreads / BM_WAVE; w++) {
                const uint32_t c = wcount[w];
                wcount[w] = tot;
                tot += c;
            }
            base = tot ? atomicAdd(&st->count, tot) : 0u;
        }
        __syncthreads();
        uint32_t at = base + wcount[wave] + incl - mine;
#pragma unroll
        for (int j = 0; j < kSelPer; j++)
            if (inmask & (1u << j)) {
                if (at < top) out[at] = rank_key(v[j], p0 + j);
                at++;
            }
        __syncthreads();
    }
}

/* One block: the `top` selected keys in descending order (bitonic sort in
 * LDS, padded with zero keys to a power of two), then their stream pairs. */
__global__ void __launch_bounds__(1024) k_select_sort(const uint64_t* __restrict__ keys, int32_t top, int32_t padded,
                                                      uint64_t seed, int32_t n, int2* __restrict__ out) {
    extern __shared__ uint64_t sk[];
    for (int i = threadIdx.x; i < padded; i += blockDim.x) sk[i] = i < top ? keys[i] : 0ull;
    __syncthreads();
    for (int size = 2; size <= padded; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < padded / 2; i += blockDim.x) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool desc = (lo & size) == 0;
                const uint64_t a = sk[lo], b = sk[hi];
                if ((a < b) == desc) {
                    sk[lo] = b;
                    sk[hi] = a;
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < top; i += blockDim.x) {
        const uint32_t p = 0xFFFFFFFFu - (uint32_t)sk[i];
        int32_t u, v;
        bm_uniform_pair(seed, (uint64_t)p, n, &u, &v);
        out[i] = make_int2(u, v);
    }
}

int grid_of(int64_t work, int threads, int64_t cap = 65536) {
    int64_t b = (work + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

MlpDev mlpdev(const BmMlpDev& d) {
    MlpDev M;
    M.W1f = reinterpret_cast<const v4s*>(d.W1f);
    M.W2f = reinterpret_cast<const v8s*>(d.W2f);
    M.W2Tf = reinterpret_cast<const v8s*>(d.W2Tf);
    M.b1 = d.params + kOffB1;
    M.b2 = d.params + kOffB2;
    M.w3 = d.params + kOffW3;
    M.b3 = d.params + kOffB3;
    return M;
}

}  // namespace

int32_t bm_mlp_param_count(void) { return kParams; }

int32_t bm_feature_chunk_edges(void) { return 256; }

hipError_t bm_launch_vertex_features(const BmDevGraph& g, const int4* chunks, int64_t nchunks, void* acc, float* vf,
                                     int32_t num_cus, hipStream_t s) {
    const int64_t cap = (int64_t)(num_cus > 0 ? num_cus : 256) * 16;
    int64_t blocks = (nchunks + 3) / 4;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_feature_chunks, dim3((unsigned)blocks), dim3(256), 0, s, g, chunks, nchunks,
                       static_cast<longlong4*>(acc));
    hipLaunchKernelGGL(k_feature_finish, dim3(grid_of((int64_t)g.n + 1, 256)), dim3(256), 0, s, g,
                       static_cast<longlong4*>(acc), reinterpret_cast<float4*>(vf));
    return hipGetLastError();
}

hipError_t bm_launch_mlp_pack(const BmMlpDev& d, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(2), dim3(256), 0, s, d.params + kOffW1, kF, kH, 0, 16,
                       static_cast<short*>(d.W1f));
    hipLaunchKernelGGL(k_mlp_pack, dim3(8), dim3(256), 0, s, d.params + kOffW2, kH, kH, 0, 32,
                       static_cast<short*>(d.W2f));
    hipLaunchKernelGGL(k_mlp_pack, dim3(8), dim3(256), 0, s, d.params + kOffW2, kH, kH, 1, 32,
                       static_cast<short*>(d.W2Tf));
    return hipGetLastError();
}

hipError_t bm_launch_mlp_forward(const BmMlpDev& d, const float* vf, const int2* list, uint64_t seed, int32_t n,
                                 int64_t k, float* score, int32_t num_cus, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    PairSrc ps{list, seed, n};
    const int64_t tiles = (k + 15) / 16;
    int64_t blocks = (tiles + kFwdWaves - 1) / kFwdWaves;
    const int64_t cap = (int64_t)(num_cus > 0 ? num_cus : 256) * 4;
    if (blocks > cap) blocks = cap;
    hipError_t e = hipFuncSetAttribute((const void*)k_mlp_forward, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)sizeof(FwdShared));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)blocks), dim3(kFwdThreads), sizeof(FwdShared), s,
                       reinterpret_cast<const float4*>(vf), ps, k, mlpdev(d), score);
    return hipGetLastError();
}

size_t bm_mlp_train_bytes(int64_t k) { return (size_t)((k + 31) / 32 * 32) * (kF + 3 * kH) * sizeof(short); }

hipError_t bm_launch_mlp_gradient(const BmMlpDev& d, const float* vf, const int2* list, uint64_t seed, int32_t n,
                                  int64_t k, const float* target, void* act, float* loss, int32_t num_cus,
                                  hipStream_t s) {
    if (k <= 0) return hipSuccess;
    TrainBuf T;
    T.B = (k + 31) / 32 * 32;
    T.XT = reinterpret_cast<short*>(act);
    T.H1T = T.XT + (size_t)kF * T.B;
    T.G1T = T.H1T + (size_t)kH * T.B;
    T.G2T = T.G1T + (size_t)kH * T.B;
    T.grad = d.grad;
    T.loss = loss;
    PairSrc ps{list, seed, n};
    const int64_t tiles = T.B / 16;
    int64_t blocks = (tiles + kFwdWaves - 1) / kFwdWaves;
    const int64_t cap = (int64_t)(num_cus > 0 ? num_cus : 256) * 2;
    if (blocks > cap) blocks = cap;
    hipError_t e = hipFuncSetAttribute((const void*)k_mlp_train_act, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)sizeof(TrainShared));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_train_act, dim3((unsigned)blocks), dim3(kFwdThreads), sizeof(TrainShared), s,
                       reinterpret_cast<const float4*>(vf), ps, k, target, mlpdev(d), T);
    int32_t slices = (int32_t)(tiles / 128);
    if (slices < 1) slices = 1;
    if (slices > 64) slices = 64;
    const int64_t items = (int64_t)kWgTiles * slices;
    hipLaunchKernelGGL(k_mlp_wgrad, dim3((unsigned)grid_of(items, 4)), dim3(256), 0, s, T, slices);
    return hipGetLastError();
}

hipError_t bm_launch_mlp_adam(const BmMlpDev& d, float lr, int64_t step, hipStream_t s) {
    const float b1c = 1.f - powf(0.9f, (float)step), b2c = 1.f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3(grid_of(kParams, 256)), dim3(256), 0, s, d.params, d.grad, d.adam_m,
                       d.adam_v, kParams, lr, b1c, b2c);
    return bm_launch_mlp_pack(d, s);
}

hipError_t bm_launch_mlp_targets(const int64_t* delta, int64_t k, float* t, hipStream_t s) {
    if (k <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_targets, dim3(grid_of(k, 256)), dim3(256), 0, s, delta, k, t);
    return hipGetLastError();
}

size_t bm_mlp_select_bytes(int64_t pool) {
    return (size_t)pool * sizeof(float) + 512 + sizeof(SelState) + kSortMax * sizeof(uint64_t);
}

int64_t bm_mlp_select_max_top(void) { return kSortMax; }

/* Scores of the pool's pairs, radix select of the `top` largest keys and a
 * one-block sort of them into pairs; scratch holds the scores, the select
 * state and the selected keys (bm_mlp_select_bytes(pool)), and the state's
 * counters and histograms are cleared by k_select_init, so the call
 * allocates nothing and needs no memset. */
hipError_t bm_launch_mlp_rank_select(const BmMlpDev& d, const float* vf, uint64_t seed, int32_t n, int64_t pool,
                                     int64_t top, void* scratch, size_t scratch_bytes, int2* out, int32_t num_cus,
                                     hipStream_t s) {
    if (pool <= 0 || top <= 0) return hipSuccess;
    if (top > kSortMax || pool >= (1LL << 32) || scratch_bytes < bm_mlp_select_bytes(pool)) return hipErrorInvalidValue;
    float* sc = static_cast<float*>(scratch);
    hipError_t e = bm_launch_mlp_forward(d, vf, nullptr, seed, n, pool, sc, num_cus, s);
    if (e != hipSuccess) return e;
    char* q = reinterpret_cast<char*>(sc + pool);
    q += (256 - (reinterpret_cast<uintptr_t>(q) & 255)) & 255;
    SelState* st = reinterpret_cast<SelState*>(q);
    uint64_t* keys = reinterpret_cast<uint64_t*>(st + 1);
    const int blocks = grid_of(pool, kSelThreads * kSelPer, (int64_t)(num_cus > 0 ? num_cus : 256) * 4);
    hipLaunchKernelGGL(k_select_init, dim3(1), dim3(256), 0, s, st, top);
    for (int pass = 0; pass < kSelPasses; pass++)
        hipLaunchKernelGGL(k_select_pass, dim3(blocks), dim3(kSelThreads), 0, s, sc, pool, st, pass);
    hipLaunchKernelGGL(k_select_compact, dim3(blocks), dim3(kSelThreads), 0, s, sc, pool, st, keys, (uint32_t)top);
    int padded = 1;
    while (padded < top) padded <<= 1;
    if (padded < 2) padded = 2;
    e = hipFuncSetAttribute((const void*)k_select_sort, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)((size_t)padded * sizeof(uint64_t)));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_select_sort, dim3(1), dim3(1024), (size_t)padded * sizeof(uint64_t), s, keys, (int32_t)top,
                       padded, seed, n, out);
    return hipGetLastError();
}

size_t bm_mlp_rank_scratch(int64_t pool) {
    size_t temp = 0;
    (void)hipcub::DeviceRadixSort::SortPairsDescending(nullptr, temp, (const float*)nullptr, (float*)nullptr,
                                                       (const int32_t*)nullptr, (int32_t*)nullptr, (int)pool);
    const size_t sort_path = temp + 256 + (size_t)pool * (2 * sizeof(float) + 2 * sizeof(int32_t));
    const size_t select_path = bm_mlp_select_bytes(pool);
    return sort_path > select_path ? sort_path : select_path;
}

hipError_t bm_launch_mlp_rank(const BmMlpDev& d, const float* vf, uint64_t seed, int32_t n, int64_t pool, int64_t top,
                              void* scratch, size_t scratch_bytes, int2* out, int32_t num_cus, hipStream_t s) {
    if (pool <= 0 || top <= 0) return hipSuccess;
    char* p = static_cast<char*>(scratch);
    float* sc = reinterpret_cast<float*>(p);
    float* sc_sorted = sc + pool;
    int32_t* idx = reinterpret_cast<int32_t*>(sc_sorted + pool);
    int32_t* idx_sorted = idx + pool;
    char* temp = reinterpret_cast<char*>(idx_sorted + pool);
    temp += (256 - (reinterpret_cast<uintptr_t>(temp) & 255)) & 255;
    size_t temp_bytes = scratch_bytes - (size_t)(temp - p);
    /* radix select of the top keys, then one block sorts them */
    if (top <= kSortMax && pool < (1LL << 32))
        return bm_launch_mlp_rank_select(d, vf, seed, n, pool, top, scratch, scratch_bytes, out, num_cus, s);
    hipError_t e = bm_launch_mlp_forward(d, vf, nullptr, seed, n, pool, sc, num_cus, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_iota, dim3(grid_of(pool, 256)), dim3(256), 0, s, idx, pool);
    e = hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, sc, sc_sorted, idx, idx_sorted, (int)pool, 0,
                                                     32, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gather_stream_pairs, dim3(grid_of(top, 256)), dim3(256), 0, s, seed, n, idx_sorted, top, out);
    return hipGetLastError();
}
