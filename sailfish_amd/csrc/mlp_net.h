// mlp_net.h -- the closure network on gfx950 matrix cores: layout, operands and the forward
// pass shared by training (mlp.hip) and inference (closure_infer.hip).
//
// Network: 45 inputs (per field ubar, vbar, zeta, h, f: the value at a coarse point and the
// 8 differences neighbour - point of its 3x3 stencil) -> 64 ReLU -> 64 ReLU -> 2 outputs
// (ubar and vbar tendency residuals).  Padded shapes: input 48, output 16.  Parameter
// vector (FP32 master copy, 8336 values):
//   W1[48][64], b1[64], W2[64][64], b2[64], W3[64][16], b3[16]
// Products on the matrix cores with BF16 operands and FP32 accumulation; biases, loss and
// Adam in FP32.  Layer l: Z = A W + b (FP32), A' = relu(bf16(Z)), rounding to nearest
// even and then the sign test on the packed values: a BF16 value is negative exactly when
// its bit pattern is a negative 16-bit integer, so one v_pk_max_i16 applies the ReLU to
// two values.  Backward: dZ = dA * [A > 0]; the products take bf16(dZ), the bias
// gradients sum the FP32 dZ.
//
// Every product keeps the samples on its column index ("transposed" products
// Z^T = W^T A^T), 16 samples per wave at a time, so that one layer's result is the next
// layer's operand without leaving the registers.  Operand layouts (wave64, lane l,
// r = l % 16, q = l / 16; 16 x 16 result D):
//   16x16x16: lane holds A[r][4q + j] and B[4q + j][r], j = 0..3
//   16x16x32: lane holds A[r][c(q, j)] and B[c(q, j)][r], j = 0..7, in the order
//             c(q, j) = 4q + j (j < 4), 16 + 4q + j - 4 (j >= 4) chosen here: a product
//             sums over K, so any order works that the A and B operands share
//   D:        lane holds D[4q + t][r], t = 0..3
// Two consecutive D tiles (result rows 32kb .. 32kb + 15 and 32kb + 16 .. 32kb + 31),
// packed to BF16 as elements 0..3 and 4..7, are then exactly the B operand of a 16x16x32
// product over those 32 rows.  The weight gradients sum over the samples, which the
// registers hold on the column index; a product with the identity moves them to the
// row index (exact: one product with 1.0 per sum).
#pragma once

#include "sf_internal.h"


namespace {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));    // 16x16x16 operand (BF16 bits)
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));   // 16x16x32 operand
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef short s16x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

constexpr int NIN = 45, IN = 48, H = 64, NOUT = 2, OUT = 16;
constexpr int OFF_W1 = 0, OFF_B1 = IN * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H, OFF_W3 = OFF_B2 + H,
              OFF_B3 = OFF_W3 + H * OUT, NPARAM = OFF_B3 + OUT;
constexpr int BT = 64;    // samples per block tile (16 per wave)
constexpr int NT = 256;   // threads per workgroup (4 waves)

// two FP32 -> packed BF16 pair (lo in bits 0..15), round to nearest even
__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t){lo, hi}, bf16x2_t));
}
__device__ __forceinline__ uint16_t f2bf(float x) { return (uint16_t)pack_bf16x2(x, 0.0f); }

__device__ __forceinline__ f32x4_t mma16(bf16x4_t a, bf16x4_t b, f32x4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4_t mma32(bf16x8_t a, bf16x8_t b, f32x4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// a D tile as BF16 (4 values, D rows 4q .. 4q + 3), plain or through the ReLU
__device__ __forceinline__ u32x2_t pack4(f32x4_t z)
{
    return (u32x2_t){pack_bf16x2(z[0], z[1]), pack_bf16x2(z[2], z[3])};
}
__device__ __forceinline__ uint32_t relu2(uint32_t p)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2_t, p), (s16x2_t){0, 0}));
}
__device__ __forceinline__ u32x2_t relu4(f32x4_t z)
{
    const u32x2_t p = pack4(z);
    return (u32x2_t){relu2(p[0]), relu2(p[1])};
}
__device__ __forceinline__ bf16x8_t cat8(u32x2_t lo, u32x2_t hi)
{
    return __builtin_bit_cast(bf16x8_t, (u32x4_t){lo[0], lo[1], hi[0], hi[1]});
}
__device__ __forceinline__ bf16x4_t as4(u32x2_t p) { return __builtin_bit_cast(bf16x4_t, p); }

// Forward operand layout (wpk, BF16 values, lane-major tiles of W^T):
//   layer 1, column tile tn = 0..3: a 16x16x32 tile over inputs 0..31 (64 lanes x 8),
//            then a 16x16x16 tile over inputs 32..47 (64 lanes x 4)
//   layer 2, (tn, kb): 16x16x32 tiles over layer-1 units 32kb .. 32kb + 31
//   layer 3, kb = 0, 1: 16x16x32 tiles
// Element j of lane (r, q) in a tile of column tile tn and K base kbase holds
// W[kbase + c(q, j)][16 tn + r] (16x16x16 tile: W[32 + 4q + j][16 tn + r]).
constexpr int WPK_L1 = 0, WPK_L1_TN = 768, WPK_L2 = 3072, WPK_L3 = 7168, NWPK = 8192;

// lane and element of K offset kk (0..31) within a 16x16x32 tile
__device__ __forceinline__ int k32_slot(int r, int kk)
{
    const int q = (kk & 15) >> 2, j = (kk & 3) + ((kk >> 4) << 2);
    return (r + 16 * q) * 8 + j;
}

// slot of parameter p in the packed layout, -1 for biases
__device__ __forceinline__ int wpk_slot(int p)
{
    if (p < OFF_B1) {
        const int k = p / H, n = p % H, tn = n >> 4, r = n & 15;
        if (k < 32) return WPK_L1 + tn * WPK_L1_TN + k32_slot(r, k);
        const int kk = k - 32;
        return WPK_L1 + tn * WPK_L1_TN + 512 + (r + 16 * (kk >> 2)) * 4 + (kk & 3);
    }
    if (p >= OFF_W2 && p < OFF_B2) {
        const int k = (p - OFF_W2) / H, n = (p - OFF_W2) % H;
        return WPK_L2 + ((n >> 4) * 2 + (k >> 5)) * 512 + k32_slot(n & 15, k & 31);
    }
    if (p >= OFF_W3 && p < OFF_B3) {
        const int k = (p - OFF_W3) / OUT, n = (p - OFF_W3) % OUT;
        return WPK_L3 + (k >> 5) * 512 + k32_slot(n, k & 31);
    }
    return -1;
}

// the forward network in registers: weight operands and the biases in the D layout
// (lane (r, q) holds b[16 tn + 4q + t])
struct Fwd {
    bf16x8_t w1a[4], w2[8], w3[2];
    bf16x4_t w1b[4];
    f32x4_t b1[4], b2[4], b3;
};

__device__ __forceinline__ void load_fwd(Fwd &W, const uint16_t *__restrict__ wpk, const float *__restrict__ w)
{
    const int lane = threadIdx.x & 63, q = lane >> 4;
#pragma unroll
    for (int tn = 0; tn < 4; tn++) {
        W.w1a[tn] = *(const bf16x8_t *)(wpk + WPK_L1 + tn * WPK_L1_TN + lane * 8);
        W.w1b[tn] = *(const bf16x4_t *)(wpk + WPK_L1 + tn * WPK_L1_TN + 512 + lane * 4);
        W.b1[tn] = *(const f32x4_t *)(w + OFF_B1 + 16 * tn + 4 * q);
        W.b2[tn] = *(const f32x4_t *)(w + OFF_B2 + 16 * tn + 4 * q);
    }
#pragma unroll
    for (int e = 0; e < 8; e++) W.w2[e] = *(const bf16x8_t *)(wpk + WPK_L2 + e * 512 + lane * 8);
#pragma unroll
    for (int kb = 0; kb < 2; kb++) W.w3[kb] = *(const bf16x8_t *)(wpk + WPK_L3 + kb * 512 + lane * 8);
    W.b3 = *(const f32x4_t *)(w + OFF_B3 + 4 * q);
}

// B operands of 16 samples from their BF16 input rows (lane (r, q): sample r's row)
__device__ __forceinline__ void load_x(const uint16_t *row, bf16x8_t &xa, bf16x4_t &xb)
{
    const int q = (threadIdx.x & 63) >> 4;
    xa = cat8(*(const u32x2_t *)(row + 4 * q), *(const u32x2_t *)(row + 16 + 4 * q));
    xb = *(const bf16x4_t *)(row + 32 + 4 * q);
}

struct Acts {
    u32x2_t a1[4], a2[4];   // BF16 activations, one 16-unit tile each (D rows 4q .. 4q + 3)
    f32x4_t z3;             // lanes q = 0: outputs 0, 1 of sample r in z3[0], z3[1]
};

// forward pass of 16 samples; Z = b + (first K-32 part) + (next part), in that order
__device__ __forceinline__ void forward16(const Fwd &W, bf16x8_t xa, bf16x4_t xb, Acts &A)
{
    f32x4_t z[4];
#pragma unroll
    for (int tn = 0; tn < 4; tn++) z[tn] = mma16(W.w1b[tn], xb, mma32(W.w1a[tn], xa, W.b1[tn]));
#pragma unroll
    for (int tn = 0; tn < 4; tn++) A.a1[tn] = relu4(z[tn]);
    const bf16x8_t h0 = cat8(A.a1[0], A.a1[1]), h1 = cat8(A.a1[2], A.a1[3]);
#pragma unroll
    for (int tn = 0; tn < 4; tn++) z[tn] = mma32(W.w2[2 * tn + 1], h1, mma32(W.w2[2 * tn], h0, W.b2[tn]));
#pragma unroll
    for (int tn = 0; tn < 4; tn++) A.a2[tn] = relu4(z[tn]);
    A.z3 = mma32(W.w3[1], cat8(A.a2[2], A.a2[3]), mma32(W.w3[0], cat8(A.a2[0], A.a2[1]), W.b3));
}

// Matrix-core operand hand-offs with explicit waits.  On gfx950 (this toolchain) a product
// whose A/B operand registers were written by a VALU instruction a few cycles before, or
// are overwritten by one shortly after the product issues, can read wrong operand values:
// observed as wrong outputs for whole 16-sample groups, depending only on the instruction
// schedule.  forward32_held keeps each layer's operands unchanged from 5 wait states
// before its products until 16 after them (layer granularity: one wait before, one after).
#define SF_SETTLE(...) asm volatile("s_nop 4" : __VA_ARGS__)
#define SF_HOLD(...)                                        \
    do {                                                    \
        __builtin_amdgcn_sched_barrier(0);                  \
        asm volatile("s_nop 7\n\ts_nop 7" ::__VA_ARGS__);  \
        __builtin_amdgcn_sched_barrier(0);                  \
    } while (0)

// forward16 for two 16-sample groups at once (inputs xa/xb and ya/yb; network outputs zx,
// zy in forward16's z3 layout), with the waits above: one before and one after each
// layer's products of both groups, which issue back to back.  Each group's sums are those
// of forward16, so its outputs are too.
__device__ __forceinline__ void forward32_held(const Fwd &W, bf16x8_t xa, bf16x4_t xb, bf16x8_t ya, bf16x4_t yb,
                                               f32x4_t &zx, f32x4_t &zy)
{
    bf16x8_t a0 = W.w1a[0], a1 = W.w1a[1], a2 = W.w1a[2], a3 = W.w1a[3];
    bf16x4_t c0 = W.w1b[0], c1 = W.w1b[1], c2 = W.w1b[2], c3 = W.w1b[3];
    SF_SETTLE("+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(c0), "+v"(c1), "+v"(c2), "+v"(c3), "+v"(xa), "+v"(xb),
              "+v"(ya), "+v"(yb));
    f32x4_t z[4], y[4];
    z[0] = mma16(c0, xb, mma32(a0, xa, W.b1[0]));
    y[0] = mma16(c0, yb, mma32(a0, ya, W.b1[0]));
    z[1] = mma16(c1, xb, mma32(a1, xa, W.b1[1]));
    y[1] = mma16(c1, yb, mma32(a1, ya, W.b1[1]));
    z[2] = mma16(c2, xb, mma32(a2, xa, W.b1[2]));
    y[2] = mma16(c2, yb, mma32(a2, ya, W.b1[2]));
    z[3] = mma16(c3, xb, mma32(a3, xa, W.b1[3]));
    y[3] = mma16(c3, yb, mma32(a3, ya, W.b1[3]));
    SF_HOLD("v"(a0), "v"(a1), "v"(a2), "v"(a3), "v"(c0), "v"(c1), "v"(c2), "v"(c3), "v"(xa), "v"(xb), "v"(ya),
            "v"(yb));
    u32x2_t p[4], s[4];
#pragma unroll
    for (int tn = 0; tn < 4; tn++) p[tn] = relu4(z[tn]), s[tn] = relu4(y[tn]);
    bf16x8_t h0 = cat8(p[0], p[1]), h1 = cat8(p[2], p[3]), k0 = cat8(s[0], s[1]), k1 = cat8(s[2], s[3]);
    bf16x8_t w0 = W.w2[0], w1 = W.w2[1], w2 = W.w2[2], w3 = W.w2[3], w4 = W.w2[4], w5 = W.w2[5], w6 = W.w2[6],
             w7 = W.w2[7];
    SF_SETTLE("+v"(w0), "+v"(w1), "+v"(w2), "+v"(w3), "+v"(w4), "+v"(w5), "+v"(w6), "+v"(w7), "+v"(h0), "+v"(h1),
              "+v"(k0), "+v"(k1));
    z[0] = mma32(w1, h1, mma32(w0, h0, W.b2[0]));
    y[0] = mma32(w1, k1, mma32(w0, k0, W.b2[0]));
    z[1] = mma32(w3, h1, mma32(w2, h0, W.b2[1]));
    y[1] = mma32(w3, k1, mma32(w2, k0, W.b2[1]));
    z[2] = mma32(w5, h1, mma32(w4, h0, W.b2[2]));
    y[2] = mma32(w5, k1, mma32(w4, k0, W.b2[2]));
    z[3] = mma32(w7, h1, mma32(w6, h0, W.b2[3]));
    y[3] = mma32(w7, k1, mma32(w6, k0, W.b2[3]));
    SF_HOLD("v"(w0), "v"(w1), "v"(w2), "v"(w3), "v"(w4), "v"(w5), "v"(w6), "v"(w7), "v"(h0), "v"(h1), "v"(k0),
            "v"(k1));
#pragma unroll
    for (int tn = 0; tn < 4; tn++) p[tn] = relu4(z[tn]), s[tn] = relu4(y[tn]);
    bf16x8_t g0 = cat8(p[0], p[1]), g1 = cat8(p[2], p[3]), e0 = cat8(s[0], s[1]), e1 = cat8(s[2], s[3]);
    bf16x8_t v0 = W.w3[0], v1 = W.w3[1];
    SF_SETTLE("+v"(v0), "+v"(v1), "+v"(g0), "+v"(g1), "+v"(e0), "+v"(e1));
    zx = mma32(v1, g1, mma32(v0, g0, W.b3));
    zy = mma32(v1, e1, mma32(v0, e0, W.b3));
    SF_HOLD("v"(v0), "v"(v1), "v"(g0), "v"(g1), "v"(e0), "v"(e1));
}

// Per-wave feature tile of 64 samples: row = sample, 48 BF16 inputs, row stride XS BF16
// pairs (104 B: conflict-free 8-B row writes); the rows are read back as B operands.
constexpr int XS = 26;

// LDS writes of a wave visible to its other lanes
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Features of a point as BF16 row `lane` of the wave's tile X: xv[0..44] normalised as
// (x - mu) * rsd with nrm = {mu[45], rsd[45]}, inputs 45..47 zero.
__device__ __forceinline__ void closure_row(uint32_t *X, int lane, const float *xraw, const float *nrm)
{
    u32x2_t *row = (u32x2_t *)(X + lane * XS);
#pragma unroll
    for (int e = 0; e < IN / 4; e++) {
        float x4[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int k = 4 * e + t;
            x4[t] = k < NIN ? (xraw[k] - nrm[k]) * nrm[NIN + k] : 0.0f;
        }
        row[e] = (u32x2_t){pack_bf16x2(x4[0], x4[1]), pack_bf16x2(x4[2], x4[3])};
    }
}

// Network outputs of the wave's 64 samples (tile X, written by closure_row): lane l gets
// outputs 0 and 1 of sample l.  The 16-sample groups run in pairs (0, 1), then (2, 3).
__device__ __forceinline__ void closure_eval(const Fwd &W, uint32_t *X, float &zu, float &zv)
{
    const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
    wave_lds_sync();
    zu = zv = 0.0f;
#pragma unroll
    for (int gq = 0; gq < 4; gq += 2) {
        const u32x2_t *xr = (const u32x2_t *)(X + (16 * gq + r) * XS), *yr = xr + 16 * XS / 2;
        f32x4_t zx, zy;
        forward32_held(W, cat8(xr[q], xr[4 + q]), as4(xr[8 + q]), cat8(yr[q], yr[4 + q]), as4(yr[8 + q]), zx, zy);
        // lanes q = 0 hold the outputs of sample r of each group
        const float tu = __shfl(zx[0], r), tv = __shfl(zx[1], r), su = __shfl(zy[0], r), sv = __shfl(zy[1], r);
        if (q == gq) zu = tu, zv = tv;
        if (q == gq + 1) zu = su, zv = sv;
    }
    wave_lds_sync();   // the tile is rewritten by the next chunk
}

// dZ = dA where the activation (BF16, >= 0) is positive, else 0
__device__ __forceinline__ f32x4_t relu_mask(f32x4_t d, u32x2_t a)
{
    f32x4_t o;
    o[0] = (a[0] & 0xFFFFu) ? d[0] : 0.0f;
    o[1] = (a[0] >> 16) ? d[1] : 0.0f;
    o[2] = (a[1] & 0xFFFFu) ? d[2] : 0.0f;
    o[3] = (a[1] >> 16) ? d[3] : 0.0f;
    return o;
}

}  // namespace

struct sf_mlp {
    float *w, *m, *v, *grad, *partial, *ploss, *loss;
    uint16_t *wb;               // BF16 parameters, parameter order
    uint16_t *wpk;              // BF16 weights, inference operand order (NWPK)
    float *norm;                // mu[45], 1 / sd[45] (FP32)
    float ynorm[4];
    int max_blocks;
};
