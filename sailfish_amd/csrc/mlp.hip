// mlp.hip -- the learned subgrid momentum closure on gfx950 matrix cores.
//
// Network: 45 inputs (per field ubar, vbar, zeta, h, f: the value at a coarse point and the
// 8 differences neighbour - point of its 3x3 stencil)
// -> 64 GELU -> 64 GELU -> 2 outputs (ubar and vbar tendency residuals).  Padded
// shapes: input 48, output 16.  Parameter vector (FP32 master copy, 8336 values):
//   W1[48][64], b1[64], W2[64][64], b2[64], W3[64][16], b3[16]
// Products run on v_mfma_f32_16x16x16_bf16 (BF16 operands, FP32 accumulation);
// biases, GELU, loss and Adam run in FP32.  A workgroup of 4 waves processes tiles
// of 64 samples: forward, MSE loss, backward; its gradient partial goes to its own
// slot, and a second kernel reduces the slots in a fixed order (deterministic) and
// applies Adam.
//
// MFMA 16x16x16 operand layout (wave64, lane l, r = l % 16, q = l / 16):
//   A (16 x 16, row-major M x K): lane holds A[r][4q .. 4q+3]
//   B (16 x 16, K x N):           lane holds B[4q .. 4q+3][r]
//   D (16 x 16):                  lane holds D[4q + t][r], t = 0..3
// Activation: the sigmoid form of GELU, x * sigmoid(1.702 x) (Hendrycks & Gimpel 2016),
// in FP32 with the hardware exponential and reciprocal: five VALU instructions, its
// derivative six (the erfc form of x * Phi(x) took ~15).  Training and inference use
// the same function.
// BF16 rounding is round-to-nearest-even (v_cvt_pk_bf16_f32).
#include "sf_internal.h"

#include <math.h>

namespace {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int NIN = 45, IN = 48, H = 64, NOUT = 2, OUT = 16;
constexpr int OFF_W1 = 0, OFF_B1 = IN * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H, OFF_W3 = OFF_B2 + H,
              OFF_B3 = OFF_W3 + H * OUT, NPARAM = OFF_B3 + OUT;
constexpr int BT = 64;    // samples per tile
constexpr int NT = 256;   // threads per workgroup (4 waves)

typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// two FP32 -> packed BF16 pair (lo in bits 0..15), round to nearest even
__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t){lo, hi}, bf16x2_t));
}
__device__ __forceinline__ uint16_t f2bf(float x) { return (uint16_t)pack_bf16x2(x, 0.0f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// sigmoid(1.702 x) = 1 / (1 + 2^(-1.702 log2(e) x)); for x -> -inf the power overflows to
// +inf and the reciprocal gives 0
__device__ __forceinline__ float sig1702(float x)
{
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(x * -2.4554669595930157f));
}
__device__ __forceinline__ float gelu(float x) { return x * sig1702(x); }
__device__ __forceinline__ float gelu_grad(float x)
{
    const float s = sig1702(x);
    return fmaf((1.702f * x) * s, 1.0f - s, s);
}

// One 16x16 output tile (tm, tn) of A(M x K) * B(K x N) accumulated onto acc; A(m, k)
// and B(k, n) are accessors returning BF16 bits.
template <typename FA, typename FB>
__device__ __forceinline__ f32x4_t mfma_tile(int tm, int tn, int K, FA A, FB B, f32x4_t acc)
{
    const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4_t a, b;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            a[t] = (short)A(tm * 16 + r, k0 + 4 * q + t);
            b[t] = (short)B(k0 + 4 * q + t, tn * 16 + r);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    return acc;
}

struct MlpShared {
    uint16_t W1[IN][H], W2[H][H], W3[H][OUT];    // BF16 weights
    float b1[H], b2[H], b3[OUT];
    uint16_t X[BT][IN];
    float Z1[BT][H], Z2[BT][H];
    uint16_t A1[BT][H], A2[BT][H];
    float Z3[BT][OUT];
};

__device__ void load_weights(MlpShared &S, const uint16_t *wb, const float *w)
{
    for (int e = threadIdx.x; e < IN * H; e += NT) S.W1[e / H][e % H] = wb[OFF_W1 + e];
    for (int e = threadIdx.x; e < H * H; e += NT) S.W2[e / H][e % H] = wb[OFF_W2 + e];
    for (int e = threadIdx.x; e < H * OUT; e += NT) S.W3[e / OUT][e % OUT] = wb[OFF_W3 + e];
    for (int e = threadIdx.x; e < H; e += NT) {
        S.b1[e] = w[OFF_B1 + e];
        S.b2[e] = w[OFF_B2 + e];
    }
    for (int e = threadIdx.x; e < OUT; e += NT) S.b3[e] = w[OFF_B3 + e];
}

// forward pass of the tile in S.X; leaves Z1, A1, Z2, A2, Z3 in LDS
__device__ void forward_tile(MlpShared &S)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    // Z1 = X W1 + b1 ; wave w: rows 16w..16w+15, all 4 column tiles
    for (int tn = 0; tn < 4; tn++) {
        const float bb = S.b1[tn * 16 + r];
        f32x4_t acc = {bb, bb, bb, bb};
        acc = mfma_tile(wave, tn, IN, [&](int m, int k) { return S.X[m][k]; },
                        [&](int k, int n) { return S.W1[k][n]; }, acc);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int m = wave * 16 + 4 * q + t, n = tn * 16 + r;
            S.Z1[m][n] = acc[t];
            S.A1[m][n] = f2bf(gelu(acc[t]));
        }
    }
    __syncthreads();
    for (int tn = 0; tn < 4; tn++) {
        const float bb = S.b2[tn * 16 + r];
        f32x4_t acc = {bb, bb, bb, bb};
        acc = mfma_tile(wave, tn, H, [&](int m, int k) { return S.A1[m][k]; },
                        [&](int k, int n) { return S.W2[k][n]; }, acc);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int m = wave * 16 + 4 * q + t, n = tn * 16 + r;
            S.Z2[m][n] = acc[t];
            S.A2[m][n] = f2bf(gelu(acc[t]));
        }
    }
    __syncthreads();
    {
        const float bb = S.b3[r];
        f32x4_t acc = {bb, bb, bb, bb};
        acc = mfma_tile(wave, 0, H, [&](int m, int k) { return S.A2[m][k]; },
                        [&](int k, int n) { return S.W3[k][n]; }, acc);
#pragma unroll
        for (int t = 0; t < 4; t++) S.Z3[wave * 16 + 4 * q + t][r] = acc[t];
    }
    __syncthreads();
}

// Inference operand layout.  The inference kernel evaluates the transposed products
// Z^T = W^T X^T: the A operand is a weight tile of W^T, the B operand the activations
// (lane r, q: sample r, units 4q..4q+3), and the D layout of one layer is then the B
// layout of the next.  Tile order: layer 1 (tn, kb) at tn * 3 + kb, layer 2 at
// 12 + tn * 4 + kb, layer 3 at 28 + kb; lane r + 16 q of tile (tn, kb) holds
// W[16 kb + 4 q + t][16 tn + r], t = 0..3, as 4 BF16 values.
constexpr int NWT = 32;                     // weight tiles
constexpr int NWPK = NWT * 64 * 4;          // BF16 values in the packed layout

// slot of parameter p in the packed layout, -1 for biases
__device__ __forceinline__ int wpk_slot(int p)
{
    int k, n, tile;
    if (p < OFF_B1) {
        k = p / H, n = p % H, tile = (n >> 4) * 3 + (k >> 4);
    } else if (p >= OFF_W2 && p < OFF_B2) {
        k = (p - OFF_W2) / H, n = (p - OFF_W2) % H, tile = 12 + (n >> 4) * 4 + (k >> 4);
    } else if (p >= OFF_W3 && p < OFF_B3) {
        k = (p - OFF_W3) / OUT, n = (p - OFF_W3) % OUT, tile = 28 + (k >> 4);
    } else {
        return -1;
    }
    return (tile * 64 + (n & 15) + 16 * ((k & 15) >> 2)) * 4 + (k & 3);
}

struct BwdShared {
    uint16_t dZ3[BT][OUT], dZ2[BT][H], dZ1[BT][H];
    float dZ3f[BT][OUT], dZ2f[BT][H], dZ1f[BT][H];
    float db1[H], db2[H], db3[OUT];
    float loss[NT];
};

// Forward + backward over ntile tiles per workgroup.  Sample s of tile t of block b is
// idx[(b * ntile + t) * 64 + s].  partial[b][NPARAM]: this block's gradient sum;
// partial_loss[b]: its sum of squared errors.  scale = 2 / (batch * NOUT).
__global__ __launch_bounds__(NT) void k_mlp_fwd_bwd(const uint16_t *__restrict__ wb, const float *__restrict__ w,
                                                   const uint16_t *__restrict__ X, const float *__restrict__ Y,
                                                   const int32_t *__restrict__ idx, int ntile, float scale,
                                                   float *__restrict__ partial, float *__restrict__ partial_loss)
{
    __shared__ MlpShared S;
    __shared__ BwdShared Bw;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    load_weights(S, wb, w);
    for (int e = threadIdx.x; e < H; e += NT) Bw.db1[e] = Bw.db2[e] = 0.0f;
    for (int e = threadIdx.x; e < OUT; e += NT) Bw.db3[e] = 0.0f;
    float my_loss = 0.0f;
    const f32x4_t z4 = {0.f, 0.f, 0.f, 0.f};
    f32x4_t gW3 = z4, gW2[4] = {z4, z4, z4, z4}, gW1[3] = {z4, z4, z4};
    for (int t = 0; t < ntile; t++) {
        const int32_t *ti = idx + ((long long)blockIdx.x * ntile + t) * BT;
        for (int e = threadIdx.x; e < BT * IN; e += NT) {
            const int s = e / IN, k = e % IN;
            S.X[s][k] = X[(long long)ti[s] * IN + k];
        }
        __syncthreads();
        forward_tile(S);
        // dZ3 and the loss
        for (int e = threadIdx.x; e < BT * OUT; e += NT) {
            const int s = e / OUT, o = e % OUT;
            float d = 0.0f;
            if (o < NOUT) {
                const float err = S.Z3[s][o] - Y[(long long)ti[s] * NOUT + o];
                my_loss += err * err;
                d = scale * err;
            }
            Bw.dZ3f[s][o] = d;
            Bw.dZ3[s][o] = f2bf(d);
        }
        __syncthreads();
        // dW3 = A2^T dZ3 (64 x 16): wave w owns row tile w
        gW3 = mfma_tile(wave, 0, BT, [&](int m, int k) { return S.A2[k][m]; },
                        [&](int k, int n) { return Bw.dZ3[k][n]; }, gW3);
        // dZ2 = (dZ3 W3^T) * gelu'(Z2)
        for (int tn = 0; tn < 4; tn++) {
            f32x4_t acc = mfma_tile(wave, tn, OUT, [&](int m, int k) { return Bw.dZ3[m][k]; },
                                    [&](int k, int n) { return S.W3[n][k]; }, z4);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int m = wave * 16 + 4 * q + u, n = tn * 16 + r;
                const float d = acc[u] * gelu_grad(S.Z2[m][n]);
                Bw.dZ2f[m][n] = d;
                Bw.dZ2[m][n] = f2bf(d);
            }
        }
        __syncthreads();
        // dW2 = A1^T dZ2 (64 x 64): wave w owns row tile w
        for (int tn = 0; tn < 4; tn++)
            gW2[tn] = mfma_tile(wave, tn, BT, [&](int m, int k) { return S.A1[k][m]; },
                                [&](int k, int n) { return Bw.dZ2[k][n]; }, gW2[tn]);
        // dZ1 = (dZ2 W2^T) * gelu'(Z1)
        for (int tn = 0; tn < 4; tn++) {
            f32x4_t acc = mfma_tile(wave, tn, H, [&](int m, int k) { return Bw.dZ2[m][k]; },
                                    [&](int k, int n) { return S.W2[n][k]; }, z4);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int m = wave * 16 + 4 * q + u, n = tn * 16 + r;
                const float d = acc[u] * gelu_grad(S.Z1[m][n]);
                Bw.dZ1f[m][n] = d;
                Bw.dZ1[m][n] = f2bf(d);
            }
        }
        __syncthreads();
        // dW1 = X^T dZ1 (48 x 64): wave w owns tiles w, w+4, w+8 of the 3 x 4 grid
        for (int c = 0; c < 3; c++) {
            const int tile = wave + 4 * c, tm = tile / 4, tn = tile % 4;
            gW1[c] = mfma_tile(tm, tn, BT, [&](int m, int k) { return S.X[k][m]; },
                               [&](int k, int n) { return Bw.dZ1[k][n]; }, gW1[c]);
        }
        // bias gradients (FP32 column sums in sample order)
        if (threadIdx.x < H) {
            float s1 = 0.0f, s2 = 0.0f;
            for (int s = 0; s < BT; s++) {
                s1 += Bw.dZ1f[s][threadIdx.x];
                s2 += Bw.dZ2f[s][threadIdx.x];
            }
            Bw.db1[threadIdx.x] += s1;
            Bw.db2[threadIdx.x] += s2;
        } else if (threadIdx.x < H + OUT) {
            const int o = threadIdx.x - H;
            float s3 = 0.0f;
            for (int s = 0; s < BT; s++) s3 += Bw.dZ3f[s][o];
            Bw.db3[o] += s3;
        }
        __syncthreads();
    }
    // write this block's partial gradient
    float *P = partial + (long long)blockIdx.x * NPARAM;
#pragma unroll
    for (int u = 0; u < 4; u++) P[OFF_W3 + (wave * 16 + 4 * q + u) * OUT + r] = gW3[u];
    for (int tn = 0; tn < 4; tn++)
#pragma unroll
        for (int u = 0; u < 4; u++) P[OFF_W2 + (wave * 16 + 4 * q + u) * H + tn * 16 + r] = gW2[tn][u];
    for (int c = 0; c < 3; c++) {
        const int tile = wave + 4 * c, tm = tile / 4, tn = tile % 4;
#pragma unroll
        for (int u = 0; u < 4; u++) P[OFF_W1 + (tm * 16 + 4 * q + u) * H + tn * 16 + r] = gW1[c][u];
    }
    for (int e = threadIdx.x; e < H; e += NT) {
        P[OFF_B1 + e] = Bw.db1[e];
        P[OFF_B2 + e] = Bw.db2[e];
    }
    for (int e = threadIdx.x; e < OUT; e += NT) P[OFF_B3 + e] = Bw.db3[e];
    Bw.loss[threadIdx.x] = my_loss;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int e = 0; e < NT; e++) s += Bw.loss[e];
        partial_loss[blockIdx.x] = s;
    }
}

// grad[p] = sum over the blocks' partials; optional Adam step on the FP32 master and
// refresh of the BF16 copy; loss_out = sum of squared errors / (batch * NOUT).
// A workgroup takes 64 parameters; its four waves each sum the partials of every
// fourth block (wave w: blocks w, w+4, w+8, ... in increasing order), and the four sums
// are added in wave order: a fixed order for any grid, so the result is deterministic.
// Reads are coalesced (64 consecutive parameters of one block per wave instruction).
constexpr int RP = 64;   // parameters per reduce workgroup
__global__ __launch_bounds__(256) void k_mlp_reduce_adam(const float *__restrict__ partial,
                                                        const float *__restrict__ partial_loss, int nblk,
                                                        float *__restrict__ grad, float *__restrict__ w,
                                                        float *__restrict__ m, float *__restrict__ v,
                                                        uint16_t *__restrict__ wb, uint16_t *__restrict__ wpk,
                                                        int adam, float lr,
                                                        float beta1, float beta2, float eps, float bc1, float bc2,
                                                        float inv_count, float *__restrict__ loss_out)
{
    __shared__ float red[4][RP];
    const int lane = threadIdx.x & 63, slice = threadIdx.x >> 6;
    const int p = blockIdx.x * RP + lane;
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss_out) {
        float s = 0.0f;
        for (int b = 0; b < nblk; b++) s += partial_loss[b];
        loss_out[0] = s * inv_count;
    }
    float g = 0.0f;
    if (p < NPARAM)
        for (int b = slice; b < nblk; b += 4) g += partial[(long long)b * NPARAM + p];
    red[slice][lane] = g;
    __syncthreads();
    if (slice != 0 || p >= NPARAM) return;
    g = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
    grad[p] = g;
    if (!adam) return;
    const float mm = beta1 * m[p] + (1.0f - beta1) * g;
    const float vv = beta2 * v[p] + (1.0f - beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / bc1, vh = vv / bc2;
    const float nw = w[p] - lr * mh / (sqrtf(vh) + eps);
    w[p] = nw;
    wb[p] = f2bf(nw);
    const int slot = wpk_slot(p);
    if (slot >= 0) wpk[slot] = wb[p];
}

// outputs Y[n][2] of BF16 inputs X[n][48]
__global__ __launch_bounds__(NT) void k_mlp_forward(const uint16_t *__restrict__ wb, const float *__restrict__ w,
                                                   const uint16_t *__restrict__ X, int n, float *__restrict__ Y)
{
    __shared__ MlpShared S;
    load_weights(S, wb, w);
    const long long s0 = (long long)blockIdx.x * BT;
    for (int e = threadIdx.x; e < BT * IN; e += NT) {
        const int s = e / IN, k = e % IN;
        S.X[s][k] = (s0 + s < n) ? X[(s0 + s) * IN + k] : (uint16_t)0;
    }
    __syncthreads();
    forward_tile(S);
    for (int e = threadIdx.x; e < BT * NOUT; e += NT) {
        const int s = e / NOUT, o = e % NOUT;
        if (s0 + s < n) Y[(s0 + s) * NOUT + o] = S.Z3[s][o];
    }
}

// ---- closure features and stress ------------------------------------------------------
// feature index: field * 9 + (dj + 1) * 3 + (di + 1), fields ubar, vbar, zeta, h, f.  The
// centre entry (dj = di = 0) is the field value, the others the difference neighbour -
// centre, formed in the state precision and then rounded to FP32: after the per-feature
// normalisation the stencil differences (the derivatives a subgrid closure needs) keep
// their resolution through the BF16 inputs, which the raw neighbour values (equal to
// the centre value in their leading digits) would lose.
template <typename T>
struct ClosureArgs {
    const T *u, *v, *z, *h, *f, *sustr, *svstr;
    T *sustr_eff, *svstr_eff;
    float *feat_out;            // optional raw features [S2][45]
    float mu_u, sd_u, mu_v, sd_v;   // output de-normalisation
    int szI, szJ;
};

template <typename T>
__device__ __forceinline__ float feature(const ClosureArgs<T> &a, int i, int j, int k)
{
    const int fld = k / 9, dj = (k % 9) / 3 - 1, di = k % 3 - 1;
    const long long gc = (long long)j * a.szI + i, g = gc + (long long)dj * a.szI + di;
    const T *p = fld == 0 ? a.u : fld == 1 ? a.v : fld == 2 ? a.z : fld == 3 ? a.h : a.f;
    return (dj == 0 && di == 0) ? (float)p[gc] : (float)(p[g] - p[gc]);
}

// raw features of every point with 1 <= i <= L-1, 1 <= j <= M-1 (others zero)
template <typename T>
__global__ __launch_bounds__(256) void k_closure_features(ClosureArgs<T> a)
{
    const long long S2 = (long long)a.szI * a.szJ;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= S2 * NIN) return;
    const long long g = e / NIN;
    const int k = (int)(e % NIN), i = (int)(g % a.szI), j = (int)(g / a.szI);
    const bool ok = i >= 1 && i <= a.szI - 2 && j >= 1 && j <= a.szJ - 2;
    a.feat_out[e] = ok ? feature(a, i, j, k) : 0.0f;
}

// closure stress: stress_eff = stress + R * D at interior u / v points, where R is the
// network's (de-normalised) tendency residual and D the depth at the velocity point.
// Each wave owns chunks of 64 consecutive grid points (lane = point while gathering the
// features and applying the stress) and keeps the network in registers; features are
// normalised as (x - mu) * rsd, rsd = 1 / sd rounded to FP32, and rounded to BF16.
// norm = {mu[45], rsd[45]}; wpk: the packed BF16 weights (NWPK values).
constexpr int XS = 26;   // feature-tile row stride in BF16 pairs (104 B: conflict-free 8-B row writes)

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x4_t gelu_bf16x4(f32x4_t z)
{
    const uint32_t lo = pack_bf16x2(gelu(z[0]), gelu(z[1])), hi = pack_bf16x2(gelu(z[2]), gelu(z[3]));
    return __builtin_bit_cast(bf16x4_t, (uint64_t)lo | ((uint64_t)hi << 32));
}

template <typename T>
__global__ __launch_bounds__(NT) void k_closure_stress(const uint64_t *__restrict__ wpk, const float *__restrict__ w,
                                                      const float *__restrict__ norm, ClosureArgs<T> a)
{
    __shared__ uint64_t Wl[NWT * 64];
    __shared__ __attribute__((aligned(16))) uint32_t Xl[NT / 64][64 * XS];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
    for (int e = threadIdx.x; e < NWT * 64; e += NT) Wl[e] = wpk[e];
    __syncthreads();
    bf16x4_t Wt[NWT];
#pragma unroll
    for (int t = 0; t < NWT; t++) Wt[t] = __builtin_bit_cast(bf16x4_t, Wl[t * 64 + lane]);
    f32x4_t B1[4], B2[4];
#pragma unroll
    for (int tn = 0; tn < 4; tn++) {
        B1[tn] = *(const f32x4_t *)(w + OFF_B1 + 16 * tn + 4 * q);
        B2[tn] = *(const f32x4_t *)(w + OFF_B2 + 16 * tn + 4 * q);
    }
    const f32x4_t B3 = *(const f32x4_t *)(w + OFF_B3 + 4 * q);
    uint32_t *X = Xl[wave];

    const long long S2 = (long long)a.szI * a.szJ;
    const int szI = a.szI, L = a.szI - 1, M = a.szJ - 1;
    const long long nchunk = (S2 + 63) / 64;
    const T *fld[5] = {a.u, a.v, a.z, a.h, a.f};
    for (long long c = (long long)blockIdx.x * (NT / 64) + wave; c < nchunk; c += (long long)gridDim.x * (NT / 64)) {
        const long long g = c * 64 + lane;
        const int j = (int)(g / szI), i = (int)(g - (long long)j * szI);
        const bool in = g < S2 && i >= 1 && i <= L - 1 && j >= 1 && j <= M - 1;
        const int gs = in ? (int)g : szI + 1;   // other lanes read an interior point; their outputs are unused
        // features of point gs -> row `lane` of the wave's tile
        float xv[IN];
#pragma unroll
        for (int k = 0; k < NIN; k++) {
            const int f = k / 9, dj = (k % 9) / 3 - 1, di = k % 3 - 1;
            const T c = fld[f][gs];
            const float x = (dj == 0 && di == 0) ? (float)c : (float)(fld[f][gs + dj * szI + di] - c);
            xv[k] = (x - norm[k]) * norm[NIN + k];
        }
        xv[45] = xv[46] = xv[47] = 0.0f;
        uint64_t *row = (uint64_t *)(X + lane * XS);
#pragma unroll
        for (int e = 0; e < IN / 4; e++)
            row[e] = (uint64_t)pack_bf16x2(xv[4 * e], xv[4 * e + 1]) |
                     ((uint64_t)pack_bf16x2(xv[4 * e + 2], xv[4 * e + 3]) << 32);
        wave_lds_sync();
        float Ru[4], Rv[4];
#pragma unroll
        for (int gq = 0; gq < 4; gq++) {
            const uint64_t *xr = (const uint64_t *)(X + (16 * gq + r) * XS);
            bf16x4_t xb[3], a1[4], a2[4];
#pragma unroll
            for (int kb = 0; kb < 3; kb++) xb[kb] = __builtin_bit_cast(bf16x4_t, xr[4 * kb + q]);
#pragma unroll
            for (int tn = 0; tn < 4; tn++) {
                f32x4_t acc = B1[tn];
#pragma unroll
                for (int kb = 0; kb < 3; kb++)
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Wt[tn * 3 + kb], xb[kb], acc, 0, 0, 0);
                a1[tn] = gelu_bf16x4(acc);
            }
#pragma unroll
            for (int tn = 0; tn < 4; tn++) {
                f32x4_t acc = B2[tn];
#pragma unroll
                for (int kb = 0; kb < 4; kb++)
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Wt[12 + tn * 4 + kb], a1[kb], acc, 0, 0, 0);
                a2[tn] = gelu_bf16x4(acc);
            }
            f32x4_t acc = B3;
#pragma unroll
            for (int kb = 0; kb < 4; kb++)
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Wt[28 + kb], a2[kb], acc, 0, 0, 0);
            Ru[gq] = acc[0];   // lanes q = 0: outputs 0 and 1 of sample r of group gq
            Rv[gq] = acc[1];
        }
        wave_lds_sync();   // the tile is rewritten by the next chunk
        float zu = 0.0f, zv = 0.0f;
#pragma unroll
        for (int gq = 0; gq < 4; gq++) {
            const float tu = __shfl(Ru[gq], r), tv = __shfl(Rv[gq], r);
            if (q == gq) zu = tu, zv = tv;
        }
        if (g < S2) {
            T su = a.sustr[g], sv = a.svstr[g];
            if (i >= 2 && i <= L - 1 && j >= 1 && j <= M - 1) {
                const float Rs = zu * a.sd_u + a.mu_u;
                const T Du = T(0.5) * ((a.h[g - 1] + a.z[g - 1]) + (a.h[g] + a.z[g]));
                su = su + T(Rs) * Du;
            }
            if (i >= 1 && i <= L - 1 && j >= 2 && j <= M - 1) {
                const float Rs = zv * a.sd_v + a.mu_v;
                const T Dv = T(0.5) * ((a.h[g - szI] + a.z[g - szI]) + (a.h[g] + a.z[g]));
                sv = sv + T(Rs) * Dv;
            }
            a.sustr_eff[g] = su;
            a.svstr_eff[g] = sv;
        }
    }
}

}  // namespace

// ------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------
struct sf_mlp {
    float *w, *m, *v, *grad, *partial, *ploss, *loss;
    uint16_t *wb;               // BF16 parameters, parameter order
    uint16_t *wpk;              // BF16 weights, inference operand order (NWPK)
    float *norm;                // mu[45], 1 / sd[45] (FP32)
    float ynorm[4];
    int max_blocks;
};

extern "C" int sf_mlp_num_params(void) { return NPARAM; }

extern "C" int sf_mlp_create(int32_t max_blocks, sf_mlp **out)
{
    SF_REQUIRE(out && max_blocks > 0, "sf_mlp_create: bad arguments");
    sf_mlp *p = new sf_mlp();
    *p = sf_mlp{};
    p->max_blocks = max_blocks;
    hipError_t r = hipMalloc((void **)&p->w, sizeof(float) * NPARAM);
    if (r == hipSuccess) r = hipMalloc((void **)&p->m, sizeof(float) * NPARAM);
    if (r == hipSuccess) r = hipMalloc((void **)&p->v, sizeof(float) * NPARAM);
    if (r == hipSuccess) r = hipMalloc((void **)&p->grad, sizeof(float) * NPARAM);
    if (r == hipSuccess) r = hipMalloc((void **)&p->partial, sizeof(float) * NPARAM * (size_t)max_blocks);
    if (r == hipSuccess) r = hipMalloc((void **)&p->ploss, sizeof(float) * (size_t)max_blocks);
    if (r == hipSuccess) r = hipMalloc((void **)&p->loss, sizeof(float));
    if (r == hipSuccess) r = hipMalloc((void **)&p->wb, sizeof(uint16_t) * NPARAM);
    if (r == hipSuccess) r = hipMalloc((void **)&p->wpk, sizeof(uint16_t) * NWPK);
    if (r == hipSuccess) r = hipMalloc((void **)&p->norm, sizeof(float) * (2 * NIN));
    if (r != hipSuccess) {
        (void)hipFree(p->w);
        (void)hipFree(p->m);
        (void)hipFree(p->v);
        (void)hipFree(p->grad);
        (void)hipFree(p->partial);
        (void)hipFree(p->ploss);
        (void)hipFree(p->loss);
        (void)hipFree(p->wb);
        (void)hipFree(p->wpk);
        (void)hipFree(p->norm);
        delete p;
        return sf::hip_fail(r, "sf_mlp_create");
    }
    *out = p;
    return SF_OK;
}

extern "C" int sf_mlp_destroy(sf_mlp *p)
{
    if (!p) return SF_OK;
    (void)hipFree(p->w);
    (void)hipFree(p->m);
    (void)hipFree(p->v);
    (void)hipFree(p->grad);
    (void)hipFree(p->partial);
    (void)hipFree(p->ploss);
    (void)hipFree(p->loss);
    (void)hipFree(p->wb);
    (void)hipFree(p->wpk);
    (void)hipFree(p->norm);
    delete p;
    return SF_OK;
}

namespace {
__global__ void k_to_bf16(const float *__restrict__ w, uint16_t *__restrict__ wb, uint16_t *__restrict__ wpk)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    wb[p] = f2bf(w[p]);
    const int slot = wpk_slot(p);
    if (slot >= 0) wpk[slot] = wb[p];
}
}  // namespace

/* Set the FP32 master parameters from host memory (NPARAM values); clears the Adam
 * moments and refreshes the BF16 copy. */
extern "C" int sf_mlp_set_params(sf_mlp *p, const float *host)
{
    SF_REQUIRE(p && host, "sf_mlp_set_params: null argument");
    hipError_t r = hipMemcpy(p->w, host, sizeof(float) * NPARAM, hipMemcpyHostToDevice);
    if (r == hipSuccess) r = hipMemset(p->m, 0, sizeof(float) * NPARAM);
    if (r == hipSuccess) r = hipMemset(p->v, 0, sizeof(float) * NPARAM);
    if (r != hipSuccess) return sf::hip_fail(r, "sf_mlp_set_params");
    hipLaunchKernelGGL(k_to_bf16, dim3((NPARAM + 255) / 256), dim3(256), 0, nullptr, p->w, p->wb, p->wpk);
    r = hipDeviceSynchronize();
    if (r != hipSuccess) return sf::hip_fail(r, "sf_mlp_set_params");
    return sf::check_launch("sf_mlp_set_params");
}

extern "C" int sf_mlp_get_params(sf_mlp *p, float *host)
{
    SF_REQUIRE(p && host, "sf_mlp_get_params: null argument");
    hipError_t r = hipMemcpy(host, p->w, sizeof(float) * NPARAM, hipMemcpyDeviceToHost);
    if (r != hipSuccess) return sf::hip_fail(r, "sf_mlp_get_params");
    return SF_OK;
}

extern "C" int sf_mlp_get_grad(sf_mlp *p, float *host, float *loss)
{
    SF_REQUIRE(p && host, "sf_mlp_get_grad: null argument");
    hipError_t r = hipMemcpy(host, p->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost);
    if (r == hipSuccess && loss) r = hipMemcpy(loss, p->loss, sizeof(float), hipMemcpyDeviceToHost);
    if (r != hipSuccess) return sf::hip_fail(r, "sf_mlp_get_grad");
    return SF_OK;
}

/* Input normalisation mu[45], sd[45] (features enter the network as (x - mu) * rsd with
 * rsd = 1 / sd rounded to FP32) and output de-normalisation (mu_u, sd_u, mu_v, sd_v). */
extern "C" int sf_mlp_set_norm(sf_mlp *p, const float *mu, const float *sd, const float *ynorm)
{
    SF_REQUIRE(p && mu && sd && ynorm, "sf_mlp_set_norm: null argument");
    float host[2 * NIN];
    for (int k = 0; k < NIN; k++) {
        SF_REQUIRE(sd[k] != 0.0f, "sf_mlp_set_norm: sd[%d] is zero", k);
        host[k] = mu[k];
        host[NIN + k] = 1.0f / sd[k];
    }
    hipError_t r = hipMemcpy(p->norm, host, sizeof(float) * 2 * NIN, hipMemcpyHostToDevice);
    if (r != hipSuccess) return sf::hip_fail(r, "sf_mlp_set_norm");
    for (int k = 0; k < 4; k++) p->ynorm[k] = ynorm[k];
    return SF_OK;
}

/* One training step on the batch idx[0 .. nblocks*ntile*64): forward + backward with
 * BF16 MFMA, deterministic reduction, then Adam (adam != 0; step counts from 1) or
 * gradient only (adam == 0; sf_mlp_get_grad reads it).  X: BF16 [n][48], Y: [n][2]. */
extern "C" int sf_mlp_train_step(sf_mlp *p, const uint16_t *X, const float *Y, const int32_t *idx, int32_t nblocks,
                                 int32_t ntile, int32_t adam, float lr, float beta1, float beta2, float eps,
                                 int32_t step, void *stream)
{
    SF_REQUIRE(p && X && Y && idx, "sf_mlp_train_step: null argument");
    SF_REQUIRE(nblocks > 0 && nblocks <= p->max_blocks && ntile > 0, "sf_mlp_train_step: %d blocks (max %d)",
               nblocks, p->max_blocks);
    SF_REQUIRE(!adam || step >= 1, "sf_mlp_train_step: Adam step %d", step);
    const long long batch = (long long)nblocks * ntile * BT;
    const float scale = 2.0f / (float)(batch * NOUT);
    hipStream_t s = sf::as_stream(stream);
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nblocks), dim3(NT), 0, s, p->wb, p->w, X, Y, idx, ntile, scale,
                       p->partial, p->ploss);
    const float bc1 = adam ? 1.0f - powf(beta1, (float)step) : 1.0f;
    const float bc2 = adam ? 1.0f - powf(beta2, (float)step) : 1.0f;
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((NPARAM + RP - 1) / RP), dim3(256), 0, s, p->partial, p->ploss,
                       nblocks, p->grad, p->w, p->m, p->v, p->wb, p->wpk, adam, lr, beta1, beta2, eps, bc1, bc2,
                       1.0f / (float)(batch * NOUT), p->loss);
    return sf::check_launch("sf_mlp_train_step");
}

/* Network outputs Y[n][2] of BF16 inputs X[n][48]. */
extern "C" int sf_mlp_forward(sf_mlp *p, const uint16_t *X, int32_t n, float *Y, void *stream)
{
    SF_REQUIRE(p && X && Y && n > 0, "sf_mlp_forward: bad arguments");
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + BT - 1) / BT), dim3(NT), 0, sf::as_stream(stream), p->wb, p->w, X,
                       n, Y);
    return sf::check_launch("sf_mlp_forward");
}

namespace {
template <typename T>
ClosureArgs<T> closure_args(sf_mlp *p, const sf_phys_fields *x, const void *f, void *sustr_eff, void *svstr_eff,
                            float *feat)
{
    ClosureArgs<T> a;
    a.u = (const T *)x->ubar[1];
    a.v = (const T *)x->vbar[1];
    a.z = (const T *)x->zeta[1];
    a.h = (const T *)x->h;
    a.f = (const T *)f;
    a.sustr = (const T *)x->sustr;
    a.svstr = (const T *)x->svstr;
    a.sustr_eff = (T *)sustr_eff;
    a.svstr_eff = (T *)svstr_eff;
    a.feat_out = feat;
    a.mu_u = p ? p->ynorm[0] : 0.f;
    a.sd_u = p ? p->ynorm[1] : 0.f;
    a.mu_v = p ? p->ynorm[2] : 0.f;
    a.sd_v = p ? p->ynorm[3] : 0.f;
    a.szI = x->szI;
    a.szJ = x->szJ;
    return a;
}
}  // namespace

/* Raw closure features of the time level t1 state of `x` (zeta[1], ubar[1], vbar[1], h)
 * and the Coriolis field f: out[S2][45] FP32, zero outside 1 <= i <= L-1, 1 <= j <= M-1. */
extern "C" int sf_closure_features(const sf_phys_fields *x, int32_t precision, const void *f, float *out,
                                   void *stream)
{
    SF_REQUIRE(x && f && out, "sf_closure_features: null argument");
    const long long n = (long long)x->szI * x->szJ * NIN;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (precision == SF_FP64)
        hipLaunchKernelGGL(k_closure_features<double>, grid, block, 0, sf::as_stream(stream),
                           closure_args<double>(nullptr, x, f, nullptr, nullptr, out));
    else
        hipLaunchKernelGGL(k_closure_features<float>, grid, block, 0, sf::as_stream(stream),
                           closure_args<float>(nullptr, x, f, nullptr, nullptr, out));
    return sf::check_launch("sf_closure_features");
}

/* Closure stresses of the t1 state: sustr_eff = sustr + R_u * D_u at interior u points,
 * svstr_eff likewise (R = network output, de-normalised); other points copy the
 * stress.  Pass sustr_eff/svstr_eff as the stresses of the next physics substep. */
extern "C" int sf_closure_stress(sf_mlp *p, const sf_phys_fields *x, int32_t precision, const void *f,
                                 void *sustr_eff, void *svstr_eff, void *stream)
{
    SF_REQUIRE(p && x && f && sustr_eff && svstr_eff, "sf_closure_stress: null argument");
    const long long S2 = (long long)x->szI * x->szJ;
    SF_REQUIRE(x->szI >= 3 && x->szJ >= 3 && S2 < (1LL << 30), "sf_closure_stress: grid %d x %d", x->szI, x->szJ);
    const long long nchunk = (S2 + 63) / 64;
    // persistent grid of 4-wave workgroups, one 64-point chunk per wave per pass; the
    // weights are staged through LDS once per workgroup
    const long long nblk = (nchunk + 3) / 4;
    const dim3 grid((unsigned)(nblk < 1024 ? nblk : 1024)), block(NT);
    const uint64_t *wpk = (const uint64_t *)p->wpk;
    if (precision == SF_FP64)
        hipLaunchKernelGGL(k_closure_stress<double>, grid, block, 0, sf::as_stream(stream), wpk, p->w, p->norm,
                           closure_args<double>(p, x, f, sustr_eff, svstr_eff, nullptr));
    else
        hipLaunchKernelGGL(k_closure_stress<float>, grid, block, 0, sf::as_stream(stream), wpk, p->w, p->norm,
                           closure_args<float>(p, x, f, sustr_eff, svstr_eff, nullptr));
    return sf::check_launch("sf_closure_stress");
}
