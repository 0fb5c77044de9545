// mlp.hip -- training of the closure network (forward + backward on the matrix cores,
// deterministic block reduction, Adam) and its parameter management.  Network, layouts
// and rounding: mlp_net.h.
#include "mlp_net.h"

#include <math.h>

namespace {

// Forward + backward over ntile groups of 16 samples per wave: wave w of block b takes
// samples idx[(b * ntile + t) * 64 + 16 w + r], t = 0..ntile-1.  partial[b][NPARAM]: the
// block's gradient sum (its waves' sums added in wave order); partial_loss[b]: its sum of
// squared errors.  scale = 2 / (batch * NOUT).
__global__ __launch_bounds__(NT) void k_mlp_fwd_bwd(const uint16_t *__restrict__ wpk, const uint16_t *__restrict__ wb,
                                                   const float *__restrict__ w, const uint16_t *__restrict__ X,
                                                   const float *__restrict__ Y, const int32_t *__restrict__ idx,
                                                   int ntile, float scale, float *__restrict__ partial,
                                                   float *__restrict__ partial_loss)
{
    __shared__ float red[NT / 64][H * H];   // per-wave gradient tiles for the block sum
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    Fwd W;
    load_fwd(W, wpk, w);
    // backward operands: rows of W3 (16x16x16 over the outputs) and of W2 (16x16x32 over
    // the layer-2 units)
    bf16x4_t w3r[4];
    bf16x8_t w2r[4][2];
#pragma unroll
    for (int tm = 0; tm < 4; tm++) {
        w3r[tm] = *(const bf16x4_t *)(wb + OFF_W3 + (16 * tm + r) * OUT + 4 * q);
#pragma unroll
        for (int kb = 0; kb < 2; kb++) {
            const uint16_t *p = wb + OFF_W2 + (16 * tm + r) * H + 32 * kb + 4 * q;
            w2r[tm][kb] = cat8(*(const u32x2_t *)p, *(const u32x2_t *)(p + 16));
        }
    }
    // identity as a 16x16x16 B operand: element j of lane (r, q) is 1 where 4q + j = r
    const uint32_t one = 0x3F80u;
    const bf16x4_t I = as4((u32x2_t){(4 * q == r ? one : 0u) | (4 * q + 1 == r ? one << 16 : 0u),
                                     (4 * q + 2 == r ? one : 0u) | (4 * q + 3 == r ? one << 16 : 0u)});
    const f32x4_t z4 = {0.f, 0.f, 0.f, 0.f};
    f32x4_t gW1[12], gW2[16], gW3[4], gb1[4], gb2[4];
#pragma unroll
    for (int e = 0; e < 12; e++) gW1[e] = z4;
#pragma unroll
    for (int e = 0; e < 16; e++) gW2[e] = z4;
#pragma unroll
    for (int e = 0; e < 4; e++) gW3[e] = gb1[e] = gb2[e] = z4;
    float gb3a = 0.0f, gb3b = 0.0f, sse = 0.0f;

    // the next group's operands are loaded while the current group computes
    const int32_t *ib = idx + (long long)blockIdx.x * ntile * BT + 16 * wave + r;
    int s = ib[0];
    bf16x8_t xa;
    bf16x4_t xb;
    load_x(X + (long long)s * IN, xa, xb);
    float y0 = Y[2 * (long long)s], y1 = Y[2 * (long long)s + 1];
    int s_next = ntile > 1 ? ib[BT] : s;
    for (int t = 0; t < ntile; t++) {
        bf16x8_t nxa;
        bf16x4_t nxb;
        load_x(X + (long long)s_next * IN, nxa, nxb);
        const float ny0 = Y[2 * (long long)s_next], ny1 = Y[2 * (long long)s_next + 1];
        const int s_after = t + 2 < ntile ? ib[(t + 2) * BT] : s_next;

        Acts A;
        forward16(W, xa, xb, A);
        // output error: lanes q = 0 hold outputs 0 and 1 of sample r
        const float e0 = q == 0 ? A.z3[0] - y0 : 0.0f, e1 = q == 0 ? A.z3[1] - y1 : 0.0f;
        sse += e0 * e0;
        sse += e1 * e1;
        const float d0 = scale * e0, d1 = scale * e1;
        gb3a += d0;
        gb3b += d1;
        const u32x2_t dz3 = {pack_bf16x2(d0, d1), 0u};
        // dZ2 = (W3 dZ3^T) * [A2 > 0]
        u32x2_t dz2[4];
#pragma unroll
        for (int tm = 0; tm < 4; tm++) {
            const f32x4_t d = relu_mask(mma16(w3r[tm], as4(dz3), z4), A.a2[tm]);
            gb2[tm] += d;
            dz2[tm] = pack4(d);
        }
        // dZ1 = (W2 dZ2^T) * [A1 > 0]
        const bf16x8_t g0 = cat8(dz2[0], dz2[1]), g1 = cat8(dz2[2], dz2[3]);
        u32x2_t dz1[4];
#pragma unroll
        for (int tm = 0; tm < 4; tm++) {
            const f32x4_t d = relu_mask(mma32(w2r[tm][1], g1, mma32(w2r[tm][0], g0, z4)), A.a1[tm]);
            gb1[tm] += d;
            dz1[tm] = pack4(d);
        }
        // weight gradients: sums over the group's samples, operands moved to the sample-
        // row layout by the identity product
        auto tr = [&](u32x2_t v) { return as4(pack4(mma16(as4(v), I, z4))); };
        const bf16x4_t t3 = tr(dz3);
#pragma unroll
        for (int tm = 0; tm < 4; tm++) gW3[tm] = mma16(tr(A.a2[tm]), t3, gW3[tm]);
        bf16x4_t td[4];
#pragma unroll
        for (int tn = 0; tn < 4; tn++) td[tn] = tr(dz2[tn]);
#pragma unroll
        for (int tm = 0; tm < 4; tm++) {
            const bf16x4_t ta = tr(A.a1[tm]);
#pragma unroll
            for (int tn = 0; tn < 4; tn++) gW2[tm * 4 + tn] = mma16(ta, td[tn], gW2[tm * 4 + tn]);
        }
#pragma unroll
        for (int tn = 0; tn < 4; tn++) td[tn] = tr(dz1[tn]);
        const u32x4_t xw = __builtin_bit_cast(u32x4_t, xa);
        const u32x2_t xk[3] = {(u32x2_t){xw[0], xw[1]}, (u32x2_t){xw[2], xw[3]}, __builtin_bit_cast(u32x2_t, xb)};
#pragma unroll
        for (int tk = 0; tk < 3; tk++) {
            const bf16x4_t tx = tr(xk[tk]);
#pragma unroll
            for (int tn = 0; tn < 4; tn++) gW1[tk * 4 + tn] = mma16(tx, td[tn], gW1[tk * 4 + tn]);
        }
        xa = nxa;
        xb = nxb;
        y0 = ny0;
        y1 = ny1;
        s_next = s_after;
    }

    // block sums: each wave stores its tiles (lane (r, q) holds rows 16 tm + 4q + t, column
    // 16 tn + r), then thread e adds entry e over the waves in order
    float *P = partial + (long long)blockIdx.x * NPARAM;
    float *R = red[wave];
    auto block_sum = [&](int n, int off) {
        for (int e = threadIdx.x; e < n; e += NT) P[off + e] = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
    };
#pragma unroll
    for (int tk = 0; tk < 3; tk++)
#pragma unroll
        for (int tn = 0; tn < 4; tn++)
#pragma unroll
            for (int t = 0; t < 4; t++) R[(16 * tk + 4 * q + t) * H + 16 * tn + r] = gW1[tk * 4 + tn][t];
    __syncthreads();
    block_sum(IN * H, OFF_W1);
    __syncthreads();
#pragma unroll
    for (int tm = 0; tm < 4; tm++)
#pragma unroll
        for (int tn = 0; tn < 4; tn++)
#pragma unroll
            for (int t = 0; t < 4; t++) R[(16 * tm + 4 * q + t) * H + 16 * tn + r] = gW2[tm * 4 + tn][t];
    __syncthreads();
    block_sum(H * H, OFF_W2);
    __syncthreads();
    // W3 gradients at 0..1023; bias partials, 16 per unit and wave (lane (r, q) holds unit
    // 16 tm + 4q + t of sample r): b1 at 1024 + r * 64 + unit, b2 at 2048 + r * 64 + unit;
    // b3 at 3072 + 2 r + output; squared errors at 3104 + lane
#pragma unroll
    for (int tm = 0; tm < 4; tm++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            R[(16 * tm + 4 * q + t) * OUT + r] = gW3[tm][t];
            R[1024 + r * H + 16 * tm + 4 * q + t] = gb1[tm][t];
            R[2048 + r * H + 16 * tm + 4 * q + t] = gb2[tm][t];
        }
    if (q == 0) {
        R[3072 + 2 * r] = gb3a;
        R[3072 + 2 * r + 1] = gb3b;
    }
    R[3104 + lane] = sse;
    __syncthreads();
    block_sum(H * OUT, OFF_W3);
    // 16 partials per wave, waves in order
    auto sum16 = [&](int base, int stride) {
        float acc = 0.0f;
#pragma unroll
        for (int wv = 0; wv < NT / 64; wv++) {
            float v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) v[u] = red[wv][base + u * stride];
#pragma unroll
            for (int u = 0; u < 16; u++) acc += v[u];
        }
        return acc;
    };
    const int e = threadIdx.x;
    if (e < H) {
        P[OFF_B1 + e] = sum16(1024 + e, H);
    } else if (e < 2 * H) {
        P[OFF_B2 + e - H] = sum16(2048 + e - H, H);
    } else if (e < 2 * H + OUT) {
        const int o = e - 2 * H;
        P[OFF_B3 + o] = o < NOUT ? sum16(3072 + o, 2) : 0.0f;
    } else if (e >= NT - 4) {   // squared errors: four partial sums of 64, then in order
        const int part = e - (NT - 4);
        float acc = sum16(3104 + 16 * part, 1);   // lanes 16 part .. 16 part + 15 of each wave
        red[0][4000 + part] = acc;
    }
    __syncthreads();
    if (e == NT - 1)
        partial_loss[blockIdx.x] = ((red[0][4000] + red[0][4001]) + red[0][4002]) + red[0][4003];
}

// grad[p] = sum over the blocks' partials; optional Adam step on the FP32 master and
// refresh of the BF16 copies; loss_out = sum of squared errors / (batch * NOUT).
// A workgroup of 1024 threads takes 64 parameters; its 16 waves each sum the partials of
// every 16th block (wave w: blocks w, w + 16, ... in increasing order, 16 loads in flight),
// and the 16 sums are added in wave order: a fixed order for any grid, so the result is
// deterministic.  Reads are coalesced (64 consecutive parameters of one block per wave
// instruction).
constexpr int RP = 64, RS = 16;   // parameters and block slices per reduce workgroup

// src[b * stride] summed over b = b0, b0 + db, ... < nb in that order, 16 loads in flight
__device__ __forceinline__ float sum_strided(const float *__restrict__ src, long long stride, int b0, int db, int nb)
{
    float acc = 0.0f;
    int b = b0;
    for (; b + 15 * db < nb; b += 16 * db) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = src[(long long)(b + u * db) * stride];
#pragma unroll
        for (int u = 0; u < 16; u++) acc += v[u];
    }
    for (; b < nb; b += db) acc += src[(long long)b * stride];
    return acc;
}

__global__ __launch_bounds__(RP * RS) void k_mlp_reduce_adam(const float *__restrict__ partial,
                                                            const float *__restrict__ partial_loss, int nblk,
                                                            float *__restrict__ grad, float *__restrict__ w,
                                                            float *__restrict__ m, float *__restrict__ v,
                                                            uint16_t *__restrict__ wb, uint16_t *__restrict__ wpk,
                                                            int adam, float lr, float beta1, float beta2, float eps,
                                                            float bc1, float bc2, float inv_count,
                                                            float *__restrict__ loss_out)
{
    __shared__ float red[RS][RP];
    const int lane = threadIdx.x & 63, slice = threadIdx.x >> 6;
    const int p = blockIdx.x * RP + lane;
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss_out) loss_out[0] = sum_strided(partial_loss, 1, 0, 1, nblk) * inv_count;
    red[slice][lane] = p < NPARAM ? sum_strided(partial + p, NPARAM, slice, RS, nblk) : 0.0f;
    __syncthreads();
    if (slice != 0 || p >= NPARAM) return;
    float g = red[0][lane];
#pragma unroll
    for (int s = 1; s < RS; s++) g += red[s][lane];
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

}  // namespace

// ------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------

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
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nblocks), dim3(NT), 0, s, p->wpk, p->wb, p->w, X, Y, idx, ntile, scale,
                       p->partial, p->ploss);
    const float bc1 = adam ? 1.0f - powf(beta1, (float)step) : 1.0f;
    const float bc2 = adam ? 1.0f - powf(beta2, (float)step) : 1.0f;
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((NPARAM + RP - 1) / RP), dim3(RP * RS), 0, s, p->partial, p->ploss,
                       nblocks, p->grad, p->w, p->m, p->v, p->wb, p->wpk, adam, lr, beta1, beta2, eps, bc1, bc2,
                       1.0f / (float)(batch * NOUT), p->loss);
    return sf::check_launch("sf_mlp_train_step");
}

