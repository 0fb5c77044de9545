// closure_infer.hip -- inference of the closure network: outputs of packed samples, the
// closure features of a state and the closure stresses.  Network, layouts and rounding:
// mlp_net.h.
#include "mlp_net.h"

namespace {

// outputs Y[n][2] of BF16 inputs X[n][48]; each wave takes 16-sample groups in turn
__global__ __launch_bounds__(NT) void k_mlp_forward(const uint16_t *__restrict__ wpk, const float *__restrict__ w,
                                                   const uint16_t *__restrict__ X, int n, float *__restrict__ Y)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    Fwd W;
    load_fwd(W, wpk, w);
    const int ngroup = (n + 15) / 16;
    for (int gi = blockIdx.x * (NT / 64) + wave; gi < ngroup; gi += gridDim.x * (NT / 64)) {
        const int s = gi * 16 + r < n ? gi * 16 + r : n - 1;
        bf16x8_t xa;
        bf16x4_t xb;
        load_x(X + (long long)s * IN, xa, xb);
        Acts A;
        forward16(W, xa, xb, A);
        if (q == 0 && gi * 16 + r < n) *(f32x2_t *)(Y + 2 * (long long)s) = (f32x2_t){A.z3[0], A.z3[1]};
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
// normalised as (x - mu) * rsd, rsd = 1 / sd rounded to FP32, and rounded to BF16.  The
// network runs as in the fused substep (closure_eval: the four groups in two pairs, waits
// around each layer's products).
// norm = {mu[45], rsd[45]}; wpk: the packed BF16 weights (NWPK values).

// element e of field base at byte offset off (< 2^32: the launcher checks the grid size),
// addressed as a 64-bit scalar base plus a 32-bit vector offset
template <typename T>
__device__ __forceinline__ T at(const T *base, uint32_t off)
{
    return *(const T *)((const char *)base + off);
}

template <typename T>
__global__ __launch_bounds__(NT) void k_closure_stress(const uint16_t *__restrict__ wpk, const float *__restrict__ w,
                                                      const float *__restrict__ norm, ClosureArgs<T> a)
{
    __shared__ __attribute__((aligned(16))) uint32_t Xl[NT / 64][64 * XS];
    __shared__ float nrml[NT / 64][2 * NIN];   // per wave: mu, then rsd (uniform LDS reads)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float *nrm = nrml[wave];
    nrm[lane] = norm[lane];
    if (lane < 2 * NIN - 64) nrm[64 + lane] = norm[64 + lane];
    Fwd W;
    load_fwd(W, wpk, w);
    uint32_t *X = Xl[wave];
    wave_lds_sync();

    const int S2 = a.szI * a.szJ;
    const int szI = a.szI, L = a.szI - 1, M = a.szJ - 1;
    const int nchunk = (S2 + 63) / 64;
    const T *fld[5] = {a.u, a.v, a.z, a.h, a.f};
    constexpr uint32_t E = sizeof(T);
    for (int c = blockIdx.x * (NT / 64) + wave; c < nchunk; c += gridDim.x * (NT / 64)) {
        const int g = c * 64 + lane;
        const int j = g / szI, i = g - j * szI;
        const bool in = g < S2 && i >= 1 && i <= L - 1 && j >= 1 && j <= M - 1;
        const uint32_t oc = (uint32_t)(in ? g : szI + 1) * E;   // other lanes read an interior point
        // the stress update's inputs, loaded ahead of the network
        const bool ou = g < S2 && i >= 2 && i <= L - 1 && j >= 1 && j <= M - 1;
        const bool ov = g < S2 && i >= 1 && i <= L - 1 && j >= 2 && j <= M - 1;
        const uint32_t og = (uint32_t)g * E, row = (uint32_t)szI * E;
        T su = T(0), sv = T(0), Du = T(0), Dv = T(0);
        if (g < S2) {
            su = at(a.sustr, og);
            sv = at(a.svstr, og);
        }
        if (ou) Du = T(0.5) * ((at(a.h, og - E) + at(a.z, og - E)) + (at(a.h, og) + at(a.z, og)));
        if (ov) Dv = T(0.5) * ((at(a.h, og - row) + at(a.z, og - row)) + (at(a.h, og) + at(a.z, og)));
        // features of the lane's point -> row `lane` of the wave's tile
        float xv[NIN];
#pragma unroll
        for (int k = 0; k < NIN; k++) {
            const int f = k / 9, dj = (k % 9) / 3 - 1, di = k % 3 - 1;
            const T cv = at(fld[f], oc);
            xv[k] = (dj == 0 && di == 0) ? (float)cv : (float)(at(fld[f], oc + dj * row + di * E) - cv);
        }
        closure_row(X, lane, xv, nrm);
        float zu, zv;
        closure_eval(W, X, zu, zv);
        if (g < S2) {
            if (ou) su = su + T(zu * a.sd_u + a.mu_u) * Du;
            if (ov) sv = sv + T(zv * a.sd_v + a.mu_v) * Dv;
            a.sustr_eff[g] = su;
            a.svstr_eff[g] = sv;
        }
    }
}

}  // namespace

/* Network outputs Y[n][2] of BF16 inputs X[n][48]. */
extern "C" int sf_mlp_forward(sf_mlp *p, const uint16_t *X, int32_t n, float *Y, void *stream)
{
    SF_REQUIRE(p && X && Y && n > 0, "sf_mlp_forward: bad arguments");
    const int nblk = (n + BT - 1) / BT;
    hipLaunchKernelGGL(k_mlp_forward, dim3(nblk < 2048 ? nblk : 2048), dim3(NT), 0, sf::as_stream(stream), p->wpk, p->w,
                       X, n, Y);
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
    // 32-bit byte offsets into the fields
    const long long esz = precision == SF_FP64 ? 8 : 4;
    SF_REQUIRE(x->szI >= 3 && x->szJ >= 3 && S2 * esz < (1LL << 32), "sf_closure_stress: grid %d x %d", x->szI,
               x->szJ);
    const long long nchunk = (S2 + 63) / 64;
    // persistent grid of 4-wave workgroups, one 64-point chunk per wave per pass; each
    // wave loads the network into its registers once
    const long long nblk = (nchunk + 3) / 4;
    const dim3 grid((unsigned)(nblk < 1024 ? nblk : 1024)), block(NT);
    const uint16_t *wpk = p->wpk;
    if (precision == SF_FP64)
        hipLaunchKernelGGL(k_closure_stress<double>, grid, block, 0, sf::as_stream(stream), wpk, p->w, p->norm,
                           closure_args<double>(p, x, f, sustr_eff, svstr_eff, nullptr));
    else
        hipLaunchKernelGGL(k_closure_stress<float>, grid, block, 0, sf::as_stream(stream), wpk, p->w, p->norm,
                           closure_args<float>(p, x, f, sustr_eff, svstr_eff, nullptr));
    return sf::check_launch("sf_closure_stress");
}

