// step2d_kernels.hip -- barotropic LF-AM3 kernels, one launch per reference kernel.
//
// These are the drop-in replacements behind the raw handles that
// nonlinear/step2d.py launches (modules/mod_operators.py:141-148).  One thread per
// rho point, flat index f = j*szI + i (xi fastest) so that a wavefront reads 64
// consecutive doubles of every field.  Stencil formulas keep the reference's
// left-associative evaluation order (expression-template Eval order of
// modules/mod_cppkernels.h) and contraction is off, so FP64 results equal the CPU
// restatement bit for bit.
//
// Reads before the start of an array (row-0 RtoV stencils that the i%szJ mask of
// computeMomentumPred admits, mod_cppkernels.cpp:162) are defined as 0.0.
#include "sf_internal.h"

namespace {

struct Grid2 {
    int szI, szJ, q5_stride;
    long long S2;
    const double *on_u, *om_v, *pn, *pm;
};

__device__ __forceinline__ double ld0(const double *__restrict__ p, long long f)
{
    return f < 0 ? 0.0 : p[f];
}

__device__ __forceinline__ bool interior(int i, int j, int szI, int szJ)
{
    return i != 0 && i != szI - 1 && j != 0 && j != szJ - 1;
}

// computeZetaRHS (mod_cppkernels.cpp:80-128)
__global__ __launch_bounds__(256) void k_zeta_rhs(Grid2 g, const double *__restrict__ zeta,
                                                  const double *__restrict__ h,
                                                  const double *__restrict__ ubar,
                                                  const double *__restrict__ vbar,
                                                  double *__restrict__ DU_avg1,
                                                  double *__restrict__ DV_avg1,
                                                  double *__restrict__ Zt_avg1, int isPredictor,
                                                  double cff1, double *__restrict__ res)
{
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.S2) return;
    const int szI = g.szI;
    const int i = (int)(f % szI), j = (int)(f / szI);
    if (!interior(i, j, szI, g.szJ)) return;
    const double D = zeta[f] + h[f];
    const double Dl = zeta[f - 1] + h[f - 1];
    const double Dd = zeta[f - szI] + h[f - szI];
    const double Dr = zeta[f + 1] + h[f + 1];
    const double Du = zeta[f + szI] + h[f + szI];
    const double DU = ubar[f] * ((D + Dl) * 0.5);
    const double DUr = ubar[f + 1] * ((Dr + D) * 0.5);
    const double DV = vbar[f] * ((D + Dd) * 0.5);
    const double DVu = vbar[f + szI] * ((Du + D) * 0.5);
    res[f] = (((DUr * g.on_u[f + 1] - DU * g.on_u[f]) + (DVu * g.om_v[f + szI] - DV * g.om_v[f])) *
              g.pm[f]) *
             g.pn[f];
    if (!isPredictor) {
        Zt_avg1[f] = Zt_avg1[f] + cff1 * zeta[f];
        DU_avg1[f] = DU_avg1[f] + (cff1 * DU) * g.on_u[f];
        DV_avg1[f] = DV_avg1[f] + (cff1 * DV) * g.om_v[f];
    }
}

// computeZetaPred (mod_cppkernels.cpp:133-150)
__global__ __launch_bounds__(256) void k_zeta_pred(Grid2 g, double Dt,
                                                   const double *__restrict__ zeta_t0,
                                                   double *__restrict__ zeta_t2,
                                                   const double *__restrict__ rzeta_t1)
{
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.S2) return;
    if (f % g.szI == 0 || f / g.szI == 0) return;
    zeta_t2[f] = zeta_t0[f] + Dt * rzeta_t1[f];
}

// computeMomentumRHSPred/Corr (mod_cppkernels.cpp:12-77): gz = (1-w)*za + w*zb
__global__ __launch_bounds__(256) void k_momentum_rhs(Grid2 g, const double *__restrict__ h,
                                                      double *__restrict__ rhs_u,
                                                      double *__restrict__ rhs_v,
                                                      const double *__restrict__ za,
                                                      const double *__restrict__ zb, double gc,
                                                      double w)
{
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.S2) return;
    const int szI = g.szI;
    const int i = (int)(f % szI), j = (int)(f / szI);
    if (!interior(i, j, szI, g.szJ)) return;
    const double a = 1 - w;
    const double gz = a * za[f] + w * zb[f];
    const double gzl = a * za[f - 1] + w * zb[f - 1];
    const double gzd = a * za[f - szI] + w * zb[f - szI];
    const double hu = (h[f] + h[f - 1]) * 0.5;
    const double hv = (h[f] + h[f - szI]) * 0.5;
    rhs_u[f] = (gc * g.on_u[f]) * (hu * (gz - gzl) + 0.5 * (gz * gz - gzl * gzl));
    rhs_v[f] = (gc * g.om_v[f]) * (hv * (gz - gzd) + 0.5 * (gz * gz - gzd * gzd));
}

// computeMomentumPred (mod_cppkernels.cpp:155-186), i%szJ mask kept (Q5)
__global__ __launch_bounds__(256) void k_momentum_pred(Grid2 g, double Dt,
                                                       const double *__restrict__ u_t1,
                                                       double *__restrict__ u_t2,
                                                       const double *__restrict__ v_t1,
                                                       double *__restrict__ v_t2,
                                                       const double *__restrict__ rhsu,
                                                       const double *__restrict__ rhsv,
                                                       const double *__restrict__ h,
                                                       const double *__restrict__ zeta_t1,
                                                       const double *__restrict__ zeta_t2)
{
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.S2) return;
    const int szI = g.szI, ms = g.q5_stride;
    if (f % ms == 0 || f / ms == 0) return;
    const double D1 = zeta_t1[f] + h[f];
    const double D2 = zeta_t2[f] + h[f];
    const double D1l = ld0(zeta_t1, f - 1) + ld0(h, f - 1);
    const double D2l = ld0(zeta_t2, f - 1) + ld0(h, f - 1);
    const double D1d = ld0(zeta_t1, f - szI) + ld0(h, f - szI);
    const double D2d = ld0(zeta_t2, f - szI) + ld0(h, f - szI);
    const double D1U = (D1 + D1l) * 0.5, D2U = (D2 + D2l) * 0.5;
    const double D1V = (D1 + D1d) * 0.5, D2V = (D2 + D2d) * 0.5;
    u_t2[f] = (u_t1[f] * D1U + ((Dt * rhsu[f]) * g.pm[f]) * g.pn[f]) / D2U;
    v_t2[f] = (v_t1[f] * D1V + ((Dt * rhsv[f]) * g.pm[f]) * g.pn[f]) / D2V;
}

// AdamsMoultonCorr3rd (mod_cppkernels.cpp:190-209)
__global__ __launch_bounds__(256) void k_am3(long long S2, double Dt, double *__restrict__ v_t2,
                                             const double *__restrict__ r0,
                                             const double *__restrict__ r1,
                                             const double *__restrict__ r2)
{
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= S2) return;
    constexpr double AM3_2 = 5.0 / 12.0, AM3_1 = 8.0 / 12.0, AM3_0 = -1.0 / 12.0;
    v_t2[f] = v_t2[f] + Dt * (AM3_2 * r2[f] + AM3_1 * r1[f] + AM3_0 * r0[f]);
}

// AdamsMoultonCorr3rd2 (mod_cppkernels.cpp:213-251)
__global__ __launch_bounds__(256) void k_am3_2(Grid2 g, double Dt, double *__restrict__ u_t2,
                                               double *__restrict__ v_t2,
                                               const double *__restrict__ ru0,
                                               const double *__restrict__ ru1,
                                               const double *__restrict__ ru2,
                                               const double *__restrict__ rv0,
                                               const double *__restrict__ rv1,
                                               const double *__restrict__ rv2,
                                               const double *__restrict__ h,
                                               const double *__restrict__ zeta_t1,
                                               const double *__restrict__ zeta_t2)
{
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.S2) return;
    const int szI = g.szI;
    constexpr double AM3_2 = 5.0 / 12.0, AM3_1 = 8.0 / 12.0, AM3_0 = -1.0 / 12.0;
    const double D1 = zeta_t1[f] + h[f], D2 = zeta_t2[f] + h[f];
    const double D1l = ld0(zeta_t1, f - 1) + ld0(h, f - 1);
    const double D2l = ld0(zeta_t2, f - 1) + ld0(h, f - 1);
    const double D1d = ld0(zeta_t1, f - szI) + ld0(h, f - szI);
    const double D2d = ld0(zeta_t2, f - szI) + ld0(h, f - szI);
    const double D1U = (D1 + D1l) * 0.5, D2U = (D2 + D2l) * 0.5;
    const double D1V = (D1 + D1d) * 0.5, D2V = (D2 + D2d) * 0.5;
    const double pmn = (Dt * g.pm[f]) * g.pn[f];
    const double u0 = u_t2[f], v0 = v_t2[f];
    u_t2[f] = (u0 * D1U + pmn * (AM3_2 * ru2[f] + AM3_1 * ru1[f] + AM3_0 * ru0[f])) / D2U;
    v_t2[f] = (v0 * D1V + pmn * (AM3_2 * rv2[f] + AM3_1 * rv1[f] + AM3_0 * rv0[f])) / D2V;
}

// setBC (mod_BC.c:6-13)
__global__ __launch_bounds__(256) void k_bc_set(double *__restrict__ var,
                                                const int32_t *__restrict__ idx, int n,
                                                double val)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) var[idx[k]] = val;
}

// Composed boundary map: a single workgroup reads every source into registers,
// waits at a barrier, then writes every destination (dst[] is unique).
constexpr int kMapThreads = 1024;
constexpr int kMapPerThread = 16;

__global__ __launch_bounds__(kMapThreads) void k_bc_map(double *__restrict__ var,
                                                        const int32_t *__restrict__ dst,
                                                        const int32_t *__restrict__ src, int n,
                                                        double clamp_val)
{
    double v[kMapPerThread];
#pragma unroll
    for (int r = 0; r < kMapPerThread; r++) {
        const int k = r * kMapThreads + (int)threadIdx.x;
        double x = 0.0;
        if (k < n) {
            const int s = src[k];
            x = s >= 0 ? var[s] : (s == -2 ? clamp_val : 0.0);
        }
        v[r] = x;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kMapPerThread; r++) {
        const int k = r * kMapThreads + (int)threadIdx.x;
        if (k < n) var[dst[k]] = v[r];
    }
}

Grid2 grid2(const sf_ctx *c)
{
    return Grid2{c->szI, c->szJ, c->q5_stride, c->S2, c->on_u, c->om_v, c->pn, c->pm};
}

}  // namespace

#define SF_CTX_OK(c) SF_REQUIRE((c) != nullptr, "%s: null context", __func__)

extern "C" int sf_compute_zeta_rhs(sf_ctx *c, const double *zeta, const double *h,
                                   const double *ubar, const double *vbar, double *DU_avg1,
                                   double *DV_avg1, double *Zt_avg1, int32_t is_predictor,
                                   double cff1, double *res, void *stream)
{
    SF_CTX_OK(c);
    SF_REQUIRE(zeta && h && ubar && vbar && res, "sf_compute_zeta_rhs: null field");
    SF_REQUIRE(is_predictor || (DU_avg1 && DV_avg1 && Zt_avg1), "sf_compute_zeta_rhs: null average");
    hipLaunchKernelGGL(k_zeta_rhs, dim3(sf::grid1d(c->S2)), dim3(sf::kBlock1D), 0, sf::as_stream(stream),
                       grid2(c), zeta, h, ubar, vbar, DU_avg1, DV_avg1, Zt_avg1, (int)is_predictor, cff1, res);
    return sf::check_launch("sf_compute_zeta_rhs");
}

extern "C" int sf_compute_zeta_pred(sf_ctx *c, double Dt, const double *zeta_t0,
                                    const double *zeta_t1, double *zeta_t2,
                                    const double *rzeta_t1, void *stream)
{
    (void)zeta_t1;
    SF_CTX_OK(c);
    SF_REQUIRE(zeta_t0 && zeta_t2 && rzeta_t1, "sf_compute_zeta_pred: null field");
    hipLaunchKernelGGL(k_zeta_pred, dim3(sf::grid1d(c->S2)), dim3(sf::kBlock1D), 0, sf::as_stream(stream),
                       grid2(c), Dt, zeta_t0, zeta_t2, rzeta_t1);
    return sf::check_launch("sf_compute_zeta_pred");
}

extern "C" int sf_compute_momentum_rhs_pred(sf_ctx *c, const double *h, double *rhs_ubar,
                                            double *rhs_vbar, const double *zeta_t0,
                                            const double *zeta_t2, double gconst, double weight,
                                            void *stream)
{
    SF_CTX_OK(c);
    SF_REQUIRE(h && rhs_ubar && rhs_vbar && zeta_t0 && zeta_t2, "sf_compute_momentum_rhs_pred: null field");
    hipLaunchKernelGGL(k_momentum_rhs, dim3(sf::grid1d(c->S2)), dim3(sf::kBlock1D), 0, sf::as_stream(stream),
                       grid2(c), h, rhs_ubar, rhs_vbar, zeta_t2, zeta_t0, gconst, weight);
    return sf::check_launch("sf_compute_momentum_rhs_pred");
}

extern "C" int sf_compute_momentum_rhs_corr(sf_ctx *c, const double *h, double *rhs_ubar,
                                            double *rhs_vbar, const double *zeta_t0,
                                            const double *zeta_t1, const double *zeta_t2,
                                            double gconst, void *stream)
{
    (void)zeta_t0;
    SF_CTX_OK(c);
    SF_REQUIRE(h && rhs_ubar && rhs_vbar && zeta_t1 && zeta_t2, "sf_compute_momentum_rhs_corr: null field");
    hipLaunchKernelGGL(k_momentum_rhs, dim3(sf::grid1d(c->S2)), dim3(sf::kBlock1D), 0, sf::as_stream(stream),
                       grid2(c), h, rhs_ubar, rhs_vbar, zeta_t2, zeta_t1, gconst, 2.0 / 5.0);
    return sf::check_launch("sf_compute_momentum_rhs_corr");
}

extern "C" int sf_compute_momentum_pred(sf_ctx *c, double Dt, const double *u_t1, double *u_t2,
                                        const double *v_t1, double *v_t2, const double *rhsu,
                                        const double *rhsv, const double *h,
                                        const double *zeta_t1, const double *zeta_t2,
                                        void *stream)
{
    SF_CTX_OK(c);
    SF_REQUIRE(u_t1 && u_t2 && v_t1 && v_t2 && rhsu && rhsv && h && zeta_t1 && zeta_t2,
               "sf_compute_momentum_pred: null field");
    hipLaunchKernelGGL(k_momentum_pred, dim3(sf::grid1d(c->S2)), dim3(sf::kBlock1D), 0, sf::as_stream(stream),
                       grid2(c), Dt, u_t1, u_t2, v_t1, v_t2, rhsu, rhsv, h, zeta_t1, zeta_t2);
    return sf::check_launch("sf_compute_momentum_pred");
}

extern "C" int sf_adams_moulton_corr3rd(sf_ctx *c, double Dt, double *v_t2, const double *rhs_t0,
                                        const double *rhs_t1, const double *rhs_t2, void *stream)
{
    SF_CTX_OK(c);
    SF_REQUIRE(v_t2 && rhs_t0 && rhs_t1 && rhs_t2, "sf_adams_moulton_corr3rd: null field");
    hipLaunchKernelGGL(k_am3, dim3(sf::grid1d(c->S2)), dim3(sf::kBlock1D), 0, sf::as_stream(stream),
                       c->S2, Dt, v_t2, rhs_t0, rhs_t1, rhs_t2);
    return sf::check_launch("sf_adams_moulton_corr3rd");
}

extern "C" int sf_adams_moulton_corr3rd2(sf_ctx *c, double Dt, double *u_t2, double *v_t2,
                                         const double *ru0, const double *ru1, const double *ru2,
                                         const double *rv0, const double *rv1, const double *rv2,
                                         const double *h, const double *zeta_t1,
                                         const double *zeta_t2, void *stream)
{
    SF_CTX_OK(c);
    SF_REQUIRE(u_t2 && v_t2 && ru0 && ru1 && ru2 && rv0 && rv1 && rv2 && h && zeta_t1 && zeta_t2,
               "sf_adams_moulton_corr3rd2: null field");
    hipLaunchKernelGGL(k_am3_2, dim3(sf::grid1d(c->S2)), dim3(sf::kBlock1D), 0, sf::as_stream(stream),
                       grid2(c), Dt, u_t2, v_t2, ru0, ru1, ru2, rv0, rv1, rv2, h, zeta_t1, zeta_t2);
    return sf::check_launch("sf_adams_moulton_corr3rd2");
}

extern "C" int sf_bc_set(double *var, const int32_t *idx, int32_t n, double val, void *stream)
{
    SF_REQUIRE(n >= 0, "sf_bc_set: n=%d < 0", n);
    if (n == 0) return SF_OK;
    SF_REQUIRE(var && idx, "sf_bc_set: null pointer");
    hipLaunchKernelGGL(k_bc_set, dim3(sf::grid1d(n)), dim3(sf::kBlock1D), 0, sf::as_stream(stream), var, idx,
                       (int)n, val);
    return sf::check_launch("sf_bc_set");
}

extern "C" int sf_bc_apply_map(double *var, const int32_t *dst, const int32_t *src, int32_t n,
                               double clamp_val, void *stream)
{
    SF_REQUIRE(n >= 0 && n <= kMapThreads * kMapPerThread, "sf_bc_apply_map: n=%d outside [0, %d]", n,
               kMapThreads * kMapPerThread);
    if (n == 0) return SF_OK;
    SF_REQUIRE(var && dst && src, "sf_bc_apply_map: null pointer");
    hipLaunchKernelGGL(k_bc_map, dim3(1), dim3(kMapThreads), 0, sf::as_stream(stream), var, dst, src, (int)n,
                       clamp_val);
    return sf::check_launch("sf_bc_apply_map");
}
