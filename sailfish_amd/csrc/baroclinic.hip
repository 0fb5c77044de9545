// baroclinic.hip -- the baroclinic terms the reference names but leaves out: the linear
// equation of state (utility/test.in:2130-2160), the baroclinic pressure gradient
// (nonlinear/rhs.py:31-32, "prsgrd()" commented out) and the tracer step
// (nonlinear/step3d_t.py, Fortran text).  One thread per water column (pressure
// gradient, tracer) or per cell (state equation); flat layout of the 3D kernels (level
// stride S2 = szI*szJ, xi fastest).  FP64 instances follow the oracle
// (oracle/sf_oracle3d.c) operation by operation; FP32 bulk instances read and write FP32
// fields and form each column's pressure-gradient and volume-flux sums in FP64.
#include "sf_internal.h"

namespace {

struct Col {
    int szI, szJ, N;
    long long S2;
    const double *pm, *pn;
};

inline Col col_of(const sf_ctx *c) { return Col{c->szI, c->szJ, c->szK, c->S2, c->pm, c->pn}; }

// ---- linear equation of state: rho = ((R0 - 1000) - R0 Tcoef (T - T0)) + R0 Scoef (S - S0)
template <class T>
__global__ __launch_bounds__(256) void k_rho_eos(long long n, const T *t, const T *s, double a, double bt,
                                                 double bs, double T0, double S0, T *rho)
{
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    rho[b] = T((a - bt * ((double)t[b] - T0)) + bs * ((double)s[b] - S0));
}

// ---- baroclinic pressure gradient, standard density Jacobian (Song 1998), added to ru/rv.
// Per interior u (v) column with neighbour p = f-1 (f-szI), from the top level down:
//   phi(N-1) = fac1 drho (dz_top + dz_top_p) + (fac2 + fac1 (rho + rho_p)) (z_w(N) - z_w_p(N))
//   phi(k)   = phi(k+1) + fac3 (C1 C3 - C2 C4)       (gamma-weighted Jacobian of rho and z_r)
//   r(k)    += (-0.5 (Hz + Hz_p) phi(k)) * (2 / (pn + pn_p))      [v: 2 / (pm + pm_p)]
template <class T>
__device__ void pgf_column(const Col &g, bool vdir, long long f, const T *rho, const T *z_r, const T *z_w,
                           const T *Hz, double grav, double rho0, T *r)
{
    const long long S2 = g.S2, W = g.szI;
    const int N = g.N, i = (int)(f % W), j = (int)(f / W), L = g.szI - 1, M = g.szJ - 1;
    if (!vdir && !(i >= 2 && i <= L - 1 && j >= 1 && j <= M - 1)) return;
    if (vdir && !(j >= 2 && j <= M - 1 && i >= 1 && i <= L - 1)) return;
    const long long p = vdir ? f - W : f - 1;
    const double met = vdir ? 2.0 / (g.pm[p] + g.pm[f]) : 2.0 / (g.pn[p] + g.pn[f]);
    const double fac1 = 0.5 * grav / rho0, fac2 = 1000.0 * grav / rho0, fac3 = 0.25 * grav / rho0;
#define R(k, x) ((double)rho[(long long)(k) * S2 + (x)])
#define ZR(k, x) ((double)z_r[(long long)(k) * S2 + (x)])
#define ZW(k, x) ((double)z_w[(long long)(k) * S2 + (x)])
#define HZ(k, x) ((double)Hz[(long long)(k) * S2 + (x)])
    int k = N - 1;
    const double c1 = (ZW(N, f) - ZR(k, f)) + (ZW(N, p) - ZR(k, p));
    double phi = (fac1 * (R(k, f) - R(k, p))) * c1;
    phi = phi + (fac2 + fac1 * (R(k, f) + R(k, p))) * (ZW(N, f) - ZW(N, p));
    r[(long long)k * S2 + f] = T((double)r[(long long)k * S2 + f] + ((-0.5 * (HZ(k, f) + HZ(k, p))) * phi) * met);
    for (k = N - 2; k >= 0; k--) {
        const double dzf = ZR(k + 1, f) - ZR(k, f), dzp = ZR(k + 1, p) - ZR(k, p);
        const double a2 = (ZR(k, f) - ZR(k, p)) + (ZR(k + 1, f) - ZR(k + 1, p));
        const double a3 = dzf - dzp;
        const double gam = ((0.125 * a2) * a3) / (dzf * dzp);
        const double C1 = (1.0 + gam) * (R(k + 1, f) - R(k + 1, p)) + (1.0 - gam) * (R(k, f) - R(k, p));
        const double C2 = (R(k + 1, f) + R(k + 1, p)) - (R(k, f) + R(k, p));
        const double C3 = (ZR(k + 1, f) + ZR(k + 1, p)) - (ZR(k, f) + ZR(k, p));
        const double C4 = (1.0 + gam) * (ZR(k + 1, f) - ZR(k + 1, p)) + (1.0 - gam) * (ZR(k, f) - ZR(k, p));
        phi = phi + fac3 * (C1 * C3 - C2 * C4);
        r[(long long)k * S2 + f] =
            T((double)r[(long long)k * S2 + f] + ((-0.5 * (HZ(k, f) + HZ(k, p))) * phi) * met);
    }
#undef R
#undef ZR
#undef ZW
#undef HZ
}

template <class T>
__global__ __launch_bounds__(256) void k_prsgrd(Col g, const T *rho, const T *z_r, const T *z_w, const T *Hz,
                                                double grav, double rho0, T *ru, T *rv)
{
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.S2) return;
    pgf_column(g, false, f, rho, z_r, z_w, Hz, grav, rho0, ru);
    pgf_column(g, true, f, rho, z_r, z_w, Hz, grav, rho0, rv);
}

// ---- tracer step: flux form, first-order upstream faces, forward in time, with the
// vertical volume flux of each cell's top formed from the horizontal volume fluxes and
// the thickness change (a uniform tracer stays uniform).  Interior columns first; a
// second launch copies them into the ghost columns (zero gradient).
template <class T>
__global__ __launch_bounds__(256) void k_step3d_t(Col g, const T *t, T *tn, const T *Huon, const T *Hvom,
                                                  const T *Hz_old, const T *Hz_new, double dt)
{
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= g.S2) return;
    const long long S2 = g.S2, W = g.szI;
    const int N = g.N, i = (int)(f % W), j = (int)(f / W), L = g.szI - 1, M = g.szJ - 1;
    if (!(i >= 1 && i <= L - 1 && j >= 1 && j <= M - 1)) return;
    const double mn = g.pm[f] * g.pn[f];
    double Wlow = 0.0, FClow = 0.0;
    for (int k = 0; k < N; k++) {
        const long long b = (long long)k * S2 + f;
        const double hu0 = Huon[b], hu1 = Huon[b + 1], hv0 = Hvom[b], hv1 = Hvom[b + W];
        const double Dh = mn * ((hu1 - hu0) + (hv1 - hv0));
        const double hzo = Hz_old[b], hzn = Hz_new[b];
        const double Wk = (Wlow - Dh) - (hzn - hzo) / dt;
        const double tc = t[b];
        const double FX0 = hu0 * (hu0 > 0.0 ? (double)t[b - 1] : tc);
        const double FX1 = hu1 * (hu1 > 0.0 ? tc : (double)t[b + 1]);
        const double FE0 = hv0 * (hv0 > 0.0 ? (double)t[b - W] : tc);
        const double FE1 = hv1 * (hv1 > 0.0 ? tc : (double)t[b + W]);
        const double FCk = k < N - 1 ? Wk * (Wk > 0.0 ? tc : (double)t[b + S2]) : Wk * tc;
        const double div = mn * ((FX1 - FX0) + (FE1 - FE0)) + (FCk - FClow);
        tn[b] = T((hzo * tc - dt * div) / hzn);
        Wlow = Wk;
        FClow = FCk;
    }
}

// ghost rows j = 0, M (interior i), then ghost columns i = 0, L (all j, clamped source row)
template <class T>
__global__ __launch_bounds__(256) void k_tracer_ghost_rows(Col g, T *tn)
{
    const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int k = blockIdx.y;
    const long long S2 = g.S2, W = g.szI;
    const int L = g.szI - 1, M = g.szJ - 1;
    if (i > L - 1) return;
    tn[(long long)k * S2 + i] = tn[(long long)k * S2 + W + i];
    tn[(long long)k * S2 + (long long)M * W + i] = tn[(long long)k * S2 + (long long)(M - 1) * W + i];
}

template <class T>
__global__ __launch_bounds__(256) void k_tracer_ghost_cols(Col g, T *tn)
{
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int k = blockIdx.y;
    const long long S2 = g.S2, W = g.szI;
    const int L = g.szI - 1, M = g.szJ - 1;
    if (j > M) return;
    const int jj = j < 1 ? 1 : (j > M - 1 ? M - 1 : j);
    tn[(long long)k * S2 + (long long)j * W] = tn[(long long)k * S2 + (long long)jj * W + 1];
    tn[(long long)k * S2 + (long long)j * W + L] = tn[(long long)k * S2 + (long long)jj * W + L - 1];
}

template <class T>
int eos_launch(sf_ctx *c, const void *t, const void *s, double R0, double T0, double S0, double Tcoef,
               double Scoef, void *rho, void *stream)
{
    const long long n = (long long)c->szK * c->S2;
    hipLaunchKernelGGL(k_rho_eos<T>, dim3(sf::grid1d(n)), dim3(256), 0, sf::as_stream(stream), n, (const T *)t,
                       (const T *)s, R0 - 1000.0, R0 * Tcoef, R0 * Scoef, T0, S0, (T *)rho);
    return sf::check_launch("sf_rho_eos");
}

template <class T>
int prsgrd_launch(sf_ctx *c, const void *rho, const void *z_r, const void *z_w, const void *Hz, double g,
                  double rho0, void *ru, void *rv, void *stream)
{
    hipLaunchKernelGGL(k_prsgrd<T>, dim3(sf::grid1d(c->S2)), dim3(256), 0, sf::as_stream(stream), col_of(c),
                       (const T *)rho, (const T *)z_r, (const T *)z_w, (const T *)Hz, g, rho0, (T *)ru, (T *)rv);
    return sf::check_launch("sf_prsgrd");
}

template <class T>
int tracer_launch(sf_ctx *c, const void *t, void *tn, const void *Huon, const void *Hvom, const void *Hz_old,
                  const void *Hz_new, double dt, void *stream)
{
    hipStream_t s = sf::as_stream(stream);
    const Col g = col_of(c);
    hipLaunchKernelGGL(k_step3d_t<T>, dim3(sf::grid1d(c->S2)), dim3(256), 0, s, g, (const T *)t, (T *)tn,
                       (const T *)Huon, (const T *)Hvom, (const T *)Hz_old, (const T *)Hz_new, dt);
    hipLaunchKernelGGL(k_tracer_ghost_rows<T>, dim3(sf::grid1d(c->szI), c->szK), dim3(256), 0, s, g, (T *)tn);
    hipLaunchKernelGGL(k_tracer_ghost_cols<T>, dim3(sf::grid1d(c->szJ), c->szK), dim3(256), 0, s, g, (T *)tn);
    return sf::check_launch("sf_step3d_t");
}

}  // namespace

extern "C" int sf_rho_eos(sf_ctx *c, int32_t precision, const void *t, const void *s, double R0, double T0,
                          double S0, double Tcoef, double Scoef, void *rho, void *stream)
{
    SF_REQUIRE(c && t && s && rho, "sf_rho_eos: null argument");
    if (precision == SF_FP64) return eos_launch<double>(c, t, s, R0, T0, S0, Tcoef, Scoef, rho, stream);
    if (precision == SF_FP32) return eos_launch<float>(c, t, s, R0, T0, S0, Tcoef, Scoef, rho, stream);
    sf::set_error("sf_rho_eos: unknown precision %d", precision);
    return SF_ERR_ARG;
}

extern "C" int sf_prsgrd(sf_ctx *c, int32_t precision, const void *rho, const void *z_r, const void *z_w,
                         const void *Hz, double g, double rho0, void *ru, void *rv, void *stream)
{
    SF_REQUIRE(c && rho && z_r && z_w && Hz && ru && rv, "sf_prsgrd: null argument");
    SF_REQUIRE(c->szK >= 2, "sf_prsgrd: needs at least 2 levels (N=%d)", c->szK);
    if (precision == SF_FP64) return prsgrd_launch<double>(c, rho, z_r, z_w, Hz, g, rho0, ru, rv, stream);
    if (precision == SF_FP32) return prsgrd_launch<float>(c, rho, z_r, z_w, Hz, g, rho0, ru, rv, stream);
    sf::set_error("sf_prsgrd: unknown precision %d", precision);
    return SF_ERR_ARG;
}

extern "C" int sf_step3d_t(sf_ctx *c, int32_t precision, const void *t, void *tn, const void *Huon,
                           const void *Hvom, const void *Hz_old, const void *Hz_new, double dt, void *stream)
{
    SF_REQUIRE(c && t && tn && Huon && Hvom && Hz_old && Hz_new, "sf_step3d_t: null argument");
    SF_REQUIRE(t != tn, "sf_step3d_t: the new tracer must not alias the old one");
    SF_REQUIRE(dt > 0.0, "sf_step3d_t: dt=%g", dt);
    if (precision == SF_FP64) return tracer_launch<double>(c, t, tn, Huon, Hvom, Hz_old, Hz_new, dt, stream);
    if (precision == SF_FP32) return tracer_launch<float>(c, t, tn, Huon, Hvom, Hz_old, Hz_new, dt, stream);
    sf::set_error("sf_step3d_t: unknown precision %d", precision);
    return SF_ERR_ARG;
}
