// phys2d_kernels.hip -- barotropic LF-AM3 substep of ROMS in its 2D form ("physics
// mode"): the terms the live Sailfish kernels leave out (2D advection, Coriolis,
// wind stress, bottom drag) and proper leapfrog/AM3 time levels, restated from the
// Fortran text in obsolete/step2d_LF_AM3.h.py (free surface :917-1014, pressure
// gradient :1107-1237, C2 advection :1244-1290 + :1497-1502, Coriolis :1503-1530,
// stresses :2241-2251 / :2595-2606, momentum step :2623-2780).
//
// Grid: the live kernels' flat C-grid (u point i between rho i-1 and i, v point j
// between rho j-1 and j) with ROMS metrics derived from pm, pn per use.  Closed
// basin: u = 0 on walls i = 1, L; v = 0 on walls j = 1, M; zero-gradient ghosts.
//
// Four launches per substep, one thread per point: zeta predictor, momentum
// predictor, zeta corrector (+ time averages), momentum corrector.  Boundary points
// are written by the same launch (a ghost recomputes its interior source).
// Neighbour reads go through L1/L2; each field is read from HBM about once per
// launch.  The FP64 instance equals the oracle (oracle/sf_oracle_phys.c) bit for
// bit; FP32 keeps the averages in FP64.
#include "sf_internal.h"

namespace {

template <typename T>
struct PF {                        // static fields
    const T *h, *pm, *pn, *fomn, *sustr, *svstr;
    int szI, szJ;
    T g, rdrg;
};

template <typename T>
struct State3 {                    // one state (zeta, ubar, vbar)
    const T *z, *u, *v;
};

template <typename T>
__device__ __forceinline__ T Dat(const PF<T> &p, const T *z, long long f) { return p.h[f] + z[f]; }

template <typename T>
__device__ __forceinline__ T DUon(const PF<T> &p, const T *z, const T *u, long long f)
{
    const T onu = T(2) * sf::recip(p.pn[f - 1] + p.pn[f]);
    return ((T(0.5) * (Dat(p, z, f - 1) + Dat(p, z, f))) * u[f]) * onu;
}

template <typename T>
__device__ __forceinline__ T DVom(const PF<T> &p, const T *z, const T *v, long long f)
{
    const long long W = p.szI;
    const T omv = T(2) * sf::recip(p.pm[f - W] + p.pm[f]);
    return ((T(0.5) * (Dat(p, z, f - W) + Dat(p, z, f))) * v[f]) * omv;
}

template <typename T>
__device__ __forceinline__ T rhs_zeta(const PF<T> &p, State3<T> s, long long f)
{
    const long long W = p.szI;
    return (DUon(p, s.z, s.u, f) - DUon(p, s.z, s.u, f + 1)) + (DVom(p, s.z, s.v, f) - DVom(p, s.z, s.v, f + W));
}

template <typename T>
__device__ T rhs_u(const PF<T> &p, State3<T> s, const T *gz, long long f)
{
    const long long W = p.szI;
    const T *z = s.z, *u = s.u, *v = s.v;
    const T onu = T(2) * sf::recip(p.pn[f - 1] + p.pn[f]);
    const T omu = T(2) * sf::recip(p.pm[f - 1] + p.pm[f]);
    const T g0 = gz[f - 1], g1 = gz[f];
    T r = ((T(0.5) * p.g) * onu) * ((p.h[f - 1] + p.h[f]) * (g0 - g1) + (g0 * g0 - g1 * g1));
    const T UFx1 = T(0.25) * (DUon(p, z, u, f) + DUon(p, z, u, f + 1)) * (u[f] + u[f + 1]);
    const T UFx0 = T(0.25) * (DUon(p, z, u, f - 1) + DUon(p, z, u, f)) * (u[f - 1] + u[f]);
    const T UFe1 = T(0.25) * (DVom(p, z, v, f + W) + DVom(p, z, v, f + W - 1)) * (u[f + W] + u[f]);
    const T UFe0 = T(0.25) * (DVom(p, z, v, f) + DVom(p, z, v, f - 1)) * (u[f] + u[f - W]);
    r = r - ((UFx1 - UFx0) + (UFe1 - UFe0));
    const T c1 = ((T(0.5) * Dat(p, z, f)) * p.fomn[f]) * (v[f] + v[f + W]);
    const T c0 = ((T(0.5) * Dat(p, z, f - 1)) * p.fomn[f - 1]) * (v[f - 1] + v[f - 1 + W]);
    r = r + T(0.5) * (c1 + c0);
    r = r - ((p.rdrg * u[f]) * omu) * onu;
    r = r + (p.sustr[f] * omu) * onu;
    return r;
}

template <typename T>
__device__ T rhs_v(const PF<T> &p, State3<T> s, const T *gz, long long f)
{
    const long long W = p.szI;
    const T *z = s.z, *u = s.u, *v = s.v;
    const T omv = T(2) * sf::recip(p.pm[f - W] + p.pm[f]);
    const T onv = T(2) * sf::recip(p.pn[f - W] + p.pn[f]);
    const T g0 = gz[f - W], g1 = gz[f];
    T r = ((T(0.5) * p.g) * omv) * ((p.h[f - W] + p.h[f]) * (g0 - g1) + (g0 * g0 - g1 * g1));
    const T VFx1 = T(0.25) * (DUon(p, z, u, f + 1) + DUon(p, z, u, f + 1 - W)) * (v[f + 1] + v[f]);
    const T VFx0 = T(0.25) * (DUon(p, z, u, f) + DUon(p, z, u, f - W)) * (v[f] + v[f - 1]);
    const T VFe1 = T(0.25) * (DVom(p, z, v, f) + DVom(p, z, v, f + W)) * (v[f] + v[f + W]);
    const T VFe0 = T(0.25) * (DVom(p, z, v, f - W) + DVom(p, z, v, f)) * (v[f - W] + v[f]);
    r = r - ((VFx1 - VFx0) + (VFe1 - VFe0));
    const T c1 = ((T(0.5) * Dat(p, z, f)) * p.fomn[f]) * (u[f] + u[f + 1]);
    const T c0 = ((T(0.5) * Dat(p, z, f - W)) * p.fomn[f - W]) * (u[f - W] + u[f - W + 1]);
    r = r - T(0.5) * (c1 + c0);
    r = r - ((p.rdrg * v[f]) * omv) * onv;
    r = r + (p.svstr[f] * omv) * onv;
    return r;
}

// closed-basin boundary sources: zeta ghosts copy the nearest interior point
__device__ __forceinline__ void zeta_src(int szI, int szJ, int &i, int &j)
{
    const int L = szI - 1, M = szJ - 1;
    i = i < 1 ? 1 : (i > L - 1 ? L - 1 : i);
    j = j < 1 ? 1 : (j > M - 1 ? M - 1 : j);
}

// Coupling with the baroclinic step (sf_phys_substep_coupled; on == 0: plain substep).
// The first substep of a baroclinic step turns the vertical integral of the 3D
// right-hand side into the forcing F = rufrc - rhs of its predictor, stores F and adds
// its Adams-Bashforth extrapolation over the baroclinic steps; every other stage adds F
// (obsolete/step2d_LF_AM3.h.py:2377-2590).  DU2/DV2 accumulate the fluxes of the
// predictor's state (w2p; set on the first substep) and of the predicted state (w2c)
// (:778-846).  All in FP64 whatever the state precision.
struct Cpl {
    const double *rufrc, *rvfrc;
    double *Fu, *Fv;
    const double *Fu1, *Fv1, *Fu2, *Fv2;
    double *DU2, *DV2;
    double w2p, w2c;
    int ab, on;
};

template <typename T>
struct Sub {
    PF<T> p;
    const T *z0, *z1, *u0, *u1, *v0, *v1;
    T *z2, *u2, *v2;
    const T *rz0, *ru0, *rv0;
    T *rz1, *ru1, *rv1;
    T *zs, *us, *vs, *gz;
    double *Zt, *DUa, *DVa;
    T dt;
    int first;
    double w1;
    Cpl c;
};

// predictor forcing of a u (or v) point: the first substep forms F from the point's
// right-hand side r (and stores it where the point is its own, not a ghost's source)
template <typename T>
__device__ __forceinline__ T pred_forcing(const Sub<T> &a, bool vdir, long long f, T r, bool store)
{
    const double *frc = vdir ? a.c.rvfrc : a.c.rufrc;
    double *Fo = vdir ? a.c.Fv : a.c.Fu;
    if (!a.first) return r + T(Fo[f]);
    const double F = frc[f] - (double)r;
    if (store) Fo[f] = F;
    const double *F1 = vdir ? a.c.Fv1 : a.c.Fu1, *F2 = vdir ? a.c.Fv2 : a.c.Fu2;
    return r + T(sf::ab_extrapolate(a.c.ab, F, F1[f], F2[f]));
}

// ---- P1: predicted zeta (+ predictor RHS), and the pressure surface gz ---------------
template <typename T>
__global__ __launch_bounds__(256) void k_phys_pred_zeta(Sub<T> a)
{
    const int szI = a.p.szI, szJ = a.p.szJ;
    const long long S2 = (long long)szI * szJ;
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= S2) return;
    int i = (int)(f % szI), j = (int)(f / szI);
    const bool ghost = i < 1 || i > szI - 2 || j < 1 || j > szJ - 2;
    int si = i, sj = j;
    zeta_src(szI, szJ, si, sj);
    const long long fs = (long long)sj * szI + si;
    State3<T> s1{a.z1, a.u1, a.v1};
    const T r = rhs_zeta(a.p, s1, fs);
    T zs;
    if (a.first) zs = a.z1[fs] + ((a.p.pm[fs] * a.p.pn[fs]) * a.dt) * r;
    else zs = a.z0[fs] + ((a.p.pm[fs] * a.p.pn[fs]) * (T(2) * a.dt)) * r;
    if (!ghost) {
        a.rz1[f] = r;
        if (a.c.on) {
            const double du = a.first ? 0.0 : a.c.DU2[f], dv = a.first ? 0.0 : a.c.DV2[f];
            a.c.DU2[f] = du + a.c.w2p * (double)DUon(a.p, a.z1, a.u1, f);
            a.c.DV2[f] = dv + a.c.w2p * (double)DVom(a.p, a.z1, a.v1, f);
        }
    }
    a.zs[f] = zs;
    const T cff4 = T(4.0 / 25.0), cff5 = T(1) - T(2) * cff4;
    a.gz[f] = a.first ? T(0.5) * (a.z1[f] + zs) : cff5 * a.z1[f] + cff4 * (a.z0[f] + zs);
}

// ---- P2: predicted ubar/vbar (+ predictor RHS) -----------------------------------
template <typename T>
__device__ T pred_u(const Sub<T> &a, long long f, bool store_rhs)
{
    const long long W = a.p.szI;
    State3<T> s1{a.z1, a.u1, a.v1};
    T r = rhs_u(a.p, s1, a.gz, f);
    if (a.c.on) r = pred_forcing(a, false, f, r, store_rhs);
    if (store_rhs) a.ru1[f] = r;
    const T cff = (a.p.pm[f] + a.p.pm[f - 1]) * (a.p.pn[f] + a.p.pn[f - 1]);
    const T fac = sf::recip(Dat(a.p, (const T *)a.zs, f) + Dat(a.p, (const T *)a.zs, f - 1));
    const T *zb = a.first ? a.z1 : a.z0;
    const T *ub = a.first ? a.u1 : a.u0;
    const T c1 = a.first ? T(0.5) * a.dt : a.dt;
    (void)W;
    return (ub[f] * (Dat(a.p, zb, f) + Dat(a.p, zb, f - 1)) + (cff * c1) * r) * fac;
}

template <typename T>
__device__ T pred_v(const Sub<T> &a, long long f, bool store_rhs)
{
    const long long W = a.p.szI;
    State3<T> s1{a.z1, a.u1, a.v1};
    T r = rhs_v(a.p, s1, a.gz, f);
    if (a.c.on) r = pred_forcing(a, true, f, r, store_rhs);
    if (store_rhs) a.rv1[f] = r;
    const T cff = (a.p.pm[f] + a.p.pm[f - W]) * (a.p.pn[f] + a.p.pn[f - W]);
    const T fac = sf::recip(Dat(a.p, (const T *)a.zs, f) + Dat(a.p, (const T *)a.zs, f - W));
    const T *zb = a.first ? a.z1 : a.z0;
    const T *vb = a.first ? a.v1 : a.v0;
    const T c1 = a.first ? T(0.5) * a.dt : a.dt;
    return (vb[f] * (Dat(a.p, zb, f) + Dat(a.p, zb, f - W)) + (cff * c1) * r) * fac;
}

template <typename T>
__global__ __launch_bounds__(256) void k_phys_pred_uv(Sub<T> a)
{
    const int szI = a.p.szI, szJ = a.p.szJ, L = szI - 1, M = szJ - 1;
    const long long S2 = (long long)szI * szJ, W = szI;
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= S2) return;
    const int i = (int)(f % szI), j = (int)(f / szI);
    // ubar: interior i in [2, L-1], j in [1, M-1]; walls/column 0 zero; ghost rows copy
    T u;
    if (i <= 1 || i >= L) u = T(0);
    else if (j >= 1 && j <= M - 1) u = pred_u(a, f, true);
    else u = pred_u(a, j == 0 ? f + W : f - W, false);
    a.us[f] = u;
    T v;
    if (j <= 1 || j >= M) v = T(0);
    else if (i >= 1 && i <= L - 1) v = pred_v(a, f, true);
    else v = pred_v(a, i == 0 ? f + 1 : f - 1, false);
    a.vs[f] = v;
}

// ---- C1: corrected zeta, time averages, and the corrector pressure surface ----------
template <typename T>
__device__ T corr_zeta(const Sub<T> &a, long long f, T *rstar)
{
    State3<T> ss{a.zs, a.us, a.vs};
    const T r = rhs_zeta(a.p, ss, f);
    *rstar = r;
    const T pmn = a.p.pm[f] * a.p.pn[f];
    if (a.first) return a.z1[f] + pmn * (a.dt * r);
    const T a1 = a.dt * T(5.0) / T(12.0), a2 = a.dt * T(8.0) / T(12.0), a3 = a.dt * T(1.0) / T(12.0);
    return a.z1[f] + pmn * ((a1 * r + a2 * a.rz1[f]) - a3 * a.rz0[f]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_phys_corr_zeta(Sub<T> a)
{
    const int szI = a.p.szI, szJ = a.p.szJ;
    const long long S2 = (long long)szI * szJ;
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= S2) return;
    int i = (int)(f % szI), j = (int)(f / szI);
    const bool ghost = i < 1 || i > szI - 2 || j < 1 || j > szJ - 2;
    int si = i, sj = j;
    zeta_src(szI, szJ, si, sj);
    const long long fs = (long long)sj * szI + si;
    T rstar;
    const T z2 = corr_zeta(a, fs, &rstar);
    a.z2[f] = z2;
    const T cf4 = T(2.0 / 5.0), cf5 = T(1) - cf4;
    a.gz[f] = cf5 * z2 + cf4 * a.zs[f];
    if (!ghost) {
        const double w = a.w1;
        a.Zt[f] = a.Zt[f] + w * (double)a.zs[f];
        a.DUa[f] = a.DUa[f] + w * (double)DUon(a.p, (const T *)a.zs, (const T *)a.us, f);
        a.DVa[f] = a.DVa[f] + w * (double)DVom(a.p, (const T *)a.zs, (const T *)a.vs, f);
        if (a.c.on) {
            a.c.DU2[f] = a.c.DU2[f] + a.c.w2c * (double)DUon(a.p, (const T *)a.zs, (const T *)a.us, f);
            a.c.DV2[f] = a.c.DV2[f] + a.c.w2c * (double)DVom(a.p, (const T *)a.zs, (const T *)a.vs, f);
        }
    }
}

// ---- C2: corrected ubar/vbar -------------------------------------------------------
template <typename T>
__device__ T corr_u(const Sub<T> &a, long long f)
{
    State3<T> ss{a.zs, a.us, a.vs};
    T r = rhs_u(a.p, ss, a.gz, f);
    if (a.c.on) r = r + T(a.c.Fu[f]);
    const T cff = (a.p.pm[f] + a.p.pm[f - 1]) * (a.p.pn[f] + a.p.pn[f - 1]);
    const T fac = sf::recip(Dat(a.p, (const T *)a.z2, f) + Dat(a.p, (const T *)a.z2, f - 1));
    const T b1 = T(0.5) * a.dt * T(5.0) / T(12.0), b2 = T(0.5) * a.dt * T(8.0) / T(12.0),
            b3 = T(0.5) * a.dt * T(1.0) / T(12.0);
    const T tend = a.first ? (T(0.5) * a.dt) * r : (b1 * r + b2 * a.ru1[f]) - b3 * a.ru0[f];
    return (a.u1[f] * (Dat(a.p, a.z1, f) + Dat(a.p, a.z1, f - 1)) + cff * tend) * fac;
}

template <typename T>
__device__ T corr_v(const Sub<T> &a, long long f)
{
    const long long W = a.p.szI;
    State3<T> ss{a.zs, a.us, a.vs};
    T r = rhs_v(a.p, ss, a.gz, f);
    if (a.c.on) r = r + T(a.c.Fv[f]);
    const T cff = (a.p.pm[f] + a.p.pm[f - W]) * (a.p.pn[f] + a.p.pn[f - W]);
    const T fac = sf::recip(Dat(a.p, (const T *)a.z2, f) + Dat(a.p, (const T *)a.z2, f - W));
    const T b1 = T(0.5) * a.dt * T(5.0) / T(12.0), b2 = T(0.5) * a.dt * T(8.0) / T(12.0),
            b3 = T(0.5) * a.dt * T(1.0) / T(12.0);
    const T tend = a.first ? (T(0.5) * a.dt) * r : (b1 * r + b2 * a.rv1[f]) - b3 * a.rv0[f];
    return (a.v1[f] * (Dat(a.p, a.z1, f) + Dat(a.p, a.z1, f - W)) + cff * tend) * fac;
}

template <typename T>
__global__ __launch_bounds__(256) void k_phys_corr_uv(Sub<T> a)
{
    const int szI = a.p.szI, szJ = a.p.szJ, L = szI - 1, M = szJ - 1;
    const long long S2 = (long long)szI * szJ, W = szI;
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= S2) return;
    const int i = (int)(f % szI), j = (int)(f / szI);
    T u;
    if (i <= 1 || i >= L) u = T(0);
    else if (j >= 1 && j <= M - 1) u = corr_u(a, f);
    else u = corr_u(a, j == 0 ? f + W : f - W);
    a.u2[f] = u;
    T v;
    if (j <= 1 || j >= M) v = T(0);
    else if (i >= 1 && i <= L - 1) v = corr_v(a, f);
    else v = corr_v(a, i == 0 ? f + 1 : f - 1);
    a.v2[f] = v;
}

// ---- open edges (sf_phys_substep_open) ----------------------------------------------
// Chapman free surface on the ghost rho points and Flather normal velocity on the edge
// faces of the open edges (flags: SF_OBC_WEST/EAST/SOUTH/NORTH), applied after the stage
// kernels have written closed-wall values (nonlinear/zetabc.py, the Chapman sketch, and
// nonlinear/barotropicVelocityBC.py:66-98, Flather; ROMS zetabc/u2dbc/v2dbc):
//   ghost  zn(g) = (zb(g) + Ce zn(i)) / (1 + Ce),  Ce = dt_s 0.5 (m(g)+m(i)) sqrt(g (h(g)+zb(g)))
//   face   u = ext_u -/+ sqrt(g/D) (0.5 (zn(g)+zn(i)) - ext_z)   (- west/south, + east/north)
// zb: the level the stage starts from, dt_s its time span, zn the stage's new surface.
// Edge points b = 1 .. n (n = szJ-2 for west/east, szI-2 for south/north); the threads of
// b = 1 and b = n also write the face value into the two ghost rows (zero gradient).
struct Open {
    int flags;
    const double *ez[4], *eu[4];
    int volcons;              // obc_volcons + conserve_mass after each velocity stage
    double *vc_work, *vc_out;
};

__device__ __forceinline__ void edge_points(int e, int b, int W, int L, int M, long long &fg, long long &fi,
                                            long long &fu)
{
    if (e == 0) { fg = (long long)b * W; fi = fg + 1; fu = fi; }
    else if (e == 1) { fg = (long long)b * W + L; fi = fg - 1; fu = fg; }
    else if (e == 2) { fg = b; fi = (long long)W + b; fu = fi; }
    else { fg = (long long)M * W + b; fi = fg - W; fu = fg; }
}

template <typename T>
__global__ __launch_bounds__(256) void k_open_zeta(Sub<T> a, Open ob, int corrector)
{
    const int e = blockIdx.y;
    if (!(ob.flags & (1 << e))) return;
    const int W = a.p.szI, L = a.p.szI - 1, M = a.p.szJ - 1;
    const int n = e < 2 ? M - 1 : L - 1;
    const int b = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b > n) return;
    long long fg, fi, fu;
    edge_points(e, b, W, L, M, fg, fi, fu);
    T *zn = corrector ? a.z2 : a.zs;
    const T *zb = corrector ? a.z1 : (a.first ? a.z1 : a.z0);
    const T dts = corrector ? a.dt : (a.first ? a.dt : T(2) * a.dt);
    const T m = e < 2 ? T(0.5) * (a.p.pm[fg] + a.p.pm[fi]) : T(0.5) * (a.p.pn[fg] + a.p.pn[fi]);
    const T Ce = (dts * m) * sqrt(a.p.g * (a.p.h[fg] + zb[fg]));
    const T z = (zb[fg] + Ce * zn[fi]) / (T(1) + Ce);
    zn[fg] = z;
    if (corrector) {
        const T cf4 = T(2.0 / 5.0), cf5 = T(1) - cf4;
        a.gz[fg] = cf5 * z + cf4 * a.zs[fg];
    } else {
        const T cff4 = T(4.0 / 25.0), cff5 = T(1) - T(2) * cff4;
        a.gz[fg] = a.first ? T(0.5) * (a.z1[fg] + z) : cff5 * a.z1[fg] + cff4 * (a.z0[fg] + z);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_open_vel(Sub<T> a, Open ob, int corrector)
{
    const int e = blockIdx.y;
    if (!(ob.flags & (1 << e))) return;
    const int W = a.p.szI, L = a.p.szI - 1, M = a.p.szJ - 1;
    const int n = e < 2 ? M - 1 : L - 1;
    const int b = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b > n) return;
    long long fg, fi, fu;
    edge_points(e, b, W, L, M, fg, fi, fu);
    const T *zn = corrector ? a.z2 : a.zs;
    T *x = e < 2 ? (corrector ? a.u2 : a.us) : (corrector ? a.v2 : a.vs);
    const T D = T(0.5) * ((a.p.h[fg] + zn[fg]) + (a.p.h[fi] + zn[fi]));
    const T c = sqrt(a.p.g / D);
    const T ez = ob.ez[e] ? T(ob.ez[e][b]) : T(0), eu = ob.eu[e] ? T(ob.eu[e][b]) : T(0);
    const T dz = T(0.5) * (zn[fg] + zn[fi]) - ez;
    const T val = (e == 0 || e == 2) ? eu - c * dz : eu + c * dz;
    x[fu] = val;
    // ghost rows (west/east) or columns (south/north) of the face
    const long long step = e < 2 ? W : 1;
    if (b == 1) x[fu - step] = val;
    if (b == n) x[fu + step] = val;
}

template <typename T>
int launch_phys(const sf_phys_fields *x, double dt, int first, double w1, void *stream,
                const sf_couple_fields *cp = nullptr, const Open *ob = nullptr)
{
    Sub<T> a;
    a.c = Cpl{};
    if (cp) {
        a.c = Cpl{cp->rufrc, cp->rvfrc, cp->Fu, cp->Fv, cp->Fu1, cp->Fv1, cp->Fu2, cp->Fv2, cp->DU_avg2,
                  cp->DV_avg2, cp->w2p, cp->w2c, cp->ab, 1};
    }
    a.p = PF<T>{(const T *)x->h, (const T *)x->pm, (const T *)x->pn, (const T *)x->fomn, (const T *)x->sustr,
                (const T *)x->svstr, x->szI, x->szJ, T(x->g), T(x->rdrg)};
    a.z0 = (const T *)x->zeta[0];
    a.z1 = (const T *)x->zeta[1];
    a.z2 = (T *)x->zeta[2];
    a.u0 = (const T *)x->ubar[0];
    a.u1 = (const T *)x->ubar[1];
    a.u2 = (T *)x->ubar[2];
    a.v0 = (const T *)x->vbar[0];
    a.v1 = (const T *)x->vbar[1];
    a.v2 = (T *)x->vbar[2];
    a.rz0 = (const T *)x->rzeta[0];
    a.rz1 = (T *)x->rzeta[1];
    a.ru0 = (const T *)x->rubar[0];
    a.ru1 = (T *)x->rubar[1];
    a.rv0 = (const T *)x->rvbar[0];
    a.rv1 = (T *)x->rvbar[1];
    a.zs = (T *)x->scratch[0];
    a.us = (T *)x->scratch[1];
    a.vs = (T *)x->scratch[2];
    a.gz = (T *)x->scratch[3];
    a.Zt = x->Zt_avg1;
    a.DUa = x->DU_avg1;
    a.DVa = x->DV_avg1;
    a.dt = T(dt);
    a.first = first;
    a.w1 = w1;
    const long long S2 = (long long)x->szI * x->szJ;
    const dim3 grid(sf::grid1d(S2)), block(256);
    hipStream_t s = sf::as_stream(stream);
    const int nedge = (x->szI > x->szJ ? x->szI : x->szJ);
    const dim3 egrid(sf::grid1d(nedge), 4);
    hipLaunchKernelGGL(k_phys_pred_zeta<T>, grid, block, 0, s, a);
    if (ob) hipLaunchKernelGGL(k_open_zeta<T>, egrid, block, 0, s, a, *ob, 0);
    const int prec = sizeof(T) == 8 ? SF_FP64 : SF_FP32;
    const bool vc = ob && ob->volcons && ob->flags;
    hipLaunchKernelGGL(k_phys_pred_uv<T>, grid, block, 0, s, a);
    if (ob) hipLaunchKernelGGL(k_open_vel<T>, egrid, block, 0, s, a, *ob, 0);
    // the volume constraint on the predicted velocities: the corrector's free surface
    // then sees no net flow through the open edges
    if (vc) {
        const int rc = sf_obc_volcons(a.zs, x->h, x->pm, x->pn, a.us, a.vs, x->szI, x->szJ, prec, ob->flags, 1,
                                      ob->vc_work, ob->vc_out, stream);
        if (rc != SF_OK) return rc;
    }
    hipLaunchKernelGGL(k_phys_corr_zeta<T>, grid, block, 0, s, a);
    if (ob) hipLaunchKernelGGL(k_open_zeta<T>, egrid, block, 0, s, a, *ob, 1);
    hipLaunchKernelGGL(k_phys_corr_uv<T>, grid, block, 0, s, a);
    if (ob) hipLaunchKernelGGL(k_open_vel<T>, egrid, block, 0, s, a, *ob, 1);
    if (vc) {
        const int rc = sf_obc_volcons(x->zeta[2], x->h, x->pm, x->pn, x->ubar[2], x->vbar[2], x->szI, x->szJ, prec,
                                      ob->flags, 1, ob->vc_work, ob->vc_out, stream);
        if (rc != SF_OK) return rc;
    }
    return sf::check_launch("sf_phys_substep");
}

}  // namespace

extern "C" int sf_phys_substep(const sf_phys_fields *x, int32_t precision, double dt, int32_t first,
                               double cff1, void *stream)
{
    SF_REQUIRE(x, "sf_phys_substep: null argument");
    SF_REQUIRE(x->szI >= 4 && x->szJ >= 4, "sf_phys_substep: grid %dx%d too small", x->szI, x->szJ);
    SF_REQUIRE(x->h && x->pm && x->pn && x->fomn && x->sustr && x->svstr, "sf_phys_substep: null grid field");
    for (int k = 0; k < 3; k++)
        SF_REQUIRE(x->zeta[k] && x->ubar[k] && x->vbar[k], "sf_phys_substep: null state slot");
    SF_REQUIRE(x->rzeta[0] && x->rzeta[1] && x->rubar[0] && x->rubar[1] && x->rvbar[0] && x->rvbar[1],
               "sf_phys_substep: null rhs slot");
    for (int k = 0; k < 4; k++) SF_REQUIRE(x->scratch[k], "sf_phys_substep: null scratch");
    SF_REQUIRE(x->Zt_avg1 && x->DU_avg1 && x->DV_avg1, "sf_phys_substep: null average");
    if (precision == SF_FP64) return launch_phys<double>(x, dt, first, cff1, stream);
    if (precision == SF_FP32) return launch_phys<float>(x, dt, first, cff1, stream);
    sf::set_error("sf_phys_substep: unknown precision %d", precision);
    return SF_ERR_ARG;
}

extern "C" int sf_phys_substep_coupled(const sf_phys_fields *x, const sf_couple_fields *cp, int32_t precision,
                                       double dt, int32_t first, double cff1, void *stream)
{
    SF_REQUIRE(x && cp, "sf_phys_substep_coupled: null argument");
    SF_REQUIRE(x->szI >= 4 && x->szJ >= 4, "sf_phys_substep_coupled: grid %dx%d too small", x->szI, x->szJ);
    SF_REQUIRE(x->h && x->pm && x->pn && x->fomn && x->sustr && x->svstr,
               "sf_phys_substep_coupled: null grid field");
    for (int k = 0; k < 3; k++)
        SF_REQUIRE(x->zeta[k] && x->ubar[k] && x->vbar[k], "sf_phys_substep_coupled: null state slot");
    SF_REQUIRE(x->rzeta[0] && x->rzeta[1] && x->rubar[0] && x->rubar[1] && x->rvbar[0] && x->rvbar[1],
               "sf_phys_substep_coupled: null rhs slot");
    for (int k = 0; k < 4; k++) SF_REQUIRE(x->scratch[k], "sf_phys_substep_coupled: null scratch");
    SF_REQUIRE(x->Zt_avg1 && x->DU_avg1 && x->DV_avg1, "sf_phys_substep_coupled: null average");
    SF_REQUIRE(cp->Fu && cp->Fv && cp->DU_avg2 && cp->DV_avg2, "sf_phys_substep_coupled: null coupling field");
    SF_REQUIRE(!first || (cp->rufrc && cp->rvfrc && cp->Fu1 && cp->Fv1 && cp->Fu2 && cp->Fv2),
               "sf_phys_substep_coupled: the first substep needs rufrc/rvfrc and the forcing history");
    SF_REQUIRE(cp->ab >= 0 && cp->ab <= 2, "sf_phys_substep_coupled: ab=%d", cp->ab);
    if (precision == SF_FP64) return launch_phys<double>(x, dt, first, cff1, stream, cp);
    if (precision == SF_FP32) return launch_phys<float>(x, dt, first, cff1, stream, cp);
    sf::set_error("sf_phys_substep_coupled: unknown precision %d", precision);
    return SF_ERR_ARG;
}

extern "C" int sf_phys_substep_open(const sf_phys_fields *x, const sf_open_bc *obc, int32_t precision, double dt,
                                    int32_t first, double cff1, double *volcons_work, double *volcons_out,
                                    void *stream)
{
    SF_REQUIRE(x && obc, "sf_phys_substep_open: null argument");
    SF_REQUIRE(x->szI >= 4 && x->szJ >= 4, "sf_phys_substep_open: grid %dx%d too small", x->szI, x->szJ);
    SF_REQUIRE(x->h && x->pm && x->pn && x->fomn && x->sustr && x->svstr, "sf_phys_substep_open: null grid field");
    for (int k = 0; k < 3; k++)
        SF_REQUIRE(x->zeta[k] && x->ubar[k] && x->vbar[k], "sf_phys_substep_open: null state slot");
    SF_REQUIRE(x->rzeta[0] && x->rzeta[1] && x->rubar[0] && x->rubar[1] && x->rvbar[0] && x->rvbar[1],
               "sf_phys_substep_open: null rhs slot");
    for (int k = 0; k < 4; k++) SF_REQUIRE(x->scratch[k], "sf_phys_substep_open: null scratch");
    SF_REQUIRE(x->Zt_avg1 && x->DU_avg1 && x->DV_avg1, "sf_phys_substep_open: null average");
    SF_REQUIRE((obc->flags & ~15) == 0, "sf_phys_substep_open: unknown edge flags 0x%x", obc->flags);
    SF_REQUIRE(!obc->volcons || (volcons_work && volcons_out), "sf_phys_substep_open: volcons needs work and out");
    Open ob;
    ob.flags = obc->flags;
    for (int e = 0; e < 4; e++) {
        ob.ez[e] = obc->zeta_ext[e];
        ob.eu[e] = obc->vel_ext[e];
    }
    ob.volcons = obc->volcons;
    ob.vc_work = volcons_work;
    ob.vc_out = volcons_out;
    if (precision == SF_FP64) return launch_phys<double>(x, dt, first, cff1, stream, nullptr, &ob);
    if (precision == SF_FP32) return launch_phys<float>(x, dt, first, cff1, stream, nullptr, &ob);
    sf::set_error("sf_phys_substep_open: unknown precision %d", precision);
    return SF_ERR_ARG;
}
