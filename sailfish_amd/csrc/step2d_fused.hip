// step2d_fused.hip -- one barotropic substep (predictor + corrector + both boundary
// passes of nonlinear/step2d.py:23-127) in one tiled kernel plus one small kernel
// over the points that computeMomentumPred masks.
//
// Layout and tiling.  A workgroup of 256 threads owns a TI x TJ tile T of rho points
// (TI along xi, the contiguous axis) and stages what it needs in LDS on the tile
// plus a 2-point halo (R2); predictor quantities are recomputed on the 1-point halo
// (R1) instead of being exchanged between workgroups.  Each HBM field is read or
// written once per substep (halo re-reads hit L2), so the kernel is bound by HBM
// bandwidth.  Workgroup ids are remapped so that each XCD walks a contiguous band
// of tiles (its L2 then serves the halo re-reads).
//
// Stages (each followed by a barrier), with the reference code they restate:
//   A  D1 = zeta_t1 + h, h, zeta_t0 on R2
//   B  FX = (ubar_t1*RtoU(D1))*on_u, FY = (vbar_t1*RtoV(D1))*om_v  (mod_cppkernels.cpp:80-128)
//   C  predicted zeta before BC on R1: zeta_t0 + Dt*rzeta          (cpp:133-150)
//   D  zeta boundary map on R1                                      (zetabc.py:5-12)
//   E  rhs_ubar/rhs_vbar (cpp:12-42), ubar/vbar (cpp:155-186) on T + right column /
//      top row, velocity boundary map (barotropicVelocityBC.py:7-31)
//   F  corrector on T: rzeta + averages (cpp:80-128), zeta AM3 (cpp:190-209)
//   G  zeta boundary map on T (zetabc.py:5-12), stores.
// The masked-points kernel then finishes the points computeMomentumPred skips
// (i%szJ mask, cpp:162): their velocity boundary values, main3d's vbar_t2 := 0
// (main3d.py:189), and the AM3 update that step2d.py:120-123 applies to ubar_t1
// through the handle bound at mod_operators.py:146.  That update is kept only at
// masked points: everywhere else the next predictor overwrites it before any read.
//
// Race freedom: inside the tiled kernel, every array that is read at halo points
// (ubar_t1, vbar_t1, ubar_t2/vbar_t2 at masked points, rzeta_t1/rubar_t1/rvbar_t1 at
// non-interior points) is either not written by the kernel or written only where no
// other workgroup reads it.
//
// FP64 arithmetic keeps the reference's operand order with contraction off, so the
// FP64 instance reproduces the one-kernel-per-reference-kernel sequence bit for bit.
// The FP32 instance keeps the same order in float and accumulates Zt/DU/DV in FP64.
//
// Not written (never read by any kernel of the path): the corrector's momentum RHS
// (rubar_t2, rvbar_t2) and the predictor's rvbar_t1.
#include "sf_internal.h"
#include "step2d_fused.h"

#include <algorithm>
#include <vector>

namespace {

constexpr int TI = SF_FUSED_TI;
constexpr int TJ = SF_FUSED_TJ;
constexpr int NT = 256;
constexpr int W2 = TI + 4;           // R2 frame width
constexpr int H2 = TJ + 4;
constexpr int N2 = W2 * H2;

// dense boundary-code values
constexpr int kKeep = -1;
constexpr int kZero = -2;
constexpr int kClamp = -3;

template <typename T>
struct Fields {
    const T *h, *on_u, *om_v, *pm, *pn;
    const T *zeta_t0, *zeta_t1;
    T *zeta_t2;
    T *ubar_t1, *ubar_t2;
    const T *vbar_t1;
    T *vbar_t2;
    const T *rzeta_t0;
    T *rzeta_t1;
    const T *rzeta_t2;
    const T *rubar_t0;
    T *rubar_t1;
    const T *rvbar_t1;
    double *Zt_avg1, *DU_avg1, *DV_avg1;
    const int32_t *bc_zeta, *bc_ubar, *bc_vbar;
};

struct Params {
    int szI, szJ, q5_stride;
    int tiles_i, tiles_j;
    int off_i, off_j;        // tile origin offset (the first tile starts at -off)
    long long S2;
    double Dt, cff1, clamp;
    int zero_vbar;
    int fold_masked;         // masked points finished by the tiled kernel (stride == szI)
};

__device__ __forceinline__ int lidx(int li, int lj) { return lj * W2 + li; }

// AdamsMoultonCorr3rd (cpp:190-209): x + Dt*(AM3_2*r2 + AM3_1*r1 + AM3_0*r0)
template <typename T>
__device__ __forceinline__ T am3(T x, T Dt, T r2, T r1, T r0)
{
    const T AM3_2 = T(5.0 / 12.0), AM3_1 = T(8.0 / 12.0), AM3_0 = T(-1.0 / 12.0);
    return x + Dt * (AM3_2 * r2 + AM3_1 * r1 + AM3_0 * r0);
}

__device__ __forceinline__ bool is_masked(long long f, int stride)
{
    return (f % stride == 0) || (f / stride == 0);
}

// NTS: outputs written with streaming (non-temporal) stores and the FP64 averages read
// with streaming loads (chosen by the launcher from the bytes one substep moves)
template <bool NTS, typename V>
__device__ __forceinline__ void stv(V *p, V v)
{
    if constexpr (NTS) __builtin_nontemporal_store(v, p);
    else *p = v;
}
template <bool NTS>
__device__ __forceinline__ double ldavg(const double *p)
{
    if constexpr (NTS) return __builtin_nontemporal_load(p);
    return *p;
}

// VC: a velocity boundary map copies from neighbours (gradient boundaries).  Stage E
// then keeps the predicted velocities before the map in two more LDS planes, and a
// destination takes its source's value from them after a barrier.
template <typename T, bool NTS, bool VC>
__global__ __launch_bounds__(NT) void k_step2d_fused(Fields<T> F, Params P)
{
    // six LDS planes on the R2 frame, reused as their contents die:
    //   P1  A-B: D1 = zeta_t1 + h          D-E: D2 = predicted zeta + h
    //   P2  A-E: h
    //   P3  B-C: FX                         E-F: PX = DU*on_u of the predicted state
    //   P4  B-C: FY                         E-F: PY = DV*om_v of the predicted state
    //   P5  C-D: predicted zeta before BC   F-G: corrected zeta before BC
    //   P6  D-E: gz = (1-w)*predicted zeta + w*zeta_t0
    __shared__ T P1[N2], P2[N2], P3[N2], P4[N2], P5[N2], P6[N2];
    __shared__ T P7[VC ? N2 : 1], P8[VC ? N2 : 1];   // VC: ubar, vbar of E before the map
    T *const sD1 = P1, *const sD2 = P1, *const sH = P2;
    T *const sFX = P3, *const sPX = P3, *const sFY = P4, *const sPY = P4;
    T *const sZq = P5, *const sZc = P5, *const sGZ = P6;

    const int szI = P.szI, szJ = P.szJ;
    const int nblk = P.tiles_i * P.tiles_j;
    int b = blockIdx.x;
    if ((nblk & 7) == 0) b = (blockIdx.x & 7) * (nblk >> 3) + (blockIdx.x >> 3);
    const int ti = b % P.tiles_i, tj = b / P.tiles_i;
    const int i0 = ti * TI - P.off_i, j0 = tj * TJ - P.off_j;
    const int gi0 = i0 - 2, gj0 = j0 - 2;   // global coords of the R2 frame origin
    const int tid = threadIdx.x;
    const T Dt = T(P.Dt);

    auto inside = [&](int i, int j) { return i >= 0 && i < szI && j >= 0 && j < szJ; };
    auto interior = [&](int i, int j) { return i > 0 && i < szI - 1 && j > 0 && j < szJ - 1; };
    auto band = [&](int i, int j) { return i <= 1 || i >= szI - 2 || j <= 1 || j >= szJ - 2; };

    // Every thread owns the same R2-frame points r = tid + k*NT in all stages, so
    // per-point inputs stay in registers; only neighbour accesses go through LDS.
    constexpr int KP = (N2 + NT - 1) / NT;
    T ub1[KP], vb1[KP], onu[KP], omv[KP], pmr[KP], pnr[KP];
    T rz0[KP], rz2[KP], up[KP], vp[KP], z0r[KP], hr[KP], zp[KP], dU1[KP], dV1[KP];
    double zt[KP], du[KP], dv[KP];

    // ---- prologue: every bulk load of the tile, issued back to back ----------------
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const int r = tid + k * NT;
        const int li = r % W2, lj = r / W2;
        const int i = gi0 + li, j = gj0 + lj;
        T z1 = T(0), hh = T(0), z0 = T(0);
        ub1[k] = vb1[k] = onu[k] = omv[k] = pmr[k] = pnr[k] = rz0[k] = rz2[k] = T(0);
        zt[k] = du[k] = dv[k] = 0.0;
        if (r < N2 && inside(i, j)) {
            const long long f = (long long)j * szI + i;
            z1 = F.zeta_t1[f];
            hh = F.h[f];
            z0 = F.zeta_t0[f];
            ub1[k] = F.ubar_t1[f];
            vb1[k] = F.vbar_t1[f];
            onu[k] = F.on_u[f];
            omv[k] = F.om_v[f];
            if (li >= 1 && li <= TI + 2 && lj >= 1 && lj <= TJ + 2) {
                pmr[k] = F.pm[f];
                pnr[k] = F.pn[f];
            }
            if (li >= 2 && li <= TI + 1 && lj >= 2 && lj <= TJ + 1) {
                rz0[k] = F.rzeta_t0[f];
                rz2[k] = F.rzeta_t2[f];
                if (interior(i, j)) {
                    zt[k] = ldavg<NTS>(F.Zt_avg1 + f);
                    du[k] = ldavg<NTS>(F.DU_avg1 + f);
                    dv[k] = ldavg<NTS>(F.DV_avg1 + f);
                }
            }
        }
        z0r[k] = z0;
        hr[k] = hh;
        if (r < N2) {
            sD1[r] = z1 + hh;
            sH[r] = hh;
        }
    }
    __syncthreads();

    // ---- B: face fluxes of the t1 state (computeZetaRHS, cpp:80-128) ---------------
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const int r = tid + k * NT;
        if (r >= N2) continue;
        const int li = r % W2, lj = r / W2;
        const int i = gi0 + li, j = gj0 + lj;
        T fx = T(0), fy = T(0);
        dU1[k] = dV1[k] = T(0);
        if (li >= 1 && lj >= 1 && inside(i, j)) {
            const T D = sD1[r];
            // ubar_t1*RtoU(D1) is also the first term of computeMomentumPred's
            // numerator (cpp:183): kept for stage E.  Row 0 reads row -1 as zero
            // (the reference reads before the array there).
            dU1[k] = ub1[k] * ((D + sD1[r - 1]) * T(0.5));
            dV1[k] = vb1[k] * ((D + sD1[r - W2]) * T(0.5));
            if (i >= 1 && j >= 1) {
                fx = dU1[k] * onu[k];
                fy = dV1[k] * omv[k];
            }
        }
        sFX[r] = fx;
        sFY[r] = fy;
    }
    __syncthreads();

    // ---- C: predicted zeta before BC on R1 (computeZetaPred, cpp:133-150) ----------
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const int r = tid + k * NT;
        if (r >= N2) continue;
        const int li = r % W2, lj = r / W2;
        if (li < 1 || li > TI + 2 || lj < 1 || lj > TJ + 2) continue;
        const int i = gi0 + li, j = gj0 + lj;
        T zq = T(0);
        if (inside(i, j)) {
            const long long f = (long long)j * szI + i;
            if (i == 0 || j == 0) {
                // computeZetaPred leaves column 0 and row 0 alone (the buffer keeps its
                // value).  Outside the own tile such points are boundary-map
                // destinations whose halo copies are never consumed: not read.
                if (li >= 2 && li <= TI + 1 && lj >= 2 && lj <= TJ + 1) zq = F.zeta_t2[f];
            } else {
                T rz;
                if (interior(i, j))
                    rz = (((sFX[r + 1] - sFX[r]) + (sFY[r + W2] - sFY[r])) * pmr[k]) * pnr[k];
                else
                    rz = F.rzeta_t1[f];   // computeZetaRHS never writes non-interior points
                zq = z0r[k] + Dt * rz;
            }
        }
        sZq[r] = zq;
    }
    __syncthreads();

    // ---- D: zeta boundary map on R1 (zetabc.py:5-12) --------------------------------
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const int r = tid + k * NT;
        if (r >= N2) continue;
        const int li = r % W2, lj = r / W2;
        if (li < 1 || li > TI + 2 || lj < 1 || lj > TJ + 2) continue;
        const int i = gi0 + li, j = gj0 + lj;
        T z = sZq[r];
        if (inside(i, j) && band(i, j)) {
            const int s = F.bc_zeta[(long long)j * szI + i];
            if (s >= 0) z = sZq[lidx(s % szI - gi0, s / szI - gj0)];
        }
        zp[k] = z;
    }
    __syncthreads();   // P1's D1 and P5's pre-map values are read above: write after
    {
        const T w = T(2.0 / 5.0);
        const T a = T(1) - w;
#pragma unroll
        for (int k = 0; k < KP; k++) {
            const int r = tid + k * NT;
            if (r >= N2) continue;
            const int li = r % W2, lj = r / W2;
            if (li < 1 || li > TI + 2 || lj < 1 || lj > TJ + 2) continue;
            sD2[r] = zp[k] + hr[k];
            sGZ[r] = a * zp[k] + w * z0r[k];
        }
    }
    __syncthreads();

    // ---- E: momentum RHS and predicted velocities on T + right column + top row
    //      (cpp:12-42, cpp:155-186, barotropicVelocityBC.py:7-31); then the corrector's
    //      face fluxes PX/PY and time averages (cpp:80-128) ---------------------------
    const T w = T(2.0 / 5.0);
    const T a = T(1) - w;
    const T gc = T(10.0);
    const double cff1 = P.cff1;
    // the boundary map at (i, j), the corrector's face fluxes of the predicted state and
    // the time averages; VC copies read the pre-map values of E in P7 / P8
    auto e_tail = [&](int k, int r, int i, int j, long long f, bool inT, T u2, T v2) {
        if (band(i, j)) {
            const int cu = F.bc_ubar[f], cv = F.bc_vbar[f];
            if (cu == kZero) u2 = T(0);
            else if (cu == kClamp) u2 = T(P.clamp);
            else if (VC && cu >= 0) u2 = P7[lidx(cu % szI - gi0, cu / szI - gj0)];
            if (cv == kZero) v2 = T(0);
            else if (cv == kClamp) v2 = T(P.clamp);
            else if (VC && cv >= 0) v2 = P8[lidx(cv % szI - gi0, cv / szI - gj0)];
        }
        up[k] = u2;
        vp[k] = v2;
        // corrector fluxes of the predicted state: DU = ubar*RtoU(D), DV = vbar*RtoV(D)
        const T D = sD2[r];
        const T Dl = sD2[r - 1];
        const T Dd = sD2[r - W2];
        const T DU = u2 * ((D + Dl) * T(0.5));
        const T DV = v2 * ((D + Dd) * T(0.5));
        if (inT && interior(i, j)) {
            stv<NTS>(F.Zt_avg1 + f, zt[k] + cff1 * (double)zp[k]);
            stv<NTS>(F.DU_avg1 + f, du[k] + (cff1 * (double)DU) * (double)onu[k]);
            stv<NTS>(F.DV_avg1 + f, dv[k] + (cff1 * (double)DV) * (double)omv[k]);
        }
        sPX[r] = DU * onu[k];
        sPY[r] = DV * omv[k];
    };
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const int r = tid + k * NT;
        if (r >= N2) continue;
        const int li = r % W2, lj = r / W2;
        up[k] = vp[k] = T(0);
        if (li < 2 || li > TI + 2 || lj < 2 || lj > TJ + 2) continue;
        const int i = gi0 + li, j = gj0 + lj;
        if (inside(i, j)) {
            const long long f = (long long)j * szI + i;
            const bool inT = li <= TI + 1 && lj <= TJ + 1;
            T u2, v2;
            if (!is_masked(f, P.q5_stride)) {
                T ru, rv;
                if (interior(i, j)) {
                    const T gz = sGZ[r];
                    const T gzl = sGZ[r - 1];
                    const T gzd = sGZ[r - W2];
                    const T hu = (sH[r] + sH[r - 1]) * T(0.5);
                    const T hv = (sH[r] + sH[r - W2]) * T(0.5);
                    ru = (gc * onu[k]) * (hu * (gz - gzl) + T(0.5) * (gz * gz - gzl * gzl));
                    rv = (gc * omv[k]) * (hv * (gz - gzd) + T(0.5) * (gz * gz - gzd * gzd));
                    if (inT) stv<NTS>(F.rubar_t1 + f, ru);
                } else {
                    ru = F.rubar_t1[f];   // never written at non-interior points
                    rv = F.rvbar_t1[f];
                }
                const T D2 = sD2[r], D2l = sD2[r - 1], D2d = sD2[r - W2];
                u2 = (dU1[k] + ((Dt * ru) * pmr[k]) * pnr[k]) / ((D2 + D2l) * T(0.5));
                v2 = (dV1[k] + ((Dt * rv) * pmr[k]) * pnr[k]) / ((D2 + D2d) * T(0.5));
            } else {
                // masked: the t2 buffer keeps its value
                u2 = F.ubar_t2[f];
                v2 = F.vbar_t2[f];
                if (interior(i, j) && inT) {
                    const T gz = sGZ[r];
                    const T gzl = sGZ[r - 1];
                    const T hu = (sH[r] + sH[r - 1]) * T(0.5);
                    stv<NTS>(F.rubar_t1 + f, (gc * onu[k]) * (hu * (gz - gzl) + T(0.5) * (gz * gz - gzl * gzl)));
                }
            }
            if constexpr (VC) {
                up[k] = u2;   // the map is applied after the barrier below
                vp[k] = v2;
                P7[r] = u2;
                P8[r] = v2;
            } else {
                e_tail(k, r, i, j, f, inT, u2, v2);
            }
        } else {
            sPX[r] = T(0);
            sPY[r] = T(0);
        }
    }
    if constexpr (VC) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KP; k++) {
            const int r = tid + k * NT;
            if (r >= N2) continue;
            const int li = r % W2, lj = r / W2;
            if (li < 2 || li > TI + 2 || lj < 2 || lj > TJ + 2) continue;
            const int i = gi0 + li, j = gj0 + lj;
            if (!inside(i, j)) continue;
            e_tail(k, r, i, j, (long long)j * szI + i, li <= TI + 1 && lj <= TJ + 1, up[k], vp[k]);
        }
    }
    __syncthreads();

    // ---- F: corrector on T: rzeta (cpp:80-128), zeta AM3 (cpp:190-209) -----------------
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const int r = tid + k * NT;
        if (r >= N2) continue;
        const int li = r % W2, lj = r / W2;
        if (li < 2 || li > TI + 1 || lj < 2 || lj > TJ + 1) continue;
        const int i = gi0 + li, j = gj0 + lj;
        if (!inside(i, j)) continue;
        const long long f = (long long)j * szI + i;
        T rz;
        if (interior(i, j)) {
            rz = (((sPX[r + 1] - sPX[r]) + (sPY[r + W2] - sPY[r])) * pmr[k]) * pnr[k];
            stv<NTS>(F.rzeta_t1 + f, rz);
        } else {
            rz = F.rzeta_t1[f];
        }
        sZc[r] = am3(zp[k], Dt, rz2[k], rz, rz0[k]);
    }
    __syncthreads();

    // ---- G: zeta boundary map on T (zetabc.py:5-12), stores ------------------------
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const int r = tid + k * NT;
        if (r >= N2) continue;
        const int li = r % W2, lj = r / W2;
        if (li < 2 || li > TI + 1 || lj < 2 || lj > TJ + 1) continue;
        const int i = gi0 + li, j = gj0 + lj;
        if (!inside(i, j)) continue;
        const long long f = (long long)j * szI + i;
        T z = sZc[r];
        if (band(i, j)) {
            const int s = F.bc_zeta[f];
            if (s >= 0) z = sZc[lidx(s % szI - gi0, s / szI - gj0)];
        }
        stv<NTS>(F.zeta_t2 + f, z);
        if (!is_masked(f, P.q5_stride)) {
            stv<NTS>(F.ubar_t2 + f, up[k]);
            stv<NTS>(F.vbar_t2 + f, P.zero_vbar ? T(0) : vp[k]);
        } else if (P.fold_masked) {
            // stride == szI: the masked points are column 0 and row 0, which no
            // other workgroup reads, so they are finished here (as k_step2d_masked)
            stv<NTS>(F.ubar_t2 + f, up[k]);
            stv<NTS>(F.vbar_t2 + f, P.zero_vbar ? T(0) : vp[k]);
            stv<NTS>(F.ubar_t1 + f, am3(ub1[k], Dt, F.rubar_t1[f], F.rubar_t0[f], vb1[k]));
        }
    }
}


// Masked points (f%stride == 0 or f < stride): velocity boundary constants,
// main3d's vbar_t2 := 0, and the AM3 update of ubar_t1 from step2d.py:120-123.
template <typename T>
__global__ __launch_bounds__(256) void k_step2d_masked(Fields<T> F, Params P)
{
    const long long ncol = (P.S2 + P.q5_stride - 1) / P.q5_stride;   // f = k*stride
    const long long n = ncol + (P.q5_stride - 1);                     // plus f = 1..stride-1
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const long long f = k < ncol ? k * P.q5_stride : k - ncol + 1;
    if (f >= P.S2) return;
    const int szI = P.szI, szJ = P.szJ;
    const int i = (int)(f % szI), j = (int)(f / szI);
    const bool band = i <= 1 || i >= szI - 2 || j <= 1 || j >= szJ - 2;
    T up = F.ubar_t2[f], vp = F.vbar_t2[f];
    if (band) {
        // a copy's source is no destination of its map, so its stored value is its
        // value before the map (masked sources keep theirs; vbar sources are only zeroed
        // under zero_vbar, which zeroes this point too)
        const int cu = F.bc_ubar[f], cv = F.bc_vbar[f];
        if (cu == kZero) up = T(0);
        else if (cu == kClamp) up = T(P.clamp);
        else if (cu >= 0) up = F.ubar_t2[cu];
        if (cv == kZero) vp = T(0);
        else if (cv == kClamp) vp = T(P.clamp);
        else if (cv >= 0) vp = F.vbar_t2[cv];
    }
    F.ubar_t2[f] = up;
    F.vbar_t2[f] = P.zero_vbar ? T(0) : vp;
    F.ubar_t1[f] = am3(F.ubar_t1[f], T(P.Dt), F.rubar_t1[f], F.rubar_t0[f], F.vbar_t1[f]);
}

// Tile origin: shift the tiling by one point where the last tile would otherwise be
// a single row or column (the boundary maps copy from the neighbouring row/column,
// which must then lie in the same tile).
void tile_origin(int szI, int szJ, int *off_i, int *off_j)
{
    *off_i = (szI % TI == 1) ? 1 : 0;
    *off_j = (szJ % TJ == 1) ? 1 : 0;
}

template <typename T>
int launch(sf_ctx *c, const sf_step2d_fields *a, double dtfast, double cff1, double clamp, int flags,
           int parts, hipStream_t s)
{
    Fields<T> F;
    F.h = (const T *)a->h;
    F.on_u = (const T *)a->on_u;
    F.om_v = (const T *)a->om_v;
    F.pm = (const T *)a->pm;
    F.pn = (const T *)a->pn;
    F.zeta_t0 = (const T *)a->zeta[0];
    F.zeta_t1 = (const T *)a->zeta[1];
    F.zeta_t2 = (T *)a->zeta[2];
    F.ubar_t1 = (T *)a->ubar[1];
    F.ubar_t2 = (T *)a->ubar[2];
    F.vbar_t1 = (const T *)a->vbar[1];
    F.vbar_t2 = (T *)a->vbar[2];
    F.rzeta_t0 = (const T *)a->rzeta[0];
    F.rzeta_t1 = (T *)a->rzeta[1];
    F.rzeta_t2 = (const T *)a->rzeta[2];
    F.rubar_t0 = (const T *)a->rubar[0];
    F.rubar_t1 = (T *)a->rubar[1];
    F.rvbar_t1 = (const T *)a->rvbar[1];
    F.Zt_avg1 = a->Zt_avg1;
    F.DU_avg1 = a->DU_avg1;
    F.DV_avg1 = a->DV_avg1;
    F.bc_zeta = c->bc_zeta;
    F.bc_ubar = c->bc_ubar;
    F.bc_vbar = c->bc_vbar;
    Params P;
    P.szI = c->szI;
    P.szJ = c->szJ;
    P.q5_stride = c->q5_stride;
    tile_origin(c->szI, c->szJ, &P.off_i, &P.off_j);
    P.tiles_i = (c->szI + P.off_i + TI - 1) / TI;
    P.tiles_j = (c->szJ + P.off_j + TJ - 1) / TJ;
    P.S2 = c->S2;
    P.Dt = dtfast;
    P.cff1 = cff1;
    P.clamp = clamp;
    P.zero_vbar = (flags & SF_SUBSTEP_ZERO_VBAR) ? 1 : 0;
    P.fold_masked = (c->q5_stride == c->szI) ? 1 : 0;
    if (parts & 1) {
        // 11 bulk fields read and 5 written, 3 FP64 averages read and written
        const bool nts = (double)c->S2 * (16.0 * sizeof(T) + 6.0 * 8) > sf::kStreamingBytes;
        const dim3 grid(P.tiles_i * P.tiles_j);
        if (c->bc_vel_copy) {
            if (nts) hipLaunchKernelGGL((k_step2d_fused<T, true, true>), grid, dim3(NT), 0, s, F, P);
            else hipLaunchKernelGGL((k_step2d_fused<T, false, true>), grid, dim3(NT), 0, s, F, P);
        } else {
            if (nts) hipLaunchKernelGGL((k_step2d_fused<T, true, false>), grid, dim3(NT), 0, s, F, P);
            else hipLaunchKernelGGL((k_step2d_fused<T, false, false>), grid, dim3(NT), 0, s, F, P);
        }
        int rc = sf::check_launch("sf_step2d_substep (tiles)");
        if (rc != SF_OK) return rc;
    }
    if ((parts & 2) && !P.fold_masked) {
        const long long nm = (c->S2 + c->q5_stride - 1) / c->q5_stride + c->q5_stride - 1;
        hipLaunchKernelGGL(k_step2d_masked<T>, dim3(sf::grid1d(nm)), dim3(256), 0, s, F, P);
        return sf::check_launch("sf_step2d_substep (masked points)");
    }
    return SF_OK;
}

__global__ void k_fill_i32(int32_t *p, long long n, int32_t v)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

__global__ void k_scatter_codes(int32_t *p, const int32_t *dst, const int32_t *src, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int s = src[k];
    p[dst[k]] = s >= 0 ? s : (s == -1 ? kZero : kClamp);
}

}  // namespace

// Host-side validation of the composed maps against what the tiled kernel assumes.
extern "C" int sf_step2d_check_maps(int32_t szI, int32_t szJ, int32_t q5_stride, const int32_t *zd,
                                    const int32_t *zs, int32_t nz, const int32_t *ud, const int32_t *us,
                                    int32_t nu, const int32_t *vd, const int32_t *vs, int32_t nv)
{
    SF_REQUIRE(szI >= 3 && szJ >= 3 && q5_stride > 0, "sf_step2d_check_maps: bad sizes");
    SF_REQUIRE((nz == 0 || (zd && zs)) && (nu == 0 || (ud && us)) && (nv == 0 || (vd && vs)),
               "sf_step2d_check_maps: null map array");
    const long long S2 = (long long)szI * szJ;
    int off_i, off_j;
    tile_origin(szI, szJ, &off_i, &off_j);
    auto tix = [&](int i) { return (i + off_i) / TI; };
    auto tjx = [&](int j) { return (j + off_j) / TJ; };
    auto band = [&](int i, int j) { return i <= 1 || i >= szI - 2 || j <= 1 || j >= szJ - 2; };
    for (int k = 0; k < nz; k++) {
        const int d = zd[k], s = zs[k];
        if (d < 0 || d >= S2 || !band(d % szI, d / szI)) {
            sf::set_error("zeta map destination %d is not on the boundary band", d);
            return SF_ERR_UNSUPPORTED;
        }
        if (s < 0 || s >= S2) {
            sf::set_error("zeta map entry %d -> %d: constants are not supported by the fused substep", d, s);
            return SF_ERR_UNSUPPORTED;
        }
        const int di = d % szI, dj = d / szI, si = s % szI, sj = s / szI;
        if (abs(si - di) > 1 || abs(sj - dj) > 1 || tix(si) != tix(di) || tjx(sj) != tjx(dj)) {
            sf::set_error("zeta map source %d of %d is not a neighbour in the same %dx%d tile", s, d, TI, TJ);
            return SF_ERR_UNSUPPORTED;
        }
        // (a source on column 0 or row 0, e.g. (0,L) <- (0,L-1), is read only inside its
        // own tile; halo copies of such destinations are never consumed)
        // every tile whose 1-point halo contains the destination must also hold the
        // source on its halo
        const int ntj = (szJ + off_j + TJ - 1) / TJ, nti = (szI + off_i + TI - 1) / TI;
        for (int tj = tjx(dj > 0 ? dj - 1 : 0); tj <= tjx(dj + 1) && tj < ntj; tj++)
            for (int ti = tix(di > 0 ? di - 1 : 0); ti <= tix(di + 1) && ti < nti; ti++) {
                const int x0 = ti * TI - off_i, y0 = tj * TJ - off_j;
                const int a0 = x0 - 1, a1 = x0 + TI, b0 = y0 - 1, b1 = y0 + TJ;
                const bool dIn = di >= a0 && di <= a1 && dj >= b0 && dj <= b1;
                const bool sIn = si >= a0 && si <= a1 && sj >= b0 && sj <= b1;
                if (dIn && !sIn) {
                    sf::set_error("zeta map source %d of %d falls outside the halo of tile (%d,%d)", s, d, ti, tj);
                    return SF_ERR_UNSUPPORTED;
                }
            }
    }
    // Velocity maps: constants anywhere on the band; copies (gradient boundaries) from a
    // neighbour that is no destination of the same map, and that lies in every stage-E
    // region (tile + right column + top row: [x0, x0+TI] x [y0, y0+TJ]) that holds the
    // destination, where the kernel reads its value before the map from LDS.
    const int32_t *dd[2] = {ud, vd}, *ss[2] = {us, vs};
    const int nn[2] = {nu, nv};
    const int nti = (szI + off_i + TI - 1) / TI, ntj = (szJ + off_j + TJ - 1) / TJ;
    std::vector<char> isdst((size_t)S2);
    for (int q = 0; q < 2; q++) {
        std::fill(isdst.begin(), isdst.end(), 0);
        for (int k = 0; k < nn[q]; k++)
            if (dd[q][k] >= 0 && dd[q][k] < S2) isdst[dd[q][k]] = 1;
        for (int k = 0; k < nn[q]; k++) {
            const int d = dd[q][k], s = ss[q][k];
            if (d < 0 || d >= S2 || !band(d % szI, d / szI)) {
                sf::set_error("velocity map destination %d is not on the boundary band", d);
                return SF_ERR_UNSUPPORTED;
            }
            if (s < 0) continue;   // a constant
            const int di = d % szI, dj = d / szI, si = s % szI, sj = s / szI;
            if (s >= S2 || abs(si - di) > 1 || abs(sj - dj) > 1 || isdst[s]) {
                sf::set_error("velocity map entry %d <- %d: the fused substep copies only from a neighbour "
                              "that the map leaves unchanged", d, s);
                return SF_ERR_UNSUPPORTED;
            }
            for (int tj = tjx(dj) - 1; tj <= tjx(dj); tj++)
                for (int ti = tix(di) - 1; ti <= tix(di); ti++) {
                    if (ti < 0 || tj < 0 || ti >= nti || tj >= ntj) continue;
                    const int x0 = ti * TI - off_i, y0 = tj * TJ - off_j;
                    const bool dIn = di >= x0 && di <= x0 + TI && dj >= y0 && dj <= y0 + TJ;
                    const bool sIn = si >= x0 && si <= x0 + TI && sj >= y0 && sj <= y0 + TJ;
                    if (dIn && !sIn) {
                        sf::set_error("velocity map source %d of %d is outside the region of tile (%d,%d)", s, d,
                                      ti, tj);
                        return SF_ERR_UNSUPPORTED;
                    }
                }
        }
    }
    if (q5_stride != szI && (szI % q5_stride) != 0) {
        sf::set_error("fused substep: the i%%szJ mask admits column-0 points on a %dx%d grid; "
                      "use the one-kernel-per-step path or SF_FLAG_FIX_Q5_MASK", szI, szJ);
        return SF_ERR_UNSUPPORTED;
    }
    return SF_OK;
}

extern "C" int sf_ctx_set_bc_maps(sf_ctx *c, const int32_t *zeta_dst, const int32_t *zeta_src, int32_t nz,
                                  const int32_t *ubar_dst, const int32_t *ubar_src, int32_t nu,
                                  const int32_t *vbar_dst, const int32_t *vbar_src, int32_t nv)
{
    SF_REQUIRE(c, "sf_ctx_set_bc_maps: null context");
    SF_REQUIRE(nz >= 0 && nu >= 0 && nv >= 0, "sf_ctx_set_bc_maps: negative size");
    SF_REQUIRE((nz == 0 || (zeta_dst && zeta_src)) && (nu == 0 || (ubar_dst && ubar_src)) &&
                   (nv == 0 || (vbar_dst && vbar_src)),
               "sf_ctx_set_bc_maps: null map array");
    c->bc_ready = 0;
    int rc = sf_step2d_check_maps(c->szI, c->szJ, c->q5_stride, zeta_dst, zeta_src, nz, ubar_dst, ubar_src, nu,
                                  vbar_dst, vbar_src, nv);
    if (rc != SF_OK) return rc;
    c->bc_vel_copy = 0;
    for (int k = 0; k < nu; k++) c->bc_vel_copy |= ubar_src[k] >= 0;
    for (int k = 0; k < nv; k++) c->bc_vel_copy |= vbar_src[k] >= 0;
    hipError_t e;
    int32_t **fields[3] = {&c->bc_zeta, &c->bc_ubar, &c->bc_vbar};
    for (auto p : fields)
        if (!*p) {
            e = hipMalloc((void **)p, sizeof(int32_t) * c->S2);
            if (e != hipSuccess) return sf::hip_fail(e, "sf_ctx_set_bc_maps: hipMalloc");
        }
    const int32_t *ds[3] = {zeta_dst, ubar_dst, vbar_dst}, *sr[3] = {zeta_src, ubar_src, vbar_src};
    const int n[3] = {nz, nu, nv};
    int32_t *tmp = nullptr;
    const int nmax = nz > nu ? (nz > nv ? nz : nv) : (nu > nv ? nu : nv);
    if (nmax > 0) {
        e = hipMalloc((void **)&tmp, sizeof(int32_t) * 2 * nmax);
        if (e != hipSuccess) return sf::hip_fail(e, "sf_ctx_set_bc_maps: hipMalloc");
    }
    for (int q = 0; q < 3; q++) {
        int32_t *dptr = *fields[q];
        hipLaunchKernelGGL(k_fill_i32, dim3(sf::grid1d(c->S2)), dim3(256), 0, 0, dptr, c->S2, kKeep);
        if (n[q] > 0) {
            e = hipMemcpy(tmp, ds[q], sizeof(int32_t) * n[q], hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(tmp + nmax, sr[q], sizeof(int32_t) * n[q], hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                (void)hipFree(tmp);
                return sf::hip_fail(e, "sf_ctx_set_bc_maps: copy");
            }
            hipLaunchKernelGGL(k_scatter_codes, dim3(sf::grid1d(n[q])), dim3(256), 0, 0, dptr, tmp, tmp + nmax, n[q]);
            e = hipDeviceSynchronize();
            if (e != hipSuccess) {
                (void)hipFree(tmp);
                return sf::hip_fail(e, "sf_ctx_set_bc_maps: scatter");
            }
        }
    }
    e = hipDeviceSynchronize();
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess) return sf::hip_fail(e, "sf_ctx_set_bc_maps");
    c->bc_ready = 1;
    return SF_OK;
}

extern "C" int sf_step2d_substep_parts(sf_ctx *c, const sf_step2d_fields *a, int32_t precision,
                                       double dtfast, double cff1, double clamp_val, int32_t flags,
                                       int32_t parts, void *stream)
{
    SF_REQUIRE(parts >= 1 && parts <= 3, "sf_step2d_substep_parts: parts=%d not in 1..3", parts);
    SF_REQUIRE(c && a, "sf_step2d_substep: null argument");
    SF_REQUIRE(c->bc_ready, "sf_step2d_substep: boundary maps not set (sf_ctx_set_bc_maps)");
    SF_REQUIRE(a->h && a->on_u && a->om_v && a->pm && a->pn, "sf_step2d_substep: null grid field");
    for (int k = 0; k < 3; k++)
        SF_REQUIRE(a->zeta[k] && a->rzeta[k], "sf_step2d_substep: null zeta/rzeta slot");
    SF_REQUIRE(a->ubar[1] && a->ubar[2] && a->vbar[1] && a->vbar[2] && a->rubar[0] && a->rubar[1] &&
                   a->rvbar[1] && a->Zt_avg1 && a->DU_avg1 && a->DV_avg1,
               "sf_step2d_substep: null state field");
    if (precision == SF_FP64)
        return launch<double>(c, a, dtfast, cff1, clamp_val, flags, parts, sf::as_stream(stream));
    if (precision == SF_FP32)
        return launch<float>(c, a, dtfast, cff1, clamp_val, flags, parts, sf::as_stream(stream));
    sf::set_error("sf_step2d_substep: unknown precision %d", precision);
    return SF_ERR_ARG;
}

extern "C" int sf_step2d_substep(sf_ctx *c, const sf_step2d_fields *a, int32_t precision, double dtfast,
                                 double cff1, double clamp_val, int32_t flags, void *stream)
{
    return sf_step2d_substep_parts(c, a, precision, dtfast, cff1, clamp_val, flags, 3, stream);
}

extern "C" int sf_step2d_tile_shape(int32_t *ti, int32_t *tj)
{
    SF_REQUIRE(ti && tj, "sf_step2d_tile_shape: null argument");
    *ti = TI;
    *tj = TJ;
    return SF_OK;
}
