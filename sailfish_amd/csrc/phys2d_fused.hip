// phys2d_fused.hip -- the physics-mode barotropic substep (ROMS LF-AM3 in 2D form,
// obsolete/step2d_LF_AM3.h.py:917-2780) as ONE launch per substep.
//
// Each workgroup owns an output tile of TI x TJ points (all fields: zeta, ubar,
// vbar, their right-hand sides and the FP64 time averages of the tile's points).
// It loads a frame of (TI+5) x (TJ+5) points (3 to the low side, 2 to the high
// side) and evaluates the four stages of the substep on shrinking regions, so that
// no intermediate state ever goes to HBM:
//
//   stage  region (frame coords, core = [3, 3+T))   writes (LDS plane)
//   L      [0, T+5)    z1                               Z1
//   A      [1, T+5)    DU1 = D*u*on_u, DV1 = D*v*om_v   DU, DV
//   P1     [1, T+4)    predicted zeta zs, its RHS rz    ZS, RZ
//   P2     [2, T+4)    predicted ubar us, vbar vs       US, VS
//   C0     [2, T+4)    DUs, DVs of the predicted state  DU, DV (reused)
//   C1     [2, T+3)    corrected zeta z2 (+ averages)   Z1 (reused as Z2)
//   C2     [3, T+3)    corrected ubar, vbar             -
//
// Seven planes of (TI+5)(TJ+5) values.  The core points come first in every stage
// with the same thread mapping, so a thread keeps its core points' values (z1, u1,
// v1, zs, the predictor right-hand sides) in registers from stage to stage.  Fields
// that are only read (h, pm, pn, fomn, the stresses, the old time level, the
// history right-hand sides) are read from global memory where they are used; after
// their first touch those reads hit L1/L2.
//
// Closed-basin boundary: a ghost rho point takes the value of its clamped interior
// source, computed by the same thread from the same stencil; u = 0 at i <= 1 and
// i = L, v = 0 at j <= 1 and j = M; the remaining ghost rows/columns of u and v
// copy their inward neighbour.  The tile origin is shifted by one point when
// sz % T == 1 so that the source of every ghost u/v point is a core point of the
// same tile (its predictor right-hand side was written by the same workgroup).
//
// Expression order is the one of phys2d_kernels.hip; with contraction off the FP64
// instance equals the oracle bit for bit and each precision equals the four-launch
// form bit for bit.
#include "sf_internal.h"

namespace {

template <typename T>
struct PFA {
    const T *h, *pm, *pn, *fomn, *sustr, *svstr;
    const T *z0, *z1, *u0, *u1, *v0, *v1;
    T *z2, *u2, *v2;
    const T *rz0, *ru0, *rv0;
    T *rz1, *ru1, *rv1;
    double *Zt, *DUa, *DVa;
    int szI, szJ, tiles_i, tiles_j, off_i, off_j;
    T g, rdrg, dt;
    int first;
    double w1;
};

// Visit the points of region [3-LO, 3+TI+HI) x [3-LO, 3+TJ+HI) of the frame: the
// TI*TJ core points first (slot k >= 0, point tid + k*NT in row-major core order),
// then the ring around the core (slot -1).
template <int TI, int TJ, int NT, int LO, int HI, typename F>
__device__ __forceinline__ void for_region(int tid, F &&fn)
{
    constexpr int CORE = TI * TJ;
    static_assert(CORE % NT == 0, "core must be a multiple of the block");
#pragma unroll
    for (int k = 0; k < CORE / NT; k++) {
        const int idx = tid + k * NT;
        fn(k, 3 + idx % TI, 3 + idx / TI);
    }
    constexpr int W = TI + LO + HI, H = TJ + LO + HI;
    constexpr int RING = W * H - CORE;
    if constexpr (RING > 0) {
        constexpr int NBOT = LO * W, WROW = LO + HI, NMID = TJ * WROW;
#pragma unroll
        for (int k = 0; k < (RING + NT - 1) / NT; k++) {
            const int t = tid + k * NT;
            if (t < RING) {
                int a, b;
                if (t < NBOT) {
                    b = 3 - LO + t / W;
                    a = 3 - LO + t % W;
                } else if (t < NBOT + NMID) {
                    const int s = t - NBOT;
                    b = 3 + s / WROW;
                    const int c = s % WROW;
                    a = c < LO ? 3 - LO + c : 3 + TI + (c - LO);
                } else {
                    const int s = t - NBOT - NMID;
                    b = 3 + TJ + s / W;
                    a = 3 - LO + s % W;
                }
                fn(-1, a, b);
            }
        }
    }
}

template <int TI, int TJ>
struct PhysLds {
    static constexpr int FW = TI + 5, FH = TJ + 5, FN = FW * FH, PLANES = 7;
};

// Element (f + d) of p, addressed as the uniform base p + d plus a 32-bit byte offset:
// every neighbour of a point then shares one offset register (global_load saddr form).
template <typename T>
__device__ __forceinline__ T ldo(const T *p, long long d, uint32_t boff)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p + d) + boff);
}
template <typename T>
__device__ __forceinline__ void sto(T *p, uint32_t boff, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + boff) = v;
}

// One output tile (ti, tj) of grid P; sm = PLANES x FN values of LDS.
template <typename T, int TI, int TJ, int NT>
__device__ __forceinline__ void phys_tile(const PFA<T> &P, int ti, int tj, T (*sm)[PhysLds<TI, TJ>::FN])
{
    constexpr int FW = TI + 5;
    constexpr int KC = TI * TJ / NT;
    constexpr uint32_t ES = sizeof(T);
    T *const sZ1 = sm[0], *const sZ2 = sm[0];
    T *const sDU = sm[1], *const sDV = sm[2];
    T *const sZS = sm[3], *const sRZ = sm[4], *const sUS = sm[5], *const sVS = sm[6];

    const int szI = P.szI, szJ = P.szJ, L = szI - 1, M = szJ - 1;
    const long long W = szI;
    const int fi0 = ti * TI - P.off_i - 3, fj0 = tj * TJ - P.off_j - 3;   // frame origin
    const int tid = threadIdx.x;
    const T dt = P.dt;
    const bool first = P.first != 0;

    auto inside = [&](int i, int j) { return i >= 0 && i <= L && j >= 0 && j <= M; };
    auto clampi = [](int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); };
    auto off = [&](int i, int j) { return (uint32_t)(j * szI + i) * ES; };
    auto offd = [&](int i, int j) { return (uint32_t)(j * szI + i) * 8u; };

    // metric factors of a u point / v point / rho point at byte offset o
    auto onu_at = [&](uint32_t o) { return T(2) / (ldo(P.pn, -1, o) + ldo(P.pn, 0, o)); };
    auto omu_at = [&](uint32_t o) { return T(2) / (ldo(P.pm, -1, o) + ldo(P.pm, 0, o)); };
    auto omv_at = [&](uint32_t o) { return T(2) / (ldo(P.pm, -W, o) + ldo(P.pm, 0, o)); };
    auto onv_at = [&](uint32_t o) { return T(2) / (ldo(P.pn, -W, o) + ldo(P.pn, 0, o)); };
    auto cffu_at = [&](uint32_t o) {
        return (ldo(P.pm, 0, o) + ldo(P.pm, -1, o)) * (ldo(P.pn, 0, o) + ldo(P.pn, -1, o));
    };
    auto cffv_at = [&](uint32_t o) {
        return (ldo(P.pm, 0, o) + ldo(P.pm, -W, o)) * (ldo(P.pn, 0, o) + ldo(P.pn, -W, o));
    };

    // core-point registers: state, predictor right-hand sides and metric factors
    T z1c[KC], u1c[KC], v1c[KC], zsc[KC], ruc[KC], rvc[KC];
    T onuc[KC], omuc[KC], cffuc[KC], omvc[KC], onvc[KC], cffvc[KC], pmnc[KC];

    // ---- L: z1 over the frame ----------------------------------------------------
    for_region<TI, TJ, NT, 3, 2>(tid, [&](int k, int a, int b) {
        const int i = fi0 + a, j = fj0 + b;
        const T v = inside(i, j) ? ldo(P.z1, 0, off(i, j)) : T(0);
        sZ1[b * FW + a] = v;
        if (k >= 0) z1c[k] = v;
    });
    __syncthreads();

    // ---- A: DU1 at u points, DV1 at v points (+ core metric factors) ---------------
    for_region<TI, TJ, NT, 2, 2>(tid, [&](int k, int a, int b) {
        const int i = fi0 + a, j = fj0 + b;
        const int o = b * FW + a;
        T du = T(0), dv = T(0), u1 = T(0), v1 = T(0);
        if (inside(i, j)) {
            const uint32_t f = off(i, j);
            u1 = ldo(P.u1, 0, f);
            v1 = ldo(P.v1, 0, f);
            const T h0 = ldo(P.h, 0, f);
            if (i >= 1) {
                const T onu = onu_at(f);
                du = ((T(0.5) * ((ldo(P.h, -1, f) + sZ1[o - 1]) + (h0 + sZ1[o]))) * u1) * onu;
                if (k >= 0) {
                    onuc[k] = onu;
                    omuc[k] = omu_at(f);
                    cffuc[k] = cffu_at(f);
                }
            }
            if (j >= 1) {
                const T omv = omv_at(f);
                dv = ((T(0.5) * ((ldo(P.h, -W, f) + sZ1[o - FW]) + (h0 + sZ1[o]))) * v1) * omv;
                if (k >= 0) {
                    omvc[k] = omv;
                    onvc[k] = onv_at(f);
                    cffvc[k] = cffv_at(f);
                }
            }
            if (k >= 0) pmnc[k] = ldo(P.pm, 0, f) * ldo(P.pn, 0, f);
        }
        sDU[o] = du;
        sDV[o] = dv;
        if (k >= 0) {
            u1c[k] = u1;
            v1c[k] = v1;
        }
    });
    __syncthreads();

    // ---- P1: predicted zeta and its right-hand side -----------------------------
    for_region<TI, TJ, NT, 2, 1>(tid, [&](int k, int a, int b) {
        const int i = fi0 + a, j = fj0 + b;
        const int o = b * FW + a;
        T zs = T(0), r = T(0);
        if (inside(i, j)) {
            const int qi = clampi(i, 1, L - 1), qj = clampi(j, 1, M - 1);
            const bool own = qi == i && qj == j;
            const int qo = o + (qj - j) * FW + (qi - i);
            const uint32_t fq = off(qi, qj);
            r = (sDU[qo] - sDU[qo + 1]) + (sDV[qo] - sDV[qo + FW]);
            const T pmn = (k >= 0 && own) ? pmnc[k] : ldo(P.pm, 0, fq) * ldo(P.pn, 0, fq);
            if (first) zs = sZ1[qo] + (pmn * dt) * r;
            else zs = ldo(P.z0, 0, fq) + (pmn * (T(2) * dt)) * r;
            if (k >= 0 && own) sto(P.rz1, fq, r);
        }
        sZS[o] = zs;
        sRZ[o] = r;
        if (k >= 0) zsc[k] = zs;
    });
    __syncthreads();

    // ---- P2: predicted ubar / vbar ------------------------------------------------
    const T cff4 = T(4.0 / 25.0), cff5 = T(1) - T(2) * cff4;
    const T c1dt = first ? T(0.5) * dt : dt;
    // pressure surface of the predictor at LDS index o, global neighbour d of offset f
    auto gz_pred = [&](int o, long long d, uint32_t f) {
        return first ? T(0.5) * (sZ1[o] + sZS[o]) : cff5 * sZ1[o] + cff4 * (ldo(P.z0, d, f) + sZS[o]);
    };
    // depth of the "before" state of the predictor
    auto Db = [&](int o, long long d, uint32_t f) {
        return first ? ldo(P.h, d, f) + sZ1[o] : ldo(P.h, d, f) + ldo(P.z0, d, f);
    };

    for_region<TI, TJ, NT, 1, 1>(tid, [&](int k, int a, int b) {
        const int i = fi0 + a, j = fj0 + b;
        const int o = b * FW + a;
        T us = T(0), vs = T(0);
        if (inside(i, j)) {
            // ubar
            if (i >= 2 && i <= L - 1) {
                const int qj = clampi(j, 1, M - 1);
                const bool own = k >= 0 && qj == j;
                const int q = o + (qj - j) * FW;
                const uint32_t f = off(i, qj);
                const T onu = own ? onuc[k] : onu_at(f);
                const T omu = own ? omuc[k] : omu_at(f);
                const T g0 = gz_pred(q - 1, -1, f), g1 = gz_pred(q, 0, f);
                const T hm = ldo(P.h, -1, f), hc = ldo(P.h, 0, f);
                T r = ((T(0.5) * P.g) * onu) * ((hm + hc) * (g0 - g1) + (g0 * g0 - g1 * g1));
                const T um = ldo(P.u1, -1, f), uc = ldo(P.u1, 0, f), up = ldo(P.u1, 1, f);
                const T un = ldo(P.u1, W, f), usn = ldo(P.u1, -W, f);
                const T UFx1 = T(0.25) * (sDU[q] + sDU[q + 1]) * (uc + up);
                const T UFx0 = T(0.25) * (sDU[q - 1] + sDU[q]) * (um + uc);
                const T UFe1 = T(0.25) * (sDV[q + FW] + sDV[q + FW - 1]) * (un + uc);
                const T UFe0 = T(0.25) * (sDV[q] + sDV[q - 1]) * (uc + usn);
                r = r - ((UFx1 - UFx0) + (UFe1 - UFe0));
                const T cr1 = ((T(0.5) * (hc + sZ1[q])) * ldo(P.fomn, 0, f)) * (ldo(P.v1, 0, f) + ldo(P.v1, W, f));
                const T cr0 = ((T(0.5) * (hm + sZ1[q - 1])) * ldo(P.fomn, -1, f)) *
                              (ldo(P.v1, -1, f) + ldo(P.v1, W - 1, f));
                r = r + T(0.5) * (cr1 + cr0);
                r = r - ((P.rdrg * uc) * omu) * onu;
                r = r + (ldo(P.sustr, 0, f) * omu) * onu;
                if (own) {
                    sto(P.ru1, f, r);
                    ruc[k] = r;
                }
                const T cff = own ? cffuc[k] : cffu_at(f);
                const T fac = T(1) / ((hc + sZS[q]) + (hm + sZS[q - 1]));
                const T ub = first ? uc : ldo(P.u0, 0, f);
                us = (ub * (Db(q, 0, f) + Db(q - 1, -1, f)) + (cff * c1dt) * r) * fac;
            }
            // vbar
            if (j >= 2 && j <= M - 1) {
                const int qi = clampi(i, 1, L - 1);
                const bool own = k >= 0 && qi == i;
                const int q = o + (qi - i);
                const uint32_t f = off(qi, j);
                const T omv = own ? omvc[k] : omv_at(f);
                const T onv = own ? onvc[k] : onv_at(f);
                const T g0 = gz_pred(q - FW, -W, f), g1 = gz_pred(q, 0, f);
                const T hs = ldo(P.h, -W, f), hc = ldo(P.h, 0, f);
                T r = ((T(0.5) * P.g) * omv) * ((hs + hc) * (g0 - g1) + (g0 * g0 - g1 * g1));
                const T vc = ldo(P.v1, 0, f), vp = ldo(P.v1, 1, f), vm = ldo(P.v1, -1, f);
                const T vn = ldo(P.v1, W, f), vsn = ldo(P.v1, -W, f);
                const T VFx1 = T(0.25) * (sDU[q + 1] + sDU[q + 1 - FW]) * (vp + vc);
                const T VFx0 = T(0.25) * (sDU[q] + sDU[q - FW]) * (vc + vm);
                const T VFe1 = T(0.25) * (sDV[q] + sDV[q + FW]) * (vc + vn);
                const T VFe0 = T(0.25) * (sDV[q - FW] + sDV[q]) * (vsn + vc);
                r = r - ((VFx1 - VFx0) + (VFe1 - VFe0));
                const T cr1 = ((T(0.5) * (hc + sZ1[q])) * ldo(P.fomn, 0, f)) * (ldo(P.u1, 0, f) + ldo(P.u1, 1, f));
                const T cr0 = ((T(0.5) * (hs + sZ1[q - FW])) * ldo(P.fomn, -W, f)) *
                              (ldo(P.u1, -W, f) + ldo(P.u1, 1 - W, f));
                r = r - T(0.5) * (cr1 + cr0);
                r = r - ((P.rdrg * vc) * omv) * onv;
                r = r + (ldo(P.svstr, 0, f) * omv) * onv;
                if (own) {
                    sto(P.rv1, f, r);
                    rvc[k] = r;
                }
                const T cff = own ? cffvc[k] : cffv_at(f);
                const T fac = T(1) / ((hc + sZS[q]) + (hs + sZS[q - FW]));
                const T vb = first ? vc : ldo(P.v0, 0, f);
                vs = (vb * (Db(q, 0, f) + Db(q - FW, -W, f)) + (cff * c1dt) * r) * fac;
            }
        }
        sUS[o] = us;
        sVS[o] = vs;
    });
    __syncthreads();

    // ---- C0: fluxes DUs, DVs of the predicted state ----------------------------------
    T duc[KC], dvc[KC];
    for_region<TI, TJ, NT, 1, 1>(tid, [&](int k, int a, int b) {
        const int i = fi0 + a, j = fj0 + b;
        const int o = b * FW + a;
        T du = T(0), dv = T(0);
        if (inside(i, j)) {
            const uint32_t f = off(i, j);
            const T h0 = ldo(P.h, 0, f);
            if (i >= 1) {
                const T onu = k >= 0 ? onuc[k] : onu_at(f);
                du = ((T(0.5) * ((ldo(P.h, -1, f) + sZS[o - 1]) + (h0 + sZS[o]))) * sUS[o]) * onu;
            }
            if (j >= 1) {
                const T omv = k >= 0 ? omvc[k] : omv_at(f);
                dv = ((T(0.5) * ((ldo(P.h, -W, f) + sZS[o - FW]) + (h0 + sZS[o]))) * sVS[o]) * omv;
            }
        }
        sDU[o] = du;
        sDV[o] = dv;
        if (k >= 0) {
            duc[k] = du;
            dvc[k] = dv;
        }
    });
    __syncthreads();

    // ---- C1: corrected zeta, time averages ------------------------------------------
    const T a1 = dt * T(5.0) / T(12.0), a2 = dt * T(8.0) / T(12.0), a3 = dt * T(1.0) / T(12.0);
    for_region<TI, TJ, NT, 1, 0>(tid, [&](int k, int a, int b) {
        const int i = fi0 + a, j = fj0 + b;
        const int o = b * FW + a;
        T z2 = T(0);
        if (inside(i, j)) {
            const int qi = clampi(i, 1, L - 1), qj = clampi(j, 1, M - 1);
            const bool own = k >= 0 && qi == i && qj == j;
            const int qo = o + (qj - j) * FW + (qi - i);
            const uint32_t fq = off(qi, qj);
            const T r = (sDU[qo] - sDU[qo + 1]) + (sDV[qo] - sDV[qo + FW]);
            const T pmn = own ? pmnc[k] : ldo(P.pm, 0, fq) * ldo(P.pn, 0, fq);
            const T z1q = own ? z1c[k] : ldo(P.z1, 0, fq);
            if (first) z2 = z1q + pmn * (dt * r);
            else z2 = z1q + pmn * ((a1 * r + a2 * sRZ[qo]) - a3 * ldo(P.rz0, 0, fq));
            if (k >= 0) {
                sto(P.z2, off(i, j), z2);
                if (own) {
                    const double w = P.w1;
                    const uint32_t fd = offd(i, j);
                    sto(P.Zt, fd, ldo(P.Zt, 0, fd) + w * (double)zsc[k]);
                    sto(P.DUa, fd, ldo(P.DUa, 0, fd) + w * (double)duc[k]);
                    sto(P.DVa, fd, ldo(P.DVa, 0, fd) + w * (double)dvc[k]);
                }
            }
        }
        // Z2 reuses the Z1 plane: every reader of Z1 finished before the last barrier
        sZ2[o] = z2;
    });
    __syncthreads();

    // ---- C2: corrected ubar / vbar on the core -----------------------------------------
    const T cf4 = T(2.0 / 5.0), cf5 = T(1) - cf4;
    const T b1 = T(0.5) * dt * T(5.0) / T(12.0), b2 = T(0.5) * dt * T(8.0) / T(12.0),
            b3 = T(0.5) * dt * T(1.0) / T(12.0);
    auto gz_corr = [&](int o) { return cf5 * sZ2[o] + cf4 * sZS[o]; };
    for_region<TI, TJ, NT, 0, 0>(tid, [&](int k, int a, int b) {
        const int i = fi0 + a, j = fj0 + b;
        if (!inside(i, j)) return;
        const int o = b * FW + a;
        const uint32_t fo = off(i, j);
        // ubar
        T u2 = T(0);
        if (i >= 2 && i <= L - 1) {
            const int qj = clampi(j, 1, M - 1);
            const bool own = qj == j;
            const int q = o + (qj - j) * FW;
            const uint32_t f = off(i, qj);
            const T onu = own ? onuc[k] : onu_at(f);
            const T omu = own ? omuc[k] : omu_at(f);
            const T g0 = gz_corr(q - 1), g1 = gz_corr(q);
            const T hm = ldo(P.h, -1, f), hc = ldo(P.h, 0, f);
            T r = ((T(0.5) * P.g) * onu) * ((hm + hc) * (g0 - g1) + (g0 * g0 - g1 * g1));
            const T uc = sUS[q];
            const T UFx1 = T(0.25) * (sDU[q] + sDU[q + 1]) * (uc + sUS[q + 1]);
            const T UFx0 = T(0.25) * (sDU[q - 1] + sDU[q]) * (sUS[q - 1] + uc);
            const T UFe1 = T(0.25) * (sDV[q + FW] + sDV[q + FW - 1]) * (sUS[q + FW] + uc);
            const T UFe0 = T(0.25) * (sDV[q] + sDV[q - 1]) * (uc + sUS[q - FW]);
            r = r - ((UFx1 - UFx0) + (UFe1 - UFe0));
            const T cr1 = ((T(0.5) * (hc + sZS[q])) * ldo(P.fomn, 0, f)) * (sVS[q] + sVS[q + FW]);
            const T cr0 = ((T(0.5) * (hm + sZS[q - 1])) * ldo(P.fomn, -1, f)) * (sVS[q - 1] + sVS[q - 1 + FW]);
            r = r + T(0.5) * (cr1 + cr0);
            r = r - ((P.rdrg * uc) * omu) * onu;
            r = r + (ldo(P.sustr, 0, f) * omu) * onu;
            const T cff = own ? cffuc[k] : cffu_at(f);
            const T fac = T(1) / ((hc + sZ2[q]) + (hm + sZ2[q - 1]));
            T tend;
            if (first) tend = (T(0.5) * dt) * r;
            else tend = (b1 * r + b2 * (own ? ruc[k] : ldo((const T *)P.ru1, 0, f))) - b3 * ldo(P.ru0, 0, f);
            const T z1q = own ? z1c[k] : ldo(P.z1, 0, f);
            const T u1q = own ? u1c[k] : ldo(P.u1, 0, f);
            u2 = (u1q * ((hc + z1q) + (hm + ldo(P.z1, -1, f))) + cff * tend) * fac;
        }
        sto(P.u2, fo, u2);
        // vbar
        T v2 = T(0);
        if (j >= 2 && j <= M - 1) {
            const int qi = clampi(i, 1, L - 1);
            const bool own = qi == i;
            const int q = o + (qi - i);
            const uint32_t f = off(qi, j);
            const T omv = own ? omvc[k] : omv_at(f);
            const T onv = own ? onvc[k] : onv_at(f);
            const T g0 = gz_corr(q - FW), g1 = gz_corr(q);
            const T hs = ldo(P.h, -W, f), hc = ldo(P.h, 0, f);
            T r = ((T(0.5) * P.g) * omv) * ((hs + hc) * (g0 - g1) + (g0 * g0 - g1 * g1));
            const T vc = sVS[q];
            const T VFx1 = T(0.25) * (sDU[q + 1] + sDU[q + 1 - FW]) * (sVS[q + 1] + vc);
            const T VFx0 = T(0.25) * (sDU[q] + sDU[q - FW]) * (vc + sVS[q - 1]);
            const T VFe1 = T(0.25) * (sDV[q] + sDV[q + FW]) * (vc + sVS[q + FW]);
            const T VFe0 = T(0.25) * (sDV[q - FW] + sDV[q]) * (sVS[q - FW] + vc);
            r = r - ((VFx1 - VFx0) + (VFe1 - VFe0));
            const T cr1 = ((T(0.5) * (hc + sZS[q])) * ldo(P.fomn, 0, f)) * (sUS[q] + sUS[q + 1]);
            const T cr0 = ((T(0.5) * (hs + sZS[q - FW])) * ldo(P.fomn, -W, f)) * (sUS[q - FW] + sUS[q - FW + 1]);
            r = r - T(0.5) * (cr1 + cr0);
            r = r - ((P.rdrg * vc) * omv) * onv;
            r = r + (ldo(P.svstr, 0, f) * omv) * onv;
            const T cff = own ? cffvc[k] : cffv_at(f);
            const T fac = T(1) / ((hc + sZ2[q]) + (hs + sZ2[q - FW]));
            T tend;
            if (first) tend = (T(0.5) * dt) * r;
            else tend = (b1 * r + b2 * (own ? rvc[k] : ldo((const T *)P.rv1, 0, f))) - b3 * ldo(P.rv0, 0, f);
            const T z1q = own ? z1c[k] : ldo(P.z1, 0, f);
            const T v1q = own ? v1c[k] : ldo(P.v1, 0, f);
            v2 = (v1q * ((hc + z1q) + (hs + ldo(P.z1, -W, f))) + cff * tend) * fac;
        }
        sto(P.v2, fo, v2);
    });
}

// Block id -> tile: each XCD (block id mod 8) walks a contiguous band of tiles so that
// neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_band(int nblk)
{
    int b = blockIdx.x;
    if ((nblk & 7) == 0) b = (blockIdx.x & 7) * (nblk >> 3) + (blockIdx.x >> 3);
    return b;
}

template <typename T, int TI, int TJ, int NT>
__global__ __launch_bounds__(NT) void k_phys_fused(PFA<T> P)
{
    __shared__ T sm[PhysLds<TI, TJ>::PLANES][PhysLds<TI, TJ>::FN];
    const int b = xcd_band(P.tiles_i * P.tiles_j);
    phys_tile<T, TI, TJ, NT>(P, b % P.tiles_i, b / P.tiles_i, sm);
}

template <typename T, int TI, int TJ>
int launch_fused(const sf_phys_fields *x, double dt, int first, double w1, void *stream)
{
    constexpr int NT = 256;
    PFA<T> a;
    a.h = (const T *)x->h;
    a.pm = (const T *)x->pm;
    a.pn = (const T *)x->pn;
    a.fomn = (const T *)x->fomn;
    a.sustr = (const T *)x->sustr;
    a.svstr = (const T *)x->svstr;
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
    a.Zt = x->Zt_avg1;
    a.DUa = x->DU_avg1;
    a.DVa = x->DV_avg1;
    a.szI = x->szI;
    a.szJ = x->szJ;
    a.off_i = (x->szI % TI == 1) ? 1 : 0;
    a.off_j = (x->szJ % TJ == 1) ? 1 : 0;
    a.tiles_i = (x->szI + a.off_i + TI - 1) / TI;
    a.tiles_j = (x->szJ + a.off_j + TJ - 1) / TJ;
    a.g = T(x->g);
    a.rdrg = T(x->rdrg);
    a.dt = T(dt);
    a.first = first;
    a.w1 = w1;
    SF_REQUIRE((long long)x->szI * x->szJ < (1LL << 29), "sf_phys_substep_fused: grid of %dx%d points exceeds 2^29",
               x->szI, x->szJ);
    const long long nb = (long long)a.tiles_i * a.tiles_j;
    SF_REQUIRE(nb < (1LL << 31), "sf_phys_substep_fused: grid too large");
    hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT>), dim3((unsigned)nb), dim3(NT), 0, sf::as_stream(stream), a);
    return sf::check_launch("sf_phys_substep_fused");
}

template <typename T>
int dispatch_tj(const sf_phys_fields *x, double dt, int first, double w1, int tj, void *stream)
{
    switch (tj) {
    case 4: return launch_fused<T, 64, 4>(x, dt, first, w1, stream);
    case 8: return launch_fused<T, 64, 8>(x, dt, first, w1, stream);
    case 16: return launch_fused<T, 64, 16>(x, dt, first, w1, stream);
    default: break;
    }
    sf::set_error("sf_phys_substep_fused: unsupported tile height %d (4, 8 or 16)", tj);
    return SF_ERR_ARG;
}

}  // namespace

extern "C" int sf_phys_fused_default_tile(int32_t precision)
{
    return precision == SF_FP64 ? 8 : 16;
}

extern "C" int sf_phys_substep_fused(const sf_phys_fields *x, int32_t precision, double dt, int32_t first,
                                     double cff1, int32_t tile_j, void *stream)
{
    SF_REQUIRE(x, "sf_phys_substep_fused: null argument");
    SF_REQUIRE(x->szI >= 4 && x->szJ >= 4, "sf_phys_substep_fused: grid %dx%d too small", x->szI, x->szJ);
    SF_REQUIRE(x->h && x->pm && x->pn && x->fomn && x->sustr && x->svstr, "sf_phys_substep_fused: null grid field");
    for (int k = 0; k < 3; k++)
        SF_REQUIRE(x->zeta[k] && x->ubar[k] && x->vbar[k], "sf_phys_substep_fused: null state slot");
    SF_REQUIRE(x->rzeta[0] && x->rzeta[1] && x->rubar[0] && x->rubar[1] && x->rvbar[0] && x->rvbar[1],
               "sf_phys_substep_fused: null rhs slot");
    SF_REQUIRE(x->Zt_avg1 && x->DU_avg1 && x->DV_avg1, "sf_phys_substep_fused: null average");
    const int tj = tile_j > 0 ? tile_j : sf_phys_fused_default_tile(precision);
    if (precision == SF_FP64) return dispatch_tj<double>(x, dt, first, cff1, tj, stream);
    if (precision == SF_FP32) return dispatch_tj<float>(x, dt, first, cff1, tj, stream);
    sf::set_error("sf_phys_substep_fused: unknown precision %d", precision);
    return SF_ERR_ARG;
}
