// phys2d_fused_impl.h -- kernels of the physics-mode barotropic substep (ROMS LF-AM3
// in 2D form, obsolete/step2d_LF_AM3.h.py:917-2780) as ONE launch per substep, for a
// single basin (k_phys_fused) and for an ensemble of basins (k_phys_ensemble).
// Included by phys2d_fused.hip (FP64 instances, C ABI) and phys2d_fused_f32.hip
// (FP32 instances).
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
// Eleven planes of (TI+5)(TJ+5) values: the seven above plus ubar, vbar (t1), pm and
// pn over the whole frame, filled in stage L, so that every stencil neighbour of
// those fields is an LDS read.  The core points come first in every stage with the
// same thread mapping, so a thread keeps its core points' values (z1, u1, v1, zs,
// the predictor right-hand sides, metric factors) in registers from stage to stage.
// The other read-only fields (h, fomn, the stresses, the old time level, the history
// right-hand sides) are read from global memory where they are used.
//
// Closed-basin boundary: a ghost rho point takes the value of its clamped interior
// source, computed by the same thread from the same stencil; u = 0 at i <= 1 and
// i = L, v = 0 at j <= 1 and j = M; the remaining ghost rows/columns of u and v
// copy their inward neighbour.  The tile origin is shifted by one point when
// sz % T == 1 so that the source of every ghost u/v point is a core point of the
// same tile (its predictor right-hand side was written by the same workgroup).
//
// Expression order is the one of phys2d_kernels.hip: the FP64 instances (contraction
// off) equal the oracle and the four-launch form bit for bit; the FP32 instances are
// compiled with contraction on and agree with the four-launch FP32 form to rounding.
#pragma once

#include "mlp_net.h"
#include "sf_internal.h"

#include <vector>

// the learned closure of a fused substep: the network and its Coriolis field
struct ClosureIn {
    const sf_mlp *mlp;
    const void *f;
};

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
    int nts;   // streaming (non-temporal) stores and average loads: set when one substep's
               // bytes exceed what the 256 MB last-level cache keeps for the next substep
    double w1;
    sf_couple_fields c;   // coupling with the baroclinic step (CPL instances only)
    // the learned closure evaluated in the tile (CLO instances only): network operands,
    // input normalisation {mu[45], 1/sd[45]}, output de-normalisation, Coriolis field
    const uint16_t *cl_wpk;
    const float *cl_w, *cl_norm;
    const T *cl_f;
    float cl_ynorm[4];
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
    static constexpr int FW = TI + 5, FH = TJ + 5, FN = FW * FH, PLANES = 11;
};

// FP64 closure instances keep the network's feature tiles in frame planes 1..6 (written
// only from stage A on, after a barrier that ends the closure phase): 27 KB less LDS, so
// two 64 x 4 workgroups fit a CU.  FP32 planes are too small; the tiles get their own array.
template <typename T, bool CLO>
constexpr bool cl_tiles_in_frame = CLO && sizeof(T) == 8;

// Element (f + d) of p, given the byte offset boff of element f: the address is the
// array's scalar base plus a 32-bit vector offset (global_load saddr form).  The
// neighbour offset boff + d*sizeof(T) is computed in 32-bit arithmetic, so it is one
// vector add per neighbour and point, shared by every array read at that neighbour.
// All field pointers point to global memory; the explicit address space keeps the
// global_load/global_store forms also for pointers read from a member table.
#define SF_GLOBAL __attribute__((address_space(1)))
#define SF_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ T ldo(const T *p, long long d, uint32_t boff)
{
    const uint32_t o = boff + (uint32_t)(d * (long long)sizeof(T));
    return *(const SF_GLOBAL T *)((const SF_GLOBAL char *)p + o);
}
// The outputs of a substep are read by no other workgroup of the launch, and the FP64
// averages are read once per point.  When a substep moves more bytes than the last-level
// cache keeps between launches, streaming (non-temporal) stores and average loads keep L2
// for the halo reads; below that size plain ones let the next substep find its inputs in
// that cache.  NTS is a template parameter: a runtime choice between the two store forms
// would be merged into one plain store by the compiler.
template <bool NTS>
__device__ __forceinline__ double ldavg(const double *p, uint32_t boff)
{
    if constexpr (NTS) return __builtin_nontemporal_load((const SF_GLOBAL double *)((const SF_GLOBAL char *)p + boff));
    return ldo(p, 0, boff);
}
template <bool NTS, typename T>
__device__ __forceinline__ void sto(T *p, uint32_t boff, T v)
{
    if constexpr (NTS) __builtin_nontemporal_store(v, (SF_GLOBAL T *)((SF_GLOBAL char *)p + boff));
    else *(SF_GLOBAL T *)((SF_GLOBAL char *)p + boff) = v;
}

// Workgroup barrier that orders LDS only: global loads issued before it stay in
// flight (the stages exchange data through LDS exclusively).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Map ring index t of region [3-LO, 3+TI+HI) x [3-LO, 3+TJ+HI) minus the core to (a, b).
template <int TI, int TJ, int LO, int HI>
__device__ __forceinline__ void ring_point(int t, int &a, int &b)
{
    constexpr int W = TI + LO + HI, NBOT = LO * W, WROW = LO + HI, NMID = TJ * WROW;
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
}
template <int TI, int TJ, int LO, int HI>
constexpr int ring_size() { return (TI + LO + HI) * (TJ + LO + HI) - TI * TJ; }

// inputs of the ubar / vbar right-hand side at a u / v point
template <typename T>
struct UVIn {
    T m1, m2, cff;              // u: onu, omu ; v: omv, onv
    T pmc, pmb, pnc, pnb;       // raw pm, pn at the point and the back neighbour (generic path)
    T hb, hc;                   // h at the back neighbour (i-1 / j-1) and the point
    T xb, xc, xf, yf, yb;       // same-component velocity: back, centre, forward, cross+, cross-
    T fc, fb;                   // fomn at the point and the back neighbour
    T oc, of, ob, obf;          // other component for Coriolis (see rhs_u / rhs_v)
    T st;                       // surface stress
    T x0, z0c, z0b;             // predictor "before" velocity and zeta (t0) at point / back
};

// The tile's arguments.  Closure instances (launched only as k_phys_fused, whose one kernel
// argument is P) read them through the kernel-argument segment pointer, which an empty asm
// statement re-derives after the closure phase; other instances use P itself.
template <typename T, bool CLO>
__device__ __forceinline__ auto arg_ptr(const PFA<T> &P)
{
    if constexpr (CLO) return (const __attribute__((address_space(4))) PFA<T> *)__builtin_amdgcn_kernarg_segment_ptr();
    else return &P;
}

template <typename T, int TI, int TJ, int NT, bool FIRST, bool NTS, bool CPL = false, bool CLO = false>
__device__ __forceinline__ void phys_tile(const PFA<T> &P, int ti, int tj, T (*sm)[PhysLds<TI, TJ>::FN],
                                          T *side, double *sideF = nullptr, uint32_t *clX = nullptr,
                                          float *clN = nullptr, T *clS = nullptr, T *clH = nullptr,
                                          T *clF = nullptr)
{
    // the arguments through a pointer that the closure instances re-derive after the closure
    // phase, so that pointers first needed from stage A on are loaded there again instead of
    // being held across the phase
    auto Pp = arg_ptr<T, CLO>(P);
    constexpr int FW = TI + 5;
    constexpr int KC = TI * TJ / NT;
    constexpr uint32_t ES = sizeof(T);
    T *const sZ1 = sm[0], *const sZ2 = sm[0];
    T *const sDU = sm[1], *const sDV = sm[2];
    T *const sZS = sm[3], *const sRZ = sm[4], *const sUS = sm[5], *const sVS = sm[6];
    T *const sU1 = sm[7], *const sV1 = sm[8];   // ubar, vbar at t1 over the frame (read-only after L)
    T *const sPM = sm[9], *const sPN = sm[10];  // pm, pn over the frame (read-only after L)
    // predictor right-hand sides of the ghost-source row (j = 1, M-1) / column (i = 1, L-1)
    T *const sRUlo = side, *const sRUhi = side + TI, *const sRVlo = side + 2 * TI, *const sRVhi = side + 2 * TI + TJ;
    // CPL: the forcing F of the same ghost-source points (formed in P2 on the first substep)
    double *const sFUlo = sideF, *const sFUhi = sideF + TI, *const sFVlo = sideF + 2 * TI,
                 *const sFVhi = sideF + 2 * TI + TJ;

    const int szI = Pp->szI, szJ = Pp->szJ, L = szI - 1, M = szJ - 1;
    const long long W = szI;
    const int fi0 = ti * TI - Pp->off_i - 3, fj0 = tj * TJ - Pp->off_j - 3;   // frame origin
    const int tid = threadIdx.x;
    const T dt = Pp->dt;
    constexpr bool first = FIRST;

    auto inside = [&](int i, int j) { return i >= 0 && i <= L && j >= 0 && j <= M; };
    auto clampi = [](int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); };
    auto off = [&](int i, int j) { return (uint32_t)(j * szI + i) * ES; };
    auto offd = [&](int i, int j) { return (uint32_t)(j * szI + i) * 8u; };
    auto core_ab = [&](int k, int &a, int &b) {
        const int idx = tid + k * NT;
        a = 3 + idx % TI;
        b = 3 + idx / TI;
    };

    // metric factors of a u point / v point at byte offset o
    auto onu_at = [&](uint32_t o) { return T(2) * sf::recip(ldo(Pp->pn, -1, o) + ldo(Pp->pn, 0, o)); };
    auto omu_at = [&](uint32_t o) { return T(2) * sf::recip(ldo(Pp->pm, -1, o) + ldo(Pp->pm, 0, o)); };
    auto omv_at = [&](uint32_t o) { return T(2) * sf::recip(ldo(Pp->pm, -W, o) + ldo(Pp->pm, 0, o)); };
    auto onv_at = [&](uint32_t o) { return T(2) * sf::recip(ldo(Pp->pn, -W, o) + ldo(Pp->pn, 0, o)); };
    auto cffu_at = [&](uint32_t o) {
        return (ldo(Pp->pm, 0, o) + ldo(Pp->pm, -1, o)) * (ldo(Pp->pn, 0, o) + ldo(Pp->pn, -1, o));
    };
    auto cffv_at = [&](uint32_t o) {
        return (ldo(Pp->pm, 0, o) + ldo(Pp->pm, -W, o)) * (ldo(Pp->pn, 0, o) + ldo(Pp->pn, -W, o));
    };

    // generic (load-on-demand) inputs of the u / v right-hand side at offset f
    auto load_u = [&](uint32_t f, int q) {
        UVIn<T> x;
        x.pmc = sPM[q];
        x.pmb = sPM[q - 1];
        x.pnc = sPN[q];
        x.pnb = sPN[q - 1];
        x.hb = ldo(Pp->h, -1, f);
        x.hc = ldo(Pp->h, 0, f);
        x.xb = sU1[q - 1];
        x.xc = sU1[q];
        x.xf = sU1[q + 1];
        x.yf = sU1[q + FW];
        x.yb = sU1[q - FW];
        x.fc = ldo(Pp->fomn, 0, f);
        x.fb = ldo(Pp->fomn, -1, f);
        x.oc = sV1[q];
        x.of = sV1[q + FW];
        x.ob = sV1[q - 1];
        x.obf = sV1[q + FW - 1];
        x.st = ldo(Pp->sustr, 0, f);
        x.x0 = first ? x.xc : ldo(Pp->u0, 0, f);
        x.z0c = first ? T(0) : ldo(Pp->z0, 0, f);
        x.z0b = first ? T(0) : ldo(Pp->z0, -1, f);
        return x;
    };
    auto load_v = [&](uint32_t f, int q) {
        UVIn<T> x;
        x.pmc = sPM[q];
        x.pmb = sPM[q - FW];
        x.pnc = sPN[q];
        x.pnb = sPN[q - FW];
        x.hb = ldo(Pp->h, -W, f);
        x.hc = ldo(Pp->h, 0, f);
        x.xb = sV1[q - FW];
        x.xc = sV1[q];
        x.xf = sV1[q + FW];
        x.yf = sV1[q + 1];
        x.yb = sV1[q - 1];
        x.fc = ldo(Pp->fomn, 0, f);
        x.fb = ldo(Pp->fomn, -W, f);
        x.oc = sU1[q];
        x.of = sU1[q + 1];
        x.ob = sU1[q - FW];
        x.obf = sU1[q + 1 - FW];
        x.st = ldo(Pp->svstr, 0, f);
        x.x0 = first ? x.xc : ldo(Pp->v0, 0, f);
        x.z0c = first ? T(0) : ldo(Pp->z0, 0, f);
        x.z0b = first ? T(0) : ldo(Pp->z0, -W, f);
        return x;
    };

    // metric factors of generic inputs (same expressions as onu_at / omu_at / cffu_at ...)
    auto metrics_u = [&](UVIn<T> &x) {
        x.m1 = T(2) * sf::recip(x.pnb + x.pnc);
        x.m2 = T(2) * sf::recip(x.pmb + x.pmc);
        x.cff = (x.pmc + x.pmb) * (x.pnc + x.pnb);
    };
    auto metrics_v = [&](UVIn<T> &x) {
        x.m1 = T(2) * sf::recip(x.pmb + x.pmc);
        x.m2 = T(2) * sf::recip(x.pnb + x.pnc);
        x.cff = (x.pmc + x.pmb) * (x.pnc + x.pnb);
    };

    // CPL: forcing of the predictor right-hand side r of a u (v) point at FP64 byte offset
    // fd: the first substep forms F = rufrc - r and adds its Adams-Bashforth
    // extrapolation over the baroclinic steps, the others add F (as pred_forcing of the
    // four-launch form); F is returned in Fk
    auto pred_forcing = [&](bool vdir, uint32_t fd, T r, double &Fk) {
        const double *Fo = vdir ? Pp->c.Fv : Pp->c.Fu;
        if constexpr (FIRST) {
            const double F = ldo(vdir ? Pp->c.rvfrc : Pp->c.rufrc, 0, fd) - (double)r;
            Fk = F;
            const double F1 = Pp->c.ab > 0 ? ldo(vdir ? Pp->c.Fv1 : Pp->c.Fu1, 0, fd) : 0.0;
            const double F2 = Pp->c.ab > 1 ? ldo(vdir ? Pp->c.Fv2 : Pp->c.Fu2, 0, fd) : 0.0;
            return r + T(sf::ab_extrapolate(Pp->c.ab, F, F1, F2));
        } else {
            Fk = ldo(Fo, 0, fd);
            return r + T(Fk);
        }
    };

    // ---- predictor momentum at LDS index q from inputs x (returns rhs in r) ----------
    const T cff4 = T(4.0 / 25.0), cff5 = T(1) - T(2) * cff4;
    const T c1dt = first ? T(0.5) * dt : dt;
    auto gzp = [&](int o, T z0) { return first ? T(0.5) * (sZ1[o] + sZS[o]) : cff5 * sZ1[o] + cff4 * (z0 + sZS[o]); };
    auto Dbp = [&](int o, T h, T z0) { return first ? h + sZ1[o] : h + z0; };
    auto pred_u = [&](int q, const UVIn<T> &x, T &r, uint32_t fd = 0, double *Fk = nullptr) {
        const T onu = x.m1, omu = x.m2;
        const T g0 = gzp(q - 1, x.z0b), g1 = gzp(q, x.z0c);
        r = ((T(0.5) * Pp->g) * onu) * ((x.hb + x.hc) * (g0 - g1) + (g0 * g0 - g1 * g1));
        const T UFx1 = T(0.25) * (sDU[q] + sDU[q + 1]) * (x.xc + x.xf);
        const T UFx0 = T(0.25) * (sDU[q - 1] + sDU[q]) * (x.xb + x.xc);
        const T UFe1 = T(0.25) * (sDV[q + FW] + sDV[q + FW - 1]) * (x.yf + x.xc);
        const T UFe0 = T(0.25) * (sDV[q] + sDV[q - 1]) * (x.xc + x.yb);
        r = r - ((UFx1 - UFx0) + (UFe1 - UFe0));
        const T cr1 = ((T(0.5) * (x.hc + sZ1[q])) * x.fc) * (x.oc + x.of);
        const T cr0 = ((T(0.5) * (x.hb + sZ1[q - 1])) * x.fb) * (x.ob + x.obf);
        r = r + T(0.5) * (cr1 + cr0);
        r = r - ((Pp->rdrg * x.xc) * omu) * onu;
        r = r + (x.st * omu) * onu;
        if constexpr (CPL) {
            double F;
            r = pred_forcing(false, fd, r, F);
            if (Fk) *Fk = F;
        }
        const T fac = sf::recip((x.hc + sZS[q]) + (x.hb + sZS[q - 1]));
        return (x.x0 * (Dbp(q, x.hc, x.z0c) + Dbp(q - 1, x.hb, x.z0b)) + (x.cff * c1dt) * r) * fac;
    };
    auto pred_v = [&](int q, const UVIn<T> &x, T &r, uint32_t fd = 0, double *Fk = nullptr) {
        const T omv = x.m1, onv = x.m2;
        const T g0 = gzp(q - FW, x.z0b), g1 = gzp(q, x.z0c);
        r = ((T(0.5) * Pp->g) * omv) * ((x.hb + x.hc) * (g0 - g1) + (g0 * g0 - g1 * g1));
        const T VFx1 = T(0.25) * (sDU[q + 1] + sDU[q + 1 - FW]) * (x.yf + x.xc);
        const T VFx0 = T(0.25) * (sDU[q] + sDU[q - FW]) * (x.xc + x.yb);
        const T VFe1 = T(0.25) * (sDV[q] + sDV[q + FW]) * (x.xc + x.xf);
        const T VFe0 = T(0.25) * (sDV[q - FW] + sDV[q]) * (x.xb + x.xc);
        r = r - ((VFx1 - VFx0) + (VFe1 - VFe0));
        const T cr1 = ((T(0.5) * (x.hc + sZ1[q])) * x.fc) * (x.oc + x.of);
        const T cr0 = ((T(0.5) * (x.hb + sZ1[q - FW])) * x.fb) * (x.ob + x.obf);
        r = r - T(0.5) * (cr1 + cr0);
        r = r - ((Pp->rdrg * x.xc) * omv) * onv;
        r = r + (x.st * omv) * onv;
        if constexpr (CPL) {
            double F;
            r = pred_forcing(true, fd, r, F);
            if (Fk) *Fk = F;
        }
        const T fac = sf::recip((x.hc + sZS[q]) + (x.hb + sZS[q - FW]));
        return (x.x0 * (Dbp(q, x.hc, x.z0c) + Dbp(q - FW, x.hb, x.z0b)) + (x.cff * c1dt) * r) * fac;
    };

    // ============ core-point registers ============
    bool cin[KC];
    uint32_t cf[KC];
    T z1c[KC], u1c[KC], v1c[KC], hc_[KC], hw_[KC], hs_[KC];
    T clf0[CLO ? KC : 1];                            // CLO: f at the core point
    T clsu[CLO ? KC : 1], clsv[CLO ? KC : 1];       // CLO: wind stress at the core point
    T pm_[KC], pn_[KC];
    T onuc[KC], omuc[KC], cffuc[KC], omvc[KC], onvc[KC], cffvc[KC], pmnc[KC];
    T zsc[KC], ruc[KC], rvc[KC], duc[KC], dvc[KC];
    // G1 (predictor) and G2 (corrector) prefetch
    T z0c_[KC], z0w_[KC], z0s_[KC], u0c_[KC], v0c_[KC], fo_[KC], fow_[KC], fos_[KC], su_[KC], sv_[KC];
    T rz0_[KC], ru0_[KC], rv0_[KC], z1w_[KC], z1s_[KC];
    double zt_[KC], dua_[KC], dva_[KC];
    // CPL: stage-A fluxes of the core points, their forcing F and DU_avg2 / DV_avg2
    double du1c[KC], dv1c[KC], fuc[KC], fvc[KC], du2_[KC], dv2_[KC];

    // fast-path predicates of a core point (its computations use only its own offset)
    auto rho_int = [&](int i, int j) { return i >= 1 && i <= L - 1 && j >= 1 && j <= M - 1; };
    auto u_int = [&](int i, int j) { return i >= 2 && i <= L - 1 && j >= 1 && j <= M - 1; };
    auto v_int = [&](int i, int j) { return i >= 1 && i <= L - 1 && j >= 2 && j <= M - 1; };

    if constexpr (CLO)   // input normalisation of the closure (read after L's barrier)
        if (tid < 2 * NIN) clN[tid] = Pp->cl_norm[tid];

    // ---- G0 + L: z1 over the frame; core statics -----------------------------------------
#pragma unroll
    for (int k = 0; k < KC; k++) {
        int a, b;
        core_ab(k, a, b);
        const int i = fi0 + a, j = fj0 + b;
        cin[k] = inside(i, j);
        cf[k] = cin[k] ? off(i, j) : 0u;
        const uint32_t f = cf[k];
        // values below are read only under the predicate they are loaded under
        z1c[k] = T(0);
        if (cin[k]) {
            z1c[k] = ldo(Pp->z1, 0, f);
            u1c[k] = ldo(Pp->u1, 0, f);
            v1c[k] = ldo(Pp->v1, 0, f);
            hc_[k] = ldo(Pp->h, 0, f);
            pm_[k] = ldo(Pp->pm, 0, f);
            pn_[k] = ldo(Pp->pn, 0, f);
            if (i >= 1) hw_[k] = ldo(Pp->h, -1, f);
            if (j >= 1) hs_[k] = ldo(Pp->h, -W, f);
        }
        if constexpr (CLO) {   // the closure's f and wind stress, in flight with the frame loads
            clf0[k] = cin[k] ? ldo(Pp->cl_f, 0, f) : T(0);
            clsu[k] = u_int(i, j) ? ldo(Pp->sustr, 0, f) : T(0);
            clsv[k] = v_int(i, j) ? ldo(Pp->svstr, 0, f) : T(0);
        }
    }
    {
        constexpr int RING = ring_size<TI, TJ, 3, 2>();
        constexpr int RS = (RING + NT - 1) / NT;
        T rz[RS], ru[RS], rv[RS], rm[RS], rn[RS];
        T rh[CLO ? RS : 1], rf[CLO ? RS : 1];   // CLO: h and f of the ring for the closure's stencils
        int ro[RS];
#pragma unroll
        for (int s = 0; s < RS; s++) {
            const int t = tid + s * NT;
            rz[s] = ru[s] = rv[s] = rm[s] = rn[s] = T(0);
            if constexpr (CLO) rh[s] = rf[s] = T(0);
            ro[s] = -1;
            if (t < RING) {
                int a, b;
                ring_point<TI, TJ, 3, 2>(t, a, b);
                const int i = fi0 + a, j = fj0 + b;
                ro[s] = b * FW + a;
                if (inside(i, j)) {
                    const uint32_t f = off(i, j);
                    rz[s] = ldo(Pp->z1, 0, f);
                    ru[s] = ldo(Pp->u1, 0, f);
                    rv[s] = ldo(Pp->v1, 0, f);
                    rm[s] = ldo(Pp->pm, 0, f);
                    rn[s] = ldo(Pp->pn, 0, f);
                    if constexpr (CLO) {
                        rh[s] = ldo(Pp->h, 0, f);
                        rf[s] = ldo(Pp->cl_f, 0, f);
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < KC; k++) {
            int a, b;
            core_ab(k, a, b);
            const int o = b * FW + a;
            sZ1[o] = z1c[k];
            sU1[o] = cin[k] ? u1c[k] : T(0);
            sV1[o] = cin[k] ? v1c[k] : T(0);
            sPM[o] = cin[k] ? pm_[k] : T(0);
            sPN[o] = cin[k] ? pn_[k] : T(0);
            if constexpr (CLO) {
                clH[o] = cin[k] ? hc_[k] : T(0);
                clF[o] = clf0[k];
            }
        }
#pragma unroll
        for (int s = 0; s < RS; s++)
            if (ro[s] >= 0) {
                sZ1[ro[s]] = rz[s];
                sU1[ro[s]] = ru[s];
                sV1[ro[s]] = rv[s];
                sPM[ro[s]] = rm[s];
                sPN[ro[s]] = rn[s];
                if constexpr (CLO) {
                    clH[ro[s]] = rh[s];
                    clF[ro[s]] = rf[s];
                }
            }
    }
    lds_barrier();

    // ---- CL: the learned closure's corrector-stage stresses at the core points ----------
    // Each wave evaluates the network on its core row(s) of 64 points (features from the
    // t1 frame planes and the h, f planes staged with them) and keeps
    // sustr + R_u D_u / svstr + R_v D_v of its points for C2; those of the ghost-source
    // rows / columns (j = 1, M-1 / i = 1, L-1)
    // also go to clS for the ghost points of C2.  The same arithmetic as sf_closure_stress.
    T csu[KC], csv[KC];
    if constexpr (CLO) {
#pragma clang fp contract(off)   // the same roundings as sf_closure_stress in both precisions
        Fwd Wn;
        load_fwd(Wn, Pp->cl_wpk, Pp->cl_w);
        uint32_t *X = clX + (tid >> 6) * (64 * XS);
#pragma unroll
        for (int k = 0; k < KC; k++) {
            int a, b;
            core_ab(k, a, b);
            const int i = fi0 + a, j = fj0 + b, o = b * FW + a;
            const bool in = rho_int(i, j);
            const uint32_t f = cf[k];
            float xv[NIN];
#pragma unroll
            for (int kk = 0; kk < NIN; kk++) {
                const int fld = kk / 9, dj = (kk % 9) / 3 - 1, di = kk % 3 - 1;
                const int d = dj * FW + di;
                const T *pl = fld == 0 ? sU1 : fld == 1 ? sV1 : fld == 2 ? sZ1 : fld == 3 ? clH : clF;
                const T cv = pl[o], nv = pl[o + d];
                xv[kk] = (dj == 0 && di == 0) ? (float)cv : (float)(nv - cv);
            }
            closure_row(X, tid & 63, xv, clN);
            float zu, zv;
            closure_eval(Wn, X, zu, zv);
            csu[k] = csv[k] = T(0);   // read in C2 only at u / v interior points
            if (u_int(i, j)) {
                const T Du = T(0.5) * ((hw_[k] + sZ1[o - 1]) + (hc_[k] + sZ1[o]));
                csu[k] = clsu[k] + T(zu * Pp->cl_ynorm[1] + Pp->cl_ynorm[0]) * Du;
                if (j == 1) clS[a - 3] = csu[k];
                if (j == M - 1) clS[TI + a - 3] = csu[k];
            }
            if (v_int(i, j)) {
                const T Dv = T(0.5) * ((hs_[k] + sZ1[o - FW]) + (hc_[k] + sZ1[o]));
                csv[k] = clsv[k] + T(zv * Pp->cl_ynorm[3] + Pp->cl_ynorm[2]) * Dv;
                if (i == 1) clS[2 * TI + b - 3] = csv[k];
                if (i == L - 1) clS[2 * TI + TJ + b - 3] = csv[k];
            }
        }
    }

    if constexpr (cl_tiles_in_frame<T, CLO>) lds_barrier();   // feature tiles read; planes 1..6 free
    if constexpr (CLO) asm volatile("" : "+s"(Pp));

    // ---- A: DU1 at u points, DV1 at v points ---------------------------------------------
    {
        constexpr int RING = ring_size<TI, TJ, 2, 2>();
        constexpr int RS = (RING + NT - 1) / NT;
        T rh[RS], rhw[RS], rhs[RS];
        int ro[RS], ri[RS], rj[RS];
#pragma unroll
        for (int s = 0; s < RS; s++) {
            const int t = tid + s * NT;
            ro[s] = -1;
            ri[s] = rj[s] = -1;
            if (t < RING) {
                int a, b;
                ring_point<TI, TJ, 2, 2>(t, a, b);
                const int i = fi0 + a, j = fj0 + b;
                ro[s] = b * FW + a;
                if (inside(i, j)) {
                    ri[s] = i;
                    rj[s] = j;
                    const uint32_t f = off(i, j);
                    rh[s] = ldo(Pp->h, 0, f);
                    if (i >= 1) rhw[s] = ldo(Pp->h, -1, f);
                    if (j >= 1) rhs[s] = ldo(Pp->h, -W, f);
                }
            }
        }
        // core: metric factors, fluxes
#pragma unroll
        for (int k = 0; k < KC; k++) {
            int a, b;
            core_ab(k, a, b);
            const int i = fi0 + a, j = fj0 + b, o = b * FW + a;
            T du = T(0), dv = T(0);
            if (cin[k]) {
                pmnc[k] = pm_[k] * pn_[k];
                if (i >= 1) {
                    const T pmw = sPM[o - 1], pnw = sPN[o - 1];
                    onuc[k] = T(2) * sf::recip(pnw + pn_[k]);
                    omuc[k] = T(2) * sf::recip(pmw + pm_[k]);
                    cffuc[k] = (pm_[k] + pmw) * (pn_[k] + pnw);
                    du = ((T(0.5) * ((hw_[k] + sZ1[o - 1]) + (hc_[k] + sZ1[o]))) * u1c[k]) * onuc[k];
                }
                if constexpr (CPL) du1c[k] = (double)du;
                if (j >= 1) {
                    const T pms = sPM[o - FW], pns = sPN[o - FW];
                    omvc[k] = T(2) * sf::recip(pms + pm_[k]);
                    onvc[k] = T(2) * sf::recip(pns + pn_[k]);
                    cffvc[k] = (pm_[k] + pms) * (pn_[k] + pns);
                    dv = ((T(0.5) * ((hs_[k] + sZ1[o - FW]) + (hc_[k] + sZ1[o]))) * v1c[k]) * omvc[k];
                }
                if constexpr (CPL) dv1c[k] = (double)dv;
            }
            sDU[o] = du;
            sDV[o] = dv;
        }
        // ring fluxes
#pragma unroll
        for (int s = 0; s < RS; s++) {
            if (ro[s] < 0) continue;
            const int o = ro[s];
            T du = T(0), dv = T(0);
            if (ri[s] >= 1) {
                const T onu = T(2) * sf::recip(sPN[o - 1] + sPN[o]);
                du = ((T(0.5) * ((rhw[s] + sZ1[o - 1]) + (rh[s] + sZ1[o]))) * sU1[o]) * onu;
            }
            if (rj[s] >= 1) {
                const T omv = T(2) * sf::recip(sPM[o - FW] + sPM[o]);
                dv = ((T(0.5) * ((rhs[s] + sZ1[o - FW]) + (rh[s] + sZ1[o]))) * sV1[o]) * omv;
            }
            sDU[o] = du;
            sDV[o] = dv;
        }
    }
        // core: issue the predictor inputs (G1)
#pragma unroll
        for (int k = 0; k < KC; k++) {
            int a, b;
            core_ab(k, a, b);
            const int i = fi0 + a, j = fj0 + b;
            const uint32_t f = cf[k];
            const bool ui = cin[k] && u_int(i, j), vi = cin[k] && v_int(i, j), ri_ = cin[k] && rho_int(i, j);
            if (ri_ || ui || vi) {
                if (!first) z0c_[k] = ldo(Pp->z0, 0, f);
                fo_[k] = ldo(Pp->fomn, 0, f);
            }
            if (ui) {
                if (!first) {
                    z0w_[k] = ldo(Pp->z0, -1, f);
                    u0c_[k] = ldo(Pp->u0, 0, f);
                }
                fow_[k] = ldo(Pp->fomn, -1, f);
                su_[k] = ldo(Pp->sustr, 0, f);
            }
            if (vi) {
                if (!first) {
                    z0s_[k] = ldo(Pp->z0, -W, f);
                    v0c_[k] = ldo(Pp->v0, 0, f);
                }
                fos_[k] = ldo(Pp->fomn, -W, f);
                sv_[k] = ldo(Pp->svstr, 0, f);
            }
        }
    lds_barrier();

    // ---- P1: predicted zeta and its right-hand side -----------------------------------
    {
        constexpr int RING = ring_size<TI, TJ, 2, 1>();
        constexpr int RS = (RING + NT - 1) / NT;
        T rz0[RS];
        int ro[RS], rq[RS];
#pragma unroll
        for (int s = 0; s < RS; s++) {
            const int t = tid + s * NT;
            ro[s] = -1;
            rq[s] = 0;
            if (t < RING) {
                int a, b;
                ring_point<TI, TJ, 2, 1>(t, a, b);
                const int i = fi0 + a, j = fj0 + b;
                ro[s] = b * FW + a;
                if (inside(i, j)) {
                    const int qi = clampi(i, 1, L - 1), qj = clampi(j, 1, M - 1);
                    rq[s] = ro[s] + (qj - j) * FW + (qi - i);
                    if (!first) rz0[s] = ldo(Pp->z0, 0, off(qi, qj));
                } else {
                    ro[s] = -2 - ro[s];   // outside: store zeros
                }
            }
        }
        auto zs_of = [&](int qo, T pmn, T z0, T &r) {
            r = (sDU[qo] - sDU[qo + 1]) + (sDV[qo] - sDV[qo + FW]);
            if (first) return sZ1[qo] + (pmn * dt) * r;
            return z0 + (pmn * (T(2) * dt)) * r;
        };
#pragma unroll
        for (int k = 0; k < KC; k++) {
            int a, b;
            core_ab(k, a, b);
            const int i = fi0 + a, j = fj0 + b, o = b * FW + a;
            T zs = T(0), r = T(0);
            if (cin[k]) {
                if (rho_int(i, j)) {
                    zs = zs_of(o, pmnc[k], z0c_[k], r);
                    sto<NTS>(Pp->rz1, cf[k], r);
                } else {
                    const int qi = clampi(i, 1, L - 1), qj = clampi(j, 1, M - 1);
                    const uint32_t fq = off(qi, qj);
                    const int oq = o + (qj - j) * FW + (qi - i);
                    const T pmn = sPM[oq] * sPN[oq];
                    const T z0 = first ? T(0) : ldo(Pp->z0, 0, fq);
                    zs = zs_of(oq, pmn, z0, r);
                }
            }
            sZS[o] = zs;
            sRZ[o] = r;
            zsc[k] = zs;
        }
#pragma unroll
        for (int s = 0; s < RS; s++) {
            if (ro[s] == -1) continue;
            if (ro[s] < -1) {
                sZS[-2 - ro[s]] = T(0);
                sRZ[-2 - ro[s]] = T(0);
                continue;
            }
            T r;
            const T zs = zs_of(rq[s], sPM[rq[s]] * sPN[rq[s]], rz0[s], r);
            sZS[ro[s]] = zs;
            sRZ[ro[s]] = r;
        }
    }
    lds_barrier();

    // ---- P2: predicted ubar / vbar --------------------------------------------------
    {
#pragma unroll
        for (int k = 0; k < KC; k++) {
            int a, b;
            core_ab(k, a, b);
            const int i = fi0 + a, j = fj0 + b, o = b * FW + a;
            T us = T(0), vs = T(0);
            if (cin[k]) {
                if (u_int(i, j)) {
                    UVIn<T> x;
                    x.m1 = onuc[k];
                    x.m2 = omuc[k];
                    x.cff = cffuc[k];
                    x.hb = hw_[k];
                    x.hc = hc_[k];
                    x.xb = sU1[o - 1];
                    x.xc = u1c[k];
                    x.xf = sU1[o + 1];
                    x.yf = sU1[o + FW];
                    x.yb = sU1[o - FW];
                    x.fc = fo_[k];
                    x.fb = fow_[k];
                    x.oc = v1c[k];
                    x.of = sV1[o + FW];
                    x.ob = sV1[o - 1];
                    x.obf = sV1[o + FW - 1];
                    x.st = su_[k];
                    x.x0 = first ? u1c[k] : u0c_[k];
                    x.z0c = z0c_[k];
                    x.z0b = z0w_[k];
                    T r;
                    double F = 0.0;
                    us = pred_u(o, x, r, offd(i, j), &F);
                    sto<NTS>(Pp->ru1, cf[k], r);
                    ruc[k] = r;
                    if (j == 1) sRUlo[a - 3] = r;
                    if (j == M - 1) sRUhi[a - 3] = r;
                    if constexpr (CPL) {
                        fuc[k] = F;
                        if constexpr (FIRST) {
                            sto<NTS>(Pp->c.Fu, offd(i, j), F);
                            if (j == 1) sFUlo[a - 3] = F;
                            if (j == M - 1) sFUhi[a - 3] = F;
                        }
                    }
                } else if (i >= 2 && i <= L - 1) {   // ghost row: source (i, 1) or (i, M-1)
                    const int qj = clampi(j, 1, M - 1);
                    T r;
                    UVIn<T> x = load_u(off(i, qj), o + (qj - j) * FW);
                    metrics_u(x);
                    us = pred_u(o + (qj - j) * FW, x, r, offd(i, qj));
                }
                if (v_int(i, j)) {
                    UVIn<T> x;
                    x.m1 = omvc[k];
                    x.m2 = onvc[k];
                    x.cff = cffvc[k];
                    x.hb = hs_[k];
                    x.hc = hc_[k];
                    x.xb = sV1[o - FW];
                    x.xc = v1c[k];
                    x.xf = sV1[o + FW];
                    x.yf = sV1[o + 1];
                    x.yb = sV1[o - 1];
                    x.fc = fo_[k];
                    x.fb = fos_[k];
                    x.oc = u1c[k];
                    x.of = sU1[o + 1];
                    x.ob = sU1[o - FW];
                    x.obf = sU1[o + 1 - FW];
                    x.st = sv_[k];
                    x.x0 = first ? v1c[k] : v0c_[k];
                    x.z0c = z0c_[k];
                    x.z0b = z0s_[k];
                    T r;
                    double F = 0.0;
                    vs = pred_v(o, x, r, offd(i, j), &F);
                    sto<NTS>(Pp->rv1, cf[k], r);
                    rvc[k] = r;
                    if (i == 1) sRVlo[b - 3] = r;
                    if (i == L - 1) sRVhi[b - 3] = r;
                    if constexpr (CPL) {
                        fvc[k] = F;
                        if constexpr (FIRST) {
                            sto<NTS>(Pp->c.Fv, offd(i, j), F);
                            if (i == 1) sFVlo[b - 3] = F;
                            if (i == L - 1) sFVhi[b - 3] = F;
                        }
                    }
                } else if (j >= 2 && j <= M - 1) {   // ghost column: source (1, j) or (L-1, j)
                    const int qi = clampi(i, 1, L - 1);
                    T r;
                    UVIn<T> x = load_v(off(qi, j), o + (qi - i));
                    metrics_v(x);
                    vs = pred_v(o + (qi - i), x, r, offd(qi, j));
                }
            }
            sUS[o] = us;
            sVS[o] = vs;
        }
        // core: issue the corrector inputs (G2)
#pragma unroll
        for (int k = 0; k < KC; k++) {
            int a, b;
            core_ab(k, a, b);
            const int i = fi0 + a, j = fj0 + b, o = b * FW + a;
            const uint32_t f = cf[k];
            if (cin[k] && rho_int(i, j)) {
                const uint32_t fd = offd(i, j);
                if (!first) rz0_[k] = ldo(Pp->rz0, 0, f);
                zt_[k] = ldavg<NTS>(Pp->Zt, fd);
                dua_[k] = ldavg<NTS>(Pp->DUa, fd);
                dva_[k] = ldavg<NTS>(Pp->DVa, fd);
                if constexpr (CPL) {
                    du2_[k] = FIRST ? 0.0 : ldavg<NTS>(Pp->c.DU_avg2, fd);
                    dv2_[k] = FIRST ? 0.0 : ldavg<NTS>(Pp->c.DV_avg2, fd);
                }
            }
            if (cin[k] && u_int(i, j)) {
                if (!first) ru0_[k] = ldo(Pp->ru0, 0, f);
                z1w_[k] = sZ1[o - 1];   // the Z1 plane holds z1 until C1
            }
            if (cin[k] && v_int(i, j)) {
                if (!first) rv0_[k] = ldo(Pp->rv0, 0, f);
                z1s_[k] = sZ1[o - FW];
            }
        }
        // ring: inputs loaded on demand
        constexpr int RING = ring_size<TI, TJ, 1, 1>();
#pragma unroll
        for (int s = 0; s < (RING + NT - 1) / NT; s++) {
            const int t = tid + s * NT;
            if (t >= RING) continue;
            int a, b;
            ring_point<TI, TJ, 1, 1>(t, a, b);
            const int i = fi0 + a, j = fj0 + b, o = b * FW + a;
            T us = T(0), vs = T(0), r;
            if (inside(i, j)) {
                if (i >= 2 && i <= L - 1) {
                    const int qj = clampi(j, 1, M - 1);
                    UVIn<T> x = load_u(off(i, qj), o + (qj - j) * FW);
                    metrics_u(x);
                    us = pred_u(o + (qj - j) * FW, x, r, offd(i, qj));
                }
                if (j >= 2 && j <= M - 1) {
                    const int qi = clampi(i, 1, L - 1);
                    UVIn<T> x = load_v(off(qi, j), o + (qi - i));
                    metrics_v(x);
                    vs = pred_v(o + (qi - i), x, r, offd(qi, j));
                }
            }
            sUS[o] = us;
            sVS[o] = vs;
        }
    }
    lds_barrier();

    // ---- C0: fluxes DUs, DVs of the predicted state ------------------------------------
    {
        constexpr int RING = ring_size<TI, TJ, 1, 1>();
        constexpr int RS = (RING + NT - 1) / NT;
        T rh[RS], rhw[RS], rhs[RS];
        int ro[RS], ri[RS], rj[RS];
#pragma unroll
        for (int s = 0; s < RS; s++) {
            const int t = tid + s * NT;
            ro[s] = -1;
            ri[s] = rj[s] = -1;
            if (t < RING) {
                int a, b;
                ring_point<TI, TJ, 1, 1>(t, a, b);
                const int i = fi0 + a, j = fj0 + b;
                ro[s] = b * FW + a;
                if (inside(i, j)) {
                    ri[s] = i;
                    rj[s] = j;
                    const uint32_t f = off(i, j);
                    rh[s] = ldo(Pp->h, 0, f);
                    if (i >= 1) rhw[s] = ldo(Pp->h, -1, f);
                    if (j >= 1) rhs[s] = ldo(Pp->h, -W, f);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < KC; k++) {
            int a, b;
            core_ab(k, a, b);
            const int i = fi0 + a, j = fj0 + b, o = b * FW + a;
            T du = T(0), dv = T(0);
            if (cin[k]) {
                if (i >= 1) du = ((T(0.5) * ((hw_[k] + sZS[o - 1]) + (hc_[k] + sZS[o]))) * sUS[o]) * onuc[k];
                if (j >= 1) dv = ((T(0.5) * ((hs_[k] + sZS[o - FW]) + (hc_[k] + sZS[o]))) * sVS[o]) * omvc[k];
            }
            sDU[o] = du;
            sDV[o] = dv;
            duc[k] = du;
            dvc[k] = dv;
        }
#pragma unroll
        for (int s = 0; s < RS; s++) {
            if (ro[s] < 0) continue;
            const int o = ro[s];
            T du = T(0), dv = T(0);
            if (ri[s] >= 1) {
                const T onu = T(2) * sf::recip(sPN[o - 1] + sPN[o]);
                du = ((T(0.5) * ((rhw[s] + sZS[o - 1]) + (rh[s] + sZS[o]))) * sUS[o]) * onu;
            }
            if (rj[s] >= 1) {
                const T omv = T(2) * sf::recip(sPM[o - FW] + sPM[o]);
                dv = ((T(0.5) * ((rhs[s] + sZS[o - FW]) + (rh[s] + sZS[o]))) * sVS[o]) * omv;
            }
            sDU[o] = du;
            sDV[o] = dv;
        }
    }
    lds_barrier();

    // ---- C1: corrected zeta, time averages ------------------------------------------
    const T a1 = dt * T(5.0) / T(12.0), a2 = dt * T(8.0) / T(12.0), a3 = dt * T(1.0) / T(12.0);
    {
        auto z2_of = [&](int qo, T pmn, T z1q, T rz0) {
            const T r = (sDU[qo] - sDU[qo + 1]) + (sDV[qo] - sDV[qo + FW]);
            if (first) return z1q + pmn * (dt * r);
            return z1q + pmn * ((a1 * r + a2 * sRZ[qo]) - a3 * rz0);
        };
        constexpr int RING = ring_size<TI, TJ, 1, 0>();
        constexpr int RS = (RING + NT - 1) / NT;
        T rz1[RS], rrz0[RS];
        int ro[RS], rq[RS];
#pragma unroll
        for (int s = 0; s < RS; s++) {
            const int t = tid + s * NT;
            ro[s] = -1;
            rq[s] = 0;
            if (t < RING) {
                int a, b;
                ring_point<TI, TJ, 1, 0>(t, a, b);
                const int i = fi0 + a, j = fj0 + b;
                ro[s] = b * FW + a;
                if (inside(i, j)) {
                    const int qi = clampi(i, 1, L - 1), qj = clampi(j, 1, M - 1);
                    rq[s] = ro[s] + (qj - j) * FW + (qi - i);
                    const uint32_t fq = off(qi, qj);
                    rz1[s] = ldo(Pp->z1, 0, fq);
                    if (!first) rrz0[s] = ldo(Pp->rz0, 0, fq);
                } else {
                    ro[s] = -2 - ro[s];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < KC; k++) {
            int a, b;
            core_ab(k, a, b);
            const int i = fi0 + a, j = fj0 + b, o = b * FW + a;
            T z2 = T(0);
            if (cin[k]) {
                if (rho_int(i, j)) {
                    z2 = z2_of(o, pmnc[k], z1c[k], rz0_[k]);
                    const double w = Pp->w1;
                    const uint32_t fd = offd(i, j);
                    sto<NTS>(Pp->Zt, fd, zt_[k] + w * (double)zsc[k]);
                    sto<NTS>(Pp->DUa, fd, dua_[k] + w * (double)duc[k]);
                    sto<NTS>(Pp->DVa, fd, dva_[k] + w * (double)dvc[k]);
                    if constexpr (CPL) {
                        sto<NTS>(Pp->c.DU_avg2, fd, (du2_[k] + Pp->c.w2p * du1c[k]) + Pp->c.w2c * (double)duc[k]);
                        sto<NTS>(Pp->c.DV_avg2, fd, (dv2_[k] + Pp->c.w2p * dv1c[k]) + Pp->c.w2c * (double)dvc[k]);
                    }
                } else {
                    const int qi = clampi(i, 1, L - 1), qj = clampi(j, 1, M - 1);
                    const uint32_t fq = off(qi, qj);
                    const int oq = o + (qj - j) * FW + (qi - i);
                    const T pmn = sPM[oq] * sPN[oq];
                    const T rz0 = first ? T(0) : ldo(Pp->rz0, 0, fq);
                    z2 = z2_of(oq, pmn, ldo(Pp->z1, 0, fq), rz0);
                }
                sto<NTS>(Pp->z2, cf[k], z2);
            }
            // Z2 reuses the Z1 plane: every reader of Z1 finished before the last barrier
            sZ2[o] = z2;
        }
#pragma unroll
        for (int s = 0; s < RS; s++) {
            if (ro[s] == -1) continue;
            if (ro[s] < -1) {
                sZ2[-2 - ro[s]] = T(0);
                continue;
            }
            sZ2[ro[s]] = z2_of(rq[s], sPM[rq[s]] * sPN[rq[s]], rz1[s], rrz0[s]);
        }
    }
    lds_barrier();

    // ---- C2: corrected ubar / vbar on the core -----------------------------------------
    const T cf4 = T(2.0 / 5.0), cf5 = T(1) - cf4;
    const T b1 = T(0.5) * dt * T(5.0) / T(12.0), b2 = T(0.5) * dt * T(8.0) / T(12.0),
            b3 = T(0.5) * dt * T(1.0) / T(12.0);
    auto gzc = [&](int o) { return cf5 * sZ2[o] + cf4 * sZS[o]; };
    // corrected ubar at LDS index q: metrics m1 = onu, m2 = omu; statics hb/hc, fc/fb, st
    auto corr_u = [&](int q, T onu, T omu, T cff, T hb, T hc, T fc, T fb, T st, T ru1, T ru0, T z1c_, T z1b,
                      T u1c_, double F = 0.0) {
        const T g0 = gzc(q - 1), g1 = gzc(q);
        T r = ((T(0.5) * Pp->g) * onu) * ((hb + hc) * (g0 - g1) + (g0 * g0 - g1 * g1));
        const T uc = sUS[q];
        const T UFx1 = T(0.25) * (sDU[q] + sDU[q + 1]) * (uc + sUS[q + 1]);
        const T UFx0 = T(0.25) * (sDU[q - 1] + sDU[q]) * (sUS[q - 1] + uc);
        const T UFe1 = T(0.25) * (sDV[q + FW] + sDV[q + FW - 1]) * (sUS[q + FW] + uc);
        const T UFe0 = T(0.25) * (sDV[q] + sDV[q - 1]) * (uc + sUS[q - FW]);
        r = r - ((UFx1 - UFx0) + (UFe1 - UFe0));
        const T cr1 = ((T(0.5) * (hc + sZS[q])) * fc) * (sVS[q] + sVS[q + FW]);
        const T cr0 = ((T(0.5) * (hb + sZS[q - 1])) * fb) * (sVS[q - 1] + sVS[q - 1 + FW]);
        r = r + T(0.5) * (cr1 + cr0);
        r = r - ((Pp->rdrg * uc) * omu) * onu;
        r = r + (st * omu) * onu;
        if constexpr (CPL) r = r + T(F);
        const T fac = sf::recip((hc + sZ2[q]) + (hb + sZ2[q - 1]));
        T tend;
        if (first) tend = (T(0.5) * dt) * r;
        else tend = (b1 * r + b2 * ru1) - b3 * ru0;
        return (u1c_ * ((hc + z1c_) + (hb + z1b)) + cff * tend) * fac;
    };
    auto corr_v = [&](int q, T omv, T onv, T cff, T hb, T hc, T fc, T fb, T st, T rv1, T rv0, T z1c_, T z1b,
                      T v1c_, double F = 0.0) {
        const T g0 = gzc(q - FW), g1 = gzc(q);
        T r = ((T(0.5) * Pp->g) * omv) * ((hb + hc) * (g0 - g1) + (g0 * g0 - g1 * g1));
        const T vc = sVS[q];
        const T VFx1 = T(0.25) * (sDU[q + 1] + sDU[q + 1 - FW]) * (sVS[q + 1] + vc);
        const T VFx0 = T(0.25) * (sDU[q] + sDU[q - FW]) * (vc + sVS[q - 1]);
        const T VFe1 = T(0.25) * (sDV[q] + sDV[q + FW]) * (vc + sVS[q + FW]);
        const T VFe0 = T(0.25) * (sDV[q - FW] + sDV[q]) * (sVS[q - FW] + vc);
        r = r - ((VFx1 - VFx0) + (VFe1 - VFe0));
        const T cr1 = ((T(0.5) * (hc + sZS[q])) * fc) * (sUS[q] + sUS[q + 1]);
        const T cr0 = ((T(0.5) * (hb + sZS[q - FW])) * fb) * (sUS[q - FW] + sUS[q - FW + 1]);
        r = r - T(0.5) * (cr1 + cr0);
        r = r - ((Pp->rdrg * vc) * omv) * onv;
        r = r + (st * omv) * onv;
        if constexpr (CPL) r = r + T(F);
        const T fac = sf::recip((hc + sZ2[q]) + (hb + sZ2[q - FW]));
        T tend;
        if (first) tend = (T(0.5) * dt) * r;
        else tend = (b1 * r + b2 * rv1) - b3 * rv0;
        return (v1c_ * ((hc + z1c_) + (hb + z1b)) + cff * tend) * fac;
    };
#pragma unroll
    for (int k = 0; k < KC; k++) {
        int a, b;
        core_ab(k, a, b);
        const int i = fi0 + a, j = fj0 + b, o = b * FW + a;
        if (!cin[k]) continue;
        T u2 = T(0), v2 = T(0);
        if (u_int(i, j)) {
            u2 = corr_u(o, onuc[k], omuc[k], cffuc[k], hw_[k], hc_[k], fo_[k], fow_[k], CLO ? csu[k] : su_[k], ruc[k],
                        ru0_[k], z1c[k], z1w_[k], u1c[k], CPL ? fuc[k] : 0.0);
        } else if (i >= 2 && i <= L - 1) {   // ghost row j = 0 or M: source row 1 or M-1
            const int qj = clampi(j, 1, M - 1);
            const uint32_t f = off(i, qj);
            const T ru1 = qj == 1 ? sRUlo[a - 3] : sRUhi[a - 3];
            const T ru0 = first ? T(0) : ldo(Pp->ru0, 0, f);
            double F = 0.0;
            if constexpr (CPL) F = FIRST ? (qj == 1 ? sFUlo[a - 3] : sFUhi[a - 3]) : ldo(Pp->c.Fu, 0, offd(i, qj));
            const T st = CLO ? clS[(qj == 1 ? 0 : TI) + a - 3] : ldo(Pp->sustr, 0, f);
            u2 = corr_u(o + (qj - j) * FW, onu_at(f), omu_at(f), cffu_at(f), ldo(Pp->h, -1, f), ldo(Pp->h, 0, f),
                        ldo(Pp->fomn, 0, f), ldo(Pp->fomn, -1, f), st, ru1, ru0, ldo(Pp->z1, 0, f), ldo(Pp->z1, -1, f),
                        sU1[o + (qj - j) * FW], F);
        }
        sto<NTS>(Pp->u2, cf[k], u2);
        if (v_int(i, j)) {
            v2 = corr_v(o, omvc[k], onvc[k], cffvc[k], hs_[k], hc_[k], fo_[k], fos_[k], CLO ? csv[k] : sv_[k], rvc[k],
                        rv0_[k], z1c[k], z1s_[k], v1c[k], CPL ? fvc[k] : 0.0);
        } else if (j >= 2 && j <= M - 1) {   // ghost column i = 0 or L: source column 1 or L-1
            const int qi = clampi(i, 1, L - 1);
            const uint32_t f = off(qi, j);
            const T rv1 = qi == 1 ? sRVlo[b - 3] : sRVhi[b - 3];
            const T rv0 = first ? T(0) : ldo(Pp->rv0, 0, f);
            double F = 0.0;
            if constexpr (CPL) F = FIRST ? (qi == 1 ? sFVlo[b - 3] : sFVhi[b - 3]) : ldo(Pp->c.Fv, 0, offd(qi, j));
            const T st = CLO ? clS[2 * TI + (qi == 1 ? 0 : TJ) + b - 3] : ldo(Pp->svstr, 0, f);
            v2 = corr_v(o + (qi - i), omv_at(f), onv_at(f), cffv_at(f), ldo(Pp->h, -W, f), ldo(Pp->h, 0, f),
                        ldo(Pp->fomn, 0, f), ldo(Pp->fomn, -W, f), st, rv1, rv0, ldo(Pp->z1, 0, f), ldo(Pp->z1, -W, f),
                        sV1[o + (qi - i)], F);
        }
        sto<NTS>(Pp->v2, cf[k], v2);
    }
}

// Block id -> tile: each XCD (block id mod 8) walks a contiguous band of tiles so that
// neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_band(int nblk)
{
    int b = blockIdx.x;
    if ((nblk & 7) == 0) b = (blockIdx.x & 7) * (nblk >> 3) + (blockIdx.x >> 3);
    return b;
}

template <typename T, int TI, int TJ, int NT, bool FIRST, bool NTS, bool CPL = false, bool CLO = false>
__global__ __launch_bounds__(NT) void k_phys_fused(PFA<T> P)
{
    __shared__ T sm[PhysLds<TI, TJ>::PLANES][PhysLds<TI, TJ>::FN];
    __shared__ T side[2 * TI + 2 * TJ];
    __shared__ double sideF[CPL ? 2 * TI + 2 * TJ : 1];
    constexpr bool IN_FRAME = cl_tiles_in_frame<T, CLO>;
    static_assert(!IN_FRAME || 6 * PhysLds<TI, TJ>::FN * sizeof(T) >= (NT / 64) * 64 * XS * sizeof(uint32_t),
                  "frame planes 1..6 hold the feature tiles");
    __shared__ __attribute__((aligned(16))) uint32_t clX[CLO && !IN_FRAME ? (NT / 64) * 64 * XS : 1];
    __shared__ float clN[CLO ? 2 * NIN : 1];
    __shared__ T clS[CLO ? 2 * TI + 2 * TJ : 1];
    __shared__ T clH[CLO ? PhysLds<TI, TJ>::FN : 1], clF[CLO ? PhysLds<TI, TJ>::FN : 1];   // h, f over the frame
    const int b = xcd_band(P.tiles_i * P.tiles_j);
    phys_tile<T, TI, TJ, NT, FIRST, NTS, CPL, CLO>(P, b % P.tiles_i, b / P.tiles_i, sm, side, sideF,
                                                   IN_FRAME ? (uint32_t *)sm[1] : clX, clN, clS, clH, clF);
}


template <typename T, int TI, int TJ, int NT>
int launch_fused(const sf_phys_fields *x, double dt, int first, double w1, void *stream,
                 const sf_couple_fields *cp = nullptr, const ClosureIn *cl = nullptr)
{
    PFA<T> a;
    a.c = cp ? *cp : sf_couple_fields{};
    a.cl_wpk = cl ? cl->mlp->wpk : nullptr;
    a.cl_w = cl ? cl->mlp->w : nullptr;
    a.cl_norm = cl ? cl->mlp->norm : nullptr;
    a.cl_f = cl ? (const T *)cl->f : nullptr;
    for (int k = 0; k < 4; k++) a.cl_ynorm[k] = cl ? cl->mlp->ynorm[k] : 0.0f;
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
    a.nts = (double)x->szI * x->szJ * (3.0 * sizeof(T) * 7 + 8.0 * 6) > sf::kStreamingBytes ? 1 : 0;
    a.w1 = w1;
    SF_REQUIRE((long long)x->szI * x->szJ < (1LL << 29), "sf_phys_substep_fused: grid of %dx%d points exceeds 2^29",
               x->szI, x->szJ);
    const long long nb = (long long)a.tiles_i * a.tiles_j;
    SF_REQUIRE(nb < (1LL << 31), "sf_phys_substep_fused: grid too large");
    const dim3 grid((unsigned)nb), block(NT);
    hipStream_t s = sf::as_stream(stream);
    if constexpr (TJ <= 8) {
        if (cl) {
            SF_REQUIRE(!cp, "sf_phys_substep_closure_fused: coupling and closure together are not supported");
            if (first && a.nts) hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT, true, true, false, true>), grid, block, 0, s, a);
            else if (first) hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT, true, false, false, true>), grid, block, 0, s, a);
            else if (a.nts) hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT, false, true, false, true>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT, false, false, false, true>), grid, block, 0, s, a);
            return sf::check_launch("sf_phys_substep_closure_fused");
        }
        if (cp) {
            if (first && a.nts) hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT, true, true, true>), grid, block, 0, s, a);
            else if (first) hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT, true, false, true>), grid, block, 0, s, a);
            else if (a.nts) hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT, false, true, true>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT, false, false, true>), grid, block, 0, s, a);
            return sf::check_launch("sf_phys_substep_coupled_fused");
        }
    }
    if (cp || cl) {
        sf::set_error("%s: tile height %d (4 or 8)",
                      cp ? "sf_phys_substep_coupled_fused" : "sf_phys_substep_closure_fused", TJ);
        return SF_ERR_ARG;
    }
    if (first && a.nts) hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT, true, true>), grid, block, 0, s, a);
    else if (first) hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT, true, false>), grid, block, 0, s, a);
    else if (a.nts) hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT, false, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_phys_fused<T, TI, TJ, NT, false, false>), grid, block, 0, s, a);
    return sf::check_launch("sf_phys_substep_fused");
}

template <typename T>
int dispatch_tj(const sf_phys_fields *x, double dt, int first, double w1, int tj, void *stream,
                const sf_couple_fields *cp = nullptr, const ClosureIn *cl = nullptr)
{
    switch (tj) {
    case 4: return launch_fused<T, 64, 4, 256>(x, dt, first, w1, stream, cp, cl);
    case 8: return launch_fused<T, 64, 8, 256>(x, dt, first, w1, stream, cp, cl);
    case 16: return launch_fused<T, 64, 16, 512>(x, dt, first, w1, stream, cp, cl);
    default: break;
    }
    sf::set_error("sf_phys_substep_fused: unsupported tile height %d (4, 8 or 16)", tj);
    return SF_ERR_ARG;
}


// ---------------------------------------------------------------------------------
// Ensembles: many basins of uneven size advance in ONE launch per substep.  The
// launch walks a device-side list of (member, tile) pairs; each workgroup takes
// tiles from the list in one of two orders and runs the same per-tile function as
// the single-grid kernel, so a member's result equals its solo run bit for bit.
// ---------------------------------------------------------------------------------
// The member table holds, for every member, its kernel arguments for each of the
// three time rotations (index 3*member + rot): no runtime indexing of slot pointers
// inside the kernel.
// order 0: block b takes list entries b, b + G, b + 2G, ...
// order 1: block b (after the XCD band remap) takes a contiguous chunk of the list
template <typename T, int TI, int TJ, int NT, bool FIRST, bool NTS>
__global__ __launch_bounds__(NT) void k_phys_ensemble(const PFA<T> *__restrict__ members,
                                                      const int2 *__restrict__ list, int ntiles, int order,
                                                      int rot, double w1)
{
    __shared__ T sm[PhysLds<TI, TJ>::PLANES][PhysLds<TI, TJ>::FN];
    __shared__ T side[2 * TI + 2 * TJ];
    const int G = gridDim.x;
    int t0, t1, step;
    if (order == 0) {
        t0 = blockIdx.x;
        t1 = ntiles;
        step = G;
    } else {
        const int b = xcd_band(G);
        t0 = (int)((long long)ntiles * b / G);
        t1 = (int)((long long)ntiles * (b + 1) / G);
        step = 1;
    }
    // the tables are read-only for the whole launch: scalar loads through the constant
    // address space keep every member pointer in scalar registers
    const SF_CONST int *clist = (const SF_CONST int *)list;
    const SF_CONST PFA<T> *cmem = (const SF_CONST PFA<T> *)members;
    for (int t = t0; t < t1; t += step) {
        const int ex = clist[2 * t], ey = clist[2 * t + 1];
        PFA<T> P;
        {
            static_assert(sizeof(PFA<T>) % 4 == 0, "argument block of whole words");
            const SF_CONST uint32_t *src = (const SF_CONST uint32_t *)(cmem + 3 * ex + rot);
            uint32_t *dst = reinterpret_cast<uint32_t *>(&P);
#pragma unroll
            for (int w = 0; w < (int)(sizeof(PFA<T>) / 4); w++) dst[w] = src[w];
        }
        P.w1 = w1;
        phys_tile<T, TI, TJ, NT, FIRST, NTS>(P, ey % P.tiles_i, ey / P.tiles_i, sm, side);
        lds_barrier();   // the next tile overwrites the LDS planes
    }
}

// One workgroup per list entry (grid = number of tiles): no tile loop, so the tile
// code is register-allocated as in the single-basin kernel.  order 0 takes entry b,
// order 1 the XCD-banded entry of block b.
template <typename T, int TI, int TJ, int NT, bool FIRST, bool NTS>
__global__ __launch_bounds__(NT) void k_phys_ensemble1(const PFA<T> *__restrict__ members,
                                                       const int2 *__restrict__ list, int order, int rot,
                                                       double w1)
{
    __shared__ T sm[PhysLds<TI, TJ>::PLANES][PhysLds<TI, TJ>::FN];
    __shared__ T side[2 * TI + 2 * TJ];
    const int t = order == 0 ? (int)blockIdx.x : xcd_band(gridDim.x);
    const SF_CONST int *clist = (const SF_CONST int *)list;
    const SF_CONST PFA<T> *cmem = (const SF_CONST PFA<T> *)members;
    const int ex = clist[2 * t], ey = clist[2 * t + 1];
    PFA<T> P;
    {
        const SF_CONST uint32_t *src = (const SF_CONST uint32_t *)(cmem + 3 * ex + rot);
        uint32_t *dst = reinterpret_cast<uint32_t *>(&P);
#pragma unroll
        for (int w = 0; w < (int)(sizeof(PFA<T>) / 4); w++) dst[w] = src[w];
    }
    P.w1 = w1;
    phys_tile<T, TI, TJ, NT, FIRST, NTS>(P, ey % P.tiles_i, ey / P.tiles_i, sm, side);
}
}  // namespace


// ---------------------------------------------------------------------------------
// Ensemble C ABI
// ---------------------------------------------------------------------------------
struct sf_ens {
    int n, precision, device;
    int tile_j, order, grid;
    int ntiles, rot;
    int nts;                    // streaming stores (see PFA::nts), set by ens_build
    void *members;              // device PFA<T>[3 n]: arguments of member k for rotation r at 3k + r
    int2 *list;                 // device (member, tile) list for tile_j
    std::vector<sf_phys_fields> host;
    std::vector<double> dt;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int graph_nsub;
};

namespace {

template <int TJ>
int ens_tiles(const sf_phys_fields &x, int &ti, int &tj, int &oi, int &oj)
{
    constexpr int TI = 64;
    oi = (x.szI % TI == 1) ? 1 : 0;
    oj = (x.szJ % TJ == 1) ? 1 : 0;
    ti = (x.szI + oi + TI - 1) / TI;
    tj = (x.szJ + oj + TJ - 1) / TJ;
    return ti * tj;
}

template <typename T>
int ens_build(sf_ens *e)
{
    std::vector<PFA<T>> hm(3 * e->n);
    std::vector<int2> list;
    double bytes = 0.0;   // one substep of the whole ensemble (see PFA::nts)
    for (int k = 0; k < e->n; k++)
        bytes += (double)e->host[k].szI * e->host[k].szJ * (3.0 * sizeof(T) * 7 + 8.0 * 6);
    const int nts = bytes > sf::kStreamingBytes ? 1 : 0;
    e->nts = nts;
    for (int k = 0; k < e->n; k++) {
        const sf_phys_fields &x = e->host[k];
        int ti = 0, tj = 0, oi = 0, oj = 0, nt = 0;
        switch (e->tile_j) {
        case 4: nt = ens_tiles<4>(x, ti, tj, oi, oj); break;
        case 8: nt = ens_tiles<8>(x, ti, tj, oi, oj); break;
        default: nt = ens_tiles<16>(x, ti, tj, oi, oj); break;
        }
        for (int rot = 0; rot < 3; rot++) {
            // level t_l lives in slot (l + rot) % 3
            const int s0 = rot % 3, s1 = (rot + 1) % 3, s2 = (rot + 2) % 3;
            PFA<T> &a = hm[3 * k + rot];
            a.h = (const T *)x.h;
            a.pm = (const T *)x.pm;
            a.pn = (const T *)x.pn;
            a.fomn = (const T *)x.fomn;
            a.sustr = (const T *)x.sustr;
            a.svstr = (const T *)x.svstr;
            a.z0 = (const T *)x.zeta[s0];
            a.z1 = (const T *)x.zeta[s1];
            a.z2 = (T *)x.zeta[s2];
            a.u0 = (const T *)x.ubar[s0];
            a.u1 = (const T *)x.ubar[s1];
            a.u2 = (T *)x.ubar[s2];
            a.v0 = (const T *)x.vbar[s0];
            a.v1 = (const T *)x.vbar[s1];
            a.v2 = (T *)x.vbar[s2];
            a.rz0 = (const T *)x.rzeta[s0];
            a.rz1 = (T *)x.rzeta[s1];
            a.ru0 = (const T *)x.rubar[s0];
            a.ru1 = (T *)x.rubar[s1];
            a.rv0 = (const T *)x.rvbar[s0];
            a.rv1 = (T *)x.rvbar[s1];
            a.Zt = x.Zt_avg1;
            a.DUa = x.DU_avg1;
            a.DVa = x.DV_avg1;
            a.szI = x.szI;
            a.szJ = x.szJ;
            a.tiles_i = ti;
            a.tiles_j = tj;
            a.off_i = oi;
            a.off_j = oj;
            a.g = T(x.g);
            a.rdrg = T(x.rdrg);
            a.dt = T(e->dt[k]);
            a.first = 0;
            a.nts = nts;
            a.w1 = 0.0;
        }
        for (int t = 0; t < nt; t++) list.push_back(make_int2(k, t));
    }
    e->ntiles = (int)list.size();
    hipError_t r = hipFree(e->members);
    if (r == hipSuccess) r = hipFree(e->list);
    e->members = nullptr;
    e->list = nullptr;
    if (r == hipSuccess) r = hipMalloc(&e->members, sizeof(PFA<T>) * hm.size());
    if (r == hipSuccess) r = hipMalloc((void **)&e->list, sizeof(int2) * list.size());
    if (r == hipSuccess) r = hipMemcpy(e->members, hm.data(), sizeof(PFA<T>) * hm.size(), hipMemcpyHostToDevice);
    if (r == hipSuccess) r = hipMemcpy(e->list, list.data(), sizeof(int2) * list.size(), hipMemcpyHostToDevice);
    if (r != hipSuccess) return sf::hip_fail(r, "sf_ens: device tables");
    return SF_OK;
}

template <typename T, int TJ>
void ens_launch_tj(sf_ens *e, int first, double w1, int rot, hipStream_t s)
{
    constexpr int TI = 64, NT = TJ == 16 ? 512 : 256;
    const int G = e->grid > 0 ? (e->grid < e->ntiles ? e->grid : e->ntiles) : e->ntiles;
    const PFA<T> *m = (const PFA<T> *)e->members;
    if (G == e->ntiles) {   // one workgroup per tile
        if (first && e->nts)
            hipLaunchKernelGGL((k_phys_ensemble1<T, TI, TJ, NT, true, true>), dim3(G), dim3(NT), 0, s, m, e->list,
                               e->order, rot, w1);
        else if (first)
            hipLaunchKernelGGL((k_phys_ensemble1<T, TI, TJ, NT, true, false>), dim3(G), dim3(NT), 0, s, m, e->list,
                               e->order, rot, w1);
        else if (e->nts)
            hipLaunchKernelGGL((k_phys_ensemble1<T, TI, TJ, NT, false, true>), dim3(G), dim3(NT), 0, s, m, e->list,
                               e->order, rot, w1);
        else
            hipLaunchKernelGGL((k_phys_ensemble1<T, TI, TJ, NT, false, false>), dim3(G), dim3(NT), 0, s, m,
                               e->list, e->order, rot, w1);
        return;
    }
    if (first && e->nts)
        hipLaunchKernelGGL((k_phys_ensemble<T, TI, TJ, NT, true, true>), dim3(G), dim3(NT), 0, s, m, e->list,
                           e->ntiles, e->order, rot, w1);
    else if (first)
        hipLaunchKernelGGL((k_phys_ensemble<T, TI, TJ, NT, true, false>), dim3(G), dim3(NT), 0, s, m, e->list,
                           e->ntiles, e->order, rot, w1);
    else if (e->nts)
        hipLaunchKernelGGL((k_phys_ensemble<T, TI, TJ, NT, false, true>), dim3(G), dim3(NT), 0, s, m, e->list,
                           e->ntiles, e->order, rot, w1);
    else
        hipLaunchKernelGGL((k_phys_ensemble<T, TI, TJ, NT, false, false>), dim3(G), dim3(NT), 0, s, m, e->list,
                           e->ntiles, e->order, rot, w1);
}

template <typename T>
void ens_launch(sf_ens *e, int first, double w1, int rot, hipStream_t s)
{
    switch (e->tile_j) {
    case 4: ens_launch_tj<T, 4>(e, first, w1, rot, s); break;
    case 8: ens_launch_tj<T, 8>(e, first, w1, rot, s); break;
    default: ens_launch_tj<T, 16>(e, first, w1, rot, s); break;
    }
}

}  // namespace

// FP32 instances, compiled in phys2d_fused_f32.hip with contraction on
int sf_phys_fused_dispatch_f32(const sf_phys_fields *x, double dt, int first, double w1, int tj, void *stream,
                               const sf_couple_fields *cp = nullptr, const ClosureIn *cl = nullptr);
void sf_ens_launch_f32(sf_ens *e, int first, double w1, int rot, hipStream_t s);
