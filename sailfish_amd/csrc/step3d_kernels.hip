// step3d_kernels.hip -- 3D kernels of the baroclinic step, one launch per reference
// kernel: addCoriolis / horizontalAdvection / verticalAdvection
// (nonlinear/rhs_kernels.cpp), set_maxflux / omega / set_depth / set_zeta /
// setLateralUVBCs / step3d_UV (nonlinear/step3D_uv.kernels.cpp).
//
// Layout: 3D fields are flat with level stride S2 = szI*szJ (k slowest, xi
// fastest, modules/mod_ocean.py:69-71).  One thread per water column (j,i), the
// level loop inside the thread; consecutive threads take consecutive i, so every
// level of every field is read with unit stride across a wavefront.  Stencil
// offsets are flat (the reference's Eval arithmetic, mod_cppkernels.h:337-340);
// reads outside an array are defined as 0.0.  FP64, reference operand order,
// contraction off: results equal the CPU restatement bit for bit.
#include "sf_internal.h"

namespace {

// a 3D (or 2D) field in global memory: pointer and element count; value type double
// (reference precision) or float (FP32 bulk form of the product-path kernels)
template <class T>
struct AG {
    const T *p;
    long long n;
    using value_type = T;
};
using A = AG<double>;
using AF = AG<float>;

// out-of-range reads are 0; the load itself always uses a clamped (valid) index, so it
// is unconditional and identical reads of one expression tree merge into one load
template <class T>
__device__ __forceinline__ T r3(AG<T> a, long long idx)
{
    const long long c = idx < 0 ? 0 : (idx >= a.n ? a.n - 1 : idx);
    const T x = a.p[c];
    return (idx < 0 || idx >= a.n) ? T(0) : x;
}
template <bool CHK, class T>
__device__ __forceinline__ T rd(AG<T> a, long long idx)
{
    if constexpr (CHK) return r3(a, idx);
    return a.p[idx];
}

struct G3 {
    int szI, szJ, N;
    long long S2;
    const double *on_u, *om_v, *pn, *pm;
};

__device__ __forceinline__ bool isRNode(const G3 &g, long long f) { return (f % g.szI) > 0 && (f / g.szI) > 0; }
__device__ __forceinline__ bool isUNode(const G3 &g, long long f) { return (f % g.szI) > 1 && (f / g.szI) > 0; }
__device__ __forceinline__ bool isVNode(const G3 &g, long long f)
{
    const long long j = f / g.szI;
    return (f % g.szI) > 0 && j > 1 && j < g.szJ - 1;
}
__device__ __forceinline__ bool isInnerCell(const G3 &g, long long f)
{
    const long long i = f % g.szI, j = f / g.szI;
    return i > 1 && i < g.szI - 2 && j > 1 && j < g.szJ - 2;
}

#define COLUMN_INDEX                                                      \
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x; \
    if (f >= g.S2) return;

// ---- addCoriolis (rhs_kernels.cpp:6-35) -------------------------------------------
template <bool CHK>
__device__ __forceinline__ double cor_UF(const G3 &g, A fomn, A v, A Hz, long long K, long long f)
{
    const long long b = K * g.S2 + f;
    return (rd<CHK>(Hz, b) * rd<CHK>(fomn, f)) * ((rd<CHK>(v, b) + rd<CHK>(v, b + g.szI)) * 0.5);
}
template <bool CHK>
__device__ __forceinline__ double cor_VF(const G3 &g, A fomn, A u, A Hz, long long K, long long f)
{
    const long long b = K * g.S2 + f;
    return (rd<CHK>(Hz, b) * rd<CHK>(fomn, f)) * ((rd<CHK>(u, b) + rd<CHK>(u, b + 1)) * 0.5);
}

__global__ __launch_bounds__(256) void k_addCoriolis(G3 g, A fomn, A u, A v, double *ru, double *rv, A Hz)
{
    COLUMN_INDEX
    if (!isRNode(g, f)) return;
    for (long long K = 0; K < g.N; K++) {
        const long long b = K * g.S2 + f;
        double rtou, rtov;
        // rho nodes have i, j >= 1: the reads at b-1 and b-szI exist; the forward ones
        // (b+1, b+szI) are checked against the array sizes once per level
        if (b + g.szI < v.n && b + 1 < u.n && b < Hz.n && f < fomn.n) {
            rtou = (cor_UF<false>(g, fomn, v, Hz, K, f) + cor_UF<false>(g, fomn, v, Hz, K, f - 1)) * 0.5;
            rtov = (cor_VF<false>(g, fomn, u, Hz, K, f) + cor_VF<false>(g, fomn, u, Hz, K, f - g.szI)) * 0.5;
        } else {
            rtou = (cor_UF<true>(g, fomn, v, Hz, K, f) + cor_UF<true>(g, fomn, v, Hz, K, f - 1)) * 0.5;
            rtov = (cor_VF<true>(g, fomn, u, Hz, K, f) + cor_VF<true>(g, fomn, u, Hz, K, f - g.szI)) * 0.5;
        }
        ru[b] = ru[b] + rtou;
        rv[b] = rv[b] - rtov;
    }
}

// ---- horizontalAdvection (rhs_kernels.cpp:70-132) ---------------------------------
template <class T>
struct HG {
    AG<T> u, v, Hu, Hv;
    long long W;
    using value_type = T;
};
using H = HG<double>;
// the same four fields staged in LDS for one level of a tile: element at flat offset
// idx of the tile view (row stride W = tile width + 4)
template <class T>
struct ALT {
    const T *p;
    using value_type = T;
};
template <bool CHK, class T>
__device__ __forceinline__ T rd(ALT<T> a, long long idx)
{
    return a.p[idx];
}
template <class T>
struct HLT {
    ALT<T> u, v, Hu, Hv;
    long long W;
    using value_type = T;
};
using AL = ALT<double>;
using HL = HLT<double>;

#define AT(a, o) rd<CHK>((a), base + (o))
// V: the value type of the accessor; literals are converted to it (exact in both types)
template <bool CHK, class AA, class V = typename AA::value_type>
__device__ __forceinline__ V dxx(AA a, long long base, long long o)
{
    return AT(a, o - 1) - V(2) * AT(a, o) + AT(a, o + 1);
}
template <bool CHK, class AA, class V = typename AA::value_type>
__device__ __forceinline__ V dee(AA a, long long base, long long o, long long W)
{
    return AT(a, o - W) - V(2) * AT(a, o) + AT(a, o + W);
}
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_uR(const HH &h, long long base, long long o) { return (AT(h.u, o) + AT(h.u, o + 1)) * V(0.5); }
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_uP(const HH &h, long long base, long long o) { return (AT(h.u, o) + AT(h.u, o - 1)) * V(0.5); }
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_vR(const HH &h, long long base, long long o) { return (AT(h.v, o) + AT(h.v, o + h.W)) * V(0.5); }
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_vP(const HH &h, long long base, long long o) { return (AT(h.v, o) + AT(h.v, o - h.W)) * V(0.5); }
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_HuR(const HH &h, long long base, long long o) { return (AT(h.Hu, o) + AT(h.Hu, o + 1)) * V(0.5); }
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_HuP(const HH &h, long long base, long long o) { return (AT(h.Hu, o) + AT(h.Hu, o - 1)) * V(0.5); }
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_HvP(const HH &h, long long base, long long o) { return (AT(h.Hv, o) + AT(h.Hv, o - h.W)) * V(0.5); }
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_HvR(const HH &h, long long base, long long o) { return (AT(h.Hv, o) + AT(h.Hv, o + h.W)) * V(0.5); }
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_uxxR(const HH &h, long long base, long long o)
{
    return (h_uR<CHK>(h, base, o) + h_uR<CHK>(h, base, o + 1) >= 0) ? dxx<CHK>(h.u, base, o) : dxx<CHK>(h.u, base, o + 1);
}
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_ueeP(const HH &h, long long base, long long o)
{
    return (h_HvP<CHK>(h, base, o) + h_HvP<CHK>(h, base, o - h.W) >= 0) ? dee<CHK>(h.u, base, o, h.W) : dee<CHK>(h.u, base, o - h.W, h.W);
}
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_HuxxR(const HH &h, long long base, long long o)
{
    return (dxx<CHK>(h.Hu, base, o) + dxx<CHK>(h.Hu, base, o + 1)) * V(0.5);
}
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_HueeP(const HH &h, long long base, long long o)
{
    return (dee<CHK>(h.Hu, base, o, h.W) + dee<CHK>(h.Hu, base, o - 1, h.W)) * V(0.5);
}
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_HvxxP(const HH &h, long long base, long long o)
{
    return (dxx<CHK>(h.Hv, base, o) + dxx<CHK>(h.Hv, base, o - h.W)) * V(0.5);
}
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V h_HveeR(const HH &h, long long base, long long o)
{
    return (dee<CHK>(h.Hv, base, o, h.W) + dee<CHK>(h.Hv, base, o + h.W, h.W)) * V(0.5);
}
constexpr double Gadv = -0.5;
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V UFxR(const HH &h, long long b, long long o)
{
    return h_uR<CHK>(h, b, o) + (V(Gadv) * h_uxxR<CHK>(h, b, o)) * (h_HuR<CHK>(h, b, o) + V(Gadv) * h_HuxxR<CHK>(h, b, o));
}
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V UFeP(const HH &h, long long b, long long o)
{
    return h_uP<CHK>(h, b, o) + (V(Gadv) * h_ueeP<CHK>(h, b, o)) * (h_HvP<CHK>(h, b, o) + V(Gadv) * h_HvxxP<CHK>(h, b, o));
}
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V VFxP(const HH &h, long long b, long long o)
{
    return h_vP<CHK>(h, b, o) + (V(Gadv) * h_ueeP<CHK>(h, b, o)) * (h_HuP<CHK>(h, b, o) + V(Gadv) * h_HueeP<CHK>(h, b, o));
}
template <bool CHK, class HH, class V = typename HH::value_type>
__device__ __forceinline__ V VFeR(const HH &h, long long b, long long o)
{
    return h_vR<CHK>(h, b, o) + (V(Gadv) * h_uxxR<CHK>(h, b, o)) * (h_HvR<CHK>(h, b, o) + V(Gadv) * h_HveeR<CHK>(h, b, o));
}
#undef AT

// Tiled form: a workgroup owns HTI x HTJ water columns and walks the levels; at each
// level it stages u, v, Huon, Hvom over the tile and its halo (3 rows below, 2 above,
// 2 columns each side: the reach of the expression tree) in LDS, so each value is
// read from global memory once per tile instead of once per use.  The staged rows are
// flat index ranges, so offsets that wrap across rows or levels (the reference's flat
// Eval arithmetic) and reads outside the arrays (0.0) are reproduced exactly; the
// arithmetic is the reference's expression tree in its operand order.
constexpr int HTI = 64, HTJ = 8, HNT = HTI * HTJ, HLW = HTI + 4, HLH = HTJ + 5, HLN = HLW * HLH;

__global__ __launch_bounds__(HNT) void k_horizontalAdvection_tiled(G3 g, H h, double *ru, double *rv, int tiles_i)
{
    __shared__ double sl[4][HLN];
    const int ti = blockIdx.x % tiles_i, tj = blockIdx.x / tiles_i;
    const int i0 = ti * HTI, j0 = tj * HTJ;
    const int c = threadIdx.x % HTI, r = threadIdx.x / HTI;
    const int i = i0 + c, j = j0 + r;
    const bool act = i < g.szI && j < g.szJ;
    const long long W = g.szI, n = h.u.n;
    const HL hl{AL{sl[0]}, AL{sl[1]}, AL{sl[2]}, AL{sl[3]}, HLW};
    const long long lb = (long long)(r + 3) * HLW + (c + 2);   // the column's element in the tile view
    const long long row0 = (long long)(j0 - 3) * W + (i0 - 2);  // flat index of tile-view element 0, level 0
    for (long long K = 0; K < g.N; K++) {
        const long long kb = K * g.S2;
        for (int e = threadIdx.x; e < HLN; e += HNT) {
            const int rr = e / HLW, cc = e - rr * HLW;
            const long long gi = kb + row0 + (long long)rr * W + cc;
            const bool ok = gi >= 0 && gi < n;
            const long long gc = ok ? gi : 0;
            const double a0 = h.u.p[gc], a1 = h.v.p[gc], a2 = h.Hu.p[gc], a3 = h.Hv.p[gc];
            sl[0][e] = ok ? a0 : 0.0;
            sl[1][e] = ok ? a1 : 0.0;
            sl[2][e] = ok ? a2 : 0.0;
            sl[3][e] = ok ? a3 : 0.0;
        }
        __syncthreads();
        if (act) {
            const double du = (UFxR<false>(hl, lb, 0) - UFxR<false>(hl, lb, -1)) +
                              (UFeP<false>(hl, lb, 0) - UFeP<false>(hl, lb, -HLW));
            const double dv = (VFxP<false>(hl, lb, 0) - VFxP<false>(hl, lb, -1)) +
                              (VFeR<false>(hl, lb, 0) - VFeR<false>(hl, lb, -HLW));
            const long long b = kb + (long long)j * W + i;
            ru[b] = ru[b] - du;
            rv[b] = rv[b] - dv;
        }
        __syncthreads();   // the next level overwrites the tile
    }
}

// the reference's threads S2..launch-1 store 0 at flat index i (rhs_kernels.cpp:86-91)
__global__ void k_hadv_tail(long long S2, long long end, double *ru, double *rv)
{
    const long long t = S2 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= end) return;
    ru[t] = 0.0;
    rv[t] = 0.0;
}

// ---- verticalAdvection (rhs_kernels.cpp:135-197) ----------------------------------
template <bool CHK = true, class T>
__device__ __forceinline__ T v4(AG<T> a, long long S2, long long K, long long f)
{
    return T(9.0 / 16.0) * (rd<CHK>(a, K * S2 + f) + rd<CHK>(a, (K + 1) * S2 + f)) -
           T(1.0 / 16.0) * (rd<CHK>(a, (K - 1) * S2 + f) + rd<CHK>(a, (K + 2) * S2 + f));
}
template <bool CHK = true, class T>
__device__ __forceinline__ T h4(AG<T> a, long long S2, long long K, long long f, long long d)
{
    const long long b = K * S2 + f;
    return T(9.0 / 16.0) * (rd<CHK>(a, b) + rd<CHK>(a, b + d)) - T(1.0 / 16.0) * (rd<CHK>(a, b - d) + rd<CHK>(a, b + 2 * d));
}

__global__ __launch_bounds__(256) void k_verticalAdvection(G3 g, A U, A V, A Wa, double *ru, double *rv,
                                                           const int32_t *BC)
{
    COLUMN_INDEX
    if (BC[f] >= 0 || !isInnerCell(g, f)) return;
    const long long S2 = g.S2, N = g.N, W = g.szI;
#define FU(k) (v4(U, S2, (k), f) * h4(Wa, S2, (k), f, 1))
#define FV(k) (v4(V, S2, (k), f) * h4(Wa, S2, (k), f, W))
    // interior levels: every read is in range (inner cell, levels K-1 .. K+3 <= N-1);
    // the flux at the upper face of level K is the lower-face flux of level K+1
    // (v4 at level K reads levels K-1 .. K+2: a sliding window of four values per
    // field, one new load per level)
    if (N - 2 > 2) {
        const double *Up = U.p + f, *Vp = V.p + f;
        double ua = Up[S2], ub = Up[2 * S2], uc = Up[3 * S2], ud = Up[4 * S2];
        double va = Vp[S2], vb = Vp[2 * S2], vc = Vp[3 * S2], vd = Vp[4 * S2];
        double fu = ((9.0 / 16.0) * (ub + uc) - (1.0 / 16.0) * (ua + ud)) * h4<false>(Wa, S2, 2, f, 1);
        double fv = ((9.0 / 16.0) * (vb + vc) - (1.0 / 16.0) * (va + vd)) * h4<false>(Wa, S2, 2, f, W);
        for (long long K = 2; K < N - 2; K++) {
            const double ue = Up[(K + 3) * S2], ve = Vp[(K + 3) * S2];
            const double fu1 = ((9.0 / 16.0) * (uc + ud) - (1.0 / 16.0) * (ub + ue)) * h4<false>(Wa, S2, K + 1, f, 1);
            const double fv1 = ((9.0 / 16.0) * (vc + vd) - (1.0 / 16.0) * (vb + ve)) * h4<false>(Wa, S2, K + 1, f, W);
            ru[K * S2 + f] = ru[K * S2 + f] - (fu1 - fu);
            rv[K * S2 + f] = rv[K * S2 + f] - (fv1 - fv);
            fu = fu1;
            fv = fv1;
            ub = uc, uc = ud, ud = ue;
            vb = vc, vc = vd, vd = ve;
        }
    }
    rv[1 * S2 + f] -= (9.0 / 16.0) * (FV(1) + FV(2)) - (1.0 / 16.0) * (FV(1) + FV(3));
    const double u0 = r3(U, f), u1 = r3(U, S2 + f), u2 = r3(U, 2 * S2 + f);
    const double uuu = ((9.0 / 16.0) * (u0 + u1) - (1.0 / 16.0) * (0.0 + u2)) * h4(Wa, S2, 1, f, 1);
    const double uuu2 = ((9.0 / 16.0) * (0.0 + u0) - (1.0 / 16.0) * (0.0 + u1)) * h4(Wa, S2, 0, f, 1);
    ru[1 * S2 + f] -= (FU(2) - uuu);
    ru[f] -= (uuu - uuu2);
    ru[(N - 2) * S2 + f] -= (9.0 / 16.0) * (FU(N - 1) + FU(N - 2)) - (1.0 / 16.0) * (FU(N - 1) + FU(N - 3));
    rv[(N - 2) * S2 + f] -= (9.0 / 16.0) * (FV(N - 1) + FV(N - 2)) - (1.0 / 16.0) * (FV(N - 1) + FV(N - 3));
    rv[f] = rv[S2 + f];
    ru[(N - 2) * S2 + f] = ru[(N - 3) * S2 + f];
    rv[(N - 2) * S2 + f] = rv[(N - 3) * S2 + f];
    ru[(N - 1) * S2 + f] = ru[(N - 2) * S2 + f];
    rv[(N - 1) * S2 + f] = rv[(N - 2) * S2 + f];
    ru[N * S2 + f] = ru[(N - 2) * S2 + f];
    rv[N * S2 + f] = rv[(N - 2) * S2 + f];
#undef FU
#undef FV
}

// ---- rhs3d in one launch (nonlinear/rhs.py:38-71) ---------------------------------
// ru, rv := 0; addCoriolis; horizontalAdvection (+ the zeros its threads past S2
// store); verticalAdvection -- evaluated per column and level in that order, so every
// element goes through the same sequence of roundings as with the four launches:
//   r = 0;  r = r + rtou (rho nodes);  r = r - du;  r = 0 (flat index in the tail
//   range);  r = r - (fu(K+1) - fu(K)) (inner cells, levels 2..N-3)
// and the special levels 0, 1, N-2 and the level copies of verticalAdvection follow
// per column after the level loop.  u, v, Huon, Hvom of each level are staged in LDS
// over the tile and its halo (as in k_horizontalAdvection_tiled); W and the vertical
// stencils are read from global memory, u/v along the column through a sliding window.
//
// The levels are split into nchunk chunks of kchunk levels (the last one takes the
// remainder and has at least 3 levels), one workgroup per tile and chunk.  A chunk seeds
// the vertical-flux window at its first interior level with the same expression the
// sliding window evaluates there, so the split changes no result.  The first chunk
// (levels 0, 1) does the bottom special levels, the last chunk (levels N-3 .. N) the
// top ones; with more than one chunk (N >= 6) the two touch disjoint levels.
//
// Each level: stage the tile, barrier, face fluxes (and, in FP64, the Coriolis terms),
// barrier, cells.  FP32 double-buffers the tile; FP64 uses one buffer (see below).  WPE
// is the wave count per SIMD the compiler targets: 6 for FP64 (at most 80 VGPRs), 4 or 8
// for FP32 (see rhs3d_launch).
template <int TJ, class T, int WPE>
__global__ __launch_bounds__(HTI * TJ) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_rhs3d_fused(G3 g, HG<T> h, A fomn, AG<T> Hz, AG<T> Wa, const int32_t *BC, T *ru, T *rv, long long tail_end,
                   int cor, int adv, int tiles_i, int ntiles, int kchunk, int nchunk)
{
    constexpr int NT = HTI * TJ, LN = HLW * (TJ + 5);
    // The face fluxes of the level are each evaluated once, into LDS -- x-faces
    // c' = -1 .. HTI-1 of every row (UFxR, VFxP; column c'+1) and e-faces of rows
    // r' = -1 .. TJ-1 (UFeP, VFeR; row r'+1) -- instead of once by each cell beside the
    // face.  FP32 (50 VGPRs instead of 64-69): 1024 x 512 x 32 222 us against 273 us.
    // FP64 also forms the Coriolis terms before the second barrier, so no cell reads the
    // tile after it and one tile buffer suffices: 46 KB LDS and 80 VGPRs, three
    // workgroups per CU instead of two (1024 x 512 x 32: 436 us against 443 us per cell
    // with two buffers; 512 x 256 x 32: 96 us against 104 us).
    constexpr bool SB = sizeof(T) == 8;   // single tile buffer
    __shared__ T sl[SB ? 1 : 2][4][LN];
    __shared__ T fxu[TJ][HTI + 1], fxv[TJ][HTI + 1], feu[TJ + 1][HTI], fev[TJ + 1][HTI];
    const int tile = blockIdx.x % ntiles, kc = blockIdx.x / ntiles;
    const int ti = tile % tiles_i, tj = tile / tiles_i;
    const int i0 = ti * HTI, j0 = tj * TJ;
    const int c = threadIdx.x % HTI, r = threadIdx.x / HTI;
    const int i = i0 + c, j = j0 + r;
    const bool act = i < g.szI && j < g.szJ;
    const long long W = g.szI, S2 = g.S2, N = g.N, n = h.u.n;
    const long long f = act ? (long long)j * W + i : 0;
    const long long lb = (long long)(r + 3) * HLW + (c + 2);
    const long long row0 = (long long)(j0 - 3) * W + (i0 - 2);
    const bool rnode = act && cor && isRNode(g, f);
    const bool vadv = act && adv && BC[f] < 0 && isInnerCell(g, f);
    const bool vint = vadv && N - 2 > 2;   // interior levels 2 .. N-3 handled in the level loop
    const long long k0 = (long long)kc * kchunk, k1 = kc == nchunk - 1 ? N : k0 + kchunk;
    const long long ks = k0 > 2 ? k0 : 2;   // first interior level of the chunk
    T fo_c = T(0), fo_w = T(0), fo_s = T(0);
    if (rnode) {
        fo_c = T(fomn.p[f]);
        fo_w = T(fomn.p[f - 1]);
        fo_s = T(fomn.p[f - W]);
    }
    // vertical-flux window of the column (v4 at level K reads levels K-1 .. K+2)
    T ub = T(0), uc = T(0), ud = T(0), vb = T(0), vc = T(0), vd = T(0), fu = T(0), fv = T(0);
    if (vint && ks < k1 && ks < N - 2) {
        const T *Up = h.u.p + f, *Vp = h.v.p + f;
        const T ua = Up[(ks - 1) * S2], va = Vp[(ks - 1) * S2];
        ub = Up[ks * S2], uc = Up[(ks + 1) * S2], ud = Up[(ks + 2) * S2];
        vb = Vp[ks * S2], vc = Vp[(ks + 1) * S2], vd = Vp[(ks + 2) * S2];
        fu = (T(9.0 / 16.0) * (ub + uc) - T(1.0 / 16.0) * (ua + ud)) * h4<false>(Wa, S2, ks, f, 1);
        fv = (T(9.0 / 16.0) * (vb + vc) - T(1.0 / 16.0) * (va + vd)) * h4<false>(Wa, S2, ks, f, W);
    }
    for (long long K = k0; K < k1; K++) {
        const long long kb = K * S2, b = kb + f;
        // FP32: level K writes one tile buffer while stragglers of level K-1 may still read
        // the other (their Coriolis terms); the barrier of level K orders the reuse two
        // levels on.  FP64 reads the tile only before the level's second barrier, so one
        // buffer serves every level.
        T(*sb)[LN] = sl[SB ? 0 : (int)((K - k0) & 1)];
        const HLT<T> hl{ALT<T>{sb[0]}, ALT<T>{sb[1]}, ALT<T>{sb[2]}, ALT<T>{sb[3]}, HLW};
        for (int e = threadIdx.x; e < LN; e += NT) {
            const int rr = e / HLW, cc = e - rr * HLW;
            const long long gi = kb + row0 + (long long)rr * W + cc;
            const bool ok = gi >= 0 && gi < n;
            const long long gc = ok ? gi : 0;
            const T a0 = h.u.p[gc], a1 = h.v.p[gc], a2 = h.Hu.p[gc], a3 = h.Hv.p[gc];
            sb[0][e] = ok ? a0 : T(0);
            sb[1][e] = ok ? a1 : T(0);
            sb[2][e] = ok ? a2 : T(0);
            sb[3][e] = ok ? a3 : T(0);
        }
        __syncthreads();
        T cu = T(0), cv = T(0);   // SB: the cell's Coriolis terms, formed before the tile is reused
        if constexpr (SB) {
            if (act && rnode) {
                const T *su = sb[0], *sv = sb[1];
                const T hzc = Hz.p[b], hzw = Hz.p[b - 1], hzs = Hz.p[b - W];
                const T UF0 = (hzc * fo_c) * ((sv[lb] + sv[lb + HLW]) * T(0.5));
                const T UF1 = (hzw * fo_w) * ((sv[lb - 1] + sv[lb - 1 + HLW]) * T(0.5));
                const T VF0 = (hzc * fo_c) * ((su[lb] + su[lb + 1]) * T(0.5));
                const T VF1 = (hzs * fo_s) * ((su[lb - HLW] + su[lb - HLW + 1]) * T(0.5));
                cu = (UF0 + UF1) * T(0.5);
                cv = (VF0 + VF1) * T(0.5);
            }
        }
        // wave r = tile row r, lane c = column c.  Pass 1: x-face c' = c-1 of row r;
        // pass 2: e-face of row r; pass 3: the e-faces of row -1 (wave 0) and the
        // last x-face c' = HTI-1 of every row (wave 1).  Each is the expression the
        // cells on either side of the face evaluate (UFxR(0) of the cell on its left
        // is UFxR(-1) of the cell on its right), at a tile position they read.
        {
            const long long p = (long long)(r + 3) * HLW + (c + 1);
            fxu[r][c] = UFxR<false>(hl, p, 0);
            fxv[r][c] = VFxP<false>(hl, p, 0);
        }
        feu[r + 1][c] = UFeP<false>(hl, lb, 0);
        fev[r + 1][c] = VFeR<false>(hl, lb, 0);
        if (r == 0) {
            const long long p = 2LL * HLW + (c + 2);
            feu[0][c] = UFeP<false>(hl, p, 0);
            fev[0][c] = VFeR<false>(hl, p, 0);
        } else if (r == 1 && c < TJ) {
            const long long p = (long long)(c + 3) * HLW + (HTI + 1);
            fxu[c][HTI] = UFxR<false>(hl, p, 0);
            fxv[c][HTI] = VFxP<false>(hl, p, 0);
        }
        __syncthreads();
        if (act) {
            T xu = T(0), xv = T(0);
            if (SB && rnode) {
                xu = xu + cu;
                xv = xv - cv;
            } else if (rnode) {
                const T *su = sb[0], *sv = sb[1];
                const T hzc = Hz.p[b], hzw = Hz.p[b - 1], hzs = Hz.p[b - W];
                const T UF0 = (hzc * fo_c) * ((sv[lb] + sv[lb + HLW]) * T(0.5));
                const T UF1 = (hzw * fo_w) * ((sv[lb - 1] + sv[lb - 1 + HLW]) * T(0.5));
                const T VF0 = (hzc * fo_c) * ((su[lb] + su[lb + 1]) * T(0.5));
                const T VF1 = (hzs * fo_s) * ((su[lb - HLW] + su[lb - HLW + 1]) * T(0.5));
                xu = xu + (UF0 + UF1) * T(0.5);
                xv = xv - (VF0 + VF1) * T(0.5);
            }
            const T du = (fxu[r][c + 1] - fxu[r][c]) + (feu[r + 1][c] - feu[r][c]);
            const T dv = (fxv[r][c + 1] - fxv[r][c]) + (fev[r + 1][c] - fev[r][c]);
            xu = xu - du;
            xv = xv - dv;
            if (K >= 1 && b < tail_end) xu = xv = T(0);
            if (vint && K >= 2 && K < N - 2) {
                const T ue = h.u.p[(K + 3) * S2 + f], ve = h.v.p[(K + 3) * S2 + f];
                const T fu1 = (T(9.0 / 16.0) * (uc + ud) - T(1.0 / 16.0) * (ub + ue)) * h4<false>(Wa, S2, K + 1, f, 1);
                const T fv1 = (T(9.0 / 16.0) * (vc + vd) - T(1.0 / 16.0) * (vb + ve)) * h4<false>(Wa, S2, K + 1, f, W);
                xu = xu - (fu1 - fu);
                xv = xv - (fv1 - fv);
                fu = fu1;
                fv = fv1;
                ub = uc, uc = ud, ud = ue;
                vb = vc, vc = vd, vd = ve;
            }
            ru[b] = xu;
            rv[b] = xv;
        }
    }
    if (!act) return;
    const bool bottom = kc == 0, top = kc == nchunk - 1;
    if (top) {
        ru[N * S2 + f] = T(0);
        rv[N * S2 + f] = T(0);
    }
    if (!vadv) return;
    // special levels and copies of verticalAdvection (same order as k_verticalAdvection)
    const AG<T> U = h.u, V = h.v;
#define FU(k) (v4(U, S2, (k), f) * h4(Wa, S2, (k), f, 1))
#define FV(k) (v4(V, S2, (k), f) * h4(Wa, S2, (k), f, W))
    if (bottom) {
        rv[1 * S2 + f] -= T(9.0 / 16.0) * (FV(1) + FV(2)) - T(1.0 / 16.0) * (FV(1) + FV(3));
        const T u0 = r3(U, f), u1 = r3(U, S2 + f), u2 = r3(U, 2 * S2 + f);
        const T uuu = (T(9.0 / 16.0) * (u0 + u1) - T(1.0 / 16.0) * (T(0) + u2)) * h4(Wa, S2, 1, f, 1);
        const T uuu2 = (T(9.0 / 16.0) * (T(0) + u0) - T(1.0 / 16.0) * (T(0) + u1)) * h4(Wa, S2, 0, f, 1);
        ru[1 * S2 + f] -= (FU(2) - uuu);
        ru[f] -= (uuu - uuu2);
    }
    if (top) {
        ru[(N - 2) * S2 + f] -= T(9.0 / 16.0) * (FU(N - 1) + FU(N - 2)) - T(1.0 / 16.0) * (FU(N - 1) + FU(N - 3));
        rv[(N - 2) * S2 + f] -= T(9.0 / 16.0) * (FV(N - 1) + FV(N - 2)) - T(1.0 / 16.0) * (FV(N - 1) + FV(N - 3));
    }
    if (bottom) rv[f] = rv[S2 + f];
    if (!top) return;
    ru[(N - 2) * S2 + f] = ru[(N - 3) * S2 + f];
    rv[(N - 2) * S2 + f] = rv[(N - 3) * S2 + f];
    ru[(N - 1) * S2 + f] = ru[(N - 2) * S2 + f];
    rv[(N - 1) * S2 + f] = rv[(N - 2) * S2 + f];
    ru[N * S2 + f] = ru[(N - 2) * S2 + f];
    rv[N * S2 + f] = rv[(N - 2) * S2 + f];
#undef FU
#undef FV
}

// ---- set_maxflux (step3D_uv.kernels.cpp:372-405) ----------------------------------
// the fluxes are stored with streaming (non-temporal) stores: no thread of the launch
// reads them back.  Both fluxes of a level are formed together, so Hz is read once.
template <class T>
__global__ __launch_bounds__(256) void k_set_maxflux(G3 g, const T *u, const T *v, T *Huon, T *Hvom, const T *Hz)
{
    COLUMN_INDEX
    const bool un = isUNode(g, f), vn = isVNode(g, f);
    if (!un && !vn) return;
    const T onu = un ? T(g.on_u[f]) : T(0), omv = vn ? T(g.om_v[f]) : T(0);
    for (long long K = 0; K < g.N; K++) {
        const long long b = K * g.S2 + f;
        const T hz = Hz[b];
        if (un) __builtin_nontemporal_store((((hz + Hz[b - 1]) * T(0.5)) * u[b]) * onu, Huon + b);
        if (vn) __builtin_nontemporal_store((((hz + Hz[b - g.szI]) * T(0.5)) * v[b]) * omv, Hvom + b);
    }
}

// ---- omega (step3D_uv.kernels.cpp:62-165) -----------------------------------------
__global__ __launch_bounds__(256) void k_omega(G3 g, double *W, A Hu, A Hv, const double *z_w,
                                               const int32_t *BC)
{
    COLUMN_INDEX
    if (BC[f] >= 0) return;
    const long long S2 = g.S2, N = g.N;
    double w = 0.0;
    W[f] = w;
    for (long long K = 1; K < N; K++) {
        const long long b = K * S2 + f;
        w = w - ((r3(Hu, b + 1) - r3(Hu, b)) + (r3(Hv, b + g.szI) - r3(Hv, b)));
        W[b] = w;
    }
    const double zb = z_w[f];
    const double h = z_w[(N - 1) * S2 + f] - zb;
    const double Ws = N > 1 ? w : W[(N - 1) * S2 + f];
    for (long long K = N - 1; K > 0; K--) {
        const long long b = K * S2 + f;
        W[b] = W[b] - (Ws * (z_w[b] - zb)) / h;
    }
    W[N * S2 + f] = 0.0;
}

// omega with the column's partial sums kept in registers (N <= NMAX): W is written once
// per level instead of written, read back and rewritten.  The values are those of
// k_omega: the same partial sums, then W = w - (Ws * (z_w - zb)) / h.  The sums and the
// correction are FP64 for either value type T of the fields.  (1024 x 512 x 32:
// 138 us against 149 us for k_omega; recomputing the partial sums in a second pass
// instead of holding them took 195 us.)
template <int NMAX, class T>
__global__ __launch_bounds__(256) void k_omega_reg(G3 g, T *W, AG<T> Hu, AG<T> Hv, const T *z_w,
                                                   const int32_t *BC)
{
    COLUMN_INDEX
    if (BC[f] >= 0) return;
    const long long S2 = g.S2;
    const int N = g.N;
    double wk[NMAX];
    double w = 0.0;
    W[f] = T(0);
#pragma unroll
    for (int K = 1; K < NMAX; K++) {
        wk[K] = 0.0;
        if (K < N) {
            const long long b = K * S2 + f;
            w = w - (((double)r3(Hu, b + 1) - (double)r3(Hu, b)) + ((double)r3(Hv, b + g.szI) - (double)r3(Hv, b)));
            wk[K] = w;
        }
    }
    const double zb = (double)z_w[f];
    const double h = (double)z_w[(long long)(N - 1) * S2 + f] - zb;
    const double Ws = w;   // N > 1
#pragma unroll
    for (int K = 1; K < NMAX; K++) {
        if (K < N) {
            const long long b = K * S2 + f;
            W[b] = T(wk[K] - (Ws * ((double)z_w[b] - zb)) / h);
        }
    }
    W[(long long)N * S2 + f] = T(0);
}

// ---- set_depth / set_zeta (step3D_uv.kernels.cpp:225-310) -------------------------
// T: the depth arrays z_w, z_r, Hz (double: the reference's arithmetic; float: the FP32
// bulk form, each value formed in FP64 as for double and rounded once).
template <class T>
__global__ __launch_bounds__(256) void k_set_depth(G3 g, int Vtransform, const double *Zt, T *z_w, T *z_r,
                                                   const double *h, double hc, T *Hz, const double *sc_r,
                                                   const double *sc_w, const double *Cs_r, const double *Cs_w)
{
    COLUMN_INDEX
    const long long S2 = g.S2, N = g.N;
    const double zt = Zt[f], hh = h[f];
    double zlow = 0.0;     // z_w of level K-1 (FP64) for Hz
    for (long long K = 0; K < N; K++) {
        double zw, zr;
        if (Vtransform == 1) {
            const double z_r0 = hc * (sc_r[K] - Cs_r[K]) + Cs_r[K] * hh;
            const double z_w0 = hc * (sc_w[K] - Cs_w[K]) + Cs_w[K] * hh;
            zw = z_w0 + zt * (1.0 + z_w0 / hh);
            zr = z_r0 + zt * (1.0 + z_r0 / hh);
        } else {
            const double cr = (hc * sc_r[K] + Cs_r[K] * hh) / (hc + hh);
            const double cw = (hc * sc_w[K] + Cs_w[K] * hh) / (hc + hh);
            zw = zt + (zt + hh) * cw;
            zr = zt + (zt + hh) * cr;
        }
        z_w[K * S2 + f] = T(zw);
        z_r[K * S2 + f] = T(zr);
        if (K > 0) Hz[(K - 1) * S2 + f] = T(zw - zlow);
        zlow = zw;
    }
    z_w[N * S2 + f] = T(zt);
    Hz[(N - 1) * S2 + f] = T(zt - zlow);
}

template <class T>
__global__ __launch_bounds__(256) void k_set_zeta(G3 g, T *zeta1, T *zeta2, const double *Zt)
{
    COLUMN_INDEX
    if (!isRNode(g, f)) return;
    zeta2[f] = T(Zt[f]);
    zeta1[f] = T(Zt[f]);
}

// ---- setLateralUVBCs (step3D_uv.kernels.cpp:314-368) ------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_lateral_uv(G3 g, double t, T *u, T *v, const int32_t *uType,
                                                    const int32_t *vType)
{
    COLUMN_INDEX
    const int bu = uType[f], bv = vType[f];
    if (bu <= 0 && bv <= 0) return;
    if (bu > 0 && !(bu & (1 << 4)) && (bu & (1 << 3))) {
        const T val = T(0.05 * sin(t * 0.0002));
        for (long long k = 0; k < g.N; k++) u[f + k * g.S2] = val;
    }
    if (bv > 0 && ((bv & (1 << 4)) || (bv & (1 << 3))))
        for (long long k = 0; k < g.N; k++) v[f + k * g.S2] = T(0);
}

// ---- step3d_UV (step3D_uv.kernels.cpp:646-707) ------------------------------------
// T: the 3D fields, the 2D barotropic velocities and the tridiagonal scratch (double:
// the reference's arithmetic, bit for bit; float: FP32 bulk form).  The grid metrics and
// DU_avg1/DV_avg1 are FP64, and the column sums (depth means, barotropic velocities) are
// formed in FP64 for either T.
template <class T>
__device__ void build_matrix(const G3 &g, bool vdir, long long f, double cff, double Dt, double lambda,
                             T *SD, T *D, T *RHS, const T *Hz, const T *Akv, const T *z_r, const T *u, const T *ru)
{
    const long long S2 = g.S2, N = g.N;
    const long long d = vdir ? g.szI : 1;
    if (vdir ? !isVNode(g, f) : !isUNode(g, f)) return;
#define HZU(K) ((Hz[(K) * S2 + f] + Hz[(K) * S2 + f - d]) * T(0.5))
#define AKVU(K) ((Akv[(K) * S2 + f] + Akv[(K) * S2 + f - d]) * T(0.5))
#define DZ(K) (((z_r[(K) * S2 + f] - z_r[((K) - 1) * S2 + f]) + (z_r[(K) * S2 + f - d] - z_r[((K) - 1) * S2 + f - d])) * T(0.5))
    const T pmU = (T(g.pm[f]) + T(g.pm[f - d])) * T(0.5);
    const T pnU = (T(g.pn[f]) + T(g.pn[f - d])) * T(0.5);
    const T cdt_mn = T(cff * Dt) * (pmU * pnU);
    const T mldt = T((-lambda) * Dt);
    SD[(N - 1) * S2 + f] = T(0);
    D[f] = HZU(0);
    RHS[f] = T(0);
    for (long long K = 0; K < N - 1; K++) SD[K * S2 + f] = (mldt * AKVU(K)) / DZ(K + 1);
    for (long long K = 1; K < N; K++) {
        RHS[K * S2 + f] = HZU(K) * (u[K * S2 + f] + cdt_mn * ru[K * S2 + f]);
        D[K * S2 + f] = (HZU(K) - SD[K * S2 + f]) - SD[(K - 1) * S2 + f];
    }
    D[(N - 1) * S2 + f] = HZU(N - 1) - SD[(N - 2) * S2 + f];
    SD[f] = (mldt * AKVU(0)) / DZ(1);
    D[f] = HZU(0) - T(2) * SD[f];
#undef HZU
#undef AKVU
#undef DZ
}

template <class T>
__device__ void solve_tri(const G3 &g, long long f, T *SD, T *D, T *RHS, T *x)
{
    const long long S2 = g.S2, N = g.N;
    if (!isUNode(g, f)) return;
    for (long long k = 1; k < N; k++) {
        const T w = SD[(k - 1) * S2 + f] / D[(k - 1) * S2 + f];
        D[k * S2 + f] -= w * SD[(k - 1) * S2 + f];
        RHS[k * S2 + f] -= w * RHS[(k - 1) * S2 + f];
    }
    x[(N - 1) * S2 + f] = RHS[(N - 1) * S2 + f] / D[(N - 1) * S2 + f];
    for (long long k = N - 2; k >= 0; k--)
        x[k * S2 + f] = (RHS[k * S2 + f] - SD[k * S2 + f] * x[(k + 1) * S2 + f]) / D[k * S2 + f];
}

template <class T>
__device__ void correct_mean(const G3 &g, bool vdir, long long f, T *x, const T *Hz, const double *Davg,
                             const double *metric, const T *xold)
{
    const long long S2 = g.S2, N = g.N;
    const long long d = vdir ? g.szI : 1;
    if (vdir ? !isVNode(g, f) : !isUNode(g, f)) return;
    double Hs = 0.0, Hx = 0.0;
    for (long long K = 0; K < N; K++) {
        const T hu = (Hz[K * S2 + f] + Hz[K * S2 + f - d]) * T(0.5);
        Hs += (double)hu;
        Hx += (double)xold[K * S2 + f] * (double)hu;
    }
    const double diff = (Hx - Davg[f] / metric[f]) / Hs;
    for (long long K = 0; K < N; K++) x[K * S2 + f] = T((double)x[K * S2 + f] - diff);
}

template <class T>
struct UVArgs {
    T *u, *v;
    const T *ru, *rv;
    T *ubar_t1, *vbar_t1, *ubar_t2, *vbar_t2;
    const T *Hz, *Akv, *z_r;
    const double *DU_avg1, *DV_avg1;
    const T *tmpU, *tmpV;
    T *SD, *D, *RHS;
};

template <class T>
__global__ __launch_bounds__(256) void k_step3d_uv(G3 g, UVArgs<T> a, double cff, double lambda, double Dt)
{
    COLUMN_INDEX
    const long long S2 = g.S2, N = g.N;
    build_matrix(g, false, f, cff, Dt, lambda, a.SD, a.D, a.RHS, a.Hz, a.Akv, a.z_r, a.tmpU, a.ru);
    solve_tri(g, f, a.SD, a.D, a.RHS, a.u);
    correct_mean(g, false, f, a.u, a.Hz, a.DU_avg1, g.on_u, a.tmpU);
    build_matrix(g, true, f, cff, Dt, lambda, a.SD, a.D, a.RHS, a.Hz, a.Akv, a.z_r, a.tmpV, a.rv);
    solve_tri(g, f, a.SD, a.D, a.RHS, a.v);
    correct_mean(g, true, f, a.v, a.Hz, a.DV_avg1, g.om_v, a.tmpV);
    if (!isUNode(g, f)) return;
    double HU = 0.0, HV = 0.0, HuU = 0.0, HvV = 0.0;
    for (long long K = 0; K < N; K++) {
        const long long b = K * S2 + f;
        const T hzu = (a.Hz[b] + a.Hz[b - 1]) * T(0.5);
        const T hzv = (a.Hz[b] + a.Hz[b - g.szI]) * T(0.5);
        HU += (double)hzu;
        HV += (double)hzv;
        HuU += (double)hzu * (double)a.u[b];
        HvV += (double)hzv * (double)a.v[b];
    }
    const double uUp = (HuU - a.DU_avg1[f] / g.on_u[f]) / HU;
    const double vUp = (HvV - a.DV_avg1[f] / g.om_v[f]) / HV;
    const double ub = (a.DU_avg1[f] / g.on_u[f]) / HU + uUp;
    const double vb = (a.DV_avg1[f] / g.om_v[f]) / HV + vUp;
    a.ubar_t2[f] = T(ub);
    a.vbar_t2[f] = T(vb);
    a.ubar_t1[f] = T(ub);
    a.vbar_t1[f] = T(vb);
}

// step3d_UV with each column's tridiagonal system held in registers (same arithmetic,
// same order as k_step3d_uv; N <= NMAX).  The matrix of a direction is built and
// eliminated level by level (build_matrix and the forward sweep of solve_tri fused),
// the back substitution stores the solution with correct_mean's shift applied, and the
// depth sums of correct_mean and adjustBarotropicVelocity run in the reference's level
// order (the barotropic sums re-read the stored velocities).  At u-but-not-v points the
// v pass solves the u pass's eliminated system a second time, as solve_tri does on the
// shared scratch.  Each thread reads tmpU / tmpV only before it writes u / v, so they may
// alias u / v.
template <class T, int NMAX>
__global__ __launch_bounds__(256) void k_step3d_uv_reg(G3 g, UVArgs<T> a, double cff, double lambda, double Dt)
{
    COLUMN_INDEX
    const long long S2 = g.S2;
    const int N = g.N;
    const bool un = isUNode(g, f), vn = isVNode(g, f);
    if (!un && !vn) return;
    T SD[NMAX], D[NMAX], R[NMAX];
    const T mldt = T((-lambda) * Dt);
    // build_matrix + forward elimination of direction vdir; Hs, Hx: correct_mean's sums.
    // The operands of level K+1 are loaded before level K is computed.
    auto build_elim = [&](bool vdir, const T *xold, const T *r, double &Hs, double &Hx) {
        const long long d = vdir ? g.szI : 1;
        const T pmU = (T(g.pm[f]) + T(g.pm[f - d])) * T(0.5);
        const T pnU = (T(g.pn[f]) + T(g.pn[f - d])) * T(0.5);
        const T cdt_mn = T(cff * Dt) * (pmU * pnU);
        T zf0 = a.z_r[f], zd0 = a.z_r[f - d];
        struct Lv {
            T h0, h1, xo, rr, z0, z1, k0, k1;
        };
        auto load = [&](int K) {
            Lv q{};
            const long long b = K * S2 + f;
            q.h0 = a.Hz[b];
            q.h1 = a.Hz[b - d];
            q.xo = xold[b];
            q.rr = r[b];
            if (K < N - 1) {
                q.z0 = a.z_r[b + S2];
                q.z1 = a.z_r[b + S2 - d];
                q.k0 = a.Akv[b];
                q.k1 = a.Akv[b - d];
            }
            return q;
        };
        Lv cur = load(0);
        Hs = 0.0;
        Hx = 0.0;
#pragma unroll
        for (int K = 0; K < NMAX; K++) {
            if (K < N) {
                Lv nxt{};
                if (K + 1 < N) nxt = load(K + 1);
                const T hz = (cur.h0 + cur.h1) * T(0.5);
                const T xo = cur.xo;
                Hs += (double)hz;
                Hx += (double)xo * (double)hz;
                T sd = T(0);
                if (K < N - 1) {
                    const T dz = ((cur.z0 - zf0) + (cur.z1 - zd0)) * T(0.5);
                    sd = (mldt * ((cur.k0 + cur.k1) * T(0.5))) / dz;
                    zf0 = cur.z0;
                    zd0 = cur.z1;
                }
                SD[K] = sd;
                if (K == 0) {
                    D[0] = hz - T(2) * sd;
                    R[0] = T(0);
                } else {
                    D[K] = (hz - sd) - SD[K - 1];
                    R[K] = hz * (xo + cdt_mn * cur.rr);
                    const T w = SD[K - 1] / D[K - 1];
                    D[K] -= w * SD[K - 1];
                    R[K] -= w * R[K - 1];
                }
                cur = nxt;
            }
        }
    };
    // the forward sweep of solve_tri on an already eliminated system
    auto elim_again = [&]() {
#pragma unroll
        for (int K = 1; K < NMAX; K++) {
            if (K < N) {
                const T w = SD[K - 1] / D[K - 1];
                D[K] -= w * SD[K - 1];
                R[K] -= w * R[K - 1];
            }
        }
    };
    // back substitution; stores x - shift at every level of x
    auto back_store = [&](T *x, double shift) {
        T xn = T(0);
#pragma unroll
        for (int K = NMAX - 1; K >= 0; K--) {
            if (K < N) {
                const T xk = K == N - 1 ? R[K] / D[K] : (R[K] - SD[K] * xn) / D[K];
                x[K * S2 + f] = T((double)xk - shift);
                xn = xk;
            }
        }
    };
    double HU = 0.0, HuU = 0.0;
    if (un) {
        double Hs, Hx;
        build_elim(false, a.tmpU, a.ru, Hs, Hx);
        const double diff = (Hx - a.DU_avg1[f] / g.on_u[f]) / Hs;
        back_store(a.u, diff);   // leaves SD, D, R as they are (reused below where !vn)
    }
    if (vn) {
        double Hs, Hx;
        build_elim(true, a.tmpV, a.rv, Hs, Hx);
        const double diff = (Hx - a.DV_avg1[f] / g.om_v[f]) / Hs;
        if (un) {
            back_store(a.v, diff);
        } else {
            for (int K = 0; K < N; K++) a.v[K * S2 + f] = T((double)a.v[K * S2 + f] - diff);
        }
    } else {
        elim_again();
        back_store(a.v, 0.0);
    }
    if (!un) return;
    double HV = 0.0, HvV = 0.0;
    for (int K = 0; K < N; K++) {
        const long long b = K * S2 + f;
        const T hzu = (a.Hz[b] + a.Hz[b - 1]) * T(0.5);
        const T hzv = (a.Hz[b] + a.Hz[b - g.szI]) * T(0.5);
        HU += (double)hzu;
        HV += (double)hzv;
        HuU += (double)hzu * (double)a.u[b];
        HvV += (double)hzv * (double)a.v[b];
    }
    const double uUp = (HuU - a.DU_avg1[f] / g.on_u[f]) / HU;
    const double vUp = (HvV - a.DV_avg1[f] / g.om_v[f]) / HV;
    const double ub = (a.DU_avg1[f] / g.on_u[f]) / HU + uUp;
    const double vb = (a.DV_avg1[f] / g.om_v[f]) / HV + vUp;
    a.ubar_t2[f] = T(ub);
    a.vbar_t2[f] = T(vb);
    a.ubar_t1[f] = T(ub);
    a.vbar_t1[f] = T(vb);
}

// ---- coupling of the 2D and 3D momentum equations (ROMS split-explicit form) ----------
// rufrc / rvfrc: vertical integral of the 3D right-hand side plus the surface and bottom
// stresses (nonlinear/rhs.py:76-89, commented in the reference; bottom stress rdrg * u of
// the bottom level), on the interior u (v) points of the physics substep, 0 elsewhere;
// summed in FP64 for either field precision.
template <class T>
__global__ __launch_bounds__(256) void k_couple_rufrc(G3 g, const T *ru, const T *rv, const T *u, const T *v,
                                                      const double *sustr, const double *svstr, double rdrg,
                                                      double *rufrc, double *rvfrc)
{
    COLUMN_INDEX
    const long long S2 = g.S2, N = g.N, W = g.szI;
    const int i = (int)(f % W), j = (int)(f / W), L = g.szI - 1, M = g.szJ - 1;
    double su = 0.0, sv = 0.0;
    if (i >= 2 && i <= L - 1 && j >= 1 && j <= M - 1) {
        for (long long K = 0; K < N; K++) su = su + (double)ru[K * S2 + f];
        const double onu = 2.0 / (g.pn[f - 1] + g.pn[f]), omu = 2.0 / (g.pm[f - 1] + g.pm[f]);
        su = su + ((sustr[f] - rdrg * (double)u[f]) * omu) * onu;
    }
    if (j >= 2 && j <= M - 1 && i >= 1 && i <= L - 1) {
        for (long long K = 0; K < N; K++) sv = sv + (double)rv[K * S2 + f];
        const double omv = 2.0 / (g.pm[f - W] + g.pm[f]), onv = 2.0 / (g.pn[f - W] + g.pn[f]);
        sv = sv + ((svstr[f] - rdrg * (double)v[f]) * omv) * onv;
    }
    rufrc[f] = su;
    rvfrc[f] = sv;
}

// step3d_UV coupled with the barotropic mode (nonlinear/step3d_uv.py:155-180, :365-544):
// the implicit vertical viscosity of k_step3d_uv (same matrix and solve), then per
// interior u (v) column, with DC(K) = on_u * Hz_u(K) and D = sum DC(K) formed in FP64:
//   CF = (sum DC u - DU_avg1) / D;  u -= CF;  ubar(t1) = ubar(t2) = DU_avg1 / D;
//   Huon = 0.5 (Huon + u DC);  FC = (sum Huon - DU_avg2) / D;  Huon -= DC FC.
template <class T>
__device__ void couple_column(const G3 &g, bool vdir, long long f, T *x, const T *Hz, const double *Davg1,
                              const double *Davg2, T *Hf, T *xbar1, T *xbar2)
{
    const long long S2 = g.S2, N = g.N, W = g.szI;
    const int i = (int)(f % W), j = (int)(f / W), L = g.szI - 1, M = g.szJ - 1;
    if (!vdir && !(i >= 2 && i <= L - 1 && j >= 1 && j <= M - 1)) return;
    if (vdir && !(j >= 2 && j <= M - 1 && i >= 1 && i <= L - 1)) return;
    const long long d = vdir ? W : 1;
    const double met = vdir ? 2.0 / (g.pm[f - W] + g.pm[f]) : 2.0 / (g.pn[f - 1] + g.pn[f]);
#define DC(K) (met * (double)((Hz[(K) * S2 + f] + Hz[(K) * S2 + f - d]) * T(0.5)))
    double Ds = 0.0, CF = 0.0;
    for (long long K = 0; K < N; K++) {
        Ds = Ds + DC(K);
        CF = CF + DC(K) * (double)x[K * S2 + f];
    }
    const double rD = 1.0 / Ds;
    CF = rD * (CF - Davg1[f]);
    const T xb = T(rD * Davg1[f]);
    xbar2[f] = xb;
    xbar1[f] = xb;
    for (long long K = 0; K < N; K++) x[K * S2 + f] = T((double)x[K * S2 + f] - CF);
    double FC = 0.0;
    for (long long K = 0; K < N; K++) {
        const T hf = T(0.5 * ((double)Hf[K * S2 + f] + (double)x[K * S2 + f] * DC(K)));
        Hf[K * S2 + f] = hf;
        FC = FC + (double)hf;
    }
    FC = rD * (FC - Davg2[f]);
    for (long long K = 0; K < N; K++) Hf[K * S2 + f] = T((double)Hf[K * S2 + f] - DC(K) * FC);
#undef DC
}

// The 3D momentum step of the coupled form (ROMS step3d_uv): per interior u (v) column,
//   Hz_u,new u_new - dt d/dz(Akv du_new/dz) = Hz_u,old u + dt pm_u pn_u AB3(ru)
//                                             - dt rdrg u(0) [bottom] + dt sustr [top],
// with SD(K) = (-lambda dt) Akv_u(K+1) / dz_u(K+1) (K = 0..N-2), D(K) = (Hz_u(K) - SD(K))
// - SD(K-1), AB3 over this step's ru and the two stored before it (this ru is stored for
// the next step), solved by the Thomas algorithm in the scratch columns; couple_column
// then replaces the column mean.  The reference's solve (build_matrix above) multiplies
// ru by Hz a second time and keeps no history: with the coupling moving the 3D flow that
// form grows by about Hz f dt per step, so the coupled form solves ROMS's equation.
template <class T>
struct UVC {
    T *u, *v;
    const T *ru, *rv, *ru1, *rv1, *ru2, *rv2;
    T *ru_keep, *rv_keep;
    T *ubar_t1, *vbar_t1, *ubar_t2, *vbar_t2;
    const T *Hz_old, *Hz, *Akv, *z_r;
    const double *DU_avg1, *DV_avg1, *DU_avg2, *DV_avg2;
    T *Huon, *Hvom;
    const double *sustr, *svstr;
    double rdrg;
    int ab;
    T *SD, *D, *RHS;
};

__device__ inline bool coupled_point(const G3 &g, bool vdir, long long f)
{
    const int i = (int)(f % g.szI), j = (int)(f / g.szI), L = g.szI - 1, M = g.szJ - 1;
    return vdir ? (j >= 2 && j <= M - 1 && i >= 1 && i <= L - 1) : (i >= 2 && i <= L - 1 && j >= 1 && j <= M - 1);
}

template <class T>
__device__ void momentum_column(const G3 &g, bool vdir, long long f, const UVC<T> &a, double lambda, double Dt)
{
    const long long S2 = g.S2, N = g.N;
    const long long d = vdir ? g.szI : 1;
    T *x = vdir ? a.v : a.u, *keep = vdir ? a.rv_keep : a.ru_keep;
    const T *r = vdir ? a.rv : a.ru, *r1 = vdir ? a.rv1 : a.ru1, *r2 = vdir ? a.rv2 : a.ru2;
    const double stress = vdir ? a.svstr[f] : a.sustr[f];
    T *SD = a.SD, *D = a.D, *RHS = a.RHS;
    const T *Hz = a.Hz, *Ho = a.Hz_old, *Akv = a.Akv, *z_r = a.z_r;
#define AVG(A, K) (((A)[(K) * S2 + f] + (A)[(K) * S2 + f - d]) * T(0.5))
#define DZ(K) (((z_r[(K) * S2 + f] - z_r[((K) - 1) * S2 + f]) + (z_r[(K) * S2 + f - d] - z_r[((K) - 1) * S2 + f - d])) * T(0.5))
    const double pmU = (g.pm[f] + g.pm[f - d]) * 0.5;
    const double pnU = (g.pn[f] + g.pn[f - d]) * 0.5;
    const T cdt_mn = T(Dt * (pmU * pnU));
    const T mldt = T((-lambda) * Dt);
    for (long long K = 0; K < N - 1; K++) SD[K * S2 + f] = (mldt * AVG(Akv, K + 1)) / DZ(K + 1);
    SD[(N - 1) * S2 + f] = T(0);
    for (long long K = 0; K < N; K++) {
        const long long b = K * S2 + f;
        const T rk = r[b];
        T rab;
        if (a.ab == 0) rab = rk;
        else if (a.ab == 1) rab = T(1.5) * rk - T(0.5) * r1[b];
        else rab = (T(23.0 / 12.0) * rk - T(16.0 / 12.0) * r1[b]) + T(5.0 / 12.0) * r2[b];
        keep[b] = rk;
        RHS[b] = AVG(Ho, K) * x[b] + cdt_mn * rab;
        D[b] = K == 0 ? AVG(Hz, 0) - SD[f] : (AVG(Hz, K) - SD[b]) - SD[b - S2];
    }
    RHS[f] = RHS[f] - T(Dt * a.rdrg) * x[f];
    RHS[(N - 1) * S2 + f] = RHS[(N - 1) * S2 + f] + T(Dt * stress);
#undef AVG
#undef DZ
    for (long long k = 1; k < N; k++) {
        const T w = SD[(k - 1) * S2 + f] / D[(k - 1) * S2 + f];
        D[k * S2 + f] -= w * SD[(k - 1) * S2 + f];
        RHS[k * S2 + f] -= w * RHS[(k - 1) * S2 + f];
    }
    x[(N - 1) * S2 + f] = RHS[(N - 1) * S2 + f] / D[(N - 1) * S2 + f];
    for (long long k = N - 2; k >= 0; k--)
        x[k * S2 + f] = (RHS[k * S2 + f] - SD[k * S2 + f] * x[(k + 1) * S2 + f]) / D[k * S2 + f];
}

template <class T>
__global__ __launch_bounds__(256) void k_step3d_uv_coupled(G3 g, UVC<T> a, double lambda, double Dt)
{
    COLUMN_INDEX
    if (coupled_point(g, false, f)) {
        momentum_column(g, false, f, a, lambda, Dt);
        couple_column(g, false, f, a.u, a.Hz, a.DU_avg1, a.DU_avg2, a.Huon, a.ubar_t1, a.ubar_t2);
    }
    if (coupled_point(g, true, f)) {
        momentum_column(g, true, f, a, lambda, Dt);
        couple_column(g, true, f, a.v, a.Hz, a.DV_avg1, a.DV_avg2, a.Hvom, a.vbar_t1, a.vbar_t2);
    }
}

// The coupled momentum step with each column's system in registers (same arithmetic
// and order as momentum_column + couple_column; N <= NMAX): the matrix is built and
// eliminated level by level with the operands of the next level loaded ahead, the
// solution stays in registers for the column sums, and the new thicknesses of the
// mass-flux correction are kept in the slots of the eliminated diagonal.
//
// Addressing: every access is an array's scalar base plus a 32-bit byte offset held in a
// VGPR (the launcher checks that (N + 1) levels fit in 2^32 bytes).  The offset of each
// level is the previous one plus the level stride, passed through an empty asm statement
// so that the compiler does not fold the unrolled levels into per-level 64-bit scalar
// bases (dozens of arrays x levels, which it spilled to VGPR lanes and restored at every
// use).
template <class T>
__device__ __forceinline__ T ldb(const T *p, uint32_t o)
{
    return *(const __attribute__((address_space(1))) T *)((const __attribute__((address_space(1))) char *)p + o);
}
template <class T>
__device__ __forceinline__ void stb(T *p, uint32_t o, T v)
{
    *(__attribute__((address_space(1))) T *)((__attribute__((address_space(1))) char *)p + o) = v;
}
__device__ __forceinline__ uint32_t next_level(uint32_t o, uint32_t stride)
{
    o += stride;
    asm("" : "+v"(o));
    return o;
}

template <class T, int NMAX>
__global__ __launch_bounds__(256) void k_step3d_uv_coupled_reg(G3 g, UVC<T> a, double lambda, double Dt)
{
    COLUMN_INDEX
    const int N = g.N;
    const uint32_t SB = (uint32_t)(g.S2 * (long long)sizeof(T));   // level stride in bytes
    const uint32_t o0 = (uint32_t)(f * (long long)sizeof(T));      // level 0 of the column
    __shared__ T sds[NMAX][256];   // the sub-diagonal of each column (thread)
    T D[NMAX], R[NMAX];
    const int tid = threadIdx.x;
    const T mldt = T((-lambda) * Dt);
    auto column = [&](bool vdir) {
        const long long d = vdir ? g.szI : 1;
        const uint32_t dB = (uint32_t)(d * (long long)sizeof(T));
        T *x = vdir ? a.v : a.u, *keep = vdir ? a.rv_keep : a.ru_keep, *Hf = vdir ? a.Hvom : a.Huon;
        const T *r = vdir ? a.rv : a.ru, *r1 = vdir ? a.rv1 : a.ru1, *r2 = vdir ? a.rv2 : a.ru2;
        const double stress = vdir ? a.svstr[f] : a.sustr[f];
        const double pmU = (g.pm[f] + g.pm[f - d]) * 0.5;
        const double pnU = (g.pn[f] + g.pn[f - d]) * 0.5;
        const T cdt_mn = T(Dt * (pmU * pnU));
        const int ab = a.ab;
        struct Lv {
            T hn0, hn1, ho0, ho1, xo, rk, rk1, rk2, z0, z1, k0, k1;
        };
        // operands of the level at byte offset ob (z_r and Akv of the level above)
        auto load = [&](uint32_t ob, bool top) {
            Lv q{};
            q.hn0 = ldb(a.Hz, ob);
            q.hn1 = ldb(a.Hz, ob - dB);
            q.ho0 = ldb(a.Hz_old, ob);
            q.ho1 = ldb(a.Hz_old, ob - dB);
            q.xo = ldb((const T *)x, ob);
            q.rk = ldb(r, ob);
            if (ab > 0) q.rk1 = ldb(r1, ob);
            if (ab > 1) q.rk2 = ldb(r2, ob);
            if (!top) {
                q.z0 = ldb(a.z_r, ob + SB);
                q.z1 = ldb(a.z_r, ob + SB - dB);
                q.k0 = ldb(a.Akv, ob + SB);
                q.k1 = ldb(a.Akv, ob + SB - dB);
            }
            return q;
        };
        T zf0 = ldb(a.z_r, o0), zd0 = ldb(a.z_r, o0 - dB);
        const T xbot = ldb((const T *)x, o0);
        T sdp = T(0);   // sub-diagonal of the level below
        uint32_t ob = o0;
        Lv cur = load(ob, N - 1 == 0);
#pragma unroll
        for (int K = 0; K < NMAX; K++) {
            if (K < N) {
                const uint32_t on = next_level(ob, SB);
                Lv nxt{};
                if (K + 1 < N) nxt = load(on, K + 1 >= N - 1);
                T sd = T(0);
                if (K < N - 1) {
                    const T dz = ((cur.z0 - zf0) + (cur.z1 - zd0)) * T(0.5);
                    sd = (mldt * ((cur.k0 + cur.k1) * T(0.5))) / dz;
                    zf0 = cur.z0;
                    zd0 = cur.z1;
                }
                sds[K][tid] = sd;
                T rab;
                if (ab == 0) rab = cur.rk;
                else if (ab == 1) rab = T(1.5) * cur.rk - T(0.5) * cur.rk1;
                else rab = (T(23.0 / 12.0) * cur.rk - T(16.0 / 12.0) * cur.rk1) + T(5.0 / 12.0) * cur.rk2;
                stb(keep, ob, cur.rk);
                T rhs = ((cur.ho0 + cur.ho1) * T(0.5)) * cur.xo + cdt_mn * rab;
                const T hn = (cur.hn0 + cur.hn1) * T(0.5);
                if (K == 0) {
                    D[0] = hn - sd;
                    R[0] = rhs - T(Dt * a.rdrg) * xbot;
                } else {
                    D[K] = (hn - sd) - sdp;
                    if (K == N - 1) rhs = rhs + T(Dt * stress);
                    R[K] = rhs;
                    const T w = sdp / D[K - 1];
                    D[K] -= w * sdp;
                    R[K] -= w * R[K - 1];
                }
                sdp = sd;
                cur = nxt;
                ob = on;
            }
        }
        // back substitution: R(K) := x(K)
#pragma unroll
        for (int K = NMAX - 1; K >= 0; K--) {
            if (K < N) R[K] = K == N - 1 ? R[K] / D[K] : (R[K] - sds[K][tid] * R[K + 1]) / D[K];
        }
        // couple_column: replace the column mean, correct the mass flux
        const double met = vdir ? 2.0 / (g.pm[f - g.szI] + g.pm[f]) : 2.0 / (g.pn[f - 1] + g.pn[f]);
        double Ds = 0.0, CF = 0.0;
        ob = o0;
#pragma unroll
        for (int K = 0; K < NMAX; K++) {
            if (K < N) {
                const T hn = (ldb(a.Hz, ob) + ldb(a.Hz, ob - dB)) * T(0.5);
                D[K] = hn;
                const double dc = met * (double)hn;
                Ds = Ds + dc;
                CF = CF + dc * (double)R[K];
                ob = next_level(ob, SB);
            }
        }
        const double rD = 1.0 / Ds;
        const double *Davg1 = vdir ? a.DV_avg1 : a.DU_avg1, *Davg2 = vdir ? a.DV_avg2 : a.DU_avg2;
        CF = rD * (CF - Davg1[f]);
        const T xb = T(rD * Davg1[f]);
        (vdir ? a.vbar_t2 : a.ubar_t2)[f] = xb;
        (vdir ? a.vbar_t1 : a.ubar_t1)[f] = xb;
        double FC = 0.0;
        ob = o0;
#pragma unroll
        for (int K = 0; K < NMAX; K++) {
            if (K < N) {
                const T xk = T((double)R[K] - CF);
                stb(x, ob, xk);
                const T hf = T(0.5 * ((double)ldb((const T *)Hf, ob) + (double)xk * (met * (double)D[K])));
                R[K] = hf;
                FC = FC + (double)hf;
                ob = next_level(ob, SB);
            }
        }
        FC = rD * (FC - Davg2[f]);
        ob = o0;
#pragma unroll
        for (int K = 0; K < NMAX; K++) {
            if (K < N) {
                stb(Hf, ob, T((double)R[K] - (met * (double)D[K]) * FC));
                ob = next_level(ob, SB);
            }
        }
    };
    if (coupled_point(g, false, f)) column(false);
    if (coupled_point(g, true, f)) column(true);
}

G3 grid3(const sf_ctx *c)
{
    return G3{c->szI, c->szJ, c->szK, c->S2, c->on_u, c->om_v, c->pn, c->pm};
}

}  // namespace

#define CTX3(c)                                                                  \
    SF_REQUIRE((c) != nullptr, "%s: null context", __func__);                   \
    SF_REQUIRE((c)->szK >= 4, "%s: needs at least 4 levels (N=%d)", __func__, (c)->szK)

#define LAUNCH_COLUMNS(kern, c, stream, ...)                                                                  \
    hipLaunchKernelGGL(kern, dim3(sf::grid1d((c)->S2)), dim3(sf::kBlock1D), 0, sf::as_stream(stream), __VA_ARGS__); \
    return sf::check_launch(__func__)

extern "C" int sf_add_coriolis(sf_ctx *c, const double *fomn, const double *u, const double *v, double *ru,
                               double *rv, const double *Hz, void *stream)
{
    CTX3(c);
    SF_REQUIRE(fomn && u && v && ru && rv && Hz, "sf_add_coriolis: null field");
    const long long n3 = (c->szK + 1LL) * c->S2;
    LAUNCH_COLUMNS(k_addCoriolis, c, stream, grid3(c), A{fomn, c->S2}, A{u, n3}, A{v, n3}, ru, rv,
                   A{Hz, (long long)c->szK * c->S2});
}

extern "C" int sf_horizontal_advection(sf_ctx *c, const double *u, const double *v, const double *Huon,
                                       const double *Hvom, double *ru, double *rv, int64_t launch_threads,
                                       void *stream)
{
    CTX3(c);
    SF_REQUIRE(u && v && Huon && Hvom && ru && rv, "sf_horizontal_advection: null field");
    const long long n3 = (c->szK + 1LL) * c->S2;
    H h{A{u, n3}, A{v, n3}, A{Huon, n3}, A{Hvom, n3}, c->szI};
    const int tiles_i = (c->szI + HTI - 1) / HTI, tiles_j = (c->szJ + HTJ - 1) / HTJ;
    hipLaunchKernelGGL(k_horizontalAdvection_tiled, dim3((unsigned)(tiles_i * tiles_j)), dim3(HNT), 0,
                       sf::as_stream(stream), grid3(c), h, ru, rv, tiles_i);
    int rc = sf::check_launch("sf_horizontal_advection");
    if (rc != SF_OK) return rc;
    const long long end = launch_threads < n3 ? (long long)launch_threads : n3;
    if (end > c->S2) {
        hipLaunchKernelGGL(k_hadv_tail, dim3(sf::grid1d(end - c->S2)), dim3(sf::kBlock1D), 0, sf::as_stream(stream),
                           c->S2, end, ru, rv);
        rc = sf::check_launch("sf_horizontal_advection (tail)");
    }
    return rc;
}

extern "C" int sf_vertical_advection(sf_ctx *c, const double *u, const double *v, const double *W, double *ru,
                                     double *rv, const int32_t *BC, void *stream)
{
    CTX3(c);
    SF_REQUIRE(u && v && W && ru && rv && BC, "sf_vertical_advection: null field");
    const long long n3 = (c->szK + 1LL) * c->S2;
    LAUNCH_COLUMNS(k_verticalAdvection, c, stream, grid3(c), A{u, n3}, A{v, n3}, A{W, n3}, ru, rv, BC);
}

template <class T>
static int rhs3d_launch(sf_ctx *c, const double *fomn, const T *u, const T *v, const T *Huon, const T *Hvom,
                        const T *W, const T *Hz, const int32_t *BC, T *ru, T *rv, int64_t launch_threads,
                        int32_t cor, int32_t adv, void *stream)
{
    const long long n3 = (c->szK + 1LL) * c->S2;
    HG<T> h{AG<T>{u, n3}, AG<T>{v, n3}, AG<T>{Huon, n3}, AG<T>{Hvom, n3}, c->szI};
    const long long end = launch_threads < n3 ? (long long)launch_threads : n3;
    constexpr int TJ = 8;   // 512-thread workgroups, 64 x 8 columns
    const int tiles_i = (c->szI + HTI - 1) / HTI, tiles_j = (c->szJ + TJ - 1) / TJ, ntiles = tiles_i * tiles_j;
    const long long cus = sf::cu_count(c->device);
    // FP32: the instance compiled for 8 waves per SIMD is faster on grids of at least four
    // tiles per CU (1024 x 512 x 32: 222-228 us against 324 us for the 4-wave instance,
    // both 50 VGPRs and 37 KB LDS -- the wave target changes the compiler's schedule, not
    // only its register limit); smaller grids run the 4-wave instance with level chunks
    // (512 x 256 x 32: 65 us in three chunks, 73 us in four).
    const bool wide32 = sizeof(T) == 4 && ntiles >= 4 * cus;
    // level chunks: the fewest that give every resident workgroup slot (three per CU) a
    // workgroup, with at least 3 levels per chunk (the bottom and top special levels each
    // stay within one chunk); the last chunk takes the remainder.  1024 x 512 x 32 runs
    // unsplit.
    const long long slots = 3 * cus;
    int nchunk = 1;
    while (!wide32 && (long long)ntiles * nchunk < slots && c->szK / (nchunk + 1) >= 3) nchunk++;
    const int kchunk = c->szK / nchunk;
    const dim3 grid((unsigned)ntiles * (unsigned)nchunk), block(HTI * TJ);
    const AG<T> hz{Hz, (long long)c->szK * c->S2}, wa{W, n3};
    if (wide32) {
        hipLaunchKernelGGL((k_rhs3d_fused<TJ, T, 8>), grid, block, 0, sf::as_stream(stream), grid3(c), h,
                           A{fomn, c->S2}, hz, wa, BC, ru, rv, end, cor ? 1 : 0, adv ? 1 : 0, tiles_i, ntiles,
                           kchunk, nchunk);
    } else {
        constexpr int WPE = sizeof(T) == 8 ? 6 : 4;   // FP64: at most 80 VGPRs, three workgroups per CU
        hipLaunchKernelGGL((k_rhs3d_fused<TJ, T, WPE>), grid, block, 0, sf::as_stream(stream), grid3(c), h,
                           A{fomn, c->S2}, hz, wa, BC, ru, rv, end, cor ? 1 : 0, adv ? 1 : 0, tiles_i, ntiles,
                           kchunk, nchunk);
    }
    return sf::check_launch("sf_rhs3d");
}

extern "C" int sf_rhs3d(sf_ctx *c, const double *fomn, const double *u, const double *v, const double *Huon,
                        const double *Hvom, const double *W, const double *Hz, const int32_t *BC, double *ru,
                        double *rv, int64_t launch_threads, int32_t cor, int32_t adv, void *stream)
{
    CTX3(c);
    SF_REQUIRE(fomn && u && v && Huon && Hvom && W && Hz && BC && ru && rv, "sf_rhs3d: null field");
    return rhs3d_launch<double>(c, fomn, u, v, Huon, Hvom, W, Hz, BC, ru, rv, launch_threads, cor, adv, stream);
}

extern "C" int sf_rhs3d_f32(sf_ctx *c, const double *fomn, const float *u, const float *v, const float *Huon,
                            const float *Hvom, const float *W, const float *Hz, const int32_t *BC, float *ru,
                            float *rv, int64_t launch_threads, int32_t cor, int32_t adv, void *stream)
{
    CTX3(c);
    SF_REQUIRE(fomn && u && v && Huon && Hvom && W && Hz && BC && ru && rv, "sf_rhs3d_f32: null field");
    return rhs3d_launch<float>(c, fomn, u, v, Huon, Hvom, W, Hz, BC, ru, rv, launch_threads, cor, adv, stream);
}

extern "C" int sf_set_maxflux(sf_ctx *c, const double *u, const double *v, double *Huon, double *Hvom,
                              const double *Hz, void *stream)
{
    CTX3(c);
    SF_REQUIRE(u && v && Huon && Hvom && Hz, "sf_set_maxflux: null field");
    LAUNCH_COLUMNS(k_set_maxflux<double>, c, stream, grid3(c), u, v, Huon, Hvom, Hz);
}

extern "C" int sf_set_maxflux_f32(sf_ctx *c, const float *u, const float *v, float *Huon, float *Hvom,
                                  const float *Hz, void *stream)
{
    CTX3(c);
    SF_REQUIRE(u && v && Huon && Hvom && Hz, "sf_set_maxflux_f32: null field");
    LAUNCH_COLUMNS(k_set_maxflux<float>, c, stream, grid3(c), u, v, Huon, Hvom, Hz);
}

template <class T>
static int omega_launch(sf_ctx *c, T *W, const T *Huon, const T *Hvom, const T *z_w, const int32_t *BC,
                        void *stream)
{
    const long long n3 = (c->szK + 1LL) * c->S2;
    const AG<T> hu{Huon, n3}, hv{Hvom, n3};
    if (c->szK <= 16) {
        LAUNCH_COLUMNS((k_omega_reg<16, T>), c, stream, grid3(c), W, hu, hv, z_w, BC);
    }
    if (c->szK <= 32) {
        LAUNCH_COLUMNS((k_omega_reg<32, T>), c, stream, grid3(c), W, hu, hv, z_w, BC);
    }
    LAUNCH_COLUMNS((k_omega_reg<64, T>), c, stream, grid3(c), W, hu, hv, z_w, BC);
}

extern "C" int sf_omega(sf_ctx *c, double *W, const double *Huon, const double *Hvom, const double *z_w,
                        const int32_t *BC, void *stream)
{
    CTX3(c);
    SF_REQUIRE(W && Huon && Hvom && z_w && BC, "sf_omega: null field");
    if (c->szK <= 64) return omega_launch<double>(c, W, Huon, Hvom, z_w, BC, stream);
    const long long n3 = (c->szK + 1LL) * c->S2;
    LAUNCH_COLUMNS(k_omega, c, stream, grid3(c), W, A{Huon, n3}, A{Hvom, n3}, z_w, BC);
}

extern "C" int sf_omega_f32(sf_ctx *c, float *W, const float *Huon, const float *Hvom, const float *z_w,
                            const int32_t *BC, void *stream)
{
    CTX3(c);
    SF_REQUIRE(W && Huon && Hvom && z_w && BC, "sf_omega_f32: null field");
    SF_REQUIRE(c->szK <= 64, "sf_omega_f32: at most 64 levels (N=%d)", c->szK);
    return omega_launch<float>(c, W, Huon, Hvom, z_w, BC, stream);
}

extern "C" int sf_set_depth(sf_ctx *c, int32_t Vtransform, const double *Zt_avg1, double *z_w, double *z_r,
                            const double *h, double hc, double *Hz, const double *sc_r, const double *sc_w,
                            const double *Cs_r, const double *Cs_w, void *stream)
{
    CTX3(c);
    SF_REQUIRE(Zt_avg1 && z_w && z_r && h && Hz && sc_r && sc_w && Cs_r && Cs_w, "sf_set_depth: null field");
    SF_REQUIRE(Vtransform == 1 || Vtransform == 2, "sf_set_depth: Vtransform=%d", Vtransform);
    LAUNCH_COLUMNS(k_set_depth<double>, c, stream, grid3(c), Vtransform, Zt_avg1, z_w, z_r, h, hc, Hz, sc_r, sc_w,
                   Cs_r, Cs_w);
}

extern "C" int sf_set_depth_f32(sf_ctx *c, int32_t Vtransform, const double *Zt_avg1, float *z_w, float *z_r,
                                const double *h, double hc, float *Hz, const double *sc_r, const double *sc_w,
                                const double *Cs_r, const double *Cs_w, void *stream)
{
    CTX3(c);
    SF_REQUIRE(Zt_avg1 && z_w && z_r && h && Hz && sc_r && sc_w && Cs_r && Cs_w, "sf_set_depth_f32: null field");
    SF_REQUIRE(Vtransform == 1 || Vtransform == 2, "sf_set_depth_f32: Vtransform=%d", Vtransform);
    LAUNCH_COLUMNS(k_set_depth<float>, c, stream, grid3(c), Vtransform, Zt_avg1, z_w, z_r, h, hc, Hz, sc_r, sc_w,
                   Cs_r, Cs_w);
}

extern "C" int sf_set_zeta(sf_ctx *c, double *zeta_t1, double *zeta_t2, const double *Zt_avg1, void *stream)
{
    SF_REQUIRE(c && zeta_t1 && zeta_t2 && Zt_avg1, "sf_set_zeta: null argument");
    LAUNCH_COLUMNS(k_set_zeta<double>, c, stream, grid3(c), zeta_t1, zeta_t2, Zt_avg1);
}

extern "C" int sf_set_zeta_f32(sf_ctx *c, float *zeta_t1, float *zeta_t2, const double *Zt_avg1, void *stream)
{
    SF_REQUIRE(c && zeta_t1 && zeta_t2 && Zt_avg1, "sf_set_zeta_f32: null argument");
    LAUNCH_COLUMNS(k_set_zeta<float>, c, stream, grid3(c), zeta_t1, zeta_t2, Zt_avg1);
}

extern "C" int sf_set_lateral_uv_bcs(sf_ctx *c, double t, double *u, double *v, const int32_t *uType,
                                     const int32_t *vType, void *stream)
{
    SF_REQUIRE(c && u && v && uType && vType, "sf_set_lateral_uv_bcs: null argument");
    LAUNCH_COLUMNS(k_lateral_uv<double>, c, stream, grid3(c), t, u, v, uType, vType);
}

extern "C" int sf_set_lateral_uv_bcs_f32(sf_ctx *c, double t, float *u, float *v, const int32_t *uType,
                                         const int32_t *vType, void *stream)
{
    SF_REQUIRE(c && u && v && uType && vType, "sf_set_lateral_uv_bcs_f32: null argument");
    LAUNCH_COLUMNS(k_lateral_uv<float>, c, stream, grid3(c), t, u, v, uType, vType);
}

// step3d_UV: columns of up to 32 levels in registers (k_step3d_uv_reg), deeper ones
// through the SD/D/RHS scratch (k_step3d_uv)
template <class T>
static int step3d_uv_launch(sf_ctx *c, const UVArgs<T> &a, double cff, double lambda, double Dt, void *stream)
{
    if (c->szK <= 16) {
        LAUNCH_COLUMNS((k_step3d_uv_reg<T, 16>), c, stream, grid3(c), a, cff, lambda, Dt);
    }
    if (c->szK <= 32) {
        LAUNCH_COLUMNS((k_step3d_uv_reg<T, 32>), c, stream, grid3(c), a, cff, lambda, Dt);
    }
    LAUNCH_COLUMNS(k_step3d_uv<T>, c, stream, grid3(c), a, cff, lambda, Dt);
}

extern "C" int sf_step3d_uv(sf_ctx *c, double *u, double *v, const double *ru, const double *rv, double *ubar_t1,
                            double *vbar_t1, double *ubar_t2, double *vbar_t2, const double *Hz, const double *Akv,
                            const double *z_r, const double *DU_avg1, const double *DV_avg1, const double *tmpU,
                            const double *tmpV, int32_t iic, int32_t ntfirst, double lambda, double Dt, double *SD,
                            double *D, double *RHS, void *stream)
{
    CTX3(c);
    SF_REQUIRE(u && v && ru && rv && ubar_t1 && vbar_t1 && ubar_t2 && vbar_t2 && Hz && Akv && z_r && DU_avg1 &&
                   DV_avg1 && tmpU && tmpV && SD && D && RHS,
               "sf_step3d_uv: null field");
    double cff;
    if (iic == ntfirst) cff = 1.0;
    else if (iic == ntfirst + 1) cff = 3.0 / 2.0;
    else cff = 23.0 / 12.0;
    UVArgs<double> a{u, v, ru, rv, ubar_t1, vbar_t1, ubar_t2, vbar_t2, Hz, Akv, z_r, DU_avg1, DV_avg1, tmpU, tmpV,
                     SD, D, RHS};
    return step3d_uv_launch(c, a, cff, lambda, Dt, stream);
}

extern "C" int sf_step3d_uv_f32(sf_ctx *c, float *u, float *v, const float *ru, const float *rv, float *ubar_t1,
                                float *vbar_t1, float *ubar_t2, float *vbar_t2, const float *Hz, const float *Akv,
                                const float *z_r, const double *DU_avg1, const double *DV_avg1, const float *tmpU,
                                const float *tmpV, int32_t iic, int32_t ntfirst, double lambda, double Dt, float *SD,
                                float *D, float *RHS, void *stream)
{
    CTX3(c);
    SF_REQUIRE(u && v && ru && rv && ubar_t1 && vbar_t1 && ubar_t2 && vbar_t2 && Hz && Akv && z_r && DU_avg1 &&
                   DV_avg1 && tmpU && tmpV && SD && D && RHS,
               "sf_step3d_uv_f32: null field");
    double cff;
    if (iic == ntfirst) cff = 1.0;
    else if (iic == ntfirst + 1) cff = 3.0 / 2.0;
    else cff = 23.0 / 12.0;
    UVArgs<float> a{u, v, ru, rv, ubar_t1, vbar_t1, ubar_t2, vbar_t2, Hz, Akv, z_r, DU_avg1, DV_avg1, tmpU, tmpV,
                    SD, D, RHS};
    return step3d_uv_launch(c, a, cff, lambda, Dt, stream);
}

namespace {

template <class T>
int couple_rufrc_launch(sf_ctx *c, const void *ru, const void *rv, const void *u, const void *v, const double *sustr,
                        const double *svstr, double rdrg, double *rufrc, double *rvfrc, void *stream)
{
    LAUNCH_COLUMNS(k_couple_rufrc<T>, c, stream, grid3(c), (const T *)ru, (const T *)rv, (const T *)u, (const T *)v,
                   sustr, svstr, rdrg, rufrc, rvfrc);
}

template <class T>
int step3d_uv_coupled_launch(sf_ctx *c, const sf_uv_coupled *p, double lambda, double Dt, void *stream)
{
    UVC<T> a{(T *)p->u, (T *)p->v, (const T *)p->ru, (const T *)p->rv, (const T *)p->ru1, (const T *)p->rv1,
             (const T *)p->ru2, (const T *)p->rv2, (T *)p->ru_keep, (T *)p->rv_keep, (T *)p->ubar_t1,
             (T *)p->vbar_t1, (T *)p->ubar_t2, (T *)p->vbar_t2, (const T *)p->Hz_old, (const T *)p->Hz,
             (const T *)p->Akv, (const T *)p->z_r, p->DU_avg1, p->DV_avg1, p->DU_avg2, p->DV_avg2, (T *)p->Huon,
             (T *)p->Hvom, p->sustr, p->svstr, p->rdrg, p->ab, (T *)p->SD, (T *)p->D, (T *)p->RHS};
    // register columns address every level with 32-bit byte offsets
    const bool fits = (unsigned long long)(c->szK + 1) * (unsigned long long)c->S2 * sizeof(T) < (1ull << 32);
    if (fits && c->szK <= 16) {
        LAUNCH_COLUMNS((k_step3d_uv_coupled_reg<T, 16>), c, stream, grid3(c), a, lambda, Dt);
    }
    if (fits && c->szK <= 32) {
        LAUNCH_COLUMNS((k_step3d_uv_coupled_reg<T, 32>), c, stream, grid3(c), a, lambda, Dt);
    }
    LAUNCH_COLUMNS(k_step3d_uv_coupled<T>, c, stream, grid3(c), a, lambda, Dt);
}

}  // namespace

extern "C" int sf_couple_rufrc(sf_ctx *c, int32_t precision, const void *ru, const void *rv, const void *u,
                               const void *v, const double *sustr, const double *svstr, double rdrg, double *rufrc,
                               double *rvfrc, void *stream)
{
    CTX3(c);
    SF_REQUIRE(ru && rv && u && v && sustr && svstr && rufrc && rvfrc, "sf_couple_rufrc: null field");
    if (precision == SF_FP64)
        return couple_rufrc_launch<double>(c, ru, rv, u, v, sustr, svstr, rdrg, rufrc, rvfrc, stream);
    if (precision == SF_FP32)
        return couple_rufrc_launch<float>(c, ru, rv, u, v, sustr, svstr, rdrg, rufrc, rvfrc, stream);
    sf::set_error("sf_couple_rufrc: unknown precision %d", precision);
    return SF_ERR_ARG;
}

extern "C" int sf_step3d_uv_coupled(sf_ctx *c, int32_t precision, const sf_uv_coupled *p, double lambda,
                                    double Dt, void *stream)
{
    CTX3(c);
    SF_REQUIRE(p != nullptr, "sf_step3d_uv_coupled: null fields");
    SF_REQUIRE(p->u && p->v && p->ru && p->rv && p->ru_keep && p->rv_keep && p->ubar_t1 && p->vbar_t1 &&
                   p->ubar_t2 && p->vbar_t2 && p->Hz_old && p->Hz && p->Akv && p->z_r && p->DU_avg1 && p->DV_avg1 &&
                   p->DU_avg2 && p->DV_avg2 && p->Huon && p->Hvom && p->sustr && p->svstr && p->SD && p->D &&
                   p->RHS,
               "sf_step3d_uv_coupled: null field");
    SF_REQUIRE(p->ab >= 0 && p->ab <= 2, "sf_step3d_uv_coupled: ab must be 0, 1 or 2 (got %d)", p->ab);
    SF_REQUIRE(p->ab < 1 || (p->ru1 && p->rv1), "sf_step3d_uv_coupled: ab %d needs ru1, rv1", p->ab);
    SF_REQUIRE(p->ab < 2 || (p->ru2 && p->rv2), "sf_step3d_uv_coupled: ab 2 needs ru2, rv2");
    SF_REQUIRE(p->ru_keep != p->ru1 && p->ru_keep != p->ru2 && p->rv_keep != p->rv1 && p->rv_keep != p->rv2,
               "sf_step3d_uv_coupled: ru_keep / rv_keep alias the history");
    if (precision == SF_FP64) return step3d_uv_coupled_launch<double>(c, p, lambda, Dt, stream);
    if (precision == SF_FP32) return step3d_uv_coupled_launch<float>(c, p, lambda, Dt, stream);
    sf::set_error("sf_step3d_uv_coupled: unknown precision %d", precision);
    return SF_ERR_ARG;
}
