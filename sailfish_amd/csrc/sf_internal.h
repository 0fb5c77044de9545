// sf_internal.h -- shared definitions of the gfx950 kernels and the C ABI wrappers.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/sailfish_hip.h"

// Floating-point contraction is off for the whole library (also passed as
// -ffp-contract=off): FP64 results must match the CPU restatement bit for bit.
#pragma clang fp contract(off)

struct sf_ctx {
    int szI, szJ, szK;
    long long S2;
    const double *on_u, *om_v, *pn, *pm;
    int device;
    int q5_stride;            // mask stride of computeMomentumPred (szJ, or szI when fixed)
    // dense boundary-code fields for the fused substep (one int32 per rho point),
    // allocated by sf_ctx_set_bc_maps; see step2d_fused.hip
    int32_t *bc_zeta, *bc_ubar, *bc_vbar;
    int bc_ready;
    int bc_vel_copy;          // a velocity map copies from a neighbour (gradient boundary)
};

namespace sf {

void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

inline int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return SF_OK;
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// 1 / x: IEEE division in FP64 (bitwise parity with the CPU restatement); the hardware
// reciprocal v_rcp_f32 (1 ulp) in FP32, one instruction instead of the ~10 of an IEEE
// FP32 division.  2 * recip(x) equals 2 / x exactly in FP64 (scaling by 2 is exact).
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }

// Adams-Bashforth extrapolation over baroclinic steps of the barotropic forcing of the
// coupled substep: F (ab 0), 3/2 F - 1/2 F1 (ab 1), (23/12 F - 16/12 F1) + 5/12 F2 (ab 2)
__device__ __forceinline__ double ab_extrapolate(int ab, double F, double F1, double F2)
{
    if (ab == 0) return F;
    if (ab == 1) return 1.5 * F - 0.5 * F1;
    return ((23.0 / 12.0) * F - (16.0 / 12.0) * F1) + (5.0 / 12.0) * F2;
}

constexpr int kBlock1D = 256;

// Bytes one substep must move before the fused 2D substeps write their outputs with
// streaming (non-temporal) stores: beyond the 256 MB last-level cache the next substep
// cannot re-read them from it, and streaming keeps L2 for the halo reads.  Measured on
// 2048 x 1024: plain stores faster at 264 MiB (FP32 physics), streaming faster at 352 MiB
// (FP64 reference discretisation) and 432 MiB (FP64 physics).
constexpr double kStreamingBytes = 320.0 * (1 << 20);

// compute units of a device, queried once per device (256 on MI355X)
inline int cu_count(int device)
{
    static int cache[64];
    if (device < 0 || device >= 64) return 256;
    if (cache[device] == 0) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || v <= 0) v = 256;
        cache[device] = v;
    }
    return cache[device];
}

inline unsigned grid1d(long long n, int block = kBlock1D)
{
    return (unsigned)((n + block - 1) / block);
}

}  // namespace sf

#define SF_REQUIRE(cond, ...)                 \
    do {                                      \
        if (!(cond)) {                        \
            sf::set_error(__VA_ARGS__);       \
            return SF_ERR_ARG;                \
        }                                     \
    } while (0)
