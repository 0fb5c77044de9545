// volcons.hip -- open-boundary volume conservation (nonlinear/obc_volcons.py:1-72, the
// ROMS obc_volcons / conserve_mass pair): the cross-section area and the net volume
// flux through the open boundary segments, summed in FP64 whatever the state
// precision, give the uniform correction ubar_xs = flux / area, which is then
// subtracted from the inflow-signed boundary velocities so that the net flux is zero.
//
// Flat C-grid: u point i lies between rho i-1 and i, v point j between rho j-1 and j.
// Segments (flags SF_OBC_WEST.. bits): west u(1, j), east u(L, j), j = 1..M-1;
// south v(i, 1), north v(i, M), i = 1..L-1.  Area term of a west point:
//   cff = 0.5 * ((zeta(0,j) + h(0,j)) + (zeta(1,j) + h(1,j))) * on_u(1,j),
// flux += cff * u(1,j); east/north enter with the opposite sign.  The sums are a
// fixed-shape block tree followed by an ordered sum of the block partials, so the
// result is reproducible run to run.
#include "sf_internal.h"

#include <mutex>

// Work space of sf_obc_volcons_noalloc: the block partials of the two FP64 sums, in a
// file-scope device array instead of caller or library allocations.  Each device holds
// its own copy until the process ends, shared by every call on that device.
static constexpr int kVolScratchDoubles = 4096;
static __device__ __align__(256) double scratch[kVolScratchDoubles];

namespace {

constexpr int VB = 256;

template <typename T>
struct VolArgs {
    const T *z, *h, *pm, *pn;
    T *u, *v;
    int szI, szJ, flags;
    int nw, ne, ns, nn;   // points per segment (0 when closed)
};

// boundary point b -> (flat index of the velocity point, area term, signed velocity)
template <typename T>
__device__ __forceinline__ void vol_point(const VolArgs<T> &a, int b, double &cff, double &fl)
{
    const int W = a.szI, L = a.szI - 1, M = a.szJ - 1;
    int seg = 0;
    if (b >= a.nw) {
        b -= a.nw;
        seg = 1;
        if (b >= a.ne) {
            b -= a.ne;
            seg = 2;
            if (b >= a.ns) {
                b -= a.ns;
                seg = 3;
            }
        }
    }
    if (seg <= 1) {   // west / east u points, rows j = 1 .. M-1
        const int j = 1 + b, i = seg == 0 ? 1 : L;
        const long long f = (long long)j * W + i;
        const double onu = 2.0 / ((double)a.pn[f - 1] + (double)a.pn[f]);
        cff = (0.5 * (((double)a.z[f - 1] + (double)a.h[f - 1]) + ((double)a.z[f] + (double)a.h[f]))) * onu;
        fl = seg == 0 ? cff * (double)a.u[f] : -(cff * (double)a.u[f]);
    } else {          // south / north v points, columns i = 1 .. L-1
        const int i = 1 + b, j = seg == 2 ? 1 : M;
        const long long f = (long long)j * W + i;
        const double omv = 2.0 / ((double)a.pm[f - W] + (double)a.pm[f]);
        cff = (0.5 * (((double)a.z[f - W] + (double)a.h[f - W]) + ((double)a.z[f] + (double)a.h[f]))) * omv;
        fl = seg == 2 ? cff * (double)a.v[f] : -(cff * (double)a.v[f]);
    }
}

template <typename T>
__global__ __launch_bounds__(VB) void k_volcons_partial(VolArgs<T> a, int n, double *__restrict__ part)
{
    __shared__ double sa[VB], sf[VB];
    const int b = blockIdx.x * VB + threadIdx.x;
    double cff = 0.0, fl = 0.0;
    if (b < n) vol_point(a, b, cff, fl);
    sa[threadIdx.x] = cff;
    sf[threadIdx.x] = fl;
    __syncthreads();
    for (int s = VB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sa[threadIdx.x] += sa[threadIdx.x + s];
            sf[threadIdx.x] += sf[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sa[0];
        part[2 * blockIdx.x + 1] = sf[0];
    }
}

// out[0] = area, out[1] = flux, out[2] = ubar_xs (0 when the area is 0)
__global__ void k_volcons_final(const double *__restrict__ part, int nblk, double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double area = 0.0, flux = 0.0;
    for (int k = 0; k < nblk; k++) {
        area += part[2 * k];
        flux += part[2 * k + 1];
    }
    out[0] = area;
    out[1] = flux;
    out[2] = area > 0.0 ? flux / area : 0.0;
}

// conserve_mass: west/south velocities -= ubar_xs, east/north += ubar_xs
template <typename T>
__global__ __launch_bounds__(VB) void k_volcons_apply(VolArgs<T> a, int n, const double *__restrict__ out)
{
    const int b = blockIdx.x * VB + threadIdx.x;
    if (b >= n) return;
    const double xs = out[2];
    const int W = a.szI, L = a.szI - 1, M = a.szJ - 1;
    int r = b;
    if (r < a.nw) {
        const long long f = (long long)(1 + r) * W + 1;
        a.u[f] = (T)((double)a.u[f] - xs);
        return;
    }
    r -= a.nw;
    if (r < a.ne) {
        const long long f = (long long)(1 + r) * W + L;
        a.u[f] = (T)((double)a.u[f] + xs);
        return;
    }
    r -= a.ne;
    if (r < a.ns) {
        const long long f = (long long)W + 1 + r;
        a.v[f] = (T)((double)a.v[f] - xs);
        return;
    }
    r -= a.ns;
    const long long f = (long long)M * W + 1 + r;
    a.v[f] = (T)((double)a.v[f] + xs);
}

template <typename T>
int run_volcons(const void *zeta, const void *h, const void *pm, const void *pn, void *ubar, void *vbar, int szI,
                int szJ, int flags, int apply, double *work, double *out, void *stream)
{
    VolArgs<T> a;
    a.z = (const T *)zeta;
    a.h = (const T *)h;
    a.pm = (const T *)pm;
    a.pn = (const T *)pn;
    a.u = (T *)ubar;
    a.v = (T *)vbar;
    a.szI = szI;
    a.szJ = szJ;
    a.flags = flags;
    a.nw = (flags & SF_OBC_WEST) ? szJ - 2 : 0;
    a.ne = (flags & SF_OBC_EAST) ? szJ - 2 : 0;
    a.ns = (flags & SF_OBC_SOUTH) ? szI - 2 : 0;
    a.nn = (flags & SF_OBC_NORTH) ? szI - 2 : 0;
    const int n = a.nw + a.ne + a.ns + a.nn;
    const int nblk = n > 0 ? (n + VB - 1) / VB : 1;
    hipStream_t s = sf::as_stream(stream);
    hipLaunchKernelGGL(k_volcons_partial<T>, dim3(nblk), dim3(VB), 0, s, a, n, work);
    hipLaunchKernelGGL(k_volcons_final, dim3(1), dim3(1), 0, s, work, nblk, out);
    if (apply && n > 0) hipLaunchKernelGGL(k_volcons_apply<T>, dim3(nblk), dim3(VB), 0, s, a, n, out);
    return sf::check_launch("sf_obc_volcons");
}

// Device address of `scratch` per device.  The array's name in host code is only a
// stand-in for the symbol; its address on a device comes from hipGetSymbolAddress with
// that device current, once per device.
std::mutex g_scratch_mu;
double *g_scratch_ptr[64];

int scratch_for_current_device(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return sf::hip_fail(e, "sf_obc_volcons_noalloc: hipGetDevice");
    SF_REQUIRE(dev >= 0 && dev < 64, "sf_obc_volcons_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return sf::hip_fail(e, "sf_obc_volcons_noalloc: hipGetSymbolAddress");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return SF_OK;
}

}  // namespace

extern "C" int sf_obc_volcons_work_size(int32_t szI, int32_t szJ)
{
    const int n = 2 * (szJ - 2) + 2 * (szI - 2);
    return 2 * ((n + VB - 1) / VB + 1);
}

extern "C" int sf_obc_volcons(const void *zeta, const void *h, const void *pm, const void *pn, void *ubar, void *vbar,
                              int32_t szI, int32_t szJ, int32_t precision, int32_t flags, int32_t apply,
                              double *work, double *out, void *stream)
{
    SF_REQUIRE(zeta && h && pm && pn && ubar && vbar && work && out, "sf_obc_volcons: null argument");
    SF_REQUIRE(szI >= 4 && szJ >= 4, "sf_obc_volcons: grid %dx%d too small", szI, szJ);
    SF_REQUIRE((flags & ~15) == 0, "sf_obc_volcons: unknown segment flags 0x%x", flags);
    if (precision == SF_FP64)
        return run_volcons<double>(zeta, h, pm, pn, ubar, vbar, szI, szJ, flags, apply, work, out, stream);
    if (precision == SF_FP32)
        return run_volcons<float>(zeta, h, pm, pn, ubar, vbar, szI, szJ, flags, apply, work, out, stream);
    sf::set_error("sf_obc_volcons: unknown precision %d", precision);
    return SF_ERR_ARG;
}

extern "C" int sf_obc_volcons_noalloc(const void *zeta, const void *h, const void *pm, const void *pn, void *ubar,
                                      void *vbar, int32_t szI, int32_t szJ, int32_t precision, int32_t flags,
                                      int32_t apply, double *out, void *stream)
{
    SF_REQUIRE(zeta && h && pm && pn && ubar && vbar && out, "sf_obc_volcons_noalloc: null argument");
    SF_REQUIRE(szI >= 4 && szJ >= 4, "sf_obc_volcons_noalloc: grid %dx%d too small", szI, szJ);
    const long long n = 2LL * (szJ - 2) + 2LL * (szI - 2);
    const long long need = 2 * ((n + VB - 1) / VB + 1);
    SF_REQUIRE(need <= kVolScratchDoubles,
               "sf_obc_volcons_noalloc: grid %dx%d needs %lld doubles of work space, the static maximum is %d "
               "(at most %d boundary points); use sf_obc_volcons with a caller work buffer",
               szI, szJ, need, kVolScratchDoubles, (kVolScratchDoubles / 2 - 1) * VB);
    double *work = nullptr;
    const int rc = scratch_for_current_device(&work);
    if (rc != SF_OK) return rc;
    // every partial slot of the launch is written before it is read: nothing to clear
    return sf_obc_volcons(zeta, h, pm, pn, ubar, vbar, szI, szJ, precision, flags, apply, work, out, stream);
}
