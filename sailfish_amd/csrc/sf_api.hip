// sf_api.hip -- context management and error reporting of the C ABI.
//
// sf_ctx replaces the per-module device globals that the reference sets once with
// `initialize` (modules/mod_cppkernels.h:40-41, :154-171) and `initModule`
// (modules/mod_operators.py:31-45): sizes and metric pointers are carried by the
// context, so kernels of different grids can run side by side.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <new>

#include "sf_internal.h"

namespace {
thread_local char g_err[512] = "";
}

namespace sf {

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what)
{
    set_error("%s: %s", what, hipGetErrorString(e));
    return SF_ERR_HIP;
}

}  // namespace sf

extern "C" const char *sf_last_error_string(void) { return g_err; }

extern "C" int sf_version(void) { return 100; }

extern "C" int sf_ctx_create(const sf_grid_desc *d, sf_ctx **out)
{
    SF_REQUIRE(d && out, "sf_ctx_create: null argument");
    SF_REQUIRE(d->szI >= 3 && d->szJ >= 3, "sf_ctx_create: grid %dx%d too small", d->szI, d->szJ);
    SF_REQUIRE(d->szK >= 0, "sf_ctx_create: szK=%d < 0", d->szK);
    SF_REQUIRE((long long)d->szI * d->szJ < (1LL << 31), "sf_ctx_create: grid too large for int32 indexing");
    SF_REQUIRE(d->on_u && d->om_v && d->pn && d->pm, "sf_ctx_create: null metric pointer");
    sf_ctx *c = new (std::nothrow) sf_ctx();
    if (!c) {
        sf::set_error("sf_ctx_create: out of host memory");
        return SF_ERR_NOMEM;
    }
    c->szI = d->szI;
    c->szJ = d->szJ;
    c->szK = d->szK;
    c->S2 = (long long)d->szI * d->szJ;
    c->on_u = d->on_u;
    c->om_v = d->om_v;
    c->pn = d->pn;
    c->pm = d->pm;
    c->q5_stride = (d->flags & SF_FLAG_FIX_Q5_MASK) ? d->szI : d->szJ;
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) {
        delete c;
        return sf::hip_fail(e, "sf_ctx_create: hipGetDevice");
    }
    *out = c;
    return SF_OK;
}

extern "C" int sf_ctx_destroy(sf_ctx *c)
{
    if (!c) return SF_OK;
    hipFree(c->bc_zeta);
    hipFree(c->bc_ubar);
    hipFree(c->bc_vbar);
    delete c;
    return SF_OK;
}
