// phys2d_fused_f32.hip -- FP32 instances of the fused physics substep and of the
// ensemble launch, compiled with floating-point contraction on: a multiply-add
// written within one expression becomes a fused multiply-add (one rounding).  The
// decision is made per source expression, so the single-basin and the ensemble
// kernels (the same tile code) round identically.  The FP64 instances in
// phys2d_fused.hip keep contraction off.
#include "sf_internal.h"

#pragma clang fp contract(on)

#include "phys2d_fused_impl.h"

int sf_phys_fused_dispatch_f32(const sf_phys_fields *x, double dt, int first, double w1, int tj, void *stream,
                               const sf_couple_fields *cp, const ClosureIn *cl)
{
    return dispatch_tj<float>(x, dt, first, w1, tj, stream, cp, cl);
}

void sf_ens_launch_f32(sf_ens *e, int first, double w1, int rot, hipStream_t s)
{
    ens_launch<float>(e, first, w1, rot, s);
}
