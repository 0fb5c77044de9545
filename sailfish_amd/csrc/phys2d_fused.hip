// phys2d_fused.hip -- C ABI of the fused physics substep and of ensembles, and the
// FP64 instances of their kernels (contraction off: bitwise parity with the CPU
// restatement).  The kernels are in phys2d_fused_impl.h; the FP32 instances are
// compiled in phys2d_fused_f32.hip.
#include "phys2d_fused_impl.h"


extern "C" int sf_phys_fused_default_tile(int32_t precision, int32_t szI, int32_t szJ)
{
    (void)precision;
    // 64 x 8 tiles, unless that leaves fewer than 4 workgroups per CU (256 CUs): small
    // grids run faster on 64 x 4 tiles (twice the workgroups to hide latency)
    const long long tiles8 = (long long)((szI + 63) / 64) * ((szJ + 7) / 8);
    return tiles8 < 1024 ? 4 : 8;
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
    const int tj = tile_j > 0 ? tile_j : sf_phys_fused_default_tile(precision, x->szI, x->szJ);
    if (precision == SF_FP64) return dispatch_tj<double>(x, dt, first, cff1, tj, stream);
    if (precision == SF_FP32) return sf_phys_fused_dispatch_f32(x, dt, first, cff1, tj, stream);
    sf::set_error("sf_phys_substep_fused: unknown precision %d", precision);
    return SF_ERR_ARG;
}

extern "C" int sf_phys_substep_coupled_fused(const sf_phys_fields *x, const sf_couple_fields *cp, int32_t precision,
                                             double dt, int32_t first, double cff1, int32_t tile_j, void *stream)
{
    SF_REQUIRE(x && cp, "sf_phys_substep_coupled_fused: null argument");
    SF_REQUIRE(x->szI >= 4 && x->szJ >= 4, "sf_phys_substep_coupled_fused: grid %dx%d too small", x->szI, x->szJ);
    SF_REQUIRE(x->h && x->pm && x->pn && x->fomn && x->sustr && x->svstr,
               "sf_phys_substep_coupled_fused: null grid field");
    for (int k = 0; k < 3; k++)
        SF_REQUIRE(x->zeta[k] && x->ubar[k] && x->vbar[k], "sf_phys_substep_coupled_fused: null state slot");
    SF_REQUIRE(x->rzeta[0] && x->rzeta[1] && x->rubar[0] && x->rubar[1] && x->rvbar[0] && x->rvbar[1],
               "sf_phys_substep_coupled_fused: null rhs slot");
    SF_REQUIRE(x->Zt_avg1 && x->DU_avg1 && x->DV_avg1, "sf_phys_substep_coupled_fused: null average");
    SF_REQUIRE(cp->Fu && cp->Fv && cp->DU_avg2 && cp->DV_avg2, "sf_phys_substep_coupled_fused: null coupling field");
    SF_REQUIRE(!first || (cp->rufrc && cp->rvfrc && cp->Fu1 && cp->Fv1 && cp->Fu2 && cp->Fv2),
               "sf_phys_substep_coupled_fused: the first substep needs rufrc/rvfrc and the forcing history");
    SF_REQUIRE(cp->ab >= 0 && cp->ab <= 2, "sf_phys_substep_coupled_fused: ab=%d", cp->ab);
    const int tj = tile_j > 0 ? tile_j : sf_phys_fused_default_tile(precision, x->szI, x->szJ);
    SF_REQUIRE(tj == 4 || tj == 8, "sf_phys_substep_coupled_fused: tile height %d (4 or 8)", tj);
    if (precision == SF_FP64) return dispatch_tj<double>(x, dt, first, cff1, tj, stream, cp);
    if (precision == SF_FP32) return sf_phys_fused_dispatch_f32(x, dt, first, cff1, tj, stream, cp);
    sf::set_error("sf_phys_substep_coupled_fused: unknown precision %d", precision);
    return SF_ERR_ARG;
}

extern "C" int sf_phys_substep_closure_fused(const sf_phys_fields *x, const sf_mlp *mlp, const void *f,
                                             int32_t precision, double dt, int32_t first, double cff1,
                                             int32_t tile_j, void *stream)
{
    SF_REQUIRE(x && mlp && f, "sf_phys_substep_closure_fused: null argument");
    SF_REQUIRE(x->szI >= 4 && x->szJ >= 4, "sf_phys_substep_closure_fused: grid %dx%d too small", x->szI, x->szJ);
    SF_REQUIRE(x->h && x->pm && x->pn && x->fomn && x->sustr && x->svstr,
               "sf_phys_substep_closure_fused: null grid field");
    for (int k = 0; k < 3; k++)
        SF_REQUIRE(x->zeta[k] && x->ubar[k] && x->vbar[k], "sf_phys_substep_closure_fused: null state slot");
    SF_REQUIRE(x->rzeta[0] && x->rzeta[1] && x->rubar[0] && x->rubar[1] && x->rvbar[0] && x->rvbar[1],
               "sf_phys_substep_closure_fused: null rhs slot");
    SF_REQUIRE(x->Zt_avg1 && x->DU_avg1 && x->DV_avg1, "sf_phys_substep_closure_fused: null average");
    // default 64 x 4: those closure instances keep two waves per SIMD (the 64 x 8 ones need
    // more registers than that allows, on any grid size)
    const int tj = tile_j > 0 ? tile_j : 4;
    SF_REQUIRE(tj == 4 || tj == 8, "sf_phys_substep_closure_fused: tile height %d (4 or 8)", tj);
    const ClosureIn cl{mlp, f};
    if (precision == SF_FP64) return dispatch_tj<double>(x, dt, first, cff1, tj, stream, nullptr, &cl);
    if (precision == SF_FP32) return sf_phys_fused_dispatch_f32(x, dt, first, cff1, tj, stream, nullptr, &cl);
    sf::set_error("sf_phys_substep_closure_fused: unknown precision %d", precision);
    return SF_ERR_ARG;
}

namespace {

int ens_rebuild(sf_ens *e)
{
    return e->precision == SF_FP64 ? ens_build<double>(e) : ens_build<float>(e);
}

void ens_drop_graph(sf_ens *e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    e->graph_nsub = 0;
}

}  // namespace

extern "C" int sf_ens_create(const sf_phys_fields *members, const double *dt, int32_t n, int32_t precision,
                             sf_ens **out)
{
    SF_REQUIRE(out && members && dt && n > 0, "sf_ens_create: bad arguments");
    SF_REQUIRE(precision == SF_FP64 || precision == SF_FP32, "sf_ens_create: unknown precision %d", precision);
    *out = nullptr;
    for (int k = 0; k < n; k++) {
        const sf_phys_fields &x = members[k];
        SF_REQUIRE(x.szI >= 4 && x.szJ >= 4, "sf_ens_create: member %d grid %dx%d too small", k, x.szI, x.szJ);
        SF_REQUIRE((long long)x.szI * x.szJ < (1LL << 29), "sf_ens_create: member %d exceeds 2^29 points", k);
        SF_REQUIRE(x.h && x.pm && x.pn && x.fomn && x.sustr && x.svstr, "sf_ens_create: member %d null grid field", k);
        for (int s = 0; s < 3; s++)
            SF_REQUIRE(x.zeta[s] && x.ubar[s] && x.vbar[s] && x.rzeta[s] && x.rubar[s] && x.rvbar[s],
                       "sf_ens_create: member %d null time slot %d", k, s);
        SF_REQUIRE(x.Zt_avg1 && x.DU_avg1 && x.DV_avg1, "sf_ens_create: member %d null average", k);
    }
    sf_ens *e = new sf_ens();
    e->n = n;
    e->precision = precision;
    (void)hipGetDevice(&e->device);
    e->tile_j = 8;
    e->order = 1;
    e->grid = 0;
    e->rot = 0;
    e->members = nullptr;
    e->list = nullptr;
    e->graph = nullptr;
    e->exec = nullptr;
    e->graph_nsub = 0;
    e->host.assign(members, members + n);
    e->dt.assign(dt, dt + n);
    const int rc = ens_rebuild(e);
    if (rc != SF_OK) {
        (void)hipFree(e->members);
        (void)hipFree(e->list);
        delete e;
        return rc;
    }
    *out = e;
    return SF_OK;
}

extern "C" int sf_ens_destroy(sf_ens *e)
{
    if (!e) return SF_OK;
    ens_drop_graph(e);
    (void)hipFree(e->members);
    (void)hipFree(e->list);
    delete e;
    return SF_OK;
}

extern "C" int sf_ens_configure(sf_ens *e, int32_t tile_j, int32_t order, int32_t grid)
{
    SF_REQUIRE(e, "sf_ens_configure: null ensemble");
    SF_REQUIRE(tile_j == 4 || tile_j == 8 || tile_j == 16, "sf_ens_configure: tile height %d (4, 8 or 16)", tile_j);
    SF_REQUIRE(order == 0 || order == 1, "sf_ens_configure: order %d (0 strided, 1 chunked)", order);
    SF_REQUIRE(grid >= 0, "sf_ens_configure: grid %d", grid);
    ens_drop_graph(e);
    e->order = order;
    e->grid = grid;
    if (tile_j != e->tile_j) {
        e->tile_j = tile_j;
        return ens_rebuild(e);
    }
    return SF_OK;
}

extern "C" int sf_ens_num_tiles(const sf_ens *e) { return e ? e->ntiles : -1; }

extern "C" int sf_ens_rotation(const sf_ens *e) { return e ? e->rot : -1; }

extern "C" int sf_ens_substep(sf_ens *e, int32_t first, double cff1, void *stream)
{
    SF_REQUIRE(e, "sf_ens_substep: null ensemble");
    e->rot = (e->rot + 1) % 3;   // cycleTimes2D of every member
    if (e->precision == SF_FP64) ens_launch<double>(e, first, cff1, e->rot, sf::as_stream(stream));
    else sf_ens_launch_f32(e, first, cff1, e->rot, sf::as_stream(stream));
    return sf::check_launch("sf_ens_substep");
}

extern "C" int sf_ens_capture(sf_ens *e, int32_t nsub, const double *cff1, void *stream)
{
    SF_REQUIRE(e && cff1, "sf_ens_capture: null argument");
    SF_REQUIRE(nsub > 0 && nsub % 3 == 0, "sf_ens_capture: nsub %d must be a positive multiple of 3", nsub);
    ens_drop_graph(e);
    hipStream_t cs = nullptr;
    hipError_t r = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (r != hipSuccess) return sf::hip_fail(r, "sf_ens_capture: stream");
    r = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (r == hipSuccess) {
        int rot = e->rot;
        for (int k = 0; k < nsub; k++) {
            rot = (rot + 1) % 3;
            if (e->precision == SF_FP64) ens_launch<double>(e, 0, cff1[k], rot, cs);
            else sf_ens_launch_f32(e, 0, cff1[k], rot, cs);
        }
        r = hipStreamEndCapture(cs, &e->graph);
    }
    if (r == hipSuccess) r = hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (r != hipSuccess) {
        ens_drop_graph(e);
        return sf::hip_fail(r, "sf_ens_capture");
    }
    e->graph_nsub = nsub;
    (void)stream;
    return SF_OK;
}

extern "C" int sf_ens_replay(sf_ens *e, void *stream)
{
    SF_REQUIRE(e && e->exec, "sf_ens_replay: no captured graph");
    hipError_t r = hipGraphLaunch(e->exec, sf::as_stream(stream));
    if (r != hipSuccess) return sf::hip_fail(r, "sf_ens_replay");
    // nsub is a multiple of 3: the rotation is unchanged
    return SF_OK;
}
