/*
 * sf_oracle3d.c -- CPU restatement (FP64) of the 3D kernels of the Sailfish ROMS port
 * (nonlinear/rhs_kernels.cpp, nonlinear/step3D_uv.kernels.cpp), written from scratch.
 *
 * TEST INFRASTRUCTURE ONLY (see sf_oracle2d.c).
 *
 * Pins (tests/test_oracle3d.py; tests/test_oracle_mutations.py plants mistakes here):
 *   sfo_addCoriolis          test_coriolis_uniform_flow
 *   sfo_horizontalAdvection  test_horizontal_advection_linear_u,
 *                            test_horizontal_advection_xi_curvature_closed_form,
 *                            test_horizontal_advection_eta_curvature_upwind_closed_form,
 *                            test_horizontal_advection_tail_zeroing
 *   sfo_verticalAdvection    test_vertical_advection_linear_profile,
 *                            test_vertical_advection_special_levels_closed_form
 *   sfo_set_maxflux          test_set_maxflux_closed_form
 *   sfo_omega                test_omega_divergence_free_is_zero_and_linear_flux
 *   sfo_set_depth            test_set_depth_rest_state_closed_form,
 *                            test_set_depth_vtransform1_and_free_surface
 *   sfo_set_zeta             test_set_zeta_closed_form
 *   sfo_setLateralUVBCs      test_lateral_uv_bcs_clamp_value
 *   sfo_step3d_UV            test_step3d_uv_tridiagonal_against_dense_solve (AB3 factors)
 *
 * 3D fields are flat arrays with level stride S2 = szI*szJ; a stencil offset
 * (dk, dj, di) from column f at level K reads index (K+dk)*S2 + f + dj*szI + di with
 * no wrapping rules beyond the flat arithmetic itself (the reference's Eval,
 * mod_cppkernels.h:337-340).  Reads outside [0, n) of an array of n values are
 * undefined in the reference; they return 0.0 here (r3()).
 *
 * Schedules fixed here where the reference races:
 *  - horizontalAdvection: threads i >= S2 of the launch write 0 into ru/rv at flat
 *    i (rhs_kernels.cpp:86-91), i.e. into level 1 of the first npad columns, where
 *    npad = launch threads - S2; those writes are applied after all columns.
 *  - step3d_UV's V pass at U-but-not-V nodes reuses the U pass's eliminated
 *    tridiagonal system (solveTri guards with isUNode, step3D_uv.kernels.cpp:525);
 *    each column owns its scratch, as in the reference.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct sfo_grid3 {
    int szI, szJ, N;
    const double *on_u, *om_v, *pn, *pm;
} sfo_grid3;

typedef struct {
    const double *p;
    ptrdiff_t n;
} arr_t;

static inline double r3(arr_t a, ptrdiff_t idx) { return (idx < 0 || idx >= a.n) ? 0.0 : a.p[idx]; }

#define S2_OF(g) ((ptrdiff_t)(g)->szI * (g)->szJ)

static inline int isRNode(const sfo_grid3 *g, ptrdiff_t f)
{
    const int i = (int)(f % g->szI), j = (int)(f / g->szI);
    return i > 0 && j > 0;
}
static inline int isUNode(const sfo_grid3 *g, ptrdiff_t f)
{
    const int i = (int)(f % g->szI), j = (int)(f / g->szI);
    return i > 1 && j > 0;
}
static inline int isVNode(const sfo_grid3 *g, ptrdiff_t f)
{
    const int i = (int)(f % g->szI), j = (int)(f / g->szI);
    return i > 0 && j > 1 && j < g->szJ - 1;
}
static inline int isInnerCell(const sfo_grid3 *g, ptrdiff_t f)
{
    const int i = (int)(f % g->szI), j = (int)(f / g->szI);
    return i > 1 && i < g->szI - 2 && j > 1 && j < g->szJ - 2;
}

/* ---------------------------------------------------------------------------
 * addCoriolis (rhs_kernels.cpp:6-35), R nodes, K = 0..N-1:
 *   UF = (Hz*fomn)*VtoR(v); VF = (Hz*fomn)*UtoR(u); ru += RtoU(UF); rv -= RtoV(VF)
 * ------------------------------------------------------------------------- */
static double cor_UF(const sfo_grid3 *g, arr_t fomn, arr_t v, arr_t Hz, ptrdiff_t K, ptrdiff_t f)
{
    const ptrdiff_t S2 = S2_OF(g), b = K * S2 + f;
    return (r3(Hz, b) * r3(fomn, f)) * ((r3(v, b) + r3(v, b + g->szI)) * 0.5);
}
static double cor_VF(const sfo_grid3 *g, arr_t fomn, arr_t u, arr_t Hz, ptrdiff_t K, ptrdiff_t f)
{
    const ptrdiff_t S2 = S2_OF(g), b = K * S2 + f;
    return (r3(Hz, b) * r3(fomn, f)) * ((r3(u, b) + r3(u, b + 1)) * 0.5);
}

void sfo_addCoriolis(const sfo_grid3 *g, const double *fomn, const double *u, const double *v,
                     double *ru, double *rv, const double *Hz)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N, n3 = (N + 1) * S2;
    arr_t F = {fomn, S2}, U = {u, n3}, V = {v, n3}, H = {Hz, N * S2};
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        if (!isRNode(g, f)) continue;
        for (ptrdiff_t K = 0; K < N; K++) {
            const ptrdiff_t b = K * S2 + f;
            const double rtou = (cor_UF(g, F, V, H, K, f) + cor_UF(g, F, V, H, K, f - 1)) * 0.5;
            const double rtov = (cor_VF(g, F, U, H, K, f) + cor_VF(g, F, U, H, K, f - g->szI)) * 0.5;
            ru[b] = ru[b] + rtou;
            rv[b] = rv[b] - rtov;
        }
    }
}

/* ---------------------------------------------------------------------------
 * horizontalAdvection (rhs_kernels.cpp:70-132), every column, K = 0..N-1.
 * Offsets below are flat (dj*szI + di) at level K.
 * ------------------------------------------------------------------------- */
typedef struct {
    arr_t u, v, Hu, Hv;
    ptrdiff_t szI;
} hadv_t;

#define AT(a, o) r3((a), base + (o))
static double dxx(arr_t a, ptrdiff_t base, ptrdiff_t o) { return AT(a, o - 1) - 2.0 * AT(a, o) + AT(a, o + 1); }
static double dee(arr_t a, ptrdiff_t base, ptrdiff_t o, ptrdiff_t W) { return AT(a, o - W) - 2.0 * AT(a, o) + AT(a, o + W); }

/* the eight flux expressions, evaluated at flat offset o from base (level K) */
static double h_uR(const hadv_t *h, ptrdiff_t base, ptrdiff_t o) { return (AT(h->u, o) + AT(h->u, o + 1)) * 0.5; }
static double h_uP(const hadv_t *h, ptrdiff_t base, ptrdiff_t o) { return (AT(h->u, o) + AT(h->u, o - 1)) * 0.5; }
static double h_vR(const hadv_t *h, ptrdiff_t base, ptrdiff_t o) { return (AT(h->v, o) + AT(h->v, o + h->szI)) * 0.5; }
static double h_vP(const hadv_t *h, ptrdiff_t base, ptrdiff_t o) { return (AT(h->v, o) + AT(h->v, o - h->szI)) * 0.5; }
static double h_HuR(const hadv_t *h, ptrdiff_t base, ptrdiff_t o) { return (AT(h->Hu, o) + AT(h->Hu, o + 1)) * 0.5; }
static double h_HuP(const hadv_t *h, ptrdiff_t base, ptrdiff_t o) { return (AT(h->Hu, o) + AT(h->Hu, o - 1)) * 0.5; }
static double h_HvP(const hadv_t *h, ptrdiff_t base, ptrdiff_t o) { return (AT(h->Hv, o) + AT(h->Hv, o - h->szI)) * 0.5; }
static double h_HvR(const hadv_t *h, ptrdiff_t base, ptrdiff_t o) { return (AT(h->Hv, o) + AT(h->Hv, o + h->szI)) * 0.5; }
/* upwind curvature terms (h:608-702) */
static double h_uxxR(const hadv_t *h, ptrdiff_t base, ptrdiff_t o)
{   /* upwindUtoR(DXXUtoU(u), uR): uR(o) + uR(o+1) >= 0 ? DXX(u)(o) : DXX(u)(o+1) */
    return (h_uR(h, base, o) + h_uR(h, base, o + 1) >= 0) ? dxx(h->u, base, o) : dxx(h->u, base, o + 1);
}
static double h_ueeP(const hadv_t *h, ptrdiff_t base, ptrdiff_t o)
{   /* upwindUtoP(DEEUtoU(u), HvomP): vel(o) + vel(o-W) >= 0 ? DEE(u)(o) : DEE(u)(o-W) */
    const ptrdiff_t W = h->szI;
    return (h_HvP(h, base, o) + h_HvP(h, base, o - W) >= 0) ? dee(h->u, base, o, W) : dee(h->u, base, o - W, W);
}
static double h_HuxxR(const hadv_t *h, ptrdiff_t base, ptrdiff_t o)
{   /* UtoR(DXXUtoU(Huon)) */
    return (dxx(h->Hu, base, o) + dxx(h->Hu, base, o + 1)) * 0.5;
}
static double h_HueeP(const hadv_t *h, ptrdiff_t base, ptrdiff_t o)
{   /* UtoP(DEEUtoU(Huon)) */
    return (dee(h->Hu, base, o, h->szI) + dee(h->Hu, base, o - 1, h->szI)) * 0.5;
}
static double h_HvxxP(const hadv_t *h, ptrdiff_t base, ptrdiff_t o)
{   /* VtoP(DXXVtoV(Hvom)) */
    return (dxx(h->Hv, base, o) + dxx(h->Hv, base, o - h->szI)) * 0.5;
}
static double h_HveeR(const hadv_t *h, ptrdiff_t base, ptrdiff_t o)
{   /* VtoR(DEEVtoV(Hvom)) */
    return (dee(h->Hv, base, o, h->szI) + dee(h->Hv, base, o + h->szI, h->szI)) * 0.5;
}
static const double Gadv = -0.5;
/* UFxR = uR + ((Gadv*uxxR)*(HuonR + Gadv*HuxxR)) and the three siblings (:121-125) */
static double UFxR(const hadv_t *h, ptrdiff_t b, ptrdiff_t o)
{
    return h_uR(h, b, o) + (Gadv * h_uxxR(h, b, o)) * (h_HuR(h, b, o) + Gadv * h_HuxxR(h, b, o));
}
static double UFeP(const hadv_t *h, ptrdiff_t b, ptrdiff_t o)
{
    return h_uP(h, b, o) + (Gadv * h_ueeP(h, b, o)) * (h_HvP(h, b, o) + Gadv * h_HvxxP(h, b, o));
}
static double VFxP(const hadv_t *h, ptrdiff_t b, ptrdiff_t o)
{
    return h_vP(h, b, o) + (Gadv * h_ueeP(h, b, o)) * (h_HuP(h, b, o) + Gadv * h_HueeP(h, b, o));
}
static double VFeR(const hadv_t *h, ptrdiff_t b, ptrdiff_t o)
{
    return h_vR(h, b, o) + (Gadv * h_uxxR(h, b, o)) * (h_HvR(h, b, o) + Gadv * h_HveeR(h, b, o));
}
#undef AT

void sfo_horizontalAdvection(const sfo_grid3 *g, const double *u, const double *v, const double *Huon,
                             const double *Hvom, double *ru, double *rv, int32_t launch_threads)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N, n3 = (N + 1) * S2, W = g->szI;
    hadv_t h = {{u, n3}, {v, n3}, {Huon, n3}, {Hvom, n3}, W};
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        for (ptrdiff_t K = 0; K < N; K++) {
            const ptrdiff_t b = K * S2 + f;
            const double du = (UFxR(&h, b, 0) - UFxR(&h, b, -1)) + (UFeP(&h, b, 0) - UFeP(&h, b, -W));
            const double dv = (VFxP(&h, b, 0) - VFxP(&h, b, -1)) + (VFeR(&h, b, 0) - VFeR(&h, b, -W));
            ru[b] = ru[b] - du;
            rv[b] = rv[b] - dv;
        }
    }
    /* threads S2 .. launch_threads-1 store 0 at flat index i (level 1 of the first
     * columns), scheduled after the columns */
    for (ptrdiff_t t = S2; t < launch_threads && t < n3; t++) {
        ru[t] = 0.0;
        rv[t] = 0.0;
    }
}

/* ---------------------------------------------------------------------------
 * verticalAdvection (rhs_kernels.cpp:135-197): inner cells that are not zeta
 * boundary points.  F(K) = UtoUW_4th(u)(K) * WtoUW_4th(W)(K) (h:723, :767).
 * ------------------------------------------------------------------------- */
static double v4(arr_t a, ptrdiff_t S2, ptrdiff_t K, ptrdiff_t f)
{   /* opRtoW_4th: 9/16 (a(K)+a(K+1)) - 1/16 (a(K-1)+a(K+2)) */
    return (9.0 / 16.0) * (r3(a, K * S2 + f) + r3(a, (K + 1) * S2 + f)) -
           (1.0 / 16.0) * (r3(a, (K - 1) * S2 + f) + r3(a, (K + 2) * S2 + f));
}
static double h4(arr_t a, ptrdiff_t S2, ptrdiff_t K, ptrdiff_t f, ptrdiff_t d)
{   /* opRtoU_4th (d=1) / opRtoV_4th (d=szI): 9/16 (a(o)+a(o+d)) - 1/16 (a(o-d)+a(o+2d)) */
    const ptrdiff_t b = K * S2 + f;
    return (9.0 / 16.0) * (r3(a, b) + r3(a, b + d)) - (1.0 / 16.0) * (r3(a, b - d) + r3(a, b + 2 * d));
}

void sfo_verticalAdvection(const sfo_grid3 *g, const double *u, const double *v, const double *Wv,
                           double *ru, double *rv, const int32_t *BC)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N, n3 = (N + 1) * S2, W = g->szI;
    arr_t U = {u, n3}, V = {v, n3}, Wa = {Wv, n3};
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        if (BC[f] >= 0 || !isInnerCell(g, f)) continue;
#define FU(k) (v4(U, S2, (k), f) * h4(Wa, S2, (k), f, 1))
#define FV(k) (v4(V, S2, (k), f) * h4(Wa, S2, (k), f, W))
        for (ptrdiff_t K = 2; K < N - 2; K++) {
            ru[K * S2 + f] = ru[K * S2 + f] - (FU(K + 1) - FU(K));
            rv[K * S2 + f] = rv[K * S2 + f] - (FV(K + 1) - FV(K));
        }
        /* K = 1 block (:169-176) */
        rv[1 * S2 + f] -= (9.0 / 16.0) * (FV(1) + FV(2)) - (1.0 / 16.0) * (FV(1) + FV(3));
        const double u0 = r3(U, 0 * S2 + f), u1 = r3(U, 1 * S2 + f), u2 = r3(U, 2 * S2 + f);
        const double uuu = ((9.0 / 16.0) * (u0 + u1) - (1.0 / 16.0) * (0.0 + u2)) * h4(Wa, S2, 1, f, 1);
        const double uuu2 = ((9.0 / 16.0) * (0.0 + u0) - (1.0 / 16.0) * (0.0 + u1)) * h4(Wa, S2, 0, f, 1);
        ru[1 * S2 + f] -= (FU(2) - uuu);
        ru[0 * S2 + f] -= (uuu - uuu2);
        /* K = 0 block (:178-194) */
        ru[(N - 2) * S2 + f] -= (9.0 / 16.0) * (FU(N - 1) + FU(N - 2)) - (1.0 / 16.0) * (FU(N - 1) + FU(N - 3));
        rv[(N - 2) * S2 + f] -= (9.0 / 16.0) * (FV(N - 1) + FV(N - 2)) - (1.0 / 16.0) * (FV(N - 1) + FV(N - 3));
        rv[0 * S2 + f] = rv[1 * S2 + f];
        ru[(N - 2) * S2 + f] = ru[(N - 3) * S2 + f];
        rv[(N - 2) * S2 + f] = rv[(N - 3) * S2 + f];
        ru[(N - 1) * S2 + f] = ru[(N - 2) * S2 + f];
        rv[(N - 1) * S2 + f] = rv[(N - 2) * S2 + f];
        ru[N * S2 + f] = ru[(N - 2) * S2 + f];
        rv[N * S2 + f] = rv[(N - 2) * S2 + f];
#undef FU
#undef FV
    }
}

/* ---------------------------------------------------------------------------
 * set_maxflux (step3D_uv.kernels.cpp:372-405)
 * ------------------------------------------------------------------------- */
void sfo_set_maxflux(const sfo_grid3 *g, const double *u, const double *v, double *Huon, double *Hvom,
                     const double *Hz)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        if (isUNode(g, f))
            for (ptrdiff_t K = 0; K < N; K++) {
                const ptrdiff_t b = K * S2 + f;
                Huon[b] = (((Hz[b] + Hz[b - 1]) * 0.5) * u[b]) * g->on_u[f];
            }
        if (isVNode(g, f))
            for (ptrdiff_t K = 0; K < N; K++) {
                const ptrdiff_t b = K * S2 + f;
                Hvom[b] = (((Hz[b] + Hz[b - g->szI]) * 0.5) * v[b]) * g->om_v[f];
            }
    }
}

/* ---------------------------------------------------------------------------
 * omega (step3D_uv.kernels.cpp:62-165)
 * ------------------------------------------------------------------------- */
void sfo_omega(const sfo_grid3 *g, double *Wv, const double *Huon, const double *Hvom, const double *z_w,
               const int32_t *BC)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N, n3 = (N + 1) * S2;
    arr_t Hu = {Huon, n3}, Hv = {Hvom, n3};
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        if (BC[f] >= 0) continue;
        Wv[f] = 0.0;
        for (ptrdiff_t K = 1; K < N; K++) {
            const ptrdiff_t b = K * S2 + f;
            Wv[b] = Wv[b - S2] - ((r3(Hu, b + 1) - r3(Hu, b)) + (r3(Hv, b + g->szI) - r3(Hv, b)));
        }
        const double zb = z_w[f];
        const double h = z_w[(N - 1) * S2 + f] - zb;
        const double Ws = Wv[(N - 1) * S2 + f];
        for (ptrdiff_t K = N - 1; K > 0; K--) {
            const ptrdiff_t b = K * S2 + f;
            Wv[b] = Wv[b] - (Ws * (z_w[b] - zb)) / h;
        }
        Wv[N * S2 + f] = 0.0;
    }
}

/* ---------------------------------------------------------------------------
 * set_depth (step3D_uv.kernels.cpp:225-291) and set_zeta (:294-310)
 * ------------------------------------------------------------------------- */
void sfo_set_depth(const sfo_grid3 *g, int Vtransform, const double *Zt, double *z_w, double *z_r,
                   const double *h, double hc, double *Hz, const double *sc_r, const double *sc_w,
                   const double *Cs_r, const double *Cs_w)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        if (Vtransform == 1) {
            for (ptrdiff_t K = 0; K < N; K++) {
                const double z_r0 = hc * (sc_r[K] - Cs_r[K]) + Cs_r[K] * h[f];
                const double z_w0 = hc * (sc_w[K] - Cs_w[K]) + Cs_w[K] * h[f];
                z_w[K * S2 + f] = z_w0 + Zt[f] * (1.0 + z_w0 / h[f]);
                z_r[K * S2 + f] = z_r0 + Zt[f] * (1.0 + z_r0 / h[f]);
            }
            z_w[N * S2 + f] = Zt[f];
        } else if (Vtransform == 2) {
            z_w[N * S2 + f] = 0.0;
            for (ptrdiff_t K = 0; K < N; K++) {
                const double cr = (hc * sc_r[K] + Cs_r[K] * h[f]) / (hc + h[f]);
                const double cw = (hc * sc_w[K] + Cs_w[K] * h[f]) / (hc + h[f]);
                z_w[K * S2 + f] = Zt[f] + (Zt[f] + h[f]) * cw;
                z_r[K * S2 + f] = Zt[f] + (Zt[f] + h[f]) * cr;
            }
            z_w[N * S2 + f] = Zt[f];
        }
        for (ptrdiff_t K = 0; K < N; K++) Hz[K * S2 + f] = z_w[(K + 1) * S2 + f] - z_w[K * S2 + f];
    }
}

void sfo_set_zeta(const sfo_grid3 *g, double *zeta1, double *zeta2, const double *Zt)
{
    const ptrdiff_t S2 = S2_OF(g);
    for (ptrdiff_t f = 0; f < S2; f++) {
        if (!isRNode(g, f)) continue;
        zeta2[f] = Zt[f];
        zeta1[f] = Zt[f];
    }
}

/* ---------------------------------------------------------------------------
 * setLateralUVBCs (step3D_uv.kernels.cpp:314-368); gradient v boundaries index the
 * 2D index field at 3D offsets in the reference (out of bounds): rejected (-1).
 * ------------------------------------------------------------------------- */
int sfo_setLateralUVBCs(const sfo_grid3 *g, double t, double *u, double *v, const int32_t *uIdx2,
                        const int32_t *vIdx2, const int32_t *uType, const int32_t *vType)
{
    (void)uIdx2;
    (void)vIdx2;
    const ptrdiff_t S2 = S2_OF(g), N = g->N;
    for (ptrdiff_t f = 0; f < S2; f++)
        if (vType[f] > 0 && !(vType[f] & (1 << 4)) && !(vType[f] & (1 << 3)) && (vType[f] & (1 << 6)))
            return -1;
    for (ptrdiff_t f = 0; f < S2; f++) {
        const int bu = uType[f], bv = vType[f];
        if (bu <= 0 && bv <= 0) continue;
        if (bu > 0 && !(bu & (1 << 4)) && (bu & (1 << 3))) {
            const double val = 0.05 * sin(t * 0.0002);
            for (ptrdiff_t k = 0; k < N; k++) u[f + k * S2] = val;
        }
        if (bv > 0 && ((bv & (1 << 4)) || (bv & (1 << 3))))
            for (ptrdiff_t k = 0; k < N; k++) v[f + k * S2] = 0.0;
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * step3d_UV (step3D_uv.kernels.cpp:646-707) with createVertViscousOpMatrix
 * (:568-639), solveTri (:520-565), correctBaroclinicMeanVel (:475-517) and
 * adjustBarotropicVelocity (:408-469).  SD/D/RHS: caller scratch, (N+1)*S2 each.
 * ------------------------------------------------------------------------- */
static void build_matrix(const sfo_grid3 *g, int vdir, ptrdiff_t f, double cff, double Dt, double lambda,
                         double *SD, double *D, double *RHS, const double *Hz, const double *Akv,
                         const double *z_r, const double *u, const double *ru)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N;
    const ptrdiff_t d = vdir ? g->szI : 1;   /* neighbour that RtoU / RtoV average with */
    if (vdir ? !isVNode(g, f) : !isUNode(g, f)) return;
#define HZU(K) ((Hz[(K) * S2 + f] + Hz[(K) * S2 + f - d]) * 0.5)
#define AKVU(K) ((Akv[(K) * S2 + f] + Akv[(K) * S2 + f - d]) * 0.5)
#define DZ(K) (((z_r[(K) * S2 + f] - z_r[((K) - 1) * S2 + f]) + (z_r[(K) * S2 + f - d] - z_r[((K) - 1) * S2 + f - d])) * 0.5)
    const double pmU = (g->pm[f] + g->pm[f - d]) * 0.5;
    const double pnU = (g->pn[f] + g->pn[f - d]) * 0.5;
    const double cdt_mn = (cff * Dt) * (pmU * pnU);
    SD[(N - 1) * S2 + f] = 0.0;
    D[f] = HZU(0);
    RHS[f] = 0.0;
    for (ptrdiff_t K = 0; K < N - 1; K++) SD[K * S2 + f] = (((-lambda) * Dt) * AKVU(K)) / DZ(K + 1);
    for (ptrdiff_t K = 1; K < N; K++) {
        RHS[K * S2 + f] = HZU(K) * (u[K * S2 + f] + cdt_mn * ru[K * S2 + f]);
        D[K * S2 + f] = (HZU(K) - SD[K * S2 + f]) - SD[(K - 1) * S2 + f];
    }
    D[(N - 1) * S2 + f] = HZU(N - 1) - SD[(N - 2) * S2 + f];
    SD[f] = (((-lambda) * Dt) * AKVU(0)) / DZ(1);
    D[f] = HZU(0) - 2.0 * SD[f];
#undef HZU
#undef AKVU
#undef DZ
}

static void solve_tri(const sfo_grid3 *g, ptrdiff_t f, double *SD, double *D, double *RHS, double *x)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N;
    if (!isUNode(g, f)) return;
    for (ptrdiff_t k = 1; k < N; k++) {
        const double w = SD[(k - 1) * S2 + f] / D[(k - 1) * S2 + f];
        D[k * S2 + f] -= w * SD[(k - 1) * S2 + f];
        RHS[k * S2 + f] -= w * RHS[(k - 1) * S2 + f];
    }
    x[(N - 1) * S2 + f] = RHS[(N - 1) * S2 + f] / D[(N - 1) * S2 + f];
    for (ptrdiff_t k = N - 2; k >= 0; k--)
        x[k * S2 + f] = (RHS[k * S2 + f] - SD[k * S2 + f] * x[(k + 1) * S2 + f]) / D[k * S2 + f];
}

static void correct_mean(const sfo_grid3 *g, int vdir, ptrdiff_t f, double *x, const double *Hz,
                         const double *Davg, const double *metric, const double *xold)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N;
    const ptrdiff_t d = vdir ? g->szI : 1;
    if (vdir ? !isVNode(g, f) : !isUNode(g, f)) return;
    double H = 0.0, Hx = 0.0;
    for (ptrdiff_t K = 0; K < N; K++) {
        const double hu = (Hz[K * S2 + f] + Hz[K * S2 + f - d]) * 0.5;
        H += hu;
        Hx += xold[K * S2 + f] * hu;
    }
    const double diff = (Hx - Davg[f] / metric[f]) / H;
    for (ptrdiff_t K = 0; K < N; K++) x[K * S2 + f] = x[K * S2 + f] - diff;
}

void sfo_step3d_UV(const sfo_grid3 *g, double *u, double *v, const double *ru, const double *rv,
                   double *ubar_t1, double *vbar_t1, double *ubar_t2, double *vbar_t2, const double *Hz,
                   const double *Akv, const double *z_r, const double *DU_avg1, const double *DV_avg1,
                   const double *tmpU, const double *tmpV, int iic, int ntfirst, double lambda, double Dt,
                   double *SD, double *D, double *RHS)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N;
    double cff;
    if (iic == ntfirst) cff = 1.0;
    else if (iic == ntfirst + 1) cff = 3.0 / 2.0;
    else cff = 23.0 / 12.0;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        build_matrix(g, 0, f, cff, Dt, lambda, SD, D, RHS, Hz, Akv, z_r, tmpU, ru);
        solve_tri(g, f, SD, D, RHS, u);
        correct_mean(g, 0, f, u, Hz, DU_avg1, g->on_u, tmpU);
        build_matrix(g, 1, f, cff, Dt, lambda, SD, D, RHS, Hz, Akv, z_r, tmpV, rv);
        solve_tri(g, f, SD, D, RHS, v);
        correct_mean(g, 1, f, v, Hz, DV_avg1, g->om_v, tmpV);
        if (!isUNode(g, f)) continue;
        /* adjustBarotropicVelocity (:408-469) */
        double HU = 0.0, HV = 0.0, HuU = 0.0, HvV = 0.0;
        for (ptrdiff_t K = 0; K < N; K++) {
            const ptrdiff_t b = K * S2 + f;
            const double hzu = (Hz[b] + Hz[b - 1]) * 0.5;
            const double hzv = (Hz[b] + Hz[b - g->szI]) * 0.5;
            HU += hzu;
            HV += hzv;
            HuU += hzu * u[b];
            HvV += hzv * v[b];
        }
        const double uUp = (HuU - DU_avg1[f] / g->on_u[f]) / HU;
        const double vUp = (HvV - DV_avg1[f] / g->om_v[f]) / HV;
        ubar_t2[f] = (DU_avg1[f] / g->on_u[f]) / HU + uUp;
        vbar_t2[f] = (DV_avg1[f] / g->om_v[f]) / HV + vUp;
        ubar_t1[f] = ubar_t2[f];
        vbar_t1[f] = vbar_t2[f];
    }
}

/* ---------------------------------------------------------------------------
 * Coupling of the 2D and 3D momentum equations (ROMS split-explicit form).
 *
 * sfo_couple_rufrc: the 2D forcing of the baroclinic step, the vertical integral of
 * the 3D right-hand side plus the surface and bottom stresses (nonlinear/rhs.py:76-89,
 * the commented vertIntegral + stress terms; bottom stress rdrg * u of the bottom
 * level):  rufrc = sum_K ru(K) + ((sustr - rdrg * u(0)) * om_u) * on_u, on the
 * interior u points (i = 2..L-1, j = 1..M-1), and the same for v (j = 2..M-1,
 * i = 1..L-1); 0 elsewhere.  Metrics as the physics substep derives them:
 * on_u = 2/(pn(i-1)+pn(i)), om_u = 2/(pm(i-1)+pm(i)), om_v = 2/(pm(j-1)+pm(j)),
 * on_v = 2/(pn(j-1)+pn(j)).
 * ------------------------------------------------------------------------- */
void sfo_couple_rufrc(const sfo_grid3 *g, const double *ru, const double *rv, const double *u, const double *v,
                      const double *sustr, const double *svstr, double rdrg, double *rufrc, double *rvfrc)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N, W = g->szI;
    const int L = g->szI - 1, M = g->szJ - 1;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        const int i = (int)(f % W), j = (int)(f / W);
        double su = 0.0, sv = 0.0;
        if (i >= 2 && i <= L - 1 && j >= 1 && j <= M - 1) {
            for (ptrdiff_t K = 0; K < N; K++) su = su + ru[K * S2 + f];
            const double onu = 2.0 / (g->pn[f - 1] + g->pn[f]), omu = 2.0 / (g->pm[f - 1] + g->pm[f]);
            su = su + ((sustr[f] - rdrg * u[f]) * omu) * onu;
        }
        if (j >= 2 && j <= M - 1 && i >= 1 && i <= L - 1) {
            for (ptrdiff_t K = 0; K < N; K++) sv = sv + rv[K * S2 + f];
            const double omv = 2.0 / (g->pm[f - W] + g->pm[f]), onv = 2.0 / (g->pn[f - W] + g->pn[f]);
            sv = sv + ((svstr[f] - rdrg * v[f]) * omv) * onv;
        }
        rufrc[f] = su;
        rvfrc[f] = sv;
    }
}

/* The 3D momentum step of the coupled form (ROMS step3d_uv with the "Couple 2D and 3D
 * momentum equations" section, nonlinear/step3d_uv.py:155-180 and :365-544; the live
 * kernel step3D_uv.kernels.cpp:568-707 is the uncoupled sfo_step3d_UV above).
 *
 * The reference's solve forms RHS = Hz_u (u + cff dt pm pn ru), but ru already carries
 * Hz (ru = Hz f v area, the flux divergences of Huon u, W u), so every tendency is
 * multiplied by the layer thickness a second time, and cff = 23/12 multiplies the
 * current ru with no history.  Its velocity never moves in the reference's loop, where
 * this stays hidden; with the coupling driving the 3D flow the extra factor Hz f dt
 * (6.7 for the 116 m bottom layers of a 1000 m column with dt = 300 s) makes the step
 * grow about 7x per step.  The coupled form therefore solves ROMS's equation
 *   Hz_u,new u_new - dt d/dz (Akv du_new/dz) = Hz_u,old u + dt pm_u pn_u AB3(ru)
 *                                              - dt rdrg u(0) [bottom] + dt sustr [top]
 * on the interior u points (i = 2..L-1, j = 1..M-1, levels 0..N-1; v: j = 2..M-1,
 * i = 1..L-1), the points the barotropic mode and rufrc cover (the other columns -- the
 * walls and the ghost rows -- are left as they are: the reference's lateral boundary
 * step does not zero the wall velocity, so a solve there would only accumulate the
 * wind stress), with
 *   SD(K) = (-lambda dt) Akv_u(K+1) / dz_u(K+1)   (K = 0..N-2, SD(N-1) = 0; Akv_u(K+1) at
 *            the interface between levels K and K+1, dz_u the z_r difference at u),
 *   D(K)  = (Hz_u,new(K) - SD(K)) - SD(K-1)       (SD(-1) = 0),
 *   AB3(ru) = ru (ab 0), 1.5 ru - 0.5 ru1 (ab 1), (23/12 ru - 16/12 ru1) + 5/12 ru2 (ab 2),
 * where ru1, ru2 are the right-hand sides of the two steps before; this step's ru is
 * stored into ru_keep for the next.  The bottom stress rdrg u(0) and the surface stress
 * sustr are the ones rufrc gives the barotropic mode.  Then per column, with
 * DC(K) = on_u Hz_u,new(K), D = sum DC(K):
 *   CF = (sum DC(K) u(K) - DU_avg1) / D;  u(K) -= CF;  ubar(t1) = ubar(t2) = DU_avg1 / D
 *   Huon(K) = 0.5 (Huon(K) + u(K) DC(K));  FC = (sum Huon(K) - DU_avg2) / D;
 *   Huon(K) -= DC(K) FC
 * so the new u carries the time-averaged barotropic transport DU_avg1 and the mass
 * flux the filtered transport DU_avg2.  (The text's u section omits the FC sum that its
 * v section has; both sections sum here, as ROMS does.) */
typedef struct {
    double *u, *v;                       /* szK+1 levels, updated (levels 0..N-1) */
    const double *ru, *rv;               /* this step's right-hand side (levels 0..N-1 read) */
    const double *ru1, *rv1, *ru2, *rv2; /* the two steps before (N levels) */
    double *ru_keep, *rv_keep;           /* receive ru, rv (N levels) */
    double *ubar_t1, *vbar_t1, *ubar_t2, *vbar_t2;
    const double *Hz_old, *Hz, *Akv, *z_r;
    const double *DU_avg1, *DV_avg1, *DU_avg2, *DV_avg2;
    double *Huon, *Hvom;
    const double *sustr, *svstr;
    double rdrg;
    int ab;
    double *SD, *D, *RHS;                /* scratch, N * S2 each */
} sfo_uv_coupled;

static double ab3(int ab, double r, double r1, double r2)
{
    if (ab == 0) return r;
    if (ab == 1) return 1.5 * r - 0.5 * r1;
    return ((23.0 / 12.0) * r - (16.0 / 12.0) * r1) + (5.0 / 12.0) * r2;
}

/* the interior u (vdir 0) or v (vdir 1) points of the coupled form */
static int coupled_point(const sfo_grid3 *g, int vdir, ptrdiff_t f)
{
    const int i = (int)(f % g->szI), j = (int)(f / g->szI), L = g->szI - 1, M = g->szJ - 1;
    return vdir ? (j >= 2 && j <= M - 1 && i >= 1 && i <= L - 1) : (i >= 2 && i <= L - 1 && j >= 1 && j <= M - 1);
}

static void momentum_column(const sfo_grid3 *g, int vdir, ptrdiff_t f, const sfo_uv_coupled *a, double lambda,
                            double Dt)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N;
    const ptrdiff_t d = vdir ? g->szI : 1;
    double *x = vdir ? a->v : a->u, *keep = vdir ? a->rv_keep : a->ru_keep;
    const double *r = vdir ? a->rv : a->ru, *r1 = vdir ? a->rv1 : a->ru1, *r2 = vdir ? a->rv2 : a->ru2;
    const double *stress = vdir ? a->svstr : a->sustr;
    double *SD = a->SD, *D = a->D, *RHS = a->RHS;
    const double *Hz = a->Hz, *Ho = a->Hz_old, *Akv = a->Akv, *z_r = a->z_r;
#define AVG(A, K) (((A)[(K) * S2 + f] + (A)[(K) * S2 + f - d]) * 0.5)
#define DZ(K) (((z_r[(K) * S2 + f] - z_r[((K) - 1) * S2 + f]) + (z_r[(K) * S2 + f - d] - z_r[((K) - 1) * S2 + f - d])) * 0.5)
    const double pmU = (g->pm[f] + g->pm[f - d]) * 0.5;
    const double pnU = (g->pn[f] + g->pn[f - d]) * 0.5;
    const double cdt_mn = Dt * (pmU * pnU);
    for (ptrdiff_t K = 0; K < N - 1; K++) SD[K * S2 + f] = (((-lambda) * Dt) * AVG(Akv, K + 1)) / DZ(K + 1);
    SD[(N - 1) * S2 + f] = 0.0;
    for (ptrdiff_t K = 0; K < N; K++) {
        const ptrdiff_t b = K * S2 + f;
        const double rk = r[b];
        const double rab = ab3(a->ab, rk, a->ab > 0 ? r1[b] : 0.0, a->ab > 1 ? r2[b] : 0.0);
        keep[b] = rk;
        RHS[b] = AVG(Ho, K) * x[b] + cdt_mn * rab;
        D[b] = K == 0 ? AVG(Hz, 0) - SD[f] : (AVG(Hz, K) - SD[b]) - SD[b - S2];
    }
    RHS[f] = RHS[f] - (Dt * a->rdrg) * x[f];
    RHS[(N - 1) * S2 + f] = RHS[(N - 1) * S2 + f] + Dt * stress[f];
#undef AVG
#undef DZ
    for (ptrdiff_t k = 1; k < N; k++) {
        const double w = SD[(k - 1) * S2 + f] / D[(k - 1) * S2 + f];
        D[k * S2 + f] -= w * SD[(k - 1) * S2 + f];
        RHS[k * S2 + f] -= w * RHS[(k - 1) * S2 + f];
    }
    x[(N - 1) * S2 + f] = RHS[(N - 1) * S2 + f] / D[(N - 1) * S2 + f];
    for (ptrdiff_t k = N - 2; k >= 0; k--)
        x[k * S2 + f] = (RHS[k * S2 + f] - SD[k * S2 + f] * x[(k + 1) * S2 + f]) / D[k * S2 + f];
}

static void couple_column(const sfo_grid3 *g, int vdir, ptrdiff_t f, double *x, const double *Hz,
                          const double *Davg1, const double *Davg2, double *Hf, double *xbar1, double *xbar2)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N, W = g->szI;
    const ptrdiff_t d = vdir ? W : 1;
    const double met = vdir ? 2.0 / (g->pm[f - W] + g->pm[f]) : 2.0 / (g->pn[f - 1] + g->pn[f]);
#define DC(K) (met * ((Hz[(K) * S2 + f] + Hz[(K) * S2 + f - d]) * 0.5))
    double Ds = 0.0, CF = 0.0;
    for (ptrdiff_t K = 0; K < N; K++) {
        Ds = Ds + DC(K);
        CF = CF + DC(K) * x[K * S2 + f];
    }
    const double rD = 1.0 / Ds;
    CF = rD * (CF - Davg1[f]);
    xbar2[f] = rD * Davg1[f];
    xbar1[f] = xbar2[f];
    for (ptrdiff_t K = 0; K < N; K++) x[K * S2 + f] = x[K * S2 + f] - CF;
    double FC = 0.0;
    for (ptrdiff_t K = 0; K < N; K++) {
        Hf[K * S2 + f] = 0.5 * (Hf[K * S2 + f] + x[K * S2 + f] * DC(K));
        FC = FC + Hf[K * S2 + f];
    }
    FC = rD * (FC - Davg2[f]);
    for (ptrdiff_t K = 0; K < N; K++) Hf[K * S2 + f] = Hf[K * S2 + f] - DC(K) * FC;
#undef DC
}

void sfo_step3d_uv_coupled(const sfo_grid3 *g, const sfo_uv_coupled *a, double lambda, double Dt)
{
    const ptrdiff_t S2 = S2_OF(g);
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        if (coupled_point(g, 0, f)) {
            momentum_column(g, 0, f, a, lambda, Dt);
            couple_column(g, 0, f, a->u, a->Hz, a->DU_avg1, a->DU_avg2, a->Huon, a->ubar_t1, a->ubar_t2);
        }
        if (coupled_point(g, 1, f)) {
            momentum_column(g, 1, f, a, lambda, Dt);
            couple_column(g, 1, f, a->v, a->Hz, a->DV_avg1, a->DV_avg2, a->Hvom, a->vbar_t1, a->vbar_t2);
        }
    }
}

/* ---------------------------------------------------------------------------
 * Baroclinic terms: linear equation of state, baroclinic pressure gradient and the
 * tracer step.  The reference names them but leaves them out (nonlinear/rhs.py:31-32
 * "prsgrd()" commented out; nonlinear/step3d_t.py Fortran text; the linear state
 * equation of utility/test.in:2130-2160).
 *
 * sfo_rho_eos: density anomaly over 1000 kg/m3 of the N levels,
 *   rho = ((R0 - 1000) - (R0 Tcoef)(T - T0)) + (R0 Scoef)(S - S0).
 * ------------------------------------------------------------------------- */
void sfo_rho_eos(const sfo_grid3 *g, const double *T, const double *S, double R0, double T0, double S0,
                 double Tcoef, double Scoef, double *rho)
{
    const ptrdiff_t n = (ptrdiff_t)g->N * S2_OF(g);
    const double a = R0 - 1000.0, bt = R0 * Tcoef, bs = R0 * Scoef;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t b = 0; b < n; b++) rho[b] = (a - bt * (T[b] - T0)) + bs * (S[b] - S0);
}

/* Baroclinic pressure gradient by the standard density Jacobian on s-levels (Song 1998,
 * the second-order scheme of ROMS prsgrd31), added to ru/rv (m4/s2) of levels 0..N-1 on
 * the interior u points (i = 2..L-1, j = 1..M-1) and v points (j = 2..M-1, i = 1..L-1).
 * With fac1 = 0.5 g/rho0, fac2 = 1000 g/rho0, fac3 = 0.25 g/rho0 and the neighbour
 * p = i-1 (u) or j-1 (v), from the top level N-1 down:
 *   phi(N-1) = fac1 (rho - rho_p)((z_w(N) - z_r) + (z_w(N) - z_r)_p)
 *              + (fac2 + fac1 (rho + rho_p)) (z_w(N) - z_w(N)_p)
 *   phi(k)   = phi(k+1) + fac3 (C1 C3 - C2 C4), with
 *     gamma = 0.125 (dz_r(k) + dz_r(k+1)) (z_r(k+1) - z_r(k) - (z_r(k+1) - z_r(k))_p) /
 *             ((z_r(k+1) - z_r(k)) (z_r(k+1) - z_r(k))_p),   dz_r(k) = z_r(k) - z_r(k)_p,
 *     C1 = (1+gamma) drho(k+1) + (1-gamma) drho(k),   C2 = rho(k+1) + rho_p(k+1) - rho(k) - rho_p(k),
 *     C3 = z_r(k+1) + z_r_p(k+1) - z_r(k) - z_r_p(k), C4 = (1+gamma) dz_r(k+1) + (1-gamma) dz_r(k)
 *   ru(k) += -0.5 (Hz(k) + Hz_p(k)) phi(k) on_u      (on_u = 2/(pn+pn_p); v: om_v = 2/(pm+pm_p)). */
static void pgf_column(const sfo_grid3 *g, int vdir, ptrdiff_t f, const double *rho, const double *z_r,
                       const double *z_w, const double *Hz, double grav, double rho0, double *r)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N, W = g->szI;
    const int i = (int)(f % W), j = (int)(f / W), L = g->szI - 1, M = g->szJ - 1;
    if (!vdir && !(i >= 2 && i <= L - 1 && j >= 1 && j <= M - 1)) return;
    if (vdir && !(j >= 2 && j <= M - 1 && i >= 1 && i <= L - 1)) return;
    const ptrdiff_t p = vdir ? f - W : f - 1;
    const double met = vdir ? 2.0 / (g->pm[p] + g->pm[f]) : 2.0 / (g->pn[p] + g->pn[f]);
    const double fac1 = 0.5 * grav / rho0, fac2 = 1000.0 * grav / rho0, fac3 = 0.25 * grav / rho0;
#define R(k, x) rho[(ptrdiff_t)(k) * S2 + (x)]
#define ZR(k, x) z_r[(ptrdiff_t)(k) * S2 + (x)]
#define ZW(k, x) z_w[(ptrdiff_t)(k) * S2 + (x)]
    ptrdiff_t k = N - 1;
    const double c1 = (ZW(N, f) - ZR(k, f)) + (ZW(N, p) - ZR(k, p));
    double phi = (fac1 * (R(k, f) - R(k, p))) * c1;
    phi = phi + (fac2 + fac1 * (R(k, f) + R(k, p))) * (ZW(N, f) - ZW(N, p));
    r[k * S2 + f] = r[k * S2 + f] + ((-0.5 * (Hz[k * S2 + f] + Hz[k * S2 + p])) * phi) * met;
    for (k = N - 2; k >= 0; k--) {
        const double dzf = ZR(k + 1, f) - ZR(k, f), dzp = ZR(k + 1, p) - ZR(k, p);
        const double a2 = (ZR(k, f) - ZR(k, p)) + (ZR(k + 1, f) - ZR(k + 1, p));
        const double a3 = dzf - dzp;
        const double gam = ((0.125 * a2) * a3) / (dzf * dzp);
        const double C1 = (1.0 + gam) * (R(k + 1, f) - R(k + 1, p)) + (1.0 - gam) * (R(k, f) - R(k, p));
        const double C2 = (R(k + 1, f) + R(k + 1, p)) - (R(k, f) + R(k, p));
        const double C3 = (ZR(k + 1, f) + ZR(k + 1, p)) - (ZR(k, f) + ZR(k, p));
        const double C4 = (1.0 + gam) * (ZR(k + 1, f) - ZR(k + 1, p)) + (1.0 - gam) * (ZR(k, f) - ZR(k, p));
        phi = phi + fac3 * (C1 * C3 - C2 * C4);
        r[k * S2 + f] = r[k * S2 + f] + ((-0.5 * (Hz[k * S2 + f] + Hz[k * S2 + p])) * phi) * met;
    }
#undef R
#undef ZR
#undef ZW
}

void sfo_prsgrd(const sfo_grid3 *g, const double *rho, const double *z_r, const double *z_w, const double *Hz,
                double grav, double rho0, double *ru, double *rv)
{
    const ptrdiff_t S2 = S2_OF(g);
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        pgf_column(g, 0, f, rho, z_r, z_w, Hz, grav, rho0, ru);
        pgf_column(g, 1, f, rho, z_r, z_w, Hz, grav, rho0, rv);
    }
}

/* Tracer step (nonlinear/step3d_t.py, flux form; first-order upstream fluxes, forward
 * in time): on every interior rho column (i = 1..L-1, j = 1..M-1), from the bottom up,
 *   Dh(k) = pm pn ((Huon(i+1) - Huon(i)) + (Hvom(j+1) - Hvom(j)))        [level k]
 *   Wt(k) = Wt(k-1) - Dh(k) - (Hz_new(k) - Hz_old(k)) / dt,  Wt(-1) = 0  [volume flux, top of k]
 *   FX(i) = Huon(i) * (Huon(i) > 0 ? t(i-1) : t(i)), FE likewise, FC(k) = Wt(k) * (Wt(k) > 0 ? t(k) : t(k+1))
 *   t_new(k) = (Hz_old(k) t(k) - dt (pm pn ((FX(i+1) - FX(i)) + (FE(j+1) - FE(j))) + FC(k) - FC(k-1)))
 *              / Hz_new(k),   FC(-1) = 0,  FC(N-1) = Wt(N-1) t(N-1)
 * The vertical flux is formed from the same horizontal volume fluxes and thickness change,
 * so a uniform tracer stays uniform; its value at the surface is whatever the column's
 * volume budget leaves.  Ghost columns copy their interior neighbour (zero gradient). */
void sfo_step3d_t(const sfo_grid3 *g, const double *t, double *tn, const double *Huon, const double *Hvom,
                  const double *Hz_old, const double *Hz_new, double dt)
{
    const ptrdiff_t S2 = S2_OF(g), N = g->N, W = g->szI;
    const int L = g->szI - 1, M = g->szJ - 1;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        const int i = (int)(f % W), j = (int)(f / W);
        if (!(i >= 1 && i <= L - 1 && j >= 1 && j <= M - 1)) continue;
        const double mn = g->pm[f] * g->pn[f];
        double Wlow = 0.0, FClow = 0.0;
        for (ptrdiff_t k = 0; k < N; k++) {
            const ptrdiff_t b = k * S2 + f;
            const double hu0 = Huon[b], hu1 = Huon[b + 1], hv0 = Hvom[b], hv1 = Hvom[b + W];
            const double Dh = mn * ((hu1 - hu0) + (hv1 - hv0));
            const double Wk = (Wlow - Dh) - (Hz_new[b] - Hz_old[b]) / dt;
            const double FX0 = hu0 * (hu0 > 0.0 ? t[b - 1] : t[b]);
            const double FX1 = hu1 * (hu1 > 0.0 ? t[b] : t[b + 1]);
            const double FE0 = hv0 * (hv0 > 0.0 ? t[b - W] : t[b]);
            const double FE1 = hv1 * (hv1 > 0.0 ? t[b] : t[b + W]);
            const double FCk = k < N - 1 ? Wk * (Wk > 0.0 ? t[b] : t[b + S2]) : Wk * t[b];
            const double div = mn * ((FX1 - FX0) + (FE1 - FE0)) + (FCk - FClow);
            tn[b] = (Hz_old[b] * t[b] - dt * div) / Hz_new[b];
            Wlow = Wk;
            FClow = FCk;
        }
    }
    /* ghost columns: zero gradient (rows, then columns) */
#pragma omp parallel for schedule(static)
    for (ptrdiff_t k = 0; k < N; k++) {
        for (int i = 1; i <= L - 1; i++) {
            tn[k * S2 + i] = tn[k * S2 + W + i];
            tn[k * S2 + (ptrdiff_t)M * W + i] = tn[k * S2 + (ptrdiff_t)(M - 1) * W + i];
        }
        for (int j = 0; j <= M; j++) {
            const int jj = j < 1 ? 1 : (j > M - 1 ? M - 1 : j);
            tn[k * S2 + (ptrdiff_t)j * W] = tn[k * S2 + (ptrdiff_t)jj * W + 1];
            tn[k * S2 + (ptrdiff_t)j * W + L] = tn[k * S2 + (ptrdiff_t)jj * W + L - 1];
        }
    }
}
