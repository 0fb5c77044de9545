/*
 * sf_oracle2d.c -- CPU restatement (FP64) of the barotropic LF-AM3 kernels of the
 * Sailfish ROMS port, written from scratch from the reference sources.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/, the smoke()
 * entry point and the cpu_baseline leg of bench.py load it to check (and time) the
 * HIP product path.  No product code may link, import or call it.
 *
 * Conventions (mod_cppkernels.h:3-22, :40-41, :255-264):
 *   flat index f = j*szI + i, i (xi) fastest; S2 = szI*szJ; a 3D field has level
 *   stride S2.  "left" = f-1, "down" = f-szI.  The reference evaluates every stencil
 *   expression left-associatively in FP64 (expression-template Eval order); each
 *   formula below keeps that operand order so that a build with
 *   -ffp-contract=off reproduces it bit for bit.
 *
 * Undefined behaviour in the reference, and the definition chosen here:
 *   - computeMomentumPred (mod_cppkernels.cpp:162) uses an i%szJ mask, so on grids
 *     with szI>szJ it evaluates row-0 points whose RtoV stencil reads f-szI<0, i.e.
 *     before the start of the array.  Such reads return 0.0 here (ld0()).
 *   - BC copy lists (mod_BC.c:16-23) race when a source is also a destination; the
 *     schedule fixed here is the serial one in list order (one legal outcome).
 *
 * Pins (tests/test_oracle2d.py; tests/test_oracle_mutations.py plants mistakes in
 * this file and checks that one of them fails):
 *   sfo_computeZetaRHS        test_zeta_rhs_linear_flux_closed_form,
 *                             test_zeta_rhs_telescoping_budget,
 *                             test_zeta_rhs_corrector_accumulates_averages
 *   sfo_computeZetaPred       test_zeta_pred_closed_form_dt_and_slots
 *   sfo_computeMomentumRHSPred test_momentum_rhs_pgf_closed_form,
 *                             test_momentum_rhs_pred_weight_and_slots
 *   sfo_computeMomentumRHSCorr test_momentum_rhs_corr_weight_and_slots
 *   sfo_computeMomentumPred   test_momentum_pred_closed_form, test_momentum_pred_mask_patterns
 *   sfo_AdamsMoultonCorr3rd   test_am3_constant_rhs_and_reference_coefficients
 *   sfo_AdamsMoultonCorr3rd2  test_am3_depth_weighted_closed_form
 *   sfo_setBC / sfo_copyBC    the hand-enumerated boundary sets of tests/test_host.py and
 *                             test_barotropic_loop_averages_of_rest_state
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#if defined(_OPENMP)
#include <omp.h>
#endif

typedef struct sfo_grid {
    int szI, szJ, szK;              /* L+1, M+1, N */
    const double *on_u, *om_v;      /* metric globals set by initialize (h:154-171) */
    const double *pn, *pm;
    int q5_stride;                  /* mask stride of computeMomentumPred: 0 -> szJ as in
                                       the reference (Q5), else the given stride (szI
                                       masks exactly column 0 and row 0) */
} sfo_grid;

/* read with the out-of-range-is-zero rule documented above */
static inline double ld0(const double *p, ptrdiff_t f) { return f < 0 ? 0.0 : p[f]; }

/* ------------------------------------------------------------------------------
 * computeZetaRHS  (mod_cppkernels.cpp:80-128)
 * interior only (:89); D = zeta+h; DU = ubar*RtoU(D); DV = vbar*RtoV(D);
 * res = divUVtoR(DU,DV,on_u,om_v,pn,pm)  (h:1045-1049): ((dU + dV)*pm)*pn.
 * Corrector: Zt_avg1 += cff1*zeta; DU_avg1 += (cff1*DU)*on_u; DV_avg1 likewise
 * (:116-125), with cff1 = weight1[iif-1] passed by the caller.
 * ----------------------------------------------------------------------------*/
void sfo_computeZetaRHS(const sfo_grid *g, const double *zeta, const double *h,
                        const double *ubar, const double *vbar,
                        double *DU_avg1, double *DV_avg1, double *Zt_avg1,
                        int isPredictor, double cff1, double *res)
{
    const int szI = g->szI, szJ = g->szJ;
    const ptrdiff_t S2 = (ptrdiff_t)szI * szJ;
    const double *on_u = g->on_u, *om_v = g->om_v, *pm = g->pm, *pn = g->pn;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        const int i = (int)(f % szI), j = (int)(f / szI);
        if (i == 0 || i == szI - 1 || j == 0 || j == szJ - 1) continue;
        const double D   = zeta[f] + h[f];
        const double Dl  = zeta[f - 1] + h[f - 1];
        const double Dd  = zeta[f - szI] + h[f - szI];
        const double Dr  = zeta[f + 1] + h[f + 1];
        const double Du  = zeta[f + szI] + h[f + szI];
        const double DU  = ubar[f] * ((D + Dl) * 0.5);
        const double DUr = ubar[f + 1] * ((Dr + D) * 0.5);
        const double DV  = vbar[f] * ((D + Dd) * 0.5);
        const double DVu = vbar[f + szI] * ((Du + D) * 0.5);
        res[f] = (((DUr * on_u[f + 1] - DU * on_u[f]) + (DVu * om_v[f + szI] - DV * om_v[f])) * pm[f]) * pn[f];
        if (!isPredictor) {
            Zt_avg1[f] = Zt_avg1[f] + cff1 * zeta[f];
            DU_avg1[f] = DU_avg1[f] + (cff1 * DU) * on_u[f];
            DV_avg1[f] = DV_avg1[f] + (cff1 * DV) * om_v[f];
        }
    }
}

/* computeZetaPred (mod_cppkernels.cpp:133-150): zeta_t2 = zeta_t0 + Dt*rzeta_t1 for
 * i%szI != 0 and i/szI != 0 (Q2: Dt, not 2*Dt). */
void sfo_computeZetaPred(const sfo_grid *g, double Dt, const double *zeta_t0,
                         const double *zeta_t1, double *zeta_t2, const double *rzeta_t1)
{
    (void)zeta_t1;
    const int szI = g->szI;
    const ptrdiff_t S2 = (ptrdiff_t)szI * g->szJ;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        if (f % szI == 0 || f / szI == 0) continue;
        zeta_t2[f] = zeta_t0[f] + Dt * rzeta_t1[f];
    }
}

/* shared body of computeMomentumRHSPred/Corr (mod_cppkernels.cpp:12-42, 45-77):
 * gz = (1-w)*za + w*zb; rhs_u = (g*on_u)*(RtoU(h)*DX(gz) + 0.5*DX(gz*gz)),
 * rhs_v = (g*om_v)*(RtoV(h)*DE(gz) + 0.5*DE(gz*gz)); interior only. */
static void momentum_rhs(const sfo_grid *g, const double *h, double *rhs_u, double *rhs_v,
                         const double *za, const double *zb, double gc, double w)
{
    const int szI = g->szI, szJ = g->szJ;
    const ptrdiff_t S2 = (ptrdiff_t)szI * szJ;
    const double *on_u = g->on_u, *om_v = g->om_v;
    const double a = 1 - w;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        const int i = (int)(f % szI), j = (int)(f / szI);
        if (i == 0 || i == szI - 1 || j == 0 || j == szJ - 1) continue;
        const double gz  = a * za[f] + w * zb[f];
        const double gzl = a * za[f - 1] + w * zb[f - 1];
        const double gzd = a * za[f - szI] + w * zb[f - szI];
        const double hu = (h[f] + h[f - 1]) * 0.5;
        const double hv = (h[f] + h[f - szI]) * 0.5;
        rhs_u[f] = (gc * on_u[f]) * (hu * (gz - gzl) + 0.5 * (gz * gz - gzl * gzl));
        rhs_v[f] = (gc * om_v[f]) * (hv * (gz - gzd) + 0.5 * (gz * gz - gzd * gzd));
    }
}

void sfo_computeMomentumRHSPred(const sfo_grid *g, const double *h, double *rhs_ubar,
                                double *rhs_vbar, const double *zeta_t0,
                                const double *zeta_t2, double gc, double weight)
{
    momentum_rhs(g, h, rhs_ubar, rhs_vbar, zeta_t2, zeta_t0, gc, weight);
}

void sfo_computeMomentumRHSCorr(const sfo_grid *g, const double *h, double *rhs_ubar,
                                double *rhs_vbar, const double *zeta_t0,
                                const double *zeta_t1, const double *zeta_t2, double gc)
{
    (void)zeta_t0;
    momentum_rhs(g, h, rhs_ubar, rhs_vbar, zeta_t2, zeta_t1, gc, 2.0 / 5.0);
}

/* computeMomentumPred (mod_cppkernels.cpp:155-186), with the i%szJ mask (Q5) unless
 * the grid selects another stride:
 * u_t2 = ((u_t1*D1U) + ((Dt*rhsu)*pm)*pn) / D2U, same for v with RtoV. */
void sfo_computeMomentumPred(const sfo_grid *g, double Dt, const double *u_t1, double *u_t2,
                             const double *v_t1, double *v_t2, const double *rhsu,
                             const double *rhsv, const double *h, const double *zeta_t1,
                             const double *zeta_t2)
{
    const int szI = g->szI, szJ = g->szJ;
    const ptrdiff_t S2 = (ptrdiff_t)szI * szJ;
    const double *pm = g->pm, *pn = g->pn;
    const int ms = g->q5_stride > 0 ? g->q5_stride : szJ;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        if (f % ms == 0 || f / ms == 0) continue;
        const double D1  = zeta_t1[f] + h[f];
        const double D2  = zeta_t2[f] + h[f];
        const double D1l = ld0(zeta_t1, f - 1) + ld0(h, f - 1);
        const double D2l = ld0(zeta_t2, f - 1) + ld0(h, f - 1);
        const double D1d = ld0(zeta_t1, f - szI) + ld0(h, f - szI);
        const double D2d = ld0(zeta_t2, f - szI) + ld0(h, f - szI);
        const double D1U = (D1 + D1l) * 0.5, D2U = (D2 + D2l) * 0.5;
        const double D1V = (D1 + D1d) * 0.5, D2V = (D2 + D2d) * 0.5;
        u_t2[f] = (u_t1[f] * D1U + ((Dt * rhsu[f]) * pm[f]) * pn[f]) / D2U;
        v_t2[f] = (v_t1[f] * D1V + ((Dt * rhsv[f]) * pm[f]) * pn[f]) / D2V;
    }
}

/* AdamsMoultonCorr3rd (mod_cppkernels.cpp:190-209), all f < S2. */
void sfo_AdamsMoultonCorr3rd(const sfo_grid *g, double Dt, double *v_t2, const double *rhs_t0,
                             const double *rhs_t1, const double *rhs_t2)
{
    const ptrdiff_t S2 = (ptrdiff_t)g->szI * g->szJ;
    const double AM3_2 = 5.0 / 12.0, AM3_1 = 8.0 / 12.0, AM3_0 = -1.0 / 12.0;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++)
        v_t2[f] = v_t2[f] + Dt * (AM3_2 * rhs_t2[f] + AM3_1 * rhs_t1[f] + AM3_0 * rhs_t0[f]);
}

/* AdamsMoultonCorr3rd2 (mod_cppkernels.cpp:213-251): the depth-weighted AM3 the
 * corrector intended; never launched by the reference (mod_operators.py:146 binds the
 * name to AdamsMoultonCorr3rd).  Kept for completeness of the kernel set. */
void sfo_AdamsMoultonCorr3rd2(const sfo_grid *g, double Dt, double *u_t2, double *v_t2,
                              const double *ru0, const double *ru1, const double *ru2,
                              const double *rv0, const double *rv1, const double *rv2,
                              const double *h, const double *zeta_t1, const double *zeta_t2)
{
    const int szI = g->szI;
    const ptrdiff_t S2 = (ptrdiff_t)szI * g->szJ;
    const double *pm = g->pm, *pn = g->pn;
    const double AM3_2 = 5.0 / 12.0, AM3_1 = 8.0 / 12.0, AM3_0 = -1.0 / 12.0;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) {
        const double D1  = zeta_t1[f] + h[f], D2 = zeta_t2[f] + h[f];
        const double D1l = ld0(zeta_t1, f - 1) + ld0(h, f - 1);
        const double D2l = ld0(zeta_t2, f - 1) + ld0(h, f - 1);
        const double D1d = ld0(zeta_t1, f - szI) + ld0(h, f - szI);
        const double D2d = ld0(zeta_t2, f - szI) + ld0(h, f - szI);
        const double D1U = (D1 + D1l) * 0.5, D2U = (D2 + D2l) * 0.5;
        const double D1V = (D1 + D1d) * 0.5, D2V = (D2 + D2d) * 0.5;
        const double pmn_u = (Dt * pm[f]) * pn[f];
        const double u0 = u_t2[f], v0 = v_t2[f];
        u_t2[f] = (u0 * D1U + pmn_u * (AM3_2 * ru2[f] + AM3_1 * ru1[f] + AM3_0 * ru0[f])) / D2U;
        v_t2[f] = (v0 * D1V + pmn_u * (AM3_2 * rv2[f] + AM3_1 * rv1[f] + AM3_0 * rv0[f])) / D2V;
    }
}

/* setBC / copyBC (mod_BC.c:6-23), serial list order. */
void sfo_setBC(double *var, const int32_t *idx, int n, double val)
{
    for (int k = 0; k < n; k++) var[idx[k]] = val;
}

void sfo_copyBC(double *var, const int32_t *idx1, const int32_t *idx2, int n)
{
    for (int k = 0; k < n; k++) var[idx1[k]] = var[idx2[k]];
}

int sfo_num_threads(void)
{
#if defined(_OPENMP)
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void sfo_set_num_threads(int n)
{
#if defined(_OPENMP)
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
