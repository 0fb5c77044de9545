/*
 * sf_oracle_phys.c -- CPU restatement (FP64) of the ROMS barotropic LF-AM3 step in
 * its 2D (no SOLVE3D) form, from the Fortran text the Sailfish port keeps in
 * obsolete/step2d_LF_AM3.h.py:
 *   free surface  :917-1014   (leapfrog predictor, AM3 corrector, zwrk weights)
 *   pressure grad.:1107-1237  (rhs = 0.5*g*on_u*((h+h)*(gz-gz) + (gz^2-gz^2)))
 *   advection     :1244-1290  (UV_C2ADVECTION branch) and :1497-1502 (divergence)
 *   Coriolis      :1503-1530
 *   bottom stress :2241-2251  (linear drag bustr = rdrg*ubar, nonlinear/set_vbc.py:8-13)
 *   surface stress:2595-2606
 *   momentum step :2623-2780  (first step FE/BE, then LF/AM3)
 * on the same flat C-grid as the live kernels (u point i between rho i-1 and i,
 * v point j between rho j-1 and j), with ROMS metrics derived from pm, pn:
 *   on_u = 2/(pn(i-1)+pn(i)), om_u = 2/(pm(i-1)+pm(i)),
 *   om_v = 2/(pm(j-1)+pm(j)), on_v = 2/(pn(j-1)+pn(j)).
 * Closed basin: u = 0 on the walls i=1 and i=L, v = 0 on j=1 and j=M; free-slip
 * ghost rows/columns (zero gradient) for u, v and zeta.  The barotropic averages
 * accumulate the predicted state with weight1[iif-1], as the live corrector does
 * (modules/mod_cppkernels.cpp:116-125).
 *
 * TEST INFRASTRUCTURE ONLY (see sf_oracle2d.c).  Operand order below is the order
 * the GPU kernels use; built with -ffp-contract=off.
 *
 * Pins (tests/test_oracle_phys.py, tests/test_oracle_volcons.py;
 * tests/test_oracle_mutations.py plants mistakes here):
 *   sfo_phys_rhs_zeta   test_rhs_zeta_divergence_closed_form
 *   sfo_phys_rhs_u/_v   test_rhs_pressure_gradient_closed_form, test_rhs_coriolis_closed_form,
 *                       test_rhs_wind_and_linear_drag_closed_form, test_rhs_c2_advection_closed_form
 *   sfo_phys_substep    test_leapfrog_predictor_of_a_divergent_flow,
 *                       test_predictor_pressure_surface_weights,
 *                       test_time_stepping_constant_wind_is_exact,
 *                       test_time_stepping_linear_drag_decay, test_averages_of_uniform_flow,
 *                       test_seiche_period, test_coriolis_acceleration_first_step,
 *                       test_volume_conserved_closed_basin, test_mirror_symmetry_without_rotation
 *   sfo_phys_bc         the closed-basin tests above (walls, ghosts)
 *   sfo_obc_volcons     tests/test_oracle_volcons.py (closed forms)
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>
#include <omp.h>

typedef struct sfo_phys {
    int szI, szJ;
    const double *h, *pm, *pn, *fomn, *sustr, *svstr;
    double g, rdrg;
} sfo_phys;

#define IDX(j, i) ((ptrdiff_t)(j) * W + (i))

/* the momentum right-hand sides at u point (j,i) and v point (j,i) of state
 * (z, u, v) with pressure surface gz; D = h + z. */
static inline double Dat(const sfo_phys *p, const double *z, ptrdiff_t f) { return p->h[f] + z[f]; }

static double DUon(const sfo_phys *p, const double *z, const double *u, int W, int j, int i)
{
    const ptrdiff_t f = IDX(j, i);
    const double onu = 2.0 / (p->pn[f - 1] + p->pn[f]);
    return ((0.5 * (Dat(p, z, f - 1) + Dat(p, z, f))) * u[f]) * onu;
}
static double DVom(const sfo_phys *p, const double *z, const double *v, int W, int j, int i)
{
    const ptrdiff_t f = IDX(j, i);
    const double omv = 2.0 / (p->pm[f - W] + p->pm[f]);
    return ((0.5 * (Dat(p, z, f - W) + Dat(p, z, f))) * v[f]) * omv;
}

double sfo_phys_rhs_zeta(const sfo_phys *p, const double *z, const double *u, const double *v, int j, int i)
{
    const int W = p->szI;
    return (DUon(p, z, u, W, j, i) - DUon(p, z, u, W, j, i + 1)) +
           (DVom(p, z, v, W, j, i) - DVom(p, z, v, W, j + 1, i));
}

double sfo_phys_rhs_u(const sfo_phys *p, const double *z, const double *u, const double *v, const double *gz,
                      int j, int i)
{
    const int W = p->szI;
    const ptrdiff_t f = IDX(j, i);
    const double onu = 2.0 / (p->pn[f - 1] + p->pn[f]);
    const double omu = 2.0 / (p->pm[f - 1] + p->pm[f]);
    /* pressure gradient */
    const double g0 = gz[f - 1], g1 = gz[f];
    double r = ((0.5 * p->g) * onu) * ((p->h[f - 1] + p->h[f]) * (g0 - g1) + (g0 * g0 - g1 * g1));
    /* C2 advection: UFx at rho i-1, i; UFe at psi (j,i), (j+1,i) */
    const double UFx1 = 0.25 * (DUon(p, z, u, W, j, i) + DUon(p, z, u, W, j, i + 1)) * (u[f] + u[f + 1]);
    const double UFx0 = 0.25 * (DUon(p, z, u, W, j, i - 1) + DUon(p, z, u, W, j, i)) * (u[f - 1] + u[f]);
    const double UFe1 = 0.25 * (DVom(p, z, v, W, j + 1, i) + DVom(p, z, v, W, j + 1, i - 1)) * (u[f + W] + u[f]);
    const double UFe0 = 0.25 * (DVom(p, z, v, W, j, i) + DVom(p, z, v, W, j, i - 1)) * (u[f] + u[f - W]);
    r = r - ((UFx1 - UFx0) + (UFe1 - UFe0));
    /* Coriolis: UFx = 0.5*D*fomn*(v(j)+v(j+1)) at rho i-1, i */
    const double c1 = ((0.5 * Dat(p, z, f)) * p->fomn[f]) * (v[f] + v[f + W]);
    const double c0 = ((0.5 * Dat(p, z, f - 1)) * p->fomn[f - 1]) * (v[f - 1] + v[f - 1 + W]);
    r = r + 0.5 * (c1 + c0);
    /* bottom drag and surface stress */
    r = r - ((p->rdrg * u[f]) * omu) * onu;
    r = r + (p->sustr[f] * omu) * onu;
    return r;
}

double sfo_phys_rhs_v(const sfo_phys *p, const double *z, const double *u, const double *v, const double *gz,
                      int j, int i)
{
    const int W = p->szI;
    const ptrdiff_t f = IDX(j, i);
    const double omv = 2.0 / (p->pm[f - W] + p->pm[f]);
    const double onv = 2.0 / (p->pn[f - W] + p->pn[f]);
    const double g0 = gz[f - W], g1 = gz[f];
    double r = ((0.5 * p->g) * omv) * ((p->h[f - W] + p->h[f]) * (g0 - g1) + (g0 * g0 - g1 * g1));
    /* C2 advection: VFx at psi (j,i), (j,i+1); VFe at rho j-1, j */
    const double VFx1 = 0.25 * (DUon(p, z, u, W, j, i + 1) + DUon(p, z, u, W, j - 1, i + 1)) * (v[f + 1] + v[f]);
    const double VFx0 = 0.25 * (DUon(p, z, u, W, j, i) + DUon(p, z, u, W, j - 1, i)) * (v[f] + v[f - 1]);
    const double VFe1 = 0.25 * (DVom(p, z, v, W, j, i) + DVom(p, z, v, W, j + 1, i)) * (v[f] + v[f + W]);
    const double VFe0 = 0.25 * (DVom(p, z, v, W, j - 1, i) + DVom(p, z, v, W, j, i)) * (v[f - W] + v[f]);
    r = r - ((VFx1 - VFx0) + (VFe1 - VFe0));
    /* Coriolis: VFe = 0.5*D*fomn*(u(i)+u(i+1)) at rho j-1, j */
    const double c1 = ((0.5 * Dat(p, z, f)) * p->fomn[f]) * (u[f] + u[f + 1]);
    const double c0 = ((0.5 * Dat(p, z, f - W)) * p->fomn[f - W]) * (u[f - W] + u[f - W + 1]);
    r = r - 0.5 * (c1 + c0);
    r = r - ((p->rdrg * v[f]) * omv) * onv;
    r = r + (p->svstr[f] * omv) * onv;
    return r;
}

/* closed-basin boundary values of one state (free-slip ghosts, zero normal flow) */
void sfo_phys_bc(const sfo_phys *p, double *z, double *u, double *v)
{
    const int W = p->szI, L = p->szI - 1, M = p->szJ - 1;
    for (int j = 1; j < M; j++) {
        z[IDX(j, 0)] = z[IDX(j, 1)];
        z[IDX(j, L)] = z[IDX(j, L - 1)];
    }
    for (int i = 0; i <= L; i++) {
        z[IDX(0, i)] = z[IDX(1, i)];
        z[IDX(M, i)] = z[IDX(M - 1, i)];
    }
    for (int j = 0; j <= M; j++) {
        u[IDX(j, 0)] = 0.0;
        u[IDX(j, 1)] = 0.0;
        u[IDX(j, L)] = 0.0;
    }
    for (int i = 2; i < L; i++) {
        u[IDX(0, i)] = u[IDX(1, i)];
        u[IDX(M, i)] = u[IDX(M - 1, i)];
    }
    for (int i = 0; i <= L; i++) {
        v[IDX(0, i)] = 0.0;
        v[IDX(1, i)] = 0.0;
        v[IDX(M, i)] = 0.0;
    }
    for (int j = 2; j < M; j++) {
        v[IDX(j, 0)] = v[IDX(j, 1)];
        v[IDX(j, L)] = v[IDX(j, L - 1)];
    }
}

/* Open boundaries (nonlinear/zetabc.py Chapman sketch, nonlinear/barotropicVelocityBC.py:66-98
 * Flather, as in ROMS zetabc/u2dbc/v2dbc): edges whose bit is set in flags (1 west,
 * 2 east, 4 south, 8 north) are open, the others stay closed walls.  After a stage has
 * written the new free surface zn (with closed-wall ghosts), every open edge's ghost
 * rho point takes the implicit Chapman value
 *   zn(g) = (zb(g) + Ce zn(i)) / (1 + Ce),  Ce = dt_s 0.5 (m(g)+m(i)) sqrt(g0 (h(g) + zb(g))),
 * (i the first interior point, zb the level the stage starts from, dt_s its time span,
 * m = pm across west/east edges and pn across south/north ones).  After the stage's
 * momentum, the normal velocity on the edge's face takes the Flather value from the new
 * surface on both sides of the face,
 *   west/south: u = ext_u - sqrt(g0/D) (0.5 (zn(g)+zn(i)) - ext_z),  east/north: + sqrt(g0/D)(...),
 * D = 0.5 ((h(g)+zn(g)) + (h(i)+zn(i))), with its two ghost-row copies (zero gradient
 * along the edge).  ext_z / ext_u: external free surface and normal velocity per edge
 * (west/east indexed by j, south/north by i; NULL = 0). */
typedef struct sfo_open {
    int flags;
    int volcons;
    const double *ext_z[4], *ext_u[4];
    double *out;    /* area, flux, ubar_xs of the corrector's volume constraint */
} sfo_open;

static inline double ext_at(const double *a, int k) { return a ? a[k] : 0.0; }

static void open_zeta(const sfo_phys *p, const sfo_open *ob, double *zn, const double *zb, double dts)
{
    const int W = p->szI, L = p->szI - 1, M = p->szJ - 1;
    for (int e = 0; e < 4; e++) {
        if (!(ob->flags & (1 << e))) continue;
        const int n = e < 2 ? M - 1 : L - 1;
        for (int b = 1; b <= n; b++) {
            ptrdiff_t fg, fi;
            if (e == 0) { fg = IDX(b, 0); fi = IDX(b, 1); }
            else if (e == 1) { fg = IDX(b, L); fi = IDX(b, L - 1); }
            else if (e == 2) { fg = IDX(0, b); fi = IDX(1, b); }
            else { fg = IDX(M, b); fi = IDX(M - 1, b); }
            const double m = e < 2 ? 0.5 * (p->pm[fg] + p->pm[fi]) : 0.5 * (p->pn[fg] + p->pn[fi]);
            const double Ce = (dts * m) * sqrt(p->g * (p->h[fg] + zb[fg]));
            zn[fg] = (zb[fg] + Ce * zn[fi]) / (1.0 + Ce);
        }
    }
}

static void open_vel(const sfo_phys *p, const sfo_open *ob, const double *zn, double *un, double *vn)
{
    const int W = p->szI, L = p->szI - 1, M = p->szJ - 1;
    for (int e = 0; e < 4; e++) {
        if (!(ob->flags & (1 << e))) continue;
        const int n = e < 2 ? M - 1 : L - 1;
        double *x = e < 2 ? un : vn;
        for (int b = 1; b <= n; b++) {
            ptrdiff_t fg, fi, fu;
            if (e == 0) { fg = IDX(b, 0); fi = IDX(b, 1); fu = IDX(b, 1); }
            else if (e == 1) { fg = IDX(b, L); fi = IDX(b, L - 1); fu = IDX(b, L); }
            else if (e == 2) { fg = IDX(0, b); fi = IDX(1, b); fu = IDX(1, b); }
            else { fg = IDX(M, b); fi = IDX(M - 1, b); fu = IDX(M, b); }
            const double D = 0.5 * ((p->h[fg] + zn[fg]) + (p->h[fi] + zn[fi]));
            const double c = sqrt(p->g / D);
            const double dz = 0.5 * (zn[fg] + zn[fi]) - ext_at(ob->ext_z[e], b);
            x[fu] = (e == 0 || e == 2) ? ext_at(ob->ext_u[e], b) - c * dz : ext_at(ob->ext_u[e], b) + c * dz;
        }
        /* ghost rows / columns of the edge's face: zero gradient along the edge */
        if (e == 0) { un[IDX(0, 1)] = un[IDX(1, 1)]; un[IDX(M, 1)] = un[IDX(M - 1, 1)]; }
        else if (e == 1) { un[IDX(0, L)] = un[IDX(1, L)]; un[IDX(M, L)] = un[IDX(M - 1, L)]; }
        else if (e == 2) { vn[IDX(1, 0)] = vn[IDX(1, 1)]; vn[IDX(1, L)] = vn[IDX(1, L - 1)]; }
        else { vn[IDX(M, 0)] = vn[IDX(M, 1)]; vn[IDX(M, L)] = vn[IDX(M, L - 1)]; }
    }
}

/* Coupling of a substep with the baroclinic step (ROMS split-explicit form,
 * obsolete/step2d_LF_AM3.h.py:2377-2590 and :778-846):
 *  - the first substep of a baroclinic step (first != 0) turns the vertical integral
 *    of the 3D right-hand side (rufrc) into the forcing F = rufrc - rhs_ubar of its
 *    predictor, stores F, and adds the Adams-Bashforth extrapolation of F over the
 *    baroclinic steps to that predictor (ab = 0: F; 1: 1.5 F - 0.5 F1;
 *    2: 23/12 F - 16/12 F1 + 5/12 F2, F1/F2 the forcing of the two previous steps);
 *  - every other stage adds F;
 *  - DU_avg2/DV_avg2 accumulate the fluxes of the predictor's state (weight w2p;
 *    the first substep sets them) and of the predicted state (weight w2c). */
typedef struct sfo_couple {
    const double *rufrc, *rvfrc;
    double *Fu, *Fv;
    const double *Fu1, *Fv1, *Fu2, *Fv2;
    double *DU2, *DV2;
    double w2p, w2c;
    int ab;
} sfo_couple;

static double ab_forcing(int ab, double F, double F1, double F2)
{
    if (ab == 0) return F;
    if (ab == 1) return 1.5 * F - 0.5 * F1;
    return ((23.0 / 12.0) * F - (16.0 / 12.0) * F1) + (5.0 / 12.0) * F2;
}

static void phys_substep(const sfo_phys *p, const sfo_phys *pc, double dt, int first, double cff1w,
                         const double *z0, const double *z1, double *z2, const double *u0, const double *u1,
                         double *u2, const double *v0, const double *v1, double *v2, const double *rz0, double *rz1,
                         const double *ru0, double *ru1, const double *rv0, double *rv1, double *zs, double *us,
                         double *vs, double *gz, double *Zt, double *DUa, double *DVa, const sfo_couple *cp,
                         const sfo_open *ob);

/* One substep.  Slots after the caller's rotation: index 0 = n-1, 1 = n, 2 = new.
 * zs/us/vs: predicted-state scratch (S2 each).  first != 0: forward-Euler
 * predictor and backward-Euler corrector (no history).  Writes z2/u2/v2, the
 * predictor RHS into rz1/ru1/rv1, and accumulates the averages. */
void sfo_phys_substep(const sfo_phys *p, double dt, int first, double cff1w,
                      const double *z0, const double *z1, double *z2, const double *u0, const double *u1,
                      double *u2, const double *v0, const double *v1, double *v2, const double *rz0, double *rz1,
                      const double *ru0, double *ru1, const double *rv0, double *rv1, double *zs, double *us,
                      double *vs, double *gz, double *Zt, double *DUa, double *DVa)
{
    phys_substep(p, p, dt, first, cff1w, z0, z1, z2, u0, u1, u2, v0, v1, v2, rz0, rz1, ru0, ru1, rv0, rv1, zs, us,
                 vs, gz, Zt, DUa, DVa, NULL, NULL);
}

/* The same substep coupled with the baroclinic step (sfo_couple above). */
void sfo_phys_substep_coupled(const sfo_phys *p, double dt, int first, double cff1w,
                              const double *z0, const double *z1, double *z2, const double *u0, const double *u1,
                              double *u2, const double *v0, const double *v1, double *v2, const double *rz0,
                              double *rz1, const double *ru0, double *ru1, const double *rv0, double *rv1,
                              double *zs, double *us, double *vs, double *gz, double *Zt, double *DUa, double *DVa,
                              const sfo_couple *cp)
{
    phys_substep(p, p, dt, first, cff1w, z0, z1, z2, u0, u1, u2, v0, v1, v2, rz0, rz1, ru0, ru1, rv0, rv1, zs, us,
                 vs, gz, Zt, DUa, DVa, cp, NULL);
}

/* The substep with the corrector stage's surface stress taken from csustr / csvstr
 * (the predictor keeps p->sustr / p->svstr): the form in which the learned closure's
 * stress enters the coarse substep. */
void sfo_phys_substep_cstress(const sfo_phys *p, const double *csustr, const double *csvstr, double dt, int first,
                              double cff1w, const double *z0, const double *z1, double *z2, const double *u0,
                              const double *u1, double *u2, const double *v0, const double *v1, double *v2,
                              const double *rz0, double *rz1, const double *ru0, double *ru1, const double *rv0,
                              double *rv1, double *zs, double *us, double *vs, double *gz, double *Zt, double *DUa,
                              double *DVa)
{
    sfo_phys pc = *p;
    pc.sustr = csustr;
    pc.svstr = csvstr;
    phys_substep(p, &pc, dt, first, cff1w, z0, z1, z2, u0, u1, u2, v0, v1, v2, rz0, rz1, ru0, ru1, rv0, rv1, zs, us,
                 vs, gz, Zt, DUa, DVa, NULL, NULL);
}

void sfo_obc_volcons(int szI, int szJ, const double *z, const double *h, const double *pm, const double *pn,
                     double *u, double *v, int flags, int apply, double *out);

/* The substep with open edges (sfo_open above); volcons != 0 applies obc_volcons +
 * conserve_mass (sfo_obc_volcons) to the velocities of each stage after their open-edge
 * values (the predicted state, then the new one), so every flux the free-surface
 * updates use has no net flow through the open edges and the volume is conserved. */
void sfo_phys_substep_open(const sfo_phys *p, double dt, int first, double cff1w,
                           const double *z0, const double *z1, double *z2, const double *u0, const double *u1,
                           double *u2, const double *v0, const double *v1, double *v2, const double *rz0,
                           double *rz1, const double *ru0, double *ru1, const double *rv0, double *rv1,
                           double *zs, double *us, double *vs, double *gz, double *Zt, double *DUa, double *DVa,
                           const sfo_open *ob, double *volcons_out)
{
    sfo_open o = *ob;
    o.out = volcons_out;
    phys_substep(p, p, dt, first, cff1w, z0, z1, z2, u0, u1, u2, v0, v1, v2, rz0, rz1, ru0, ru1, rv0, rv1, zs, us,
                 vs, gz, Zt, DUa, DVa, NULL, &o);
}

/* pc: the parameters of the corrector's momentum right-hand sides (p itself, or p with
 * the corrector-stage surface stress of sfo_phys_substep_cstress) */
static void phys_substep(const sfo_phys *p, const sfo_phys *pc, double dt, int first, double cff1w,
                         const double *z0, const double *z1, double *z2, const double *u0, const double *u1,
                         double *u2, const double *v0, const double *v1, double *v2, const double *rz0, double *rz1,
                         const double *ru0, double *ru1, const double *rv0, double *rv1, double *zs, double *us,
                         double *vs, double *gz, double *Zt, double *DUa, double *DVa, const sfo_couple *cp,
                         const sfo_open *ob)
{
    const int W = p->szI, L = p->szI - 1, M = p->szJ - 1;
    const ptrdiff_t S2 = (ptrdiff_t)W * p->szJ;
    /* ---- predictor: zeta ---- */
    const double cff4 = 4.0 / 25.0, cff5 = 1.0 - 2.0 * cff4;
    memcpy(zs, z1, sizeof(double) * S2);
#pragma omp parallel for schedule(static)
    for (int j = 1; j < M; j++)
        for (int i = 1; i < L; i++) {
            const ptrdiff_t f = IDX(j, i);
            const double r = sfo_phys_rhs_zeta(p, z1, u1, v1, j, i);
            rz1[f] = r;
            if (first) zs[f] = z1[f] + ((p->pm[f] * p->pn[f]) * dt) * r;
            else zs[f] = z0[f] + ((p->pm[f] * p->pn[f]) * (2.0 * dt)) * r;
            if (cp) {   /* DU_avg2/DV_avg2: fluxes of the predictor's state */
                const double a = first ? 0.0 : cp->DU2[f], b = first ? 0.0 : cp->DV2[f];
                cp->DU2[f] = a + cp->w2p * DUon(p, z1, u1, W, j, i);
                cp->DV2[f] = b + cp->w2p * DVom(p, z1, v1, W, j, i);
            }
        }
    memcpy(us, u1, sizeof(double) * S2);
    memcpy(vs, v1, sizeof(double) * S2);
    sfo_phys_bc(p, zs, us, vs);   /* zeta ghosts of the predicted surface */
    if (ob) open_zeta(p, ob, zs, first ? z1 : z0, first ? dt : 2.0 * dt);
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++)
        gz[f] = first ? 0.5 * (z1[f] + zs[f]) : cff5 * z1[f] + cff4 * (z0[f] + zs[f]);
    /* ---- predictor: momentum ---- */
    const double *zb = first ? z1 : z0, *ub = first ? u1 : u0, *vb = first ? v1 : v0;
    const double c1 = first ? 0.5 * dt : dt;
#pragma omp parallel for schedule(static)
    for (int j = 1; j < M; j++)
        for (int i = 2; i < L; i++) {
            const ptrdiff_t f = IDX(j, i);
            double r = sfo_phys_rhs_u(p, z1, u1, v1, gz, j, i);
            if (cp) {
                if (first) {
                    const double F = cp->rufrc[f] - r;
                    cp->Fu[f] = F;
                    r = r + ab_forcing(cp->ab, F, cp->Fu1[f], cp->Fu2[f]);
                } else {
                    r = r + cp->Fu[f];
                }
            }
            ru1[f] = r;
            const double cff = (p->pm[f] + p->pm[f - 1]) * (p->pn[f] + p->pn[f - 1]);
            const double fac = 1.0 / (Dat(p, zs, f) + Dat(p, zs, f - 1));
            us[f] = (ub[f] * (Dat(p, zb, f) + Dat(p, zb, f - 1)) + (cff * c1) * r) * fac;
        }
#pragma omp parallel for schedule(static)
    for (int j = 2; j < M; j++)
        for (int i = 1; i < L; i++) {
            const ptrdiff_t f = IDX(j, i);
            double r = sfo_phys_rhs_v(p, z1, u1, v1, gz, j, i);
            if (cp) {
                if (first) {
                    const double F = cp->rvfrc[f] - r;
                    cp->Fv[f] = F;
                    r = r + ab_forcing(cp->ab, F, cp->Fv1[f], cp->Fv2[f]);
                } else {
                    r = r + cp->Fv[f];
                }
            }
            rv1[f] = r;
            const double cff = (p->pm[f] + p->pm[f - W]) * (p->pn[f] + p->pn[f - W]);
            const double fac = 1.0 / (Dat(p, zs, f) + Dat(p, zs, f - W));
            vs[f] = (vb[f] * (Dat(p, zb, f) + Dat(p, zb, f - W)) + (cff * c1) * r) * fac;
        }
    sfo_phys_bc(p, zs, us, vs);
    if (ob) {
        /* the closed-wall pass above rewrote the ghost surface: restore the open edges' */
        open_zeta(p, ob, zs, first ? z1 : z0, first ? dt : 2.0 * dt);
        open_vel(p, ob, zs, us, vs);
        double tmp[3];
        if (ob->volcons) sfo_obc_volcons(p->szI, p->szJ, zs, p->h, p->pm, p->pn, us, vs, ob->flags, 1, tmp);
    }
    /* ---- corrector: zeta ---- */
    const double a1 = dt * 5.0 / 12.0, a2 = dt * 8.0 / 12.0, a3 = dt * 1.0 / 12.0;
    memcpy(z2, z1, sizeof(double) * S2);
#pragma omp parallel for schedule(static)
    for (int j = 1; j < M; j++)
        for (int i = 1; i < L; i++) {
            const ptrdiff_t f = IDX(j, i);
            const double r = sfo_phys_rhs_zeta(p, zs, us, vs, j, i);
            const double pmn = p->pm[f] * p->pn[f];
            if (first) z2[f] = z1[f] + pmn * (dt * r);
            else z2[f] = z1[f] + pmn * ((a1 * r + a2 * rz1[f]) - a3 * rz0[f]);
            /* averages of the predicted state (live corrector rule) */
            Zt[f] = Zt[f] + cff1w * zs[f];
            DUa[f] = DUa[f] + cff1w * DUon(p, zs, us, W, j, i);
            DVa[f] = DVa[f] + cff1w * DVom(p, zs, vs, W, j, i);
            if (cp) {   /* DU_avg2/DV_avg2: fluxes of the predicted state */
                cp->DU2[f] = cp->DU2[f] + cp->w2c * DUon(p, zs, us, W, j, i);
                cp->DV2[f] = cp->DV2[f] + cp->w2c * DVom(p, zs, vs, W, j, i);
            }
        }
    memcpy(u2, u1, sizeof(double) * S2);
    memcpy(v2, v1, sizeof(double) * S2);
    sfo_phys_bc(p, z2, u2, v2);
    if (ob) open_zeta(p, ob, z2, z1, dt);
    const double cf4 = 2.0 / 5.0, cf5 = 1.0 - cf4;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S2; f++) gz[f] = cf5 * z2[f] + cf4 * zs[f];
    /* ---- corrector: momentum ---- */
    const double b1 = 0.5 * dt * 5.0 / 12.0, b2 = 0.5 * dt * 8.0 / 12.0, b3 = 0.5 * dt * 1.0 / 12.0;
#pragma omp parallel for schedule(static)
    for (int j = 1; j < M; j++)
        for (int i = 2; i < L; i++) {
            const ptrdiff_t f = IDX(j, i);
            double r = sfo_phys_rhs_u(pc, zs, us, vs, gz, j, i);
            if (cp) r = r + cp->Fu[f];
            const double cff = (p->pm[f] + p->pm[f - 1]) * (p->pn[f] + p->pn[f - 1]);
            const double fac = 1.0 / (Dat(p, z2, f) + Dat(p, z2, f - 1));
            const double tend = first ? (0.5 * dt) * r : (b1 * r + b2 * ru1[f]) - b3 * ru0[f];
            u2[f] = (u1[f] * (Dat(p, z1, f) + Dat(p, z1, f - 1)) + cff * tend) * fac;
        }
#pragma omp parallel for schedule(static)
    for (int j = 2; j < M; j++)
        for (int i = 1; i < L; i++) {
            const ptrdiff_t f = IDX(j, i);
            double r = sfo_phys_rhs_v(pc, zs, us, vs, gz, j, i);
            if (cp) r = r + cp->Fv[f];
            const double cff = (p->pm[f] + p->pm[f - W]) * (p->pn[f] + p->pn[f - W]);
            const double fac = 1.0 / (Dat(p, z2, f) + Dat(p, z2, f - W));
            const double tend = first ? (0.5 * dt) * r : (b1 * r + b2 * rv1[f]) - b3 * rv0[f];
            v2[f] = (v1[f] * (Dat(p, z1, f) + Dat(p, z1, f - W)) + cff * tend) * fac;
        }
    sfo_phys_bc(p, z2, u2, v2);
    if (ob) {
        open_zeta(p, ob, z2, z1, dt);
        open_vel(p, ob, z2, u2, v2);
        if (ob->volcons) sfo_obc_volcons(p->szI, p->szJ, z2, p->h, p->pm, p->pn, u2, v2, ob->flags, 1, ob->out);
    }
}

/* obc_volcons + conserve_mass (nonlinear/obc_volcons.py:1-72) in FP64: sequential sums
 * over the open segments in the order west, east, south, north (points ascending).
 * flags: 1 west, 2 east, 4 south, 8 north.  out[0] area, out[1] flux, out[2] ubar_xs.
 * apply != 0 corrects the boundary velocities in place. */
void sfo_obc_volcons(int szI, int szJ, const double *z, const double *h, const double *pm, const double *pn,
                     double *u, double *v, int flags, int apply, double *out)
{
    const int W = szI, L = szI - 1, M = szJ - 1;
    double area = 0.0, flux = 0.0;
    for (int seg = 0; seg < 4; seg++) {
        if (!(flags & (1 << seg))) continue;
        const int n = seg < 2 ? M - 1 : L - 1;
        for (int b = 0; b < n; b++) {
            double cff, fl;
            if (seg < 2) {
                const int j = 1 + b, i = seg == 0 ? 1 : L;
                const ptrdiff_t f = (ptrdiff_t)j * W + i;
                const double onu = 2.0 / (pn[f - 1] + pn[f]);
                cff = (0.5 * ((z[f - 1] + h[f - 1]) + (z[f] + h[f]))) * onu;
                fl = seg == 0 ? cff * u[f] : -(cff * u[f]);
            } else {
                const int i = 1 + b, j = seg == 2 ? 1 : M;
                const ptrdiff_t f = (ptrdiff_t)j * W + i;
                const double omv = 2.0 / (pm[f - W] + pm[f]);
                cff = (0.5 * ((z[f - W] + h[f - W]) + (z[f] + h[f]))) * omv;
                fl = seg == 2 ? cff * v[f] : -(cff * v[f]);
            }
            area += cff;
            flux += fl;
        }
    }
    out[0] = area;
    out[1] = flux;
    out[2] = area > 0.0 ? flux / area : 0.0;
    if (!apply) return;
    const double xs = out[2];
    if (flags & 1)
        for (int j = 1; j < M; j++) u[(ptrdiff_t)j * W + 1] -= xs;
    if (flags & 2)
        for (int j = 1; j < M; j++) u[(ptrdiff_t)j * W + L] += xs;
    if (flags & 4)
        for (int i = 1; i < L; i++) v[(ptrdiff_t)W + i] -= xs;
    if (flags & 8)
        for (int i = 1; i < L; i++) v[(ptrdiff_t)M * W + i] += xs;
}
