/*
 * sailfish_hip.h -- C ABI of the MI355X (gfx950) split-explicit hot path.
 *
 * Every entry point takes device pointers, plain sizes/scalars and an explicit
 * hipStream_t (passed as void* so that this header needs no HIP include).  Nothing
 * synchronises the stream.  Return value: SF_OK (0) or a negative SF_ERR_* code;
 * sf_last_error_string() describes the last failure of the calling thread.
 *
 * Each declaration names the reference interface it replaces (path:line in the
 * Sailfish sources).  The reference passes grid sizes through device globals set
 * once by `initialize` (modules/mod_cppkernels.h:154-171); here they live in an
 * sf_ctx, so any number of grids (ensembles) can coexist in one process.
 */
#ifndef SAILFISH_HIP_H
#define SAILFISH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SF_OK = 0,
    SF_ERR_ARG = -1,          /* invalid argument (null pointer, size, range)       */
    SF_ERR_HIP = -2,          /* a HIP runtime call failed                           */
    SF_ERR_UNSUPPORTED = -3,  /* valid input that this entry point does not handle   */
    SF_ERR_NOMEM = -4
};

enum { SF_FP64 = 0, SF_FP32 = 1 };

/* sf_grid_desc.flags */
enum {
    /* computeMomentumPred masks with i%szJ (modules/mod_cppkernels.cpp:162); on grids
     * with szI != szJ that freezes interior columns.  With this flag the mask is
     * i%szI, i.e. exactly column 0 and row 0 (what the square-grid case does). */
    SF_FLAG_FIX_Q5_MASK = 1
};

typedef struct sf_ctx sf_ctx;

/* Horizontal/vertical sizes and the metric fields the reference keeps as device
 * globals (_on_u, _om_v, _pn, _pm: mod_cppkernels.h:40-41).  szI = L+1 (xi, fastest),
 * szJ = M+1, szK = N.  Metric pointers are FP64 device arrays of szI*szJ. */
typedef struct sf_grid_desc {
    int32_t szI, szJ, szK;
    int32_t flags;            /* SF_FLAG_* */
    const double *on_u, *om_v, *pn, *pm;
} sf_grid_desc;

/* Replaces `initialize` (mod_cppkernels.h:154-171) + `initModule`
 * (modules/mod_operators.py:31-45). */
int sf_ctx_create(const sf_grid_desc *desc, sf_ctx **out);
int sf_ctx_destroy(sf_ctx *ctx);
const char *sf_last_error_string(void);
/* Library/ABI version: major*10000 + minor*100 + patch. */
int sf_version(void);

/* ---------------- barotropic LF-AM3 kernels, one to one, FP64 ----------------- */

/* computeZetaRHS (modules/mod_cppkernels.cpp:80-128).  cff1 = weight1[iif-1]
 * (read on the host instead of from device memory). */
int sf_compute_zeta_rhs(sf_ctx *ctx, const double *zeta, const double *h,
                        const double *ubar, const double *vbar, double *DU_avg1,
                        double *DV_avg1, double *Zt_avg1, int32_t is_predictor,
                        double cff1, double *res, void *stream);

/* computeZetaPred (modules/mod_cppkernels.cpp:133-150). */
int sf_compute_zeta_pred(sf_ctx *ctx, double Dt, const double *zeta_t0,
                         const double *zeta_t1, double *zeta_t2, const double *rzeta_t1,
                         void *stream);

/* computeMomentumRHSPred (modules/mod_cppkernels.cpp:12-42). */
int sf_compute_momentum_rhs_pred(sf_ctx *ctx, const double *h, double *rhs_ubar,
                                 double *rhs_vbar, const double *zeta_t0,
                                 const double *zeta_t2, double g, double weight,
                                 void *stream);

/* computeMomentumRHSCorr (modules/mod_cppkernels.cpp:45-77). */
int sf_compute_momentum_rhs_corr(sf_ctx *ctx, const double *h, double *rhs_ubar,
                                 double *rhs_vbar, const double *zeta_t0,
                                 const double *zeta_t1, const double *zeta_t2, double g,
                                 void *stream);

/* computeMomentumPred (modules/mod_cppkernels.cpp:155-186). */
int sf_compute_momentum_pred(sf_ctx *ctx, double Dt, const double *u_t1, double *u_t2,
                             const double *v_t1, double *v_t2, const double *rhsu,
                             const double *rhsv, const double *h, const double *zeta_t1,
                             const double *zeta_t2, void *stream);

/* AdamsMoultonCorr3rd (modules/mod_cppkernels.cpp:190-209). */
int sf_adams_moulton_corr3rd(sf_ctx *ctx, double Dt, double *v_t2, const double *rhs_t0,
                             const double *rhs_t1, const double *rhs_t2, void *stream);

/* AdamsMoultonCorr3rd2 (modules/mod_cppkernels.cpp:213-251). */
int sf_adams_moulton_corr3rd2(sf_ctx *ctx, double Dt, double *u_t2, double *v_t2,
                              const double *rhsu_t0, const double *rhsu_t1,
                              const double *rhsu_t2, const double *rhsv_t0,
                              const double *rhsv_t1, const double *rhsv_t2,
                              const double *h, const double *zeta_t1,
                              const double *zeta_t2, void *stream);

/* ---------------- lateral boundary conditions ---------------------------------- */

/* setBC (modules/mod_BC.c:6-13): var[idx[k]] = val. */
int sf_bc_set(double *var, const int32_t *idx, int32_t n, double val, void *stream);

/* A composed boundary map: the net effect of an ordered list of setBC/copyBC calls
 * (modules/mod_BC.c:6-23, issued by nonlinear/zetabc.py:5-12 and
 * nonlinear/barotropicVelocityBC.py:7-31) run in list order.  For each k:
 * src[k] >= 0 -> var[dst[k]] = pre(var[src[k]]); -1 -> 0.0; -2 -> clamp_val.
 * dst[] entries are unique; every source is read before any destination is written.
 * n <= 16384. */
int sf_bc_apply_map(double *var, const int32_t *dst, const int32_t *src, int32_t n,
                    double clamp_val, void *stream);

/* ---------------- 3D kernels, one to one, FP64 ------------------------------------
 * 3D arrays are flat (level, eta, xi) with level stride szI*szJ: u, v, ru, rv, W,
 * Huon, Hvom, Akv, z_w have szK+1 levels, Hz and z_r have szK.  szK >= 4. */

/* addCoriolis (nonlinear/rhs_kernels.cpp:6-35). */
int sf_add_coriolis(sf_ctx *ctx, const double *fomn, const double *u, const double *v,
                    double *ru, double *rv, const double *Hz, void *stream);

/* horizontalAdvection (nonlinear/rhs_kernels.cpp:70-132).  launch_threads is the
 * thread count of the reference launch ((S2/512+1)*512, nonlinear/rhs.py:62): its
 * threads past S2 store 0 into ru/rv at their flat index, which is reproduced after
 * the columns are done. */
int sf_horizontal_advection(sf_ctx *ctx, const double *u, const double *v, const double *Huon,
                            const double *Hvom, double *ru, double *rv, int64_t launch_threads,
                            void *stream);

/* verticalAdvection (nonlinear/rhs_kernels.cpp:135-197); BC = zeta bcIdxFieldIdx2. */
int sf_vertical_advection(sf_ctx *ctx, const double *u, const double *v, const double *W,
                          double *ru, double *rv, const int32_t *BC, void *stream);

/* rhs3d as ONE launch (nonlinear/rhs.py:38-71 with rhs_kernels.cpp:6-197): ru, rv
 * := 0 over all szK+1 levels, then addCoriolis (cor != 0), horizontalAdvection with
 * the zeros of its threads past S2 (launch_threads as for sf_horizontal_advection) and
 * verticalAdvection (adv != 0), each element rounded in the same sequence as by the
 * four separate launches, so the result equals them bit for bit. */
int sf_rhs3d(sf_ctx *ctx, const double *fomn, const double *u, const double *v, const double *Huon,
             const double *Hvom, const double *W, const double *Hz, const int32_t *BC, double *ru, double *rv,
             int64_t launch_threads, int32_t cor, int32_t adv, void *stream);

/* set_maxflux (nonlinear/step3D_uv.kernels.cpp:372-405). */
int sf_set_maxflux(sf_ctx *ctx, const double *u, const double *v, double *Huon, double *Hvom,
                   const double *Hz, void *stream);

/* omega (nonlinear/step3D_uv.kernels.cpp:62-165); BC = zeta bcIdxFieldIdx2. */
int sf_omega(sf_ctx *ctx, double *W, const double *Huon, const double *Hvom, const double *z_w,
             const int32_t *BC, void *stream);

/* FP32 bulk forms of the three kernels above (the same call sites:
 * nonlinear/rhs.py:38-71, nonlinear/main3d.py:141,321 and step3D_uv.kernels.cpp:62-165,
 * 372-405): the 3D fields in FP32, the 2D grid metrics (fomn and the context's on_u,
 * om_v) in FP64, stencil arithmetic in FP32.  sf_omega_f32 forms the column sums and the
 * surface correction in FP64 and needs szK <= 64.  Error against the FP64 forms is
 * stated with the tests that check it. */
int sf_rhs3d_f32(sf_ctx *ctx, const double *fomn, const float *u, const float *v, const float *Huon,
                 const float *Hvom, const float *W, const float *Hz, const int32_t *BC, float *ru, float *rv,
                 int64_t launch_threads, int32_t cor, int32_t adv, void *stream);
int sf_set_maxflux_f32(sf_ctx *ctx, const float *u, const float *v, float *Huon, float *Hvom, const float *Hz,
                       void *stream);
int sf_omega_f32(sf_ctx *ctx, float *W, const float *Huon, const float *Hvom, const float *z_w,
                 const int32_t *BC, void *stream);

/* set_depth (nonlinear/step3D_uv.kernels.cpp:225-291). */
int sf_set_depth(sf_ctx *ctx, int32_t Vtransform, const double *Zt_avg1, double *z_w, double *z_r,
                 const double *h, double hc, double *Hz, const double *sc_r, const double *sc_w,
                 const double *Cs_r, const double *Cs_w, void *stream);

/* set_zeta (nonlinear/step3D_uv.kernels.cpp:294-310). */
int sf_set_zeta(sf_ctx *ctx, double *zeta_t1, double *zeta_t2, const double *Zt_avg1, void *stream);

/* setLateralUVBCs (nonlinear/step3D_uv.kernels.cpp:314-368) for closed/clamped walls
 * (uType/vType = bcIdxFieldType of the u and v index sets). */
int sf_set_lateral_uv_bcs(sf_ctx *ctx, double t, double *u, double *v, const int32_t *uType,
                          const int32_t *vType, void *stream);

/* FP32 bulk forms of set_depth, set_zeta and setLateralUVBCs (same call sites,
 * nonlinear/main3d.py:62-63,151,293,309): z_w, z_r, Hz, zeta and u, v in FP32; Zt_avg1,
 * h and the s-coordinate curves in FP64.  Every depth is formed in FP64 as by
 * sf_set_depth (Hz as the FP64 difference of the FP64 z_w) and rounded once. */
int sf_set_depth_f32(sf_ctx *ctx, int32_t Vtransform, const double *Zt_avg1, float *z_w, float *z_r,
                     const double *h, double hc, float *Hz, const double *sc_r, const double *sc_w,
                     const double *Cs_r, const double *Cs_w, void *stream);
int sf_set_zeta_f32(sf_ctx *ctx, float *zeta_t1, float *zeta_t2, const double *Zt_avg1, void *stream);
int sf_set_lateral_uv_bcs_f32(sf_ctx *ctx, double t, float *u, float *v, const int32_t *uType,
                              const int32_t *vType, void *stream);

/* step3d_UV (nonlinear/step3D_uv.kernels.cpp:646-707) with its scratch arrays SD, D,
 * RHS ((szK+1)*S2 doubles each, nonlinear/main3d.py:48-58).  sf_step3d_uv_f32 (declared
 * below) is its FP32 bulk form for the same call site (main3d.py:304-306).  With at
 * most SF_STEP3D_REG_LEVELS levels each column's system stays in registers (the scratch
 * is not touched) and tmpU / tmpV may be u / v themselves, so the copies main3d.py:48-52
 * makes before the call are not needed; deeper grids use the scratch and need the copies. */
#define SF_STEP3D_REG_LEVELS 32
int sf_step3d_uv(sf_ctx *ctx, double *u, double *v, const double *ru, const double *rv,
                 double *ubar_t1, double *vbar_t1, double *ubar_t2, double *vbar_t2,
                 const double *Hz, const double *Akv, const double *z_r, const double *DU_avg1,
                 const double *DV_avg1, const double *tmpU, const double *tmpV, int32_t iic,
                 int32_t ntfirst, double lambda, double Dt, double *SD, double *D, double *RHS,
                 void *stream);

/* FP32 bulk form of sf_step3d_uv: u, v, ru, rv, the 2D ubar/vbar, Hz, Akv, z_r, tmpU,
 * tmpV and the scratch in FP32; DU_avg1/DV_avg1 (barotropic averages) in FP64; the
 * column sums (depth means, barotropic velocities) formed in FP64. */
int sf_step3d_uv_f32(sf_ctx *ctx, float *u, float *v, const float *ru, const float *rv, float *ubar_t1,
                     float *vbar_t1, float *ubar_t2, float *vbar_t2, const float *Hz, const float *Akv,
                     const float *z_r, const double *DU_avg1, const double *DV_avg1, const float *tmpU,
                     const float *tmpV, int32_t iic, int32_t ntfirst, double lambda, double Dt, float *SD, float *D,
                     float *RHS, void *stream);

/* ---------------- fused barotropic substep --------------------------------------- */

/* Pointers of one substep, after the caller's time-level rotation (cycleTimes2D,
 * modules/mod_ocean.py:120-155): index k of each 3-slot array is time level t_k.
 * Bulk fields have the precision passed to sf_step2d_substep (double or float);
 * the averages Zt/DU/DV_avg1 are always double.  Slots the substep never touches
 * may be null: ubar[0], vbar[0], rubar[2], rvbar[0], rvbar[2]. */
typedef struct sf_step2d_fields {
    const void *h, *on_u, *om_v, *pm, *pn;
    void *zeta[3], *ubar[3], *vbar[3];
    void *rzeta[3], *rubar[3], *rvbar[3];
    double *Zt_avg1, *DU_avg1, *DV_avg1;
} sf_step2d_fields;

/* sf_step2d_substep flags */
enum {
    SF_SUBSTEP_ZERO_VBAR = 1   /* vbar_t2 := 0 after the corrector (nonlinear/main3d.py:189) */
};

/* Boundary maps of the fused substep, as composed by the host from the BC lists of
 * nonlinear/zetabc.py:5-12 and nonlinear/barotropicVelocityBC.py:7-31 (same encoding
 * as sf_bc_apply_map; host arrays).  Rejects (SF_ERR_UNSUPPORTED) maps the fused
 * kernel cannot apply inside a tile: zeta sources must be neighbours in the same
 * tile; velocity maps hold constants (closed/clamped walls) or copies (gradient walls)
 * from a neighbour that the map leaves unchanged and that lies in the same tile region
 * as the destination.  (The reference's vbar gradient copy, which reads ubar's source
 * list at barotropicVelocityBC.py:31, is rejected: its sources are not neighbours.) */
int sf_ctx_set_bc_maps(sf_ctx *ctx, const int32_t *zeta_dst, const int32_t *zeta_src, int32_t nz,
                       const int32_t *ubar_dst, const int32_t *ubar_src, int32_t nu,
                       const int32_t *vbar_dst, const int32_t *vbar_src, int32_t nv);

/* Host-only check (no GPU work) of composed boundary maps against the fused
 * substep's tiling, for a szI x szJ grid whose computeMomentumPred mask stride is
 * q5_stride.  SF_OK if sf_ctx_set_bc_maps would accept them. */
int sf_step2d_check_maps(int32_t szI, int32_t szJ, int32_t q5_stride,
                         const int32_t *zeta_dst, const int32_t *zeta_src, int32_t nz,
                         const int32_t *ubar_dst, const int32_t *ubar_src, int32_t nu,
                         const int32_t *vbar_dst, const int32_t *vbar_src, int32_t nv);

/* One barotropic substep: step2dPredictor + step2dCorrector
 * (nonlinear/step2d.py:23-127) in one fused kernel plus a small kernel over the
 * points that computeMomentumPred masks.  precision: SF_FP64 or SF_FP32.
 * cff1 = weight1[iif-1]; clamp_val = value of clamped velocity boundaries. */
int sf_step2d_substep(sf_ctx *ctx, const sf_step2d_fields *fields, int32_t precision,
                      double dtfast, double cff1, double clamp_val, int32_t flags, void *stream);

/* The two launches of sf_step2d_substep issued separately (parts: 1 = tiled kernel,
 * 2 = masked-points kernel, 3 = both), so that a profiler or benchmark can put
 * events between them.  Issuing 1 then 2 equals sf_step2d_substep. */
int sf_step2d_substep_parts(sf_ctx *ctx, const sf_step2d_fields *fields, int32_t precision,
                            double dtfast, double cff1, double clamp_val, int32_t flags,
                            int32_t parts, void *stream);

/* Tile shape (xi, eta) of the fused substep kernel. */
int sf_step2d_tile_shape(int32_t *ti, int32_t *tj);

/* ---------------- physics-mode barotropic substep -------------------------------- */

/* Fields of one physics-mode substep (ROMS LF-AM3 in 2D form with C2 advection,
 * Coriolis, wind stress and linear bottom drag, as written in
 * obsolete/step2d_LF_AM3.h.py:917-2780; closed basin).  Bulk arrays have the
 * precision passed to sf_phys_substep; the averages are double.  Slot k of each
 * state array is time level t_k after the caller's rotation (t0 = n-1, t1 = n,
 * t2 = n+1 is written).  scratch[0..3]: predicted zeta, ubar, vbar and the pressure
 * surface (szI*szJ each, bulk precision). */
typedef struct sf_phys_fields {
    int32_t szI, szJ;
    double g, rdrg;
    const void *h, *pm, *pn, *fomn, *sustr, *svstr;
    void *zeta[3], *ubar[3], *vbar[3];
    void *rzeta[3], *rubar[3], *rvbar[3];
    void *scratch[4];
    double *Zt_avg1, *DU_avg1, *DV_avg1;
} sf_phys_fields;

/* One physics-mode substep: predictor (leapfrog; forward Euler when first != 0) and
 * corrector (AM3; backward Euler when first != 0).  cff1 = weight of this substep
 * in the time averages.  No reference kernel: the live reference has no 2D
 * advection/Coriolis/stress terms (they exist only as Fortran text). */
int sf_phys_substep(const sf_phys_fields *fields, int32_t precision, double dt, int32_t first,
                    double cff1, void *stream);

/* The same substep as ONE launch: each workgroup evaluates predictor and corrector
 * on an LDS-staged tile of 64 x tile_j points (tile_j = 4, 8 or 16; 0 = default)
 * with a 3/2-point halo, so no intermediate state goes to memory.  scratch[] is
 * not used.  Results equal sf_phys_substep bit for bit in both precisions.  Same
 * reference lineage as sf_phys_substep. */
int sf_phys_substep_fused(const sf_phys_fields *fields, int32_t precision, double dt, int32_t first,
                          double cff1, int32_t tile_j, void *stream);

/* Default tile height of sf_phys_substep_fused for a precision and grid size (8, or 4
 * when 64 x 8 tiles would give fewer than 1024 workgroups). */
int sf_phys_fused_default_tile(int32_t precision, int32_t szI, int32_t szJ);

/* ---------------- baroclinic terms: state equation, pressure gradient, tracers ------- */

/* Linear equation of state (utility/test.in:2130-2160) on the szK levels of T and S:
 * rho = ((R0 - 1000) - R0 Tcoef (T - T0)) + R0 Scoef (S - S0), the density anomaly over
 * 1000 kg/m3.  Fields in `precision` (SF_FP64 / SF_FP32), arithmetic in FP64. */
int sf_rho_eos(sf_ctx *ctx, int32_t precision, const void *T, const void *S, double R0, double T0, double S0,
               double Tcoef, double Scoef, void *rho, void *stream);

/* Baroclinic pressure gradient (the prsgrd call nonlinear/rhs.py:31-32 leaves commented
 * out): the standard density Jacobian on s-levels (Song 1998; ROMS prsgrd31), ADDED to
 * ru/rv (levels 0 .. szK-1) on the interior u points (i = 2 .. szI-2, j = 1 .. szJ-2) and
 * v points (j = 2 .. szJ-2, i = 1 .. szI-2); rho from sf_rho_eos, z_r/Hz with szK levels,
 * z_w with szK+1, g and rho0 the gravity and Boussinesq density.  Column sums in FP64. */
int sf_prsgrd(sf_ctx *ctx, int32_t precision, const void *rho, const void *z_r, const void *z_w, const void *Hz,
              double g, double rho0, void *ru, void *rv, void *stream);

/* Tracer step (nonlinear/step3d_t.py, flux form): first-order upstream horizontal and
 * vertical advection by the mass fluxes Huon/Hvom, forward over dt, with each cell's
 * vertical volume flux formed from the horizontal ones and the thickness change
 * Hz_old -> Hz_new (a uniform tracer stays uniform); interior columns i = 1 .. szI-2,
 * j = 1 .. szJ-2, ghost columns copied from their interior neighbours.  t and tn (new)
 * have szK levels and must not alias. */
int sf_step3d_t(sf_ctx *ctx, int32_t precision, const void *t, void *tn, const void *Huon, const void *Hvom,
                const void *Hz_old, const void *Hz_new, double dt, void *stream);

/* ---------------- coupled split-explicit step (ROMS 2D/3D coupling) ---------------- */

/* Forcing of the barotropic mode by the baroclinic step: rufrc = sum over the szK levels
 * of ru + ((sustr - rdrg * u(level 0)) * om_u) * on_u on the interior u points (i = 2 ..
 * szI-2, j = 1 .. szJ-2), rvfrc likewise on the interior v points (j = 2 .. szJ-2,
 * i = 1 .. szI-2), 0 elsewhere; FP64 sums of FP64 or FP32 (precision) 3D fields.  The
 * reference computes ru, rv here but leaves this integral commented out
 * (nonlinear/rhs.py:76-89). */
int sf_couple_rufrc(sf_ctx *ctx, int32_t precision, const void *ru, const void *rv, const void *u, const void *v,
                    const double *sustr, const double *svstr, double rdrg, double *rufrc, double *rvfrc,
                    void *stream);

/* Coupling fields of one physics-mode substep (ROMS step2d with SOLVE3D,
 * obsolete/step2d_LF_AM3.h.py:778-846 and :2377-2590), all FP64.  The first substep of
 * a baroclinic step (first != 0) forms the forcing F = rufrc - rhs of its predictor into
 * Fu/Fv and adds its Adams-Bashforth extrapolation over baroclinic steps (ab = 0: F;
 * 1: 1.5 F - 0.5 F1; 2: 23/12 F - 16/12 F1 + 5/12 F2; Fu1/Fu2: F of the two previous
 * baroclinic steps); every other stage adds F.  DU_avg2/DV_avg2 accumulate w2p x the
 * fluxes of the predictor's state (set, not accumulated, on the first substep) and
 * w2c x those of the predicted state. */
typedef struct sf_couple_fields {
    const double *rufrc, *rvfrc;
    double *Fu, *Fv;
    const double *Fu1, *Fv1, *Fu2, *Fv2;
    double *DU_avg2, *DV_avg2;
    double w2p, w2c;
    int32_t ab;
} sf_couple_fields;

/* sf_phys_substep (four launches) coupled with the baroclinic step.  Same reference
 * lineage as sf_phys_substep. */
int sf_phys_substep_coupled(const sf_phys_fields *fields, const sf_couple_fields *cpl, int32_t precision,
                            double dt, int32_t first, double cff1, void *stream);

/* sf_phys_substep_fused coupled with the baroclinic step: the coupling of
 * sf_phys_substep_coupled inside the single-launch tile kernel (the forcing of the
 * ghost-source rows kept in LDS, the DU_avg2 terms from the fluxes the tile already
 * forms).  Tile height 4 or 8 (0: sf_phys_fused_default_tile).  FP64 results equal
 * sf_phys_substep_coupled bit for bit; FP32 agrees to rounding (contraction on, as
 * sf_phys_substep_fused). */
int sf_phys_substep_coupled_fused(const sf_phys_fields *fields, const sf_couple_fields *cpl, int32_t precision,
                                  double dt, int32_t first, double cff1, int32_t tile_j, void *stream);

/* Fields of the coupled 3D momentum step.  3D fields, ubar/vbar, the right-hand-side
 * history and the scratch in the step's precision; the averages and the stresses FP64. */
typedef struct sf_uv_coupled {
    void *u, *v;                          /* szK+1 levels; levels 0..szK-1 updated */
    const void *ru, *rv;                  /* this step's right-hand side (rhs3d + pressure gradient) */
    const void *ru1, *rv1, *ru2, *rv2;    /* right-hand sides of the two steps before (szK levels) */
    void *ru_keep, *rv_keep;              /* receive ru, rv (szK levels); must not alias ru1..rv2 */
    void *ubar_t1, *vbar_t1, *ubar_t2, *vbar_t2;
    const void *Hz_old, *Hz, *Akv, *z_r;  /* thicknesses before and after set_depth; Akv szK+1 levels */
    const double *DU_avg1, *DV_avg1, *DU_avg2, *DV_avg2;
    void *Huon, *Hvom;                    /* mass fluxes, corrected to DU_avg2 / DV_avg2 */
    const double *sustr, *svstr;          /* kinematic surface stress at u / v points */
    double rdrg;                          /* linear bottom drag */
    int32_t ab;                           /* 0: first step, 1: second, 2: later (Adams-Bashforth order) */
    void *SD, *D, *RHS;                   /* scratch, szK * szI * szJ elements each */
} sf_uv_coupled;

/* The 3D momentum step of the coupled form (nonlinear/step3d_uv.py:155-180 and
 * :365-544, ROMS step3d_uv with "Couple 2D and 3D momentum equations"; replaces the live
 * kernel's solve, tmpU mean correction and adjustBarotropicVelocity,
 * step3D_uv.kernels.cpp:408-707, and makes the vbar_t2 = 0 work-around of
 * nonlinear/main3d.py:189 unnecessary).  On the interior u points (i = 2..szI-2,
 * j = 1..szJ-2; v: j = 2..szJ-2, i = 1..szI-2) it solves
 *   Hz_u,new u_new - dt d/dz(Akv du_new/dz) = Hz_u,old u + dt pm_u pn_u AB3(ru)
 *                                             - dt rdrg u(0) [bottom] + dt sustr [top]
 * (the reference's solve multiplies ru by Hz a second time and has no right-hand-side
 * history, which is unstable once the coupling moves the 3D flow), then replaces the
 * column mean: u -= (sum DC u - DU_avg1) / D with DC = on_u Hz_u, ubar(t1) = ubar(t2) =
 * DU_avg1 / D, and the mass flux Huon = 0.5 (Huon + u DC) corrected to sum to DU_avg2;
 * v likewise.  Other columns are left as they are.  Column sums in FP64. */
int sf_step3d_uv_coupled(sf_ctx *ctx, int32_t precision, const sf_uv_coupled *fields, double lambda, double dt,
                         void *stream);

/* ---------------- open boundaries of the physics substep ---------------------------- */

/* Open edges of a physics-mode basin (flags: SF_OBC_WEST/EAST/SOUTH/NORTH below; the
 * other edges stay closed walls).  zeta_ext[e] / vel_ext[e]: external free surface and
 * normal barotropic velocity along edge e (0 west, 1 east, 2 south, 3 north; west/east
 * indexed by j, south/north by i, szJ or szI doubles; NULL = 0).  volcons != 0:
 * obc_volcons + conserve_mass (sf_obc_volcons) on the open edges' velocities. */
typedef struct sf_open_bc {
    int32_t flags;
    int32_t volcons;
    const double *zeta_ext[4], *vel_ext[4];
} sf_open_bc;

/* sf_phys_substep (four launches) with open edges: after each stage, the ghost free
 * surface of every open edge takes the implicit Chapman radiation value (nonlinear/zetabc.py,
 * Chapman) and the normal velocity on the edge face the Flather value
 * (nonlinear/barotropicVelocityBC.py:66-98).  volcons != 0: after each of the two
 * velocity stages (predicted and new state) the volume constraint
 * (nonlinear/obc_volcons.py:1-72, sf_obc_volcons) removes the net flow through the open
 * edges, so the free-surface updates conserve the volume (volcons_work:
 * sf_obc_volcons_work_size doubles; volcons_out: area, flux, ubar_xs of the last). */
int sf_phys_substep_open(const sf_phys_fields *fields, const sf_open_bc *obc, int32_t precision, double dt,
                         int32_t first, double cff1, double *volcons_work, double *volcons_out, void *stream);

/* ---------------- ensembles of physics-mode basins ---------------------------------- */

/* An ensemble of basins of any sizes advanced together: ONE launch per substep walks a
 * device list of (member, tile) pairs.  No reference counterpart (the reference runs
 * one grid per process, mod_cppkernels.h:154-171); each member's result equals its
 * solo sf_phys_substep_fused run bit for bit.  members[k].zeta[s] etc. are the three
 * time slots of member k with slot s = level t_s at creation; the library owns the
 * rotation from then on (sf_ens_substep performs cycleTimes2D before each substep:
 * level t_k lives in slot (k + rot) % 3, rot = sf_ens_rotation()).  scratch[] unused.
 * dt[k]: substep of member k.  The ensemble does not own member memory. */
typedef struct sf_ens sf_ens;
int sf_ens_create(const sf_phys_fields *members, const double *dt, int32_t n, int32_t precision,
                  sf_ens **out);
int sf_ens_destroy(sf_ens *ens);
/* Schedule: tile height (4, 8, 16), tile order (0 strided over blocks, 1 contiguous
 * chunk per block with XCD banding) and grid size (0 = one block per tile). */
int sf_ens_configure(sf_ens *ens, int32_t tile_j, int32_t order, int32_t grid);
int sf_ens_num_tiles(const sf_ens *ens);
int sf_ens_rotation(const sf_ens *ens);
/* One substep of every member (first != 0: Euler start), cff1 = average weight. */
int sf_ens_substep(sf_ens *ens, int32_t first, double cff1, void *stream);
/* Capture nsub (multiple of 3) non-first substeps with weights cff1[0..nsub) into a
 * HIP graph; sf_ens_replay launches it on a stream.  Reconfiguring drops the graph. */
int sf_ens_capture(sf_ens *ens, int32_t nsub, const double *cff1, void *stream);
int sf_ens_replay(sf_ens *ens, void *stream);

/* ---------------- open-boundary volume conservation --------------------------------- */

#define SF_OBC_WEST 1
#define SF_OBC_EAST 2
#define SF_OBC_SOUTH 4
#define SF_OBC_NORTH 8

/* obc_volcons + conserve_mass (nonlinear/obc_volcons.py:1-72): FP64 sums of the open
 * segments' cross-section area and net volume flux of the state (zeta, ubar, vbar) in
 * `precision`; out[0] = area, out[1] = flux, out[2] = ubar_xs = flux / area (device,
 * FP64).  apply != 0: the boundary velocities are corrected by -/+ ubar_xs (west,
 * south / east, north).  work: sf_obc_volcons_work_size(szI, szJ) doubles. */
int sf_obc_volcons_work_size(int32_t szI, int32_t szJ);
int sf_obc_volcons(const void *zeta, const void *h, const void *pm, const void *pn, void *ubar, void *vbar,
                   int32_t szI, int32_t szJ, int32_t precision, int32_t flags, int32_t apply, double *work,
                   double *out, void *stream);

/* sf_obc_volcons without the work argument and without any allocation: the block
 * partials live in a static device array of 4096 doubles (one per device, held until
 * the process ends), so sf_obc_volcons_work_size(szI, szJ) must be <= 4096, i.e. at
 * most 524 032 open-boundary points; larger grids get SF_ERR_ARG.  Same kernels and
 * results as sf_obc_volcons.  Calls on one device share that array: calls that overlap
 * in time on one device (different streams, or host threads) race.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (HIP graph
 * capture, tight serving loops). */
int sf_obc_volcons_noalloc(const void *zeta, const void *h, const void *pm, const void *pn, void *ubar,
                           void *vbar, int32_t szI, int32_t szJ, int32_t precision, int32_t flags, int32_t apply,
                           double *out, void *stream);

/* ---------------- learned subgrid momentum closure ---------------------------------- */

/* The closure network (45 -> 64 ReLU -> 64 ReLU -> 2) trained and evaluated on matrix
 * cores (BF16 operands, FP32 accumulation).  No reference counterpart (the learned
 * closure is new); it enters the physics substep as an extra momentum stress through
 * the sustr/svstr fields (obsolete/step2d_LF_AM3.h.py:2595-2606 is where ROMS adds its
 * surface stress).  Parameter vector layout: W1[48][64], b1[64], W2[64][64], b2[64],
 * W3[64][16], b3[16] (padding rows/columns zero). */
typedef struct sf_mlp sf_mlp;
int sf_mlp_num_params(void);
/* max_blocks: the largest number of workgroups of one training step. */
int sf_mlp_create(int32_t max_blocks, sf_mlp **out);
int sf_mlp_destroy(sf_mlp *mlp);
int sf_mlp_set_params(sf_mlp *mlp, const float *host_params);
int sf_mlp_get_params(sf_mlp *mlp, float *host_params);
/* gradient and loss of the last sf_mlp_train_step (host copies) */
int sf_mlp_get_grad(sf_mlp *mlp, float *host_grad, float *host_loss);
int sf_mlp_set_norm(sf_mlp *mlp, const float *mu45, const float *sd45, const float *ynorm4);
/* One step on the batch idx[0 .. nblocks*ntile*64) of X (BF16 bits [n][48]) and Y
 * ([n][2]): forward + backward + deterministic gradient reduction, then Adam when
 * adam != 0 (step counts from 1).  nblocks workgroups of 4 waves each take ntile tiles
 * of 64 samples; nblocks <= 256 keeps one workgroup per CU. */
int sf_mlp_train_step(sf_mlp *mlp, const uint16_t *X, const float *Y, const int32_t *idx, int32_t nblocks,
                      int32_t ntile, int32_t adam, float lr, float beta1, float beta2, float eps, int32_t step,
                      void *stream);
int sf_mlp_forward(sf_mlp *mlp, const uint16_t *X, int32_t n, float *Y, void *stream);
/* Raw features [S2][45] of the t1 state of a physics-mode field set (f: Coriolis): per
 * field (ubar, vbar, zeta, h, f) the 3x3 stencil at the point, the centre as its value and
 * the 8 neighbours as differences neighbour - centre (state precision, rounded to FP32). */
int sf_closure_features(const sf_phys_fields *fields, int32_t precision, const void *f, float *out, void *stream);
/* sustr_eff = sustr + R_u * D_u, svstr_eff = svstr + R_v * D_v from the t1 state. */
int sf_closure_stress(sf_mlp *mlp, const sf_phys_fields *fields, int32_t precision, const void *f,
                      void *sustr_eff, void *svstr_eff, void *stream);
/* One physics substep (sf_phys_substep_fused) with the closure evaluated inside the tile
 * kernel: the corrector stage's surface stress is sf_closure_stress's sustr_eff /
 * svstr_eff of the substep's t1 state, the predictor keeps sustr / svstr.  f: the
 * Coriolis field in the state precision.  Tile height 4 or 8 (0: 4, the height whose
 * closure instances keep two waves per SIMD). */
int sf_phys_substep_closure_fused(const sf_phys_fields *fields, const sf_mlp *mlp, const void *f,
                                  int32_t precision, double dt, int32_t first, double cff1, int32_t tile_j,
                                  void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SAILFISH_HIP_H */
