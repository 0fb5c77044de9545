"""Mutation check of the oracle's pins.

Each entry below plants one plausible mistake in a copy of the oracle's C sources (a
dropped term, a flipped sign, a swapped time slot or metric factor, a wrong
coefficient), builds that copy, and runs the oracle's CPU pin tests against it
(tests/test_oracle2d.py, test_oracle3d.py, test_oracle_phys.py,
test_oracle_volcons.py) with the environment variable SF_ORACLE_LIB pointing at the
mutant.  Every mutant must fail at least one pin; the unmutated copy, built the same
way, must pass all of them.  CPU only: the mutants run nowhere else.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
ORACLE = ROOT / "oracle"
PIN_FILES = ["tests/test_oracle2d.py", "tests/test_oracle3d.py", "tests/test_oracle_phys.py",
             "tests/test_oracle_volcons.py", "tests/test_oracle_coupled.py", "tests/test_oracle_open.py",
             "tests/test_oracle_baroclinic.py"]

# (name, source file, text to replace, replacement): the text occurs exactly once
MUTANTS = [
    # --- faithful 2D kernels (modules/mod_cppkernels.cpp) ---
    ("zeta_pred_2dt", "sf_oracle2d.c",
     "zeta_t2[f] = zeta_t0[f] + Dt * rzeta_t1[f];", "zeta_t2[f] = zeta_t0[f] + (2.0 * Dt) * rzeta_t1[f];"),
    ("zeta_pred_reads_t1", "sf_oracle2d.c",
     "zeta_t2[f] = zeta_t0[f] + Dt * rzeta_t1[f];", "zeta_t2[f] = zeta_t1[f] + Dt * rzeta_t1[f];"),
    ("rhs_corr_slots_swapped", "sf_oracle2d.c",
     "momentum_rhs(g, h, rhs_ubar, rhs_vbar, zeta_t2, zeta_t1, gc, 2.0 / 5.0);",
     "momentum_rhs(g, h, rhs_ubar, rhs_vbar, zeta_t1, zeta_t2, gc, 2.0 / 5.0);"),
    ("rhs_pred_slots_swapped", "sf_oracle2d.c",
     "momentum_rhs(g, h, rhs_ubar, rhs_vbar, zeta_t2, zeta_t0, gc, weight);",
     "momentum_rhs(g, h, rhs_ubar, rhs_vbar, zeta_t0, zeta_t2, gc, weight);"),
    ("du_avg_without_on_u", "sf_oracle2d.c",
     "DU_avg1[f] = DU_avg1[f] + (cff1 * DU) * on_u[f];", "DU_avg1[f] = DU_avg1[f] + (cff1 * DU);"),
    ("dv_avg_with_on_u", "sf_oracle2d.c",
     "DV_avg1[f] = DV_avg1[f] + (cff1 * DV) * om_v[f];", "DV_avg1[f] = DV_avg1[f] + (cff1 * DV) * on_u[f];"),
    ("zt_avg_unweighted", "sf_oracle2d.c",
     "Zt_avg1[f] = Zt_avg1[f] + cff1 * zeta[f];", "Zt_avg1[f] = Zt_avg1[f] + zeta[f];"),
    ("momentum_pred_pm_pm", "sf_oracle2d.c",
     "u_t2[f] = (u_t1[f] * D1U + ((Dt * rhsu[f]) * pm[f]) * pn[f]) / D2U;",
     "u_t2[f] = (u_t1[f] * D1U + ((Dt * rhsu[f]) * pm[f]) * pm[f]) / D2U;"),
    ("momentum_pred_depths_swapped", "sf_oracle2d.c",
     "D2V = (D2 + D2d) * 0.5;\n        u_t2[f] = (u_t1[f] * D1U + ((Dt * rhsu[f]) * pm[f]) * pn[f]) / D2U;",
     "D2V = (D2 + D2d) * 0.5;\n        u_t2[f] = (u_t1[f] * D2U + ((Dt * rhsu[f]) * pm[f]) * pn[f]) / D1U;"),
    ("pgf_sign", "sf_oracle2d.c",
     "rhs_u[f] = (gc * on_u[f]) * (hu * (gz - gzl) + 0.5 * (gz * gz - gzl * gzl));",
     "rhs_u[f] = (gc * on_u[f]) * (hu * (gzl - gz) + 0.5 * (gz * gz - gzl * gzl));"),
    ("pgf_v_metric", "sf_oracle2d.c",
     "rhs_v[f] = (gc * om_v[f]) * (hv * (gz - gzd) + 0.5 * (gz * gz - gzd * gzd));",
     "rhs_v[f] = (gc * on_u[f]) * (hv * (gz - gzd) + 0.5 * (gz * gz - gzd * gzd));"),
    ("divergence_face_index", "sf_oracle2d.c",
     "(DVu * om_v[f + szI] - DV * om_v[f])", "(DVu * om_v[f] - DV * om_v[f + szI])"),
    ("am3_weight", "sf_oracle2d.c",
     "const double AM3_2 = 5.0 / 12.0, AM3_1 = 8.0 / 12.0, AM3_0 = -1.0 / 12.0;\n#pragma omp parallel for schedule(static)\n    for (ptrdiff_t f = 0; f < S2; f++)\n        v_t2[f]",
     "const double AM3_2 = 5.0 / 12.0, AM3_1 = 7.0 / 12.0, AM3_0 = -1.0 / 12.0;\n#pragma omp parallel for schedule(static)\n    for (ptrdiff_t f = 0; f < S2; f++)\n        v_t2[f]"),
    ("am3_2_slot", "sf_oracle2d.c",
     "(AM3_2 * ru2[f] + AM3_1 * ru1[f] + AM3_0 * ru0[f])", "(AM3_2 * ru2[f] + AM3_1 * ru1[f] + AM3_0 * ru1[f])"),
    # --- physics mode (obsolete/step2d_LF_AM3.h.py) ---
    ("wind_sign", "sf_oracle_phys.c",
     "r = r + (p->sustr[f] * omu) * onu;", "r = r - (p->sustr[f] * omu) * onu;"),
    ("drag_sign", "sf_oracle_phys.c",
     "r = r - ((p->rdrg * u[f]) * omu) * onu;", "r = r + ((p->rdrg * u[f]) * omu) * onu;"),
    ("drag_v_metric", "sf_oracle_phys.c",
     "r = r - ((p->rdrg * v[f]) * omv) * onv;", "r = r - ((p->rdrg * v[f]) * omv) * omv;"),
    ("coriolis_u_sign", "sf_oracle_phys.c",
     "r = r + 0.5 * (c1 + c0);", "r = r - 0.5 * (c1 + c0);"),
    ("coriolis_v_sign", "sf_oracle_phys.c",
     "r = r - 0.5 * (c1 + c0);", "r = r + 0.5 * (c1 + c0);"),
    ("c2_advection_factor", "sf_oracle_phys.c",
     "const double UFx1 = 0.25 *", "const double UFx1 = 0.5 *"),
    ("c2_advection_shear_face", "sf_oracle_phys.c",
     "(u[f + W] + u[f]);", "(u[f + 1] + u[f]);"),
    ("pgf_phys_without_g", "sf_oracle_phys.c",
     "double r = ((0.5 * p->g) * onu)", "double r = (0.5 * onu)"),
    ("pgf_phys_sign", "sf_oracle_phys.c",
     "((p->h[f - 1] + p->h[f]) * (g0 - g1) + (g0 * g0 - g1 * g1))",
     "((p->h[f - 1] + p->h[f]) * (g1 - g0) + (g0 * g0 - g1 * g1))"),
    ("leapfrog_dt", "sf_oracle_phys.c",
     "else zs[f] = z0[f] + ((p->pm[f] * p->pn[f]) * (2.0 * dt)) * r;",
     "else zs[f] = z0[f] + ((p->pm[f] * p->pn[f]) * dt) * r;"),
    ("am3_zeta_weight", "sf_oracle_phys.c",
     "a2 = dt * 8.0 / 12.0", "a2 = dt * 7.0 / 12.0"),
    ("corrector_stress_ignored", "sf_oracle_phys.c",
     "double r = sfo_phys_rhs_u(pc, zs, us, vs, gz, j, i);", "double r = sfo_phys_rhs_u(p, zs, us, vs, gz, j, i);"),
    ("am3_momentum_weight", "sf_oracle_phys.c",
     "b1 = 0.5 * dt * 5.0 / 12.0", "b1 = 0.5 * dt * 4.0 / 12.0"),
    ("predictor_surface_weight", "sf_oracle_phys.c",
     "const double cff4 = 4.0 / 25.0", "const double cff4 = 3.0 / 25.0"),
    ("first_step_predictor_half", "sf_oracle_phys.c",
     "const double c1 = first ? 0.5 * dt : dt;", "const double c1 = first ? dt : dt;"),
    ("phys_du_average_velocity_only", "sf_oracle_phys.c",
     "DUa[f] = DUa[f] + cff1w * DUon(p, zs, us, W, j, i);", "DUa[f] = DUa[f] + cff1w * us[f];"),
    ("phys_zt_average_old_state", "sf_oracle_phys.c",
     "Zt[f] = Zt[f] + cff1w * zs[f];", "Zt[f] = Zt[f] + cff1w * z1[f];"),
    ("volcons_east_sign", "sf_oracle_phys.c",
     "fl = seg == 0 ? cff * u[f] : -(cff * u[f]);", "fl = cff * u[f];"),
    # --- 3D kernels (nonlinear/rhs_kernels.cpp, step3D_uv.kernels.cpp) ---
    ("hadv_curvature_sign", "sf_oracle3d.c",
     "static const double Gadv = -0.5;", "static const double Gadv = 0.5;"),
    ("hadv_upwind_xi_direction", "sf_oracle3d.c",
     "? dxx(h->u, base, o) : dxx(h->u, base, o + 1);", "? dxx(h->u, base, o + 1) : dxx(h->u, base, o);"),
    ("hadv_upwind_eta_direction", "sf_oracle3d.c",
     "? dee(h->u, base, o, W) : dee(h->u, base, o - W, W);", "? dee(h->u, base, o - W, W) : dee(h->u, base, o, W);"),
    ("hadv_vfe_own_curvature", "sf_oracle3d.c",
     "return h_vR(h, b, o) + (Gadv * h_uxxR(h, b, o))", "return h_vR(h, b, o) + (Gadv * h_HveeR(h, b, o))"),
    ("vadv_bottom_without_uuu2", "sf_oracle3d.c",
     "ru[0 * S2 + f] -= (uuu - uuu2);", "ru[0 * S2 + f] -= uuu;"),
    ("vadv_4th_order_weight", "sf_oracle3d.c",
     "return (9.0 / 16.0) * (r3(a, K * S2 + f)", "return (8.0 / 16.0) * (r3(a, K * S2 + f)"),
    ("vadv_level1_v", "sf_oracle3d.c",
     "rv[1 * S2 + f] -= (9.0 / 16.0) * (FV(1) + FV(2)) - (1.0 / 16.0) * (FV(1) + FV(3));",
     "rv[1 * S2 + f] -= (9.0 / 16.0) * (FV(1) + FV(2)) - (1.0 / 16.0) * (FV(2) + FV(3));"),
    ("lateral_clamp_amplitude", "sf_oracle3d.c",
     "const double val = 0.05 * sin(t * 0.0002);", "const double val = 0.5 * sin(t * 0.0002);"),
    ("coriolis3d_v_sign", "sf_oracle3d.c",
     "rv[b] = rv[b] - rtov;", "rv[b] = rv[b] + rtov;"),
    ("omega_surface_correction_sign", "sf_oracle3d.c",
     "Wv[b] = Wv[b] - (Ws * (z_w[b] - zb)) / h;", "Wv[b] = Wv[b] + (Ws * (z_w[b] - zb)) / h;"),
    ("maxflux_metric", "sf_oracle3d.c",
     "Huon[b] = (((Hz[b] + Hz[b - 1]) * 0.5) * u[b]) * g->on_u[f];",
     "Huon[b] = (((Hz[b] + Hz[b - 1]) * 0.5) * u[b]) * g->om_v[f];"),
    ("set_depth_vtransform2", "sf_oracle3d.c",
     "const double cw = (hc * sc_w[K] + Cs_w[K] * h[f]) / (hc + h[f]);",
     "const double cw = (hc * sc_w[K] + Cs_w[K] * h[f]) / h[f];"),
    ("set_zeta_one_slot", "sf_oracle3d.c",
     "        zeta2[f] = Zt[f];\n", "        zeta2[f] = zeta1[f];\n"),
    ("step3d_ab3_second_factor", "sf_oracle3d.c",
     "S2_OF(g), N = g->N;\n    double cff;\n    if (iic == ntfirst) cff = 1.0;\n    else if (iic == ntfirst + 1) cff = 3.0 / 2.0;",
     "S2_OF(g), N = g->N;\n    double cff;\n    if (iic == ntfirst) cff = 1.0;\n    else if (iic == ntfirst + 1) cff = 1.0;"),
    ("step3d_viscous_sign", "sf_oracle3d.c",
     "SD[K * S2 + f] = (((-lambda) * Dt) * AKVU(K)) / DZ(K + 1);",
     "SD[K * S2 + f] = ((lambda * Dt) * AKVU(K)) / DZ(K + 1);"),
    ("step3d_mean_correction_sign", "sf_oracle3d.c",
     "for (ptrdiff_t K = 0; K < N; K++) x[K * S2 + f] = x[K * S2 + f] - diff;",
     "for (ptrdiff_t K = 0; K < N; K++) x[K * S2 + f] = x[K * S2 + f] + diff;"),
    # --- coupled split-explicit form (ROMS 2D/3D coupling) ---
    ("cpl_ab3_middle_coefficient", "sf_oracle_phys.c",
     "(16.0 / 12.0) * F1", "(15.0 / 12.0) * F1"),
    ("cpl_ab2_coefficient", "sf_oracle_phys.c",
     "if (ab == 1) return 1.5 * F - 0.5 * F1;", "if (ab == 1) return 1.5 * F - 0.5 * F;"),
    ("cpl_forcing_sign", "sf_oracle_phys.c",
     "const double F = cp->rufrc[f] - r;", "const double F = r - cp->rufrc[f];"),
    ("cpl_corrector_without_forcing", "sf_oracle_phys.c",
     "if (cp) r = r + cp->Fu[f];", "if (cp) r = r + 0.0 * cp->Fu[f];"),
    ("cpl_avg2_corrector_state", "sf_oracle_phys.c",
     "cp->DU2[f] = cp->DU2[f] + cp->w2c * DUon(p, zs, us, W, j, i);",
     "cp->DU2[f] = cp->DU2[f] + cp->w2c * DUon(p, z1, u1, W, j, i);"),
    ("cpl_avg2_not_reset", "sf_oracle_phys.c",
     "const double a = first ? 0.0 : cp->DU2[f]", "const double a = cp->DU2[f]"),
    ("cpl_rufrc_stress_sign", "sf_oracle3d.c",
     "su = su + ((sustr[f] - rdrg * u[f]) * omu) * onu;", "su = su + ((sustr[f] + rdrg * u[f]) * omu) * onu;"),
    ("cpl_rufrc_skips_bottom", "sf_oracle3d.c",
     "for (ptrdiff_t K = 0; K < N; K++) su = su + ru[K * S2 + f];",
     "for (ptrdiff_t K = 1; K < N; K++) su = su + ru[K * S2 + f];"),
    ("cpl_huon_half_dropped", "sf_oracle3d.c",
     "Hf[K * S2 + f] = 0.5 * (Hf[K * S2 + f] + x[K * S2 + f] * DC(K));",
     "Hf[K * S2 + f] = (Hf[K * S2 + f] + x[K * S2 + f] * DC(K));"),
    ("cpl_transport_avg1_for_avg2", "sf_oracle3d.c",
     "FC = rD * (FC - Davg2[f]);", "FC = rD * (FC - Davg1[f]);"),
    ("cpl_ubar_without_depth", "sf_oracle3d.c",
     "xbar2[f] = rD * Davg1[f];", "xbar2[f] = Davg1[f];"),
    ("cpl3d_reference_rhs_form", "sf_oracle3d.c",
     "RHS[b] = AVG(Ho, K) * x[b] + cdt_mn * rab;", "RHS[b] = AVG(Hz, K) * (x[b] + cdt_mn * rab);"),
    ("cpl3d_new_thickness_for_old", "sf_oracle3d.c",
     "RHS[b] = AVG(Ho, K) * x[b] + cdt_mn * rab;", "RHS[b] = AVG(Hz, K) * x[b] + cdt_mn * rab;"),
    ("cpl3d_akv_interface_below", "sf_oracle3d.c",
     "AVG(Akv, K + 1)) / DZ(K + 1);", "AVG(Akv, K)) / DZ(K + 1);"),
    ("cpl3d_ab3_last_term", "sf_oracle3d.c",
     "(5.0 / 12.0) * r2;", "(5.0 / 12.0) * r1;"),
    ("cpl3d_ab2_coefficient", "sf_oracle3d.c",
     "if (ab == 1) return 1.5 * r - 0.5 * r1;", "if (ab == 1) return 1.5 * r - 0.5 * r;"),
    ("cpl3d_bottom_drag_dropped", "sf_oracle3d.c",
     "RHS[f] = RHS[f] - (Dt * a->rdrg) * x[f];", "RHS[f] = RHS[f] - 0.0 * x[f];"),
    ("cpl3d_surface_stress_sign", "sf_oracle3d.c",
     "RHS[(N - 1) * S2 + f] = RHS[(N - 1) * S2 + f] + Dt * stress[f];",
     "RHS[(N - 1) * S2 + f] = RHS[(N - 1) * S2 + f] - Dt * stress[f];"),
    ("cpl3d_keeps_extrapolation", "sf_oracle3d.c",
     "keep[b] = rk;", "keep[b] = rab;"),
    ("cpl3d_walls_solved", "sf_oracle3d.c",
     "        if (coupled_point(g, 0, f)) {", "        if (isUNode(g, f)) {"),
    # --- open edges (Chapman, Flather, volume constraint) ---
    ("open_flather_sign", "sf_oracle_phys.c",
     "x[fu] = (e == 0 || e == 2) ? ext_at(ob->ext_u[e], b) - c * dz : ext_at(ob->ext_u[e], b) + c * dz;",
     "x[fu] = (e == 0 || e == 2) ? ext_at(ob->ext_u[e], b) + c * dz : ext_at(ob->ext_u[e], b) - c * dz;"),
    ("open_flather_celerity", "sf_oracle_phys.c",
     "const double c = sqrt(p->g / D);", "const double c = sqrt(p->g * D);"),
    ("open_flather_ext_level_ignored", "sf_oracle_phys.c",
     "const double dz = 0.5 * (zn[fg] + zn[fi]) - ext_at(ob->ext_z[e], b);",
     "const double dz = 0.5 * (zn[fg] + zn[fi]);"),
    ("open_chapman_explicit_interior", "sf_oracle_phys.c",
     "zn[fg] = (zb[fg] + Ce * zn[fi]) / (1.0 + Ce);", "zn[fg] = (zb[fg] + Ce * zb[fi]) / (1.0 + Ce);"),
    ("open_chapman_frozen", "sf_oracle_phys.c",
     "zn[fg] = (zb[fg] + Ce * zn[fi]) / (1.0 + Ce);", "zn[fg] = zb[fg];"),
    ("open_volcons_corrector_only", "sf_oracle_phys.c",
     "if (ob->volcons) sfo_obc_volcons(p->szI, p->szJ, zs, p->h, p->pm, p->pn, us, vs, ob->flags, 1, tmp);",
     "(void)tmp;"),
    # --- baroclinic terms (EOS, pressure gradient, tracer step) ---
    ("eos_temperature_sign", "sf_oracle3d.c",
     "rho[b] = (a - bt * (T[b] - T0)) + bs * (S[b] - S0);", "rho[b] = (a + bt * (T[b] - T0)) + bs * (S[b] - S0);"),
    ("eos_salinity_dropped", "sf_oracle3d.c",
     "rho[b] = (a - bt * (T[b] - T0)) + bs * (S[b] - S0);", "rho[b] = (a - bt * (T[b] - T0));"),
    ("pgf_background_density", "sf_oracle3d.c",
     "fac2 = 1000.0 * grav / rho0", "fac2 = 0.0 * grav / rho0"),
    ("pgf_jacobian_weight", "sf_oracle3d.c",
     "const double C1 = (1.0 + gam) * (R(k + 1, f) - R(k + 1, p)) + (1.0 - gam) * (R(k, f) - R(k, p));",
     "const double C1 = (1.0 - gam) * (R(k + 1, f) - R(k + 1, p)) + (1.0 + gam) * (R(k, f) - R(k, p));"),
    ("pgf_jacobian_sign", "sf_oracle3d.c",
     "phi = phi + fac3 * (C1 * C3 - C2 * C4);", "phi = phi + fac3 * (C1 * C3 + C2 * C4);"),
    ("pgf_top_half_depth", "sf_oracle3d.c",
     "double phi = (fac1 * (R(k, f) - R(k, p))) * c1;", "double phi = (fac3 * (R(k, f) - R(k, p))) * c1;"),
    ("tracer_downstream", "sf_oracle3d.c",
     "const double FX0 = hu0 * (hu0 > 0.0 ? t[b - 1] : t[b]);", "const double FX0 = hu0 * (hu0 > 0.0 ? t[b] : t[b - 1]);"),
    ("tracer_thickness_change_ignored", "sf_oracle3d.c",
     "const double Wk = (Wlow - Dh) - (Hz_new[b] - Hz_old[b]) / dt;", "const double Wk = (Wlow - Dh);"),
    ("tracer_vertical_flux_dropped", "sf_oracle3d.c",
     "const double div = mn * ((FX1 - FX0) + (FE1 - FE0)) + (FCk - FClow);",
     "const double div = mn * ((FX1 - FX0) + (FE1 - FE0));"),
]


def _mutated_sources(tmp: Path, name: str, src: str, old: str, new: str):
    out = []
    for p in sorted(ORACLE.glob("sf_oracle*.c")):
        text = p.read_text()
        if p.name == src:
            assert text.count(old) == 1, f"mutant {name}: pattern occurs {text.count(old)} times in {src}"
            text = text.replace(old, new)
        q = tmp / p.name
        q.write_text(text)
        out.append(q)
    return out


def _build_and_pin(tmp: Path, name: str, src=None, old=None, new=None):
    from oracle import build
    d = tmp / name
    d.mkdir()
    sources = _mutated_sources(d, name, src, old, new)
    lib = build.build(verbose=False, sources=sources, out=d / f"liboracle_{name}.so", opt="-O1")
    env = dict(os.environ, SF_ORACLE_LIB=str(lib), OMP_NUM_THREADS="1")
    res = subprocess.run([sys.executable, "-m", "pytest", *PIN_FILES, "-x", "-q", "-m", "not gpu",
                          "-p", "no:cacheprovider"], cwd=ROOT, env=env, capture_output=True, text=True,
                         timeout=600)
    return name, res.returncode, res.stdout[-2000:]


def test_every_planted_mistake_fails_a_pin():
    with tempfile.TemporaryDirectory(prefix="sf_oracle_mut_") as td:
        tmp = Path(td)
        jobs = [("unmutated", None, None, None)] + MUTANTS
        with cf.ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
            results = list(ex.map(lambda m: _build_and_pin(tmp, *m), jobs))
    control = results[0]
    assert control[1] == 0, f"the unmutated oracle fails its pins:\n{control[2]}"
    survived = [name for name, rc, _ in results[1:] if rc == 0]
    crashed = [(name, rc) for name, rc, _ in results[1:] if rc not in (0, 1)]
    assert not crashed, f"pin runs that did not end in a test failure: {crashed}"
    assert not survived, f"planted mistakes that pass every pin: {survived}"


@pytest.mark.parametrize("name,src,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_pattern_is_unique(name, src, old, new):
    assert (ORACLE / src).read_text().count(old) == 1
