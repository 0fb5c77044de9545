"""Pins of the oracle's baroclinic terms (oracle/sf_oracle3d.c sfo_rho_eos, sfo_prsgrd,
sfo_step3d_t) by closed forms and invariants:

- linear equation of state at T0, S0 and its temperature slope;
- pressure gradient of a density that grows linearly along x on flat levels:
  ru(k) = -Hz (g/rho0) (drho/dx dx) |z_r(k)| dy (the hydrostatic integral);
- pressure gradient of a uniform density under a sloping free surface:
  ru(k) = -Hz_u (g/rho0) (1000 + rho) dzeta on_u at every level;
- a density linear in z on sloping s-levels: the Jacobian increments vanish, so the
  gradient per unit thickness is the same at every level below the top;
- the tracer step keeps a uniform tracer uniform under any fluxes and thickness change,
  conserves the tracer content when every column's volume budget closes, and makes no
  new extrema (upstream fluxes)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from oracle import oracle3d as O3
from sailfish_amd.mod_grid import update_metrics
from sailfish_amd.set_coords import s_coordinates

P = O.ptr
GRAV, RHO0 = 10.0, 1025.0


def model(szI, szJ, N, dx, dy, h, zeta=None):
    pm = np.full((szJ, szI), 1 / dx)
    pn = np.full((szJ, szI), 1 / dy)
    m = update_metrics(pm, pn, np.zeros((szJ, szI)))
    g = O.OGrid(szI, szJ, N, dict(h=h, pm=pm, pn=pn, on_u=m["on_u"], om_v=m["om_v"]))
    sc_r, sc_w, Cs_r, Cs_w = s_coordinates(N, 4, 5.0, 0.4)
    BC = -np.ones(g.S2, np.int32)
    M = O3.OModel3(g, N, m["fomn"], sc_r, sc_w, Cs_r, Cs_w, 50.0, 2, BC, BC, BC)
    st = O.OState(g)
    if zeta is not None:
        st.Zt_avg1[:] = np.asarray(zeta, float).ravel()
    M.set_depth(st)
    return g, M


def test_linear_eos():
    szI, szJ, N = 6, 5, 4
    g, M = model(szI, szJ, N, 1e3, 1e3, np.full((szJ, szI), 100.0))
    n = N * g.S2
    T = np.full(n, 14.0)
    S = np.full(n, 35.0)
    rho = np.zeros(n)
    O.lib().sfo_rho_eos(M.cp, P(T), P(S), 1027.0, 14.0, 35.0, 1.7e-4, 7.6e-4, P(rho))
    assert np.all(rho == 27.0)
    T2, S2_ = T + 2.0, S - 1.0
    O.lib().sfo_rho_eos(M.cp, P(T2), P(S2_), 1027.0, 14.0, 35.0, 1.7e-4, 7.6e-4, P(rho))
    assert rho == pytest.approx(27.0 - 1027.0 * 1.7e-4 * 2.0 - 1027.0 * 7.6e-4, rel=1e-14)


def _pgf(M, g, rho):
    n = (g.N + 1) * g.S2
    ru, rv = np.zeros(n), np.zeros(n)
    O.lib().sfo_prsgrd(M.cp, P(rho), P(M.z_r), P(M.z_w), P(M.Hz), GRAV, RHO0, P(ru), P(rv))
    return ru.reshape(g.N + 1, g.szJ, g.szI), rv.reshape(g.N + 1, g.szJ, g.szI)


def test_pgf_of_a_density_gradient_on_flat_levels():
    szI, szJ, N, dx, dy, c = 14, 10, 8, 2e3, 3e3, 2e-4
    g, M = model(szI, szJ, N, dx, dy, np.full((szJ, szI), 200.0))
    x = (np.arange(szI) - 0.5) * dx
    rho = np.broadcast_to(c * x, (N, szJ, szI)).ravel().copy()
    ru, rv = _pgf(M, g, rho)
    Hz = M.Hz.reshape(N, szJ, szI)
    z_r = M.z_r.reshape(N, szJ, szI)
    j, i = szJ // 2, szI // 2
    exp = -Hz[:, j, i] * (GRAV / RHO0) * (c * dx) * np.abs(z_r[:, j, i]) * dy
    assert np.allclose(ru[:N, j, i], exp, rtol=1e-10)
    assert np.abs(rv[:N, 2:-1, 1:-1]).max() <= 1e-12 * np.abs(exp).max()     # nothing along y


def test_pgf_of_a_sloping_free_surface():
    szI, szJ, N, dx, dy, s, rc = 14, 10, 6, 1e3, 1e3, 1e-5, 26.0
    x = (np.arange(szI) - 0.5) * dx
    zeta = np.broadcast_to(s * x, (szJ, szI))
    g, M = model(szI, szJ, N, dx, dy, np.full((szJ, szI), 80.0), zeta=zeta)
    rho = np.full(N * g.S2, rc)
    ru, _ = _pgf(M, g, rho)
    Hz = M.Hz.reshape(N, szJ, szI)
    j, i = szJ // 2, szI // 2
    exp = -0.5 * (Hz[:, j, i] + Hz[:, j, i - 1]) * (GRAV / RHO0) * (1000.0 + rc) * (s * dx) * dy
    assert np.allclose(ru[:N, j, i], exp, rtol=1e-9)


def test_pgf_jacobian_increments_vanish_for_rho_linear_in_z():
    szI, szJ, N, dx = 16, 10, 10, 1e3
    x = (np.arange(szI) - 0.5) * dx
    h = np.broadcast_to(50.0 + 400.0 * x / x[-1], (szJ, szI)).copy()      # sloping bottom: sloping levels
    g, M = model(szI, szJ, N, dx, dx, h)
    rho = (25.0 + 0.004 * M.z_r).copy()
    ru, _ = _pgf(M, g, rho)
    Hz = M.Hz.reshape(N, szJ, szI)
    j = szJ // 2
    for i in range(2, szI - 1):
        phi = ru[:N, j, i] / (-0.5 * (Hz[:, j, i] + Hz[:, j, i - 1]) * dx)
        assert np.ptp(phi) <= 1e-7 * max(abs(phi).max(), 1e-30)
    # and the levels do slope here (the check is not vacuous)
    z_r = M.z_r.reshape(N, szJ, szI)
    assert abs(z_r[0, j, 5] - z_r[0, j, 4]) > 1.0


def _tracer(M, g, t, Hu, Hv, Hz0, Hz1, dt):
    tn = np.zeros_like(t)
    O.lib().sfo_step3d_t(M.cp, P(t), P(tn), P(Hu), P(Hv), P(Hz0), P(Hz1), dt)
    return tn


def _fluxes(g, seed, N, exchange=False):
    rng = np.random.default_rng(seed)
    S2, W = g.S2, g.szI
    n = (N + 1) * S2
    Hu, Hv = 50.0 * rng.standard_normal(n), 50.0 * rng.standard_normal(n)
    if exchange:      # every column's transport sums to zero over the levels
        Hu = Hu.reshape(N + 1, S2)
        Hv = Hv.reshape(N + 1, S2)
        Hu[:N] -= Hu[:N].mean(axis=0)
        Hv[:N] -= Hv[:N].mean(axis=0)
        Hu, Hv = Hu.ravel(), Hv.ravel()
    Hu3, Hv3 = Hu.reshape(N + 1, g.szJ, W), Hv.reshape(N + 1, g.szJ, W)
    Hu3[:, :, :2] = 0.0       # closed walls: no flux through the faces of columns 1 and L
    Hu3[:, :, W - 1] = 0.0
    Hv3[:, :2, :] = 0.0
    Hv3[:, g.szJ - 1, :] = 0.0
    return Hu, Hv


def test_tracer_uniform_stays_uniform():
    szI, szJ, N, dt = 18, 12, 7, 30.0
    g, M = model(szI, szJ, N, 1e3, 1e3, np.full((szJ, szI), 60.0))
    rng = np.random.default_rng(1)
    Hu, Hv = _fluxes(g, 2, N)
    Hz0 = M.Hz.copy()
    Hz1 = Hz0 * (1.0 + 0.01 * rng.standard_normal(Hz0.size))
    t = np.full((N + 1) * g.S2, 12.5)
    tn = _tracer(M, g, t, Hu, Hv, Hz0, Hz1, dt)
    inner = tn.reshape(N + 1, szJ, szI)[:N]
    assert np.abs(inner - 12.5).max() <= 1e-12 * 12.5


def test_tracer_conserved_when_columns_close_and_no_new_extrema():
    szI, szJ, N, dt, dx = 20, 14, 8, 20.0, 1e3
    g, M = model(szI, szJ, N, dx, dx, np.full((szJ, szI), 60.0))
    rng = np.random.default_rng(5)
    Hu, Hv = _fluxes(g, 6, N, exchange=True)
    Hz = M.Hz.copy()
    t = (10.0 + rng.random((N + 1) * g.S2))
    tn = _tracer(M, g, t, Hu, Hv, Hz, Hz, dt)
    Hz3 = Hz.reshape(N, szJ, szI)
    content = lambda a: (a.reshape(N + 1, szJ, szI)[:N, 1:-1, 1:-1] * Hz3[:, 1:-1, 1:-1]).sum() * dx * dx
    assert abs(content(tn) - content(t)) <= 1e-12 * content(t)
    inner_t = t.reshape(N + 1, szJ, szI)[:N]
    inner_n = tn.reshape(N + 1, szJ, szI)[:N, 1:-1, 1:-1]
    assert inner_n.max() <= inner_t.max() + 1e-12 and inner_n.min() >= inner_t.min() - 1e-12
    assert np.abs(inner_n - inner_t[:, 1:-1, 1:-1]).max() > 1e-6      # the fluxes did move the tracer
