"""Pins of the oracle's 3D kernels (oracle/sf_oracle3d.c): closed forms on analytic
fields (uniform and linear), exact identities (column sums, telescoping), and a
dense linear solve for the tridiagonal step.  No golden vectors exist upstream."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from oracle import oracle3d as O3
from sailfish_amd.mod_grid import update_metrics
from sailfish_amd.set_coords import s_coordinates

P = O.ptr


def model(szI=16, szJ=12, N=8, dx=1e3, dy=1.2e3, h=None, f0=1e-4, hc=50.0, Vt=2):
    pm = np.full((szJ, szI), 1 / dx)
    pn = np.full((szJ, szI), 1 / dy)
    f = np.full((szJ, szI), f0)
    m = update_metrics(pm, pn, f)
    if h is None:
        h = np.full((szJ, szI), 100.0)
    g = O.OGrid(szI, szJ, N, dict(h=h, pm=pm, pn=pn, on_u=m["on_u"], om_v=m["om_v"]))
    sc_r, sc_w, Cs_r, Cs_w = s_coordinates(N, 4, 5.0, 0.4)
    BC = -np.ones(szI * szJ, np.int32)
    M = O3.OModel3(g, N, m["fomn"], sc_r, sc_w, Cs_r, Cs_w, hc, Vt, BC, BC, BC)
    return g, M, m


def test_set_depth_rest_state_closed_form():
    """zeta = 0: z_w = h*(hc*s + h*C)/(hc + h) (Vtransform 2, set_coords.py:290-300),
    z_w(N) = 0, column thickness sum = h."""
    rng = np.random.default_rng(0)
    h = 20 + 200 * rng.random((12, 16))
    g, M, _ = model(h=h)
    s = O.OState(g)
    M.set_depth(s)
    N, S2 = M.N, g.S2
    zw = M.z_w.reshape(N + 1, 12, 16)
    for k in range(N):
        want = h * (M.hc * M.sc_w[k] + h * M.Cs_w[k]) / (M.hc + h)
        assert np.allclose(zw[k], want, rtol=1e-13, atol=1e-12)
    assert np.all(zw[N] == 0.0)
    Hz = M.Hz.reshape(N, 12, 16)
    assert np.all(Hz > 0) and np.allclose(Hz.sum(0), h, rtol=1e-12)
    assert np.allclose(zw[0], -h, rtol=1e-13)


def test_set_depth_vtransform1_and_free_surface():
    """Vtransform 1 with zeta != 0: z = Zo + zeta*(1 + Zo/h), Zo = hc*(s - C) + C*h."""
    rng = np.random.default_rng(1)
    h = 60 + 100 * rng.random((12, 16))
    g, M, _ = model(h=h, Vt=1, hc=20.0)
    s = O.OState(g)
    s.Zt_avg1[:] = 0.3 * rng.standard_normal(g.S2)
    M.set_depth(s)
    N = M.N
    zeta = s.Zt_avg1.reshape(12, 16)
    zw = M.z_w.reshape(N + 1, 12, 16)
    for k in range(N):
        Zo = M.hc * (M.sc_w[k] - M.Cs_w[k]) + M.Cs_w[k] * h
        assert np.allclose(zw[k], Zo + zeta * (1 + Zo / h), rtol=1e-13, atol=1e-12)
    assert np.array_equal(zw[N], zeta)
    assert np.allclose(M.Hz.reshape(N, 12, 16).sum(0), h + zeta, rtol=1e-12)


def test_omega_divergence_free_is_zero_and_linear_flux():
    """Uniform fluxes: W = 0.  Huon = alpha*i at every level: W(K) = -alpha*K before
    the surface correction, then + alpha*(N-1)*(z_w(K)-z_w(0))/(z_w(N-1)-z_w(0))
    (step3D_uv.kernels.cpp:105-163)."""
    g, M, _ = model()
    s = O.OState(g)
    M.set_depth(s)
    N, S2, szI = M.N, g.S2, g.szI
    M.Huon[:] = 3.0
    M.Hvom[:] = -2.0
    O3.lib().sfo_omega(M.cp, P(M.W), P(M.Huon), P(M.Hvom), P(M.z_w), P(M.BC))
    W = M.W.reshape(N + 1, 12, 16)
    assert np.all(W[:, 1:-1, 1:-1] == 0.0)
    alpha = 0.7
    I = np.tile(np.arange(szI), 12)
    M.Huon[:] = np.tile(alpha * I, N + 1)
    M.Hvom[:] = 0.0
    O3.lib().sfo_omega(M.cp, P(M.W), P(M.Huon), P(M.Hvom), P(M.z_w), P(M.BC))
    W = M.W.reshape(N + 1, 12, 16)
    zw = M.z_w.reshape(N + 1, 12, 16)
    for K in range(1, N):
        want = -alpha * K + alpha * (N - 1) * (zw[K] - zw[0]) / (zw[N - 1] - zw[0])
        assert np.allclose(W[K][1:-1, 1:-2], want[1:-1, 1:-2], rtol=1e-12, atol=1e-12)
    assert np.all(W[0] == 0) and np.all(W[N] == 0)


def test_set_maxflux_closed_form():
    g, M, m = model()
    s = O.OState(g)
    M.set_depth(s)
    N, S2 = M.N, g.S2
    M.u[2][:] = 0.2
    M.v[2][:] = -0.1
    O3.lib().sfo_set_maxflux(M.cp, P(M.u[2]), P(M.v[2]), P(M.Huon), P(M.Hvom), P(M.Hz))
    Hz = M.Hz.reshape(N, 12, 16)
    Hu = M.Huon.reshape(N + 1, 12, 16)
    Hv = M.Hvom.reshape(N + 1, 12, 16)
    want_u = 0.5 * (Hz[:, 1:, 2:] + Hz[:, 1:, 1:-1]) * 0.2 * m["on_u"][1:, 2:]
    want_v = 0.5 * (Hz[:, 2:-1, 1:] + Hz[:, 1:-2, 1:]) * -0.1 * m["om_v"][2:-1, 1:]
    assert np.allclose(Hu[:N, 1:, 2:], want_u, rtol=1e-14)
    assert np.allclose(Hv[:N, 2:-1, 1:], want_v, rtol=1e-14)


def test_coriolis_uniform_flow():
    """u = U, v = V, uniform Hz: ru += Hz*fomn*V, rv -= Hz*fomn*U on R nodes."""
    g, M, m = model()
    N, S2 = M.N, g.S2
    M.Hz[:] = 7.0
    M.u[2][:] = 0.3
    M.v[2][:] = -0.2
    ru, rv = np.zeros((N + 1) * S2), np.zeros((N + 1) * S2)
    O3.lib().sfo_addCoriolis(M.cp, P(M.fomn), P(M.u[2]), P(M.v[2]), P(ru), P(rv), P(M.Hz))
    F = m["fomn"][0, 0]
    ru = ru.reshape(N + 1, 12, 16)
    rv = rv.reshape(N + 1, 12, 16)
    assert np.allclose(ru[:N, 1:-1, 1:-1], 7.0 * F * -0.2, rtol=1e-14)
    assert np.allclose(rv[:N, 1:-1, 1:-1], -7.0 * F * 0.3, rtol=1e-14)
    assert np.all(ru[:, 0, :] == 0) and np.all(ru[:, :, 0] == 0) and np.all(ru[N] == 0)


def test_horizontal_advection_linear_u():
    """u = a*i, v = 0, uniform Huon/Hvom: curvature terms vanish and the xi flux is
    uR = a*(i+1/2), so ru -= a on interior points (rhs_kernels.cpp:121-128); rv = 0.
    A mirrored stencil would give +a."""
    g, M, _ = model(szI=20, szJ=14, N=6)
    N, S2, szI = M.N, g.S2, g.szI
    a = 0.01
    I = np.tile(np.arange(szI, dtype=float), 14)
    M.u[2][:] = np.tile(a * I, N + 1)
    M.v[2][:] = 0.0
    M.Huon[:] = 4.0
    M.Hvom[:] = 3.0
    ru, rv = np.zeros((N + 1) * S2), np.zeros((N + 1) * S2)
    O3.lib().sfo_horizontalAdvection(M.cp, P(M.u[2]), P(M.v[2]), P(M.Huon), P(M.Hvom), P(ru), P(rv), S2)
    ru = ru.reshape(N + 1, 14, 20)
    rv = rv.reshape(N + 1, 14, 20)
    assert np.allclose(ru[1:N, 3:-3, 3:-3], -a, rtol=1e-12)
    assert np.all(np.abs(rv[1:N, 3:-3, 3:-3]) == 0)


def test_horizontal_advection_tail_zeroing():
    """Threads past S2 of the reference launch zero ru/rv at their flat index."""
    g, M, _ = model(szI=20, szJ=14, N=6)
    N, S2 = M.N, g.S2
    ru, rv = np.ones((N + 1) * S2), np.ones((N + 1) * S2)
    O3.lib().sfo_horizontalAdvection(M.cp, P(M.u[2]), P(M.v[2]), P(M.Huon), P(M.Hvom), P(ru), P(rv), S2 + 37)
    assert np.all(ru[S2:S2 + 37] == 0) and ru[S2 + 37] != 0


def test_vertical_advection_linear_profile():
    """u = b*K, v = c*K, W uniform: the 4th-order centred flux difference is W*b, so
    ru(K) -= W*b for K = 2..N-3; levels N-2..N copy level N-3; rv(0) = rv(1)
    (rhs_kernels.cpp:158-194)."""
    g, M, _ = model(szI=16, szJ=12, N=10)
    N, S2 = M.N, g.S2
    b, c, Wc = 0.02, -0.03, 0.5
    lev = np.repeat(np.arange(N + 1, dtype=float), S2)
    M.u[2][:] = b * lev
    M.v[2][:] = c * lev
    M.W[:] = Wc
    ru, rv = np.zeros((N + 1) * S2), np.zeros((N + 1) * S2)
    O3.lib().sfo_verticalAdvection(M.cp, P(M.u[2]), P(M.v[2]), P(M.W), P(ru), P(rv), P(M.BC))
    ru = ru.reshape(N + 1, 12, 16)[:, 2:-2, 2:-2]
    rv = rv.reshape(N + 1, 12, 16)[:, 2:-2, 2:-2]
    assert np.allclose(ru[2:N - 2], -Wc * b, rtol=1e-12)
    assert np.allclose(rv[2:N - 2], -Wc * c, rtol=1e-12)
    for K in (N - 2, N - 1, N):
        assert np.array_equal(ru[K], ru[N - 3])
    assert np.array_equal(rv[0], rv[1])


@pytest.mark.parametrize("iic,cff", [(0, 1.0), (1, 1.5), (5, 23.0 / 12.0)])
def test_step3d_uv_tridiagonal_against_dense_solve(iic, cff):
    """One column of step3d_UV against numpy: assemble the reference's matrix
    (createVertViscousOpMatrix, :568-639) and solve densely; the Thomas sweep
    (:520-565) plus the depth-mean correction (:475-517) must agree to 1e-12.  The
    time-step factor is 1, 3/2, 23/12 at iic = ntfirst, ntfirst+1, later (:652-660)."""
    rng = np.random.default_rng(3)
    szI, szJ, N = 8, 7, 9
    g, M, m = model(szI, szJ, N, h=80 + 40 * rng.random((7, 8)))
    s = O.OState(g)
    M.set_depth(s)
    S2 = g.S2
    n3 = (N + 1) * S2
    tmpU = 0.1 * rng.standard_normal(n3)
    tmpV = 0.1 * rng.standard_normal(n3)
    M.u[2][:] = tmpU
    M.v[2][:] = tmpV
    M.ru[2][:] = rng.standard_normal(n3)
    M.rv[2][:] = rng.standard_normal(n3)
    M.AKv[:] = 1e-2 * (1 + rng.random(n3))
    s.DU_avg1[:] = 50 * rng.standard_normal(S2)
    s.DV_avg1[:] = 50 * rng.standard_normal(S2)
    dt, lam = 300.0, 1.0
    O3.lib().sfo_step3d_UV(M.cp, P(M.u[2]), P(M.v[2]), P(M.ru[2]), P(M.rv[2]), P(s.ubar[1]), P(s.vbar[1]),
                           P(s.ubar[2]), P(s.vbar[2]), P(M.Hz), P(M.AKv), P(M.z_r), P(s.DU_avg1), P(s.DV_avg1),
                           P(tmpU), P(tmpV), iic, 0, lam, dt, P(M.eqSD), P(M.eqD), P(M.eqRHS))
    f = 3 * szI + 4      # an interior U and V node
    Hz = M.Hz.reshape(N, S2)[:, [f, f - 1]]
    HzU = 0.5 * (Hz[:, 0] + Hz[:, 1])
    Ak = M.AKv.reshape(N + 1, S2)[:, [f, f - 1]]
    AkU = 0.5 * (Ak[:, 0] + Ak[:, 1])
    zr = M.z_r.reshape(N, S2)[:, [f, f - 1]]
    dz = 0.5 * ((zr[1:, 0] - zr[:-1, 0]) + (zr[1:, 1] - zr[:-1, 1]))     # dz[K] between K and K+1
    pmn = 0.25 * (g.f["pm"][f] + g.f["pm"][f - 1]) * (g.f["pn"][f] + g.f["pn"][f - 1])
    FC = -lam * dt * AkU[:N - 1] / dz
    A = np.zeros((N, N))
    rhs = np.zeros(N)
    for K in range(N):
        A[K, K] = HzU[K] - (FC[K] if K < N - 1 else 0) - (FC[K - 1] if K > 0 else 0)
        if K < N - 1:
            A[K, K + 1] = A[K + 1, K] = FC[K]
        if K > 0:
            rhs[K] = HzU[K] * (tmpU.reshape(N + 1, S2)[K, f] + cff * dt * pmn * M.ru[2].reshape(N + 1, S2)[K, f])
    A[0, 0] = HzU[0] - 2 * FC[0]
    x = np.linalg.solve(A, rhs)
    Hsum = HzU.sum()
    diff = ((tmpU.reshape(N + 1, S2)[:N, f] * HzU).sum() - s.DU_avg1[f] / g.f["on_u"][f]) / Hsum
    got = M.u[2].reshape(N + 1, S2)[:N, f]
    assert np.allclose(got, x - diff, rtol=1e-10, atol=1e-12)
    # barotropic coupling: ubar_t2 = DU_avg1/on_u/H + (sum Hz*u - DU_avg1/on_u)/H
    HuU = (HzU * got).sum()
    ub = s.DU_avg1[f] / g.f["on_u"][f] / Hsum + (HuU - s.DU_avg1[f] / g.f["on_u"][f]) / Hsum
    assert abs(s.ubar[2][f] - ub) <= 1e-12 * max(1, abs(ub)) and s.ubar[1][f] == s.ubar[2][f]


# ------------------------------------------------------------------------------------
# Closed forms of the curvature (upwind) terms, the special levels of the vertical
# advection, and the lateral boundary values, derived by hand from the reference text
# (nonlinear/rhs_kernels.cpp:70-197, modules/mod_cppkernels.h:608-702 and :915-960,
# nonlinear/step3D_uv.kernels.cpp:294-368).
# ------------------------------------------------------------------------------------

def _hadv(szI, szJ, N, u2d, Hu, Hv, ru0=0.5, rv0=-0.25):
    g, M, _ = model(szI=szI, szJ=szJ, N=N)
    S2 = g.S2
    M.u[2][:] = np.tile(np.asarray(u2d, float).ravel(), N + 1)
    M.v[2][:] = 0.0
    M.Huon[:] = Hu
    M.Hvom[:] = Hv
    ru, rv = np.full((N + 1) * S2, ru0), np.full((N + 1) * S2, rv0)
    O3.lib().sfo_horizontalAdvection(M.cp, P(M.u[2]), P(M.v[2]), P(M.Huon), P(M.Hvom), P(ru), P(rv), S2)
    return ru.reshape(N + 1, szJ, szI), rv.reshape(N + 1, szJ, szI)


def test_horizontal_advection_xi_curvature_closed_form():
    """u = a*i^4*(1 + e*j), v = 0, uniform Huon = Hu, Hvom = Hv.  uR > 0 selects
    DXX(u)(i) = a*g*(12 i^2 + 2) (upwindUtoR, h:608-625), so with Gadv = -1/2:
      DX(UFxR) = a*g*4*(i^3 + i) - 6*a*g*Hu*(2i - 1)
      DE(UFeP) = DE(uP) = a*e*(i^4 + (i-1)^4)/2          (uP = UtoP(u), Q10)
      DE(VFeR) = -(1/2)*Hv*a*e*(12 i^2 + 2)               (u curvature in VFeR, Q11)
    ru = ru0 - DX(UFxR) - DE(UFeP), rv = rv0 - DE(VFeR) on interior levels and points."""
    szI, szJ, N, a, e, Hu, Hv = 16, 12, 4, 1e-4, 0.1, 3.0, -2.0
    I, J = np.meshgrid(np.arange(szI, dtype=float), np.arange(szJ, dtype=float))
    gj = 1 + e * J
    ru, rv = _hadv(szI, szJ, N, a * I ** 4 * gj, Hu, Hv)
    want_u = 0.5 - (a * gj * 4 * (I ** 3 + I) - 6 * a * gj * Hu * (2 * I - 1) + a * e * (I ** 4 + (I - 1) ** 4) / 2)
    want_v = -0.25 + 0.5 * Hv * a * e * (12 * I ** 2 + 2)
    sl = (slice(0, N), slice(3, -3), slice(3, -3))
    assert np.allclose(ru[sl], np.broadcast_to(want_u, ru.shape)[sl], rtol=1e-10, atol=1e-12)
    assert np.allclose(rv[sl], np.broadcast_to(want_v, rv.shape)[sl], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("Hv", [2.0, -2.0])
def test_horizontal_advection_eta_curvature_upwind_closed_form(Hv):
    """u = a*j^4*(1 + d*i), v = 0: DXX(u) = 0, DEE(u) = a*G(i)*(12 j^2 + 2), G = 1 + d*i.
    upwindUtoP (h:660-677) takes DEE(u) at j when Hvom >= 0, at j-1 otherwise: s = j or
    j - 1.  Then
      DX(UFxR) = a*j^4*d
      DE(UFeP) = a*(1 + d*(i - 1/2))*(4j^3 - 6j^2 + 4j - 1) - (1/2)*Hv*a*G(i)*12*(s^2 - (s-1)^2)
    and VFxP = -(1/2)*uetaP*Hu (the u curvature in VFxP, Q11) gives
      DX(VFxP) = -(1/2)*Hu*a*d*(12 s^2 + 2)."""
    szI, szJ, N, a, d, Hu = 14, 16, 4, 1e-4, 0.1, 3.0
    I, J = np.meshgrid(np.arange(szI, dtype=float), np.arange(szJ, dtype=float))
    G = 1 + d * I
    ru, rv = _hadv(szI, szJ, N, a * J ** 4 * G, Hu, Hv)
    s = J if Hv >= 0 else J - 1
    dcurv = 12 * (s ** 2 - (s - 1) ** 2)          # DE of 12 s^2
    de_uf = a * (1 + d * (I - 0.5)) * (4 * J ** 3 - 6 * J ** 2 + 4 * J - 1) - 0.5 * Hv * a * G * dcurv
    want_u = 0.5 - (a * J ** 4 * d + de_uf)
    want_v = -0.25 + 0.5 * Hu * a * d * (12 * s ** 2 + 2)
    sl = (slice(0, N), slice(3, -3), slice(3, -3))
    assert np.allclose(ru[sl], np.broadcast_to(want_u, ru.shape)[sl], rtol=1e-10, atol=1e-12)
    assert np.allclose(rv[sl], np.broadcast_to(want_v, rv.shape)[sl], rtol=1e-10, atol=1e-12)


def test_vertical_advection_special_levels_closed_form():
    """u = b*(K + c), v = bv*(K + cv), W uniform: F(K) = W*b*(K + 1/2 + c) for K >= 1.
    The ad-hoc levels of rhs_kernels.cpp:169-194 then give
      ru(1) = ru0 - W*b*(33 - c)/16      (F(2) - uuu, uuu = W*b*(17c + 7)/16)
      ru(0) = ru0 - W*b*(9c + 8)/16      (uuu - uuu2, uuu2 = W*b*(8c - 1)/16)
      rv(1) = rv0 - W*bv*(16cv + 31)/16, rv(0) = rv(1)
    and levels N-2, N-1, N copy level N-3 (the N-2 update is overwritten)."""
    g, M, _ = model(szI=16, szJ=12, N=10)
    N, S2 = M.N, g.S2
    b, c, bv, cv, Wc, ru0, rv0 = 0.02, 1.5, -0.03, 0.75, 0.5, 0.125, -0.375
    lev = np.repeat(np.arange(N + 1, dtype=float), S2)
    M.u[2][:] = b * (lev + c)
    M.v[2][:] = bv * (lev + cv)
    M.W[:] = Wc
    ru, rv = np.full((N + 1) * S2, ru0), np.full((N + 1) * S2, rv0)
    O3.lib().sfo_verticalAdvection(M.cp, P(M.u[2]), P(M.v[2]), P(M.W), P(ru), P(rv), P(M.BC))
    ru = ru.reshape(N + 1, 12, 16)[:, 2:-2, 2:-2]
    rv = rv.reshape(N + 1, 12, 16)[:, 2:-2, 2:-2]
    assert np.allclose(ru[1], ru0 - Wc * b * (33 - c) / 16, rtol=1e-13)
    assert np.allclose(ru[0], ru0 - Wc * b * (9 * c + 8) / 16, rtol=1e-13)
    assert np.allclose(ru[2:N - 2], ru0 - Wc * b, rtol=1e-13)
    assert np.allclose(rv[1], rv0 - Wc * bv * (16 * cv + 31) / 16, rtol=1e-13)
    assert np.array_equal(rv[0], rv[1])
    for K in (N - 2, N - 1, N):
        assert np.array_equal(ru[K], ru[N - 3]) and np.array_equal(rv[K], rv[N - 3])


def test_set_zeta_closed_form():
    """set_zeta (:294-310): zeta_t1 = zeta_t2 = Zt_avg1 on rho nodes (i > 0, j > 0);
    row 0 and column 0 keep their values."""
    g, M, _ = model(szI=9, szJ=7, N=4)
    S2 = g.S2
    rng = np.random.default_rng(8)
    Zt = rng.standard_normal(S2)
    z1, z2 = np.full(S2, np.nan), np.full(S2, np.nan)
    O3.lib().sfo_set_zeta(M.cp, P(z1), P(z2), P(Zt))
    f = np.arange(S2)
    r = (f % 9 > 0) & (f // 9 > 0)
    assert np.array_equal(z1[r], Zt[r]) and np.array_equal(z2[r], Zt[r])
    assert np.all(np.isnan(z1[~r])) and np.all(np.isnan(z2[~r]))


def test_lateral_uv_bcs_clamp_value():
    """setLateralUVBCs (:314-368): clamped u boundary points take 0.05*sin(2e-4*t) at
    every level 0..N-1 (t = pi/(2*2e-4) gives 0.05, t = 1000 gives 0.05*sin(0.2));
    closed v points take 0; other points are untouched."""
    g, M, _ = model(szI=9, szJ=7, N=4)
    N, S2 = M.N, g.S2
    uType = np.zeros(S2, np.int32)
    vType = np.zeros(S2, np.int32)
    cu = np.array([9 * 2 + 1, 9 * 3 + 1, 9 * 4 + 1])
    cv = np.array([9 * 1 + 3, 9 * 1 + 4])
    uType[cu] = (1 << 3) + 1            # bcClamped (mod_boundary.py:28-57)
    vType[cv] = 1 << 4                  # bcClosed
    BC = -np.ones(S2, np.int32)
    for t, val in ((np.pi / (2 * 2e-4), 0.05), (1000.0, 0.05 * np.sin(0.2))):
        u = np.full((N + 1) * S2, 7.0)
        v = np.full((N + 1) * S2, 7.0)
        rc = O3.lib().sfo_setLateralUVBCs(M.cp, t, P(u), P(v), P(BC), P(BC), P(uType), P(vType))
        assert rc == 0
        u, v = u.reshape(N + 1, S2), v.reshape(N + 1, S2)
        assert np.allclose(u[:N, cu], val, rtol=1e-15) and np.all(u[N, cu] == 7.0)
        assert np.all(v[:N, cv] == 0.0)
        mask = np.ones(S2, bool)
        mask[cu] = False
        assert np.all(u[:, mask] == 7.0)
