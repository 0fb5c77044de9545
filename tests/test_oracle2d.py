"""Pins of the CPU oracle's 2D kernels (oracle/sf_oracle2d.c) that do not reuse its
formulas: closed forms on analytic fields, the telescoping (flux-form) budget of
the divergence, the exact mask patterns, and deadness of the buffers the fused GPU
substep does not write.  The reference ships no golden vectors for this path."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from sailfish_amd import configs
from sailfish_amd.mod_comptimes import CompTimes
from sailfish_amd.mod_grid import update_metrics
from sailfish_amd.set_weights import set_weights, weight_moments

G = 10.0


def grid(szI, szJ, dx=1e3, dy=1.5e3, h=None, rng=None, vary=False, fix_q5=False):
    pm = np.full((szJ, szI), 1.0 / dx)
    pn = np.full((szJ, szI), 1.0 / dy)
    if vary:
        pm = pm * (1 + 0.2 * rng.random((szJ, szI)))
        pn = pn * (1 + 0.2 * rng.random((szJ, szI)))
    f = np.zeros((szJ, szI))
    m = update_metrics(pm, pn, f)
    if h is None:
        h = np.full((szJ, szI), 50.0)
    return O.OGrid(szI, szJ, 1, dict(h=h, pm=pm, pn=pn, on_u=m["on_u"], om_v=m["om_v"]), fix_q5=fix_q5)


def P(a):
    return O.ptr(a)


def test_metrics_closed_form():
    """updateMetrics on a uniform grid: on_u = dy except row 0, om_v = dx except
    column 0 (mod_grid.py:222-232), fomn = f*dx*dy."""
    pm = np.full((5, 7), 1 / 2e3)
    pn = np.full((5, 7), 1 / 3e3)
    m = update_metrics(pm, pn, np.full((5, 7), 1e-4))
    assert np.all(m["on_u"][1:] == 3e3) and np.all(m["on_u"][0] == 0)
    assert np.all(m["om_v"][:, 1:] == 2e3) and np.all(m["om_v"][:, 0] == 0)
    assert np.allclose(m["fomn"], 1e-4 * 2e3 * 3e3, rtol=1e-15)


def test_zeta_rhs_linear_flux_closed_form():
    """ubar = a*i (linear in xi), vbar = b*j, zeta = 0, h = H uniform: the divergence
    on interior points is (a*H*dy + b*H*dx)/(dx*dy) exactly up to rounding; a swapped
    metric, index or sign changes it."""
    szI, szJ, dx, dy, H = 12, 9, 1e3, 1.5e3, 50.0
    g = grid(szI, szJ, dx, dy)
    a, b = 0.01, -0.003
    I, J = np.meshgrid(np.arange(szI), np.arange(szJ))
    ubar = np.ascontiguousarray((a * I).ravel())
    vbar = np.ascontiguousarray((b * J).ravel())
    zeta = np.zeros(szI * szJ)
    res = np.full(szI * szJ, np.nan)
    z3 = [np.zeros(szI * szJ) for _ in range(3)]
    O.lib().sfo_computeZetaRHS(g.cp, P(zeta), P(g.h), P(ubar), P(vbar), P(z3[0]), P(z3[1]), P(z3[2]), 1, 0.0, P(res))
    r = res.reshape(szJ, szI)
    want = (a * H * dy + b * H * dx) / (dx * dy)
    inner = r[1:-1, 1:-1]
    # row 1 reads om_v(1,i) and om_v(2,i): both dx; column 1 reads on_u(j,1..2): dy
    assert np.allclose(inner, want, rtol=1e-12, atol=0)
    assert np.all(np.isnan(r[0])) and np.all(np.isnan(r[-1])) and np.all(np.isnan(r[:, 0])) and np.all(np.isnan(r[:, -1]))


def test_zeta_rhs_telescoping_budget():
    """Flux form: sum over interior points of rzeta/(pm*pn) equals the net boundary
    flux DU*on_u at i=L minus i=1 plus DV*om_v at j=M minus j=1 (random fields,
    non-uniform metrics).  Pins the face indexing of divUVtoR (h:1045-1049)."""
    rng = np.random.default_rng(1)
    szI, szJ = 17, 11
    g = grid(szI, szJ, rng=rng, vary=True, h=40 + 10 * rng.random((11, 17)))
    zeta = 0.1 * rng.standard_normal(szI * szJ)
    ubar = rng.standard_normal(szI * szJ)
    vbar = rng.standard_normal(szI * szJ)
    res = np.zeros(szI * szJ)
    z3 = [np.zeros(szI * szJ) for _ in range(3)]
    O.lib().sfo_computeZetaRHS(g.cp, P(zeta), P(g.h), P(ubar), P(vbar), P(z3[0]), P(z3[1]), P(z3[2]), 1, 0.0, P(res))
    pm, pn = g.f["pm"].reshape(szJ, szI), g.f["pn"].reshape(szJ, szI)
    on_u, om_v = g.f["on_u"].reshape(szJ, szI), g.f["om_v"].reshape(szJ, szI)
    D = (zeta + g.h).reshape(szJ, szI)
    U, V = ubar.reshape(szJ, szI), vbar.reshape(szJ, szI)
    lhs = np.sum(res.reshape(szJ, szI)[1:-1, 1:-1] / (pm * pn)[1:-1, 1:-1])
    fx = lambda j, i: U[j, i] * 0.5 * (D[j, i] + D[j, i - 1]) * on_u[j, i]
    fy = lambda j, i: V[j, i] * 0.5 * (D[j, i] + D[j - 1, i]) * om_v[j, i]
    rhs = sum(fx(j, szI - 1) - fx(j, 1) for j in range(1, szJ - 1)) + \
        sum(fy(szJ - 1, i) - fy(1, i) for i in range(1, szI - 1))
    assert abs(lhs - rhs) <= 1e-10 * (abs(rhs) + np.abs(res).sum())


def test_momentum_rhs_pgf_closed_form():
    """zeta linear in xi (gz = s*x) with uniform h: the pressure-gradient RHS is
    g*on_u*(H*s*dx + 0.5*s^2*dx^2*(2i-1)) at u-point i, and the eta component is 0.
    The sign of the reference (Q1: +g...) is part of the pin."""
    szI, szJ, dx, dy, H = 10, 7, 1e3, 2e3, 30.0
    g = grid(szI, szJ, dx, dy, h=np.full((7, 10), H))
    s = 1e-5
    I, J = np.meshgrid(np.arange(szI), np.arange(szJ))
    z = np.ascontiguousarray((s * I * dx).ravel())
    ru = np.full(szI * szJ, np.nan)
    rv = np.full(szI * szJ, np.nan)
    O.lib().sfo_computeMomentumRHSPred(g.cp, P(g.h), P(ru), P(rv), P(z), P(z), G, 0.4)
    ru = ru.reshape(szJ, szI)[1:-1, 1:-1]
    rv = rv.reshape(szJ, szI)[1:-1, 1:-1]
    ii = np.arange(1, szI - 1)[None, :]
    want = G * dy * (H * s * dx + 0.5 * s * s * dx * dx * (2 * ii - 1))
    assert np.allclose(ru, np.broadcast_to(want, ru.shape), rtol=1e-12)
    assert np.all(rv == 0.0)


def test_momentum_pred_mask_patterns():
    """computeMomentumPred writes exactly the points with f%szJ != 0 and f/szJ != 0
    (cpp:162, Q5), or with the fixed mask exactly i != 0 and j != 0."""
    szI, szJ = 12, 4
    for fix in (False, True):
        g = grid(szI, szJ, fix_q5=fix)
        n = szI * szJ
        z = np.zeros(n)
        u1 = np.ones(n)
        u2 = np.full(n, np.nan)
        v2 = np.full(n, np.nan)
        zero = np.zeros(n)
        O.lib().sfo_computeMomentumPred(g.cp, 1.0, P(u1), P(u2), P(u1), P(v2), P(zero), P(zero), P(g.h), P(z), P(z))
        f = np.arange(n)
        st = szI if fix else szJ
        written = ~np.isnan(u2)
        assert np.array_equal(written, (f % st != 0) & (f // st != 0))
        # rhs = 0 and zeta_t1 == zeta_t2 away from row 0: u_t2 = u_t1 (within 1 ulp)
        ok = written & (f >= szI)
        assert np.allclose(u2[ok], 1.0, rtol=2e-16, atol=0)


def test_am3_constant_rhs_and_reference_coefficients():
    """x + Dt*(5/12 r2 + 8/12 r1 - 1/12 r0): constant r gives x + Dt*r; distinct
    r0, r1, r2 pin which slot gets which weight."""
    g = grid(5, 4)
    n = 20
    x = np.ones(n)
    r0, r1, r2 = np.full(n, 12.0), np.full(n, 24.0), np.full(n, 36.0)
    O.lib().sfo_AdamsMoultonCorr3rd(g.cp, 0.5, P(x), P(r0), P(r1), P(r2))
    assert np.allclose(x, 1 + 0.5 * (5 * 3 + 8 * 2 - 1 * 1), rtol=1e-15)


def test_weights_self_reported_invariants():
    """set_weights.py:205-231: sum w1 = 1, sum w2 = 1, first moment 1, second moment
    >= 1.0001, secondary centroid ~ 1/2; ROMS reports nfast = 42 for ndtfast = 30."""
    for nd, nf_expected in ((30, 42), (40, None), (20, None)):
        ct = CompTimes(300.0, nd)
        w1, w2, nf = set_weights(ct)
        c1, c2, shift, s1, s2 = weight_moments(w1, w2, nf, nd)
        assert abs(s1 - 1) < 1e-14 and abs(s2 - 1) < 1e-14 and abs(c1 - 1) < 1e-12
        assert c2 >= 1.0001 and 0.45 < shift < 0.55
        # the -Fgamma*xi term makes the first weights negative; weights after the first
        # positive one are clipped at zero (set_weights.py:103-104)
        first_pos = int(np.argmax(w1 > 0))
        assert np.all(w1[nf:] == 0) and np.all(w1[first_pos:] >= 0)
        if nf_expected:
            assert nf == nf_expected


def _model(cfg, fix_q5=True):
    from tests.helpers2d import oracle_model
    ct = CompTimes(cfg["dtfast"] * 4, 4)
    set_weights(ct)
    return oracle_model(cfg, ct.weight1, ct.nfast, fix_q5=fix_q5)


@pytest.mark.parametrize("fix_q5", [True, False])
def test_buffers_skipped_by_fused_kernel_are_dead(fix_q5):
    """Scrambling the values the fused substep does not write (interior rvbar in
    every slot, the corrector's interior rubar slot, and ubar_t1 at unmasked points
    after the corrector's AM3 update) leaves every read state identical: zeta slots,
    ubar_t2, vbar slots, rzeta, the averages."""
    cfg = configs.random_smooth(40, 20, seed=3)
    t1, g1, s1, b1 = _model(cfg, fix_q5)
    t2, g2, s2, b2 = _model(cfg, fix_q5)
    rng = np.random.default_rng(0)
    n = g1.S2
    f = np.arange(n)
    st = g1.szI if fix_q5 else g1.szJ
    unmasked = (f % st != 0) & (f // st != 0)
    i, j = f % g1.szI, f // g1.szI
    inner = (i > 0) & (i < g1.szI - 1) & (j > 0) & (j < g1.szJ - 1)
    ni = int(inner.sum())
    for _ in range(6):
        for t, s in ((t1, s1), (t2, s2)):
            t.iif += 1
            t.time2D += t.dtfast
            s.cycleTimes2D()
            O.step2d_predictor(t, g1 if s is s1 else g2, s, b1 if s is s1 else b2)
        # second model: scramble the dead buffers between predictor and corrector
        # interior values only: no kernel ever writes the non-interior ones, which stay
        # at their initial value in both paths and are read as such
        s2.v("rvbar", 0)[inner] = rng.standard_normal(ni)
        s2.v("rvbar", 2)[inner] = rng.standard_normal(ni)
        for t, s, g, b in ((t1, s1, g1, b1), (t2, s2, g2, b2)):
            O.step2d_corrector(t, g, s, b)
        s2.v("rubar", 2)[inner] = rng.standard_normal(ni)
        s2.v("rvbar", 1)[inner] = rng.standard_normal(ni)
        u1 = s2.v("ubar", 1)
        u1[unmasked] = rng.standard_normal(int(unmasked.sum()))
    for name in ("zeta", "rzeta", "vbar"):
        for k in range(3):
            a, b = s1.v(name, k), s2.v(name, k)
            assert np.array_equal(a, b) or (np.isnan(a).any() and np.array_equal(np.isnan(a), np.isnan(b)))
    assert np.array_equal(s1.v("ubar", 2), s2.v("ubar", 2))
    for n_ in ("Zt_avg1", "DU_avg1", "DV_avg1"):
        assert np.array_equal(getattr(s1, n_), getattr(s2, n_))
