"""Pins of the coupled split-explicit oracle (oracle/sf_oracle_phys.c
sfo_phys_substep_coupled, oracle/sf_oracle3d.c sfo_couple_rufrc and
sfo_step3d_uv_coupled, oracle/oracle_coupled.py) by closed forms and invariants:

- rufrc of level-wise constant right-hand sides with wind and bottom drag;
- a coupled substep whose forcing is zero is the plain substep, bit for bit;
- the first predictor of a baroclinic step adds the Adams-Bashforth extrapolation of
  the forcing (all three orders), every other stage the forcing itself;
- the DU_avg2 weights of a barotropic loop sum to one, and a nearly frozen state
  averages to its own flux;
- after the coupled step3d_UV each column carries the averaged transport DU_avg1
  exactly, ubar = DU_avg1 / D, the mass flux sums to DU_avg2, and u differs from the
  uncoupled solve by a shift that is the same at every level."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from oracle import oracle3d as O3
from oracle import oracle_coupled as OC
from oracle import oracle_phys as OP
from sailfish_amd.mod_grid import update_metrics
from sailfish_amd.set_coords import s_coordinates
from sailfish_amd.set_weights import set_weights
from sailfish_amd.mod_comptimes import CompTimes

P = O.ptr


def flat(szI, szJ, dx, dy, H, f0=0.0):
    h = np.full((szJ, szI), H)
    return dict(szI=szI, szJ=szJ, h=h, pm=np.full_like(h, 1 / dx), pn=np.full_like(h, 1 / dy),
                f=np.full_like(h, f0), zeta0=np.zeros_like(h))


def grid3(cfg, N):
    m = update_metrics(cfg["pm"], cfg["pn"], cfg["f"])
    g = O.OGrid(cfg["szI"], cfg["szJ"], N, dict(h=cfg["h"], pm=cfg["pm"], pn=cfg["pn"], on_u=m["on_u"],
                                                 om_v=m["om_v"]))
    sc_r, sc_w, Cs_r, Cs_w = s_coordinates(N, 4, 5.0, 0.4)
    BC = -np.ones(g.S2, np.int32)
    return g, O3.OModel3(g, N, m["fomn"], sc_r, sc_w, Cs_r, Cs_w, 50.0, 2, BC, BC, BC), m


def test_rufrc_closed_form():
    szI, szJ, N, dx, dy = 12, 9, 5, 1e3, 1.5e3
    cfg = flat(szI, szJ, dx, dy, 80.0)
    g, M, _ = grid3(cfg, N)
    S2 = g.S2
    c = np.array([1.0, -2.0, 0.5, 3.0, 7.0])            # level constants (a dropped level shows)
    ru = np.concatenate([np.full(S2, ck) for ck in c] + [np.full(S2, 100.0)])    # level N is not summed
    rv = -2.0 * ru
    u = np.full((N + 1) * S2, 0.3)                       # bottom level 0.3, others ignored
    u[S2:] = 5.0
    v = np.full((N + 1) * S2, -0.2)
    v[S2:] = 9.0
    sustr, svstr, rdrg = np.full(S2, 2e-4), np.full(S2, -1e-4), 3e-4
    rufrc, rvfrc = np.zeros(S2), np.zeros(S2)
    O.lib().sfo_couple_rufrc(M.cp, P(ru), P(rv), P(u), P(v), P(sustr), P(svstr), rdrg, P(rufrc), P(rvfrc))
    rufrc, rvfrc = rufrc.reshape(szJ, szI), rvfrc.reshape(szJ, szI)
    ru_exp = c.sum() + (2e-4 - rdrg * 0.3) * dx * dy
    rv_exp = -2.0 * c.sum() + (-1e-4 - rdrg * (-0.2)) * dx * dy
    assert np.allclose(rufrc[1:szJ - 1, 2:szI - 1], ru_exp, rtol=1e-13)
    assert np.allclose(rvfrc[2:szJ - 1, 1:szI - 1], rv_exp, rtol=1e-13)
    # zero off the interior u / v points
    mask_u = np.zeros((szJ, szI), bool)
    mask_u[1:szJ - 1, 2:szI - 1] = True
    mask_v = np.zeros((szJ, szI), bool)
    mask_v[2:szJ - 1, 1:szI - 1] = True
    assert not rufrc[~mask_u].any() and not rvfrc[~mask_v].any()


def _couple(o, rufrc, rvfrc, F, F1, F2, DU2, DV2, w2p=0.0, w2c=0.0, ab=0):
    return OC.sfo_couple(P(rufrc), P(rvfrc), P(F[0]), P(F[1]), P(F1[0]), P(F1[1]), P(F2[0]), P(F2[1]), P(DU2),
                         P(DV2), w2p, w2c, ab)


def _state(o):
    return {n: getattr(o.st, n).copy() for n in ("zeta", "ubar", "vbar", "rzeta", "rubar", "rvbar")}


def _coupled_substep(o, dt, first, cp):
    v = o.st.v
    sc = o.scr
    O.lib().sfo_phys_substep_coupled(C.byref(o.c), dt, int(first), 0.0,
                                     P(v("zeta", 0)), P(v("zeta", 1)), P(v("zeta", 2)),
                                     P(v("ubar", 0)), P(v("ubar", 1)), P(v("ubar", 2)),
                                     P(v("vbar", 0)), P(v("vbar", 1)), P(v("vbar", 2)),
                                     P(v("rzeta", 0)), P(v("rzeta", 1)), P(v("rubar", 0)), P(v("rubar", 1)),
                                     P(v("rvbar", 0)), P(v("rvbar", 1)), P(sc[0]), P(sc[1]), P(sc[2]), P(sc[3]),
                                     P(o.st.Zt_avg1), P(o.st.DU_avg1), P(o.st.DV_avg1), C.byref(cp))


def test_zero_forcing_is_the_plain_substep():
    from sailfish_amd import configs
    cfg = configs.random_smooth(30, 20, seed=4)
    m = update_metrics(cfg["pm"], cfg["pn"], cfg["f"])
    a, b = OP.OPhys(cfg, m["fomn"], rdrg=3e-4), OP.OPhys(cfg, m["fomn"], rdrg=3e-4)
    S2 = a.szI * a.szJ
    # plain first substep: its predictor right-hand sides are stored in rubar/rvbar t1
    a.st.cycleTimes2D()
    a.substep(2.0, True, 0.0)
    rufrc, rvfrc = a.st.v("rubar", 1).copy(), a.st.v("rvbar", 1).copy()
    F, F1, F2 = [np.full(S2, 7.0), np.full(S2, 7.0)], [np.zeros(S2)] * 2, [np.zeros(S2)] * 2
    DU2, DV2 = np.zeros(S2), np.zeros(S2)
    b.st.cycleTimes2D()
    _coupled_substep(b, 2.0, True, _couple(b, rufrc, rvfrc, F, F1, F2, DU2, DV2, ab=2))
    sa, sb = _state(a), _state(b)
    for n in ("zeta", "ubar", "vbar", "rzeta", "rubar", "rvbar"):
        assert np.array_equal(sa[n], sb[n]), n
    # F = rufrc - rhs = 0 on the interior u / v points
    W = a.szI
    Fu, Fv = F[0].reshape(-1, W), F[1].reshape(-1, W)
    assert not Fu[1:-1, 2:-1].any() and not Fv[2:-1, 1:-1].any()


@pytest.mark.parametrize("ab,expect", [(0, lambda R, A, B: R), (1, lambda R, A, B: 1.5 * R - 0.5 * A),
                                        (2, lambda R, A, B: (23 * R - 16 * A + 5 * B) / 12)])
def test_first_predictor_extrapolates_the_forcing(ab, expect):
    """Flat basin at rest, no rotation, wind or drag: every 2D right-hand side is 0, so
    the first predictor's right-hand side (stored in rubar t1) is the extrapolated
    forcing, F = rufrc is stored, and the corrector adds F: at the basin centre
    ubar(t2) = (pm+pm)(pn+pn) (dt/2) R / (2h) = dt R / (dx dy h)."""
    szI, szJ, dx, dy, H, dt = 24, 18, 1e3, 2e3, 50.0, 3.0
    cfg = flat(szI, szJ, dx, dy, H)
    o = OP.OPhys(cfg, np.zeros((szJ, szI)), rdrg=0.0)
    S2 = szI * szJ
    R, A, B, Rv = 40.0, 25.0, -10.0, -30.0
    rufrc, rvfrc = np.full(S2, R), np.full(S2, Rv)
    F = [np.zeros(S2), np.zeros(S2)]
    F1, F2 = [np.full(S2, A), np.full(S2, A)], [np.full(S2, B), np.full(S2, B)]
    # stale filtered transports from an earlier step: the first substep sets them anew
    DU2, DV2 = np.full(S2, 123.0), np.full(S2, -77.0)
    w2p, w2c = 0.3, 0.7
    o.st.cycleTimes2D()
    _coupled_substep(o, dt, True, _couple(o, rufrc, rvfrc, F, F1, F2, DU2, DV2, w2p, w2c, ab=ab))
    jc, ic = szJ // 2, szI // 2
    # DU_avg2 = w2p DUon(rest) + w2c DUon(predicted): the Euler predictor moves the centre
    # to ubar = dt G / (dx dy h) (G the extrapolated forcing), so DU_avg2 = w2c dt G / dx
    G, Gv = expect(R, A, B), expect(Rv, A, B)
    assert DU2.reshape(szJ, szI)[jc, ic] == pytest.approx(w2c * dt * G / dx, rel=1e-12)
    assert DV2.reshape(szJ, szI)[jc, ic] == pytest.approx(w2c * dt * Gv / dy, rel=1e-12)
    ru1 = o.st.v("rubar", 1).reshape(szJ, szI)
    rv1 = o.st.v("rvbar", 1).reshape(szJ, szI)
    assert ru1[jc, ic] == pytest.approx(expect(R, A, B), rel=1e-14)
    assert rv1[jc, ic] == pytest.approx(expect(Rv, A, B), rel=1e-14)
    assert F[0].reshape(szJ, szI)[jc, ic] == R and F[1].reshape(szJ, szI)[jc, ic] == Rv
    u2 = o.st.v("ubar", 2).reshape(szJ, szI)
    v2 = o.st.v("vbar", 2).reshape(szJ, szI)
    assert u2[jc, ic] == pytest.approx(dt * R / (dx * dy * H), rel=1e-12)
    assert v2[jc, ic] == pytest.approx(dt * Rv / (dx * dy * H), rel=1e-12)
    # a later substep (leapfrog + AM3): the stored predictor right-hand side is F itself
    F1b, F2b = [np.full(S2, 1e3)] * 2, [np.full(S2, -1e3)] * 2
    o.st.cycleTimes2D()
    _coupled_substep(o, dt, False, _couple(o, rufrc * 0, rvfrc * 0, F, F1b, F2b, DU2, DV2, ab=ab))
    rhs_u = o.st.v("rubar", 1).reshape(szJ, szI)[jc, ic]
    # the rest state has moved by one substep: the 2D rhs at the centre is tiny against F
    assert rhs_u == pytest.approx(R, rel=1e-6)


def test_avg2_weights_sum_to_one():
    for ndtfast in (20, 30, 40):
        ct = CompTimes(dt=300.0, ndtfast=ndtfast)
        set_weights(ct)
        tot = sum(sum(OC.avg2_weights(ct.weight2, iif)) for iif in range(1, ct.nfast + 2))
        assert tot == pytest.approx(1.0, abs=1e-12)
        # the first substep's predictor weight is -1/12 of the second filter weight
        assert OC.avg2_weights(ct.weight2, 1)[0] == pytest.approx(-ct.weight2[1] / 12)
        assert OC.avg2_weights(ct.weight2, 3) == pytest.approx(
            ((8 * ct.weight2[2] - ct.weight2[3]) / 12, 5 * ct.weight2[2] / 12))


def test_avg2_of_a_nearly_frozen_state_is_its_flux():
    """With a tiny substep the state does not move: DU_avg2 = sum of weights x DUon."""
    from sailfish_amd import configs
    cfg = configs.random_smooth(26, 18, seed=9)
    m = update_metrics(cfg["pm"], cfg["pn"], cfg["f"])
    o = OP.OPhys(cfg, m["fomn"], rdrg=3e-4)
    S2 = o.szI * o.szJ
    ct = CompTimes(dt=300.0, ndtfast=30)
    set_weights(ct)
    zero = np.zeros(S2)
    F = [np.zeros(S2), np.zeros(S2)]
    DU2, DV2 = np.zeros(S2), np.zeros(S2)
    z, u, v = (o.st.v(n, 1).copy() for n in ("zeta", "ubar", "vbar"))
    for iif in range(1, ct.nfast + 2):
        o.st.cycleTimes2D()
        w2p, w2c = OC.avg2_weights(ct.weight2, iif)
        # zero forcing: rufrc = 0 gives F = -rhs; dt 1e-9 s keeps the state frozen
        _coupled_substep(o, 1e-9, iif == 1, _couple(o, zero, zero, F, [zero] * 2, [zero] * 2, DU2, DV2, w2p, w2c))
    W = o.szI
    h, pn, pm = (np.asarray(cfg[k]).reshape(-1) for k in ("h", "pn", "pm"))
    # interior u points (the walls' velocities are zeroed by the substep)
    fu = np.arange(S2).reshape(-1, W)[1:-1, 2:-1].ravel()
    fv = np.arange(S2).reshape(-1, W)[2:-1, 1:-1].ravel()
    duon = 0.5 * ((h[fu - 1] + z[fu - 1]) + (h[fu] + z[fu])) * u[fu] * 2 / (pn[fu - 1] + pn[fu])
    dvom = 0.5 * ((h[fv - W] + z[fv - W]) + (h[fv] + z[fv])) * v[fv] * 2 / (pm[fv - W] + pm[fv])
    assert np.allclose(DU2[fu], duon, rtol=1e-6, atol=1e-9 * np.abs(duon).max())
    assert np.allclose(DV2[fv], dvom, rtol=1e-6, atol=1e-9 * np.abs(dvom).max())


def test_coupled_step3d_carries_the_averaged_transports():
    szI, szJ, N = 14, 11, 6
    rng = np.random.default_rng(3)
    from sailfish_amd import configs
    cfg = configs.random_smooth(szI, szJ, seed=6)
    g, M, m = grid3(cfg, N)
    S2 = g.S2
    Zt = 0.1 * rng.standard_normal(S2)
    st = O.OState(g)
    st.Zt_avg1[:] = Zt
    M.set_depth(st)
    n3 = (N + 1) * S2
    u, v = 0.1 * rng.standard_normal(n3), 0.1 * rng.standard_normal(n3)
    ru, rv = rng.standard_normal(n3), rng.standard_normal(n3)
    M.AKv[:] = 1e-3 * (1 + rng.random(n3))
    Hu0, Hv0 = 100 * rng.standard_normal(n3), 100 * rng.standard_normal(n3)
    DU1, DV1, DU2, DV2 = (1e4 * rng.standard_normal(S2) for _ in range(4))
    ub = [np.zeros(S2) for _ in range(4)]
    L = O.lib()

    def run(coupled):
        uu, vv, Hu, Hv = u.copy(), v.copy(), Hu0.copy(), Hv0.copy()
        tmpU, tmpV = u.copy(), v.copy()
        if coupled:
            L.sfo_step3d_uv_coupled(M.cp, P(uu), P(vv), P(ru), P(rv), P(ub[0]), P(ub[1]), P(ub[2]), P(ub[3]),
                                    P(M.Hz), P(M.AKv), P(M.z_r), P(DU1), P(DV1), P(DU2), P(DV2), P(Hu), P(Hv),
                                    P(tmpU), P(tmpV), 3, 0, 1.0, 60.0, P(M.eqSD), P(M.eqD), P(M.eqRHS))
        else:
            bars = [np.zeros(S2) for _ in range(4)]     # kept alive across the call
            L.sfo_step3d_UV(M.cp, P(uu), P(vv), P(ru), P(rv), P(bars[0]), P(bars[1]), P(bars[2]), P(bars[3]),
                            P(M.Hz), P(M.AKv), P(M.z_r), P(DU1), P(DV1), P(tmpU), P(tmpV), 3, 0,
                            1.0, 60.0, P(M.eqSD), P(M.eqD), P(M.eqRHS))
        return uu.reshape(N + 1, szJ, szI), vv.reshape(N + 1, szJ, szI), Hu.reshape(N + 1, szJ, szI), \
            Hv.reshape(N + 1, szJ, szI)

    uc, vc, Huc, Hvc = run(True)
    up, vp, _, _ = run(False)
    Hz = M.Hz.reshape(N, szJ, szI)
    pm, pn = cfg["pm"], cfg["pn"]
    D1u, D1v = DU1.reshape(szJ, szI), DV1.reshape(szJ, szI)
    D2u, D2v = DU2.reshape(szJ, szI), DV2.reshape(szJ, szI)
    for j in range(1, szJ - 1):
        for i in range(2, szI - 1):
            DC = 2 / (pn[j, i - 1] + pn[j, i]) * 0.5 * (Hz[:, j, i] + Hz[:, j, i - 1])
            assert (DC * uc[:N, j, i]).sum() == pytest.approx(D1u[j, i], rel=1e-12, abs=1e-9)
            assert ub[2].reshape(szJ, szI)[j, i] == pytest.approx(D1u[j, i] / DC.sum(), rel=1e-13)
            assert ub[0].reshape(szJ, szI)[j, i] == ub[2].reshape(szJ, szI)[j, i]
            assert Huc[:N, j, i].sum() == pytest.approx(D2u[j, i], rel=1e-12, abs=1e-9)
            # Huon(K) = 0.5 (Huon0 + u DC) - DC FC with one FC per column
            fc = (0.5 * (Hu0.reshape(N + 1, szJ, szI)[:N, j, i] + uc[:N, j, i] * DC) - Huc[:N, j, i]) / DC
            assert np.ptp(fc) <= 1e-9 * (abs(fc).max() + 1)
            # one shift of the uncoupled solve's velocity per column
            assert np.ptp(uc[:N, j, i] - up[:N, j, i]) <= 1e-12
            assert Huc[N, j, i] == Hu0.reshape(N + 1, szJ, szI)[N, j, i]      # top level untouched
    for j in range(2, szJ - 1):
        for i in range(1, szI - 1):
            DC = 2 / (pm[j - 1, i] + pm[j, i]) * 0.5 * (Hz[:, j, i] + Hz[:, j - 1, i])
            assert (DC * vc[:N, j, i]).sum() == pytest.approx(D1v[j, i], rel=1e-12, abs=1e-9)
            assert ub[3].reshape(szJ, szI)[j, i] == pytest.approx(D1v[j, i] / DC.sum(), rel=1e-13)
            assert Hvc[:N, j, i].sum() == pytest.approx(D2v[j, i], rel=1e-12, abs=1e-9)
