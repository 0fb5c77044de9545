"""GPU parity of the baroclinic terms (sf_rho_eos, sf_prsgrd, sf_step3d_t) against the
oracle: each kernel bit for bit in FP64 on a sloping basin with a random state; the
coupled split-explicit step with the pressure gradient and the tracer step over three
baroclinic steps, bit for bit; the FP32 bulk form against FP64; and the 1024 x 512 x 32
workload with its stratification T = 4 + 16 exp(z/500) for two coupled baroclinic steps
(finite, tracers within their initial range: the upstream step makes no new extrema)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from sailfish_amd import configs
from tests.helpers2d import bitwise_equal

pytestmark = pytest.mark.gpu


def _cfg(szI, szJ, N, seed):
    cfg = configs.random_smooth(szI, szJ, seed=seed)
    rng = np.random.default_rng(seed + 1)
    sm = lambda: configs.gaussian_smooth(rng.standard_normal((szJ, szI)), 1.5)
    cfg["u0"] = np.stack([0.05 * sm() for _ in range(N + 1)])
    cfg["v0"] = np.stack([0.05 * sm() for _ in range(N + 1)])
    cfg["dtfast"] = 2.0
    return cfg


def test_kernels_fp64_against_oracle():
    import torch
    from oracle import oracle as O
    from oracle import oracle3d as O3
    from sailfish_amd.baroclinic import Tracers
    from sailfish_amd.main3d import initial_depths
    from sailfish_amd.model2d import build_model3d
    szI, szJ, N = 45, 29, 9
    print(f"[case] baroclinic kernels FP64 {szI}x{szJ}x{N}", file=sys.stderr, flush=True)
    cfg = _cfg(szI, szJ, N, seed=17)
    ct, GRID, OCEAN, _B = build_model3d(cfg, N=N, ndtfast=6, dt=12.0)
    initial_depths(GRID, OCEAN)
    rng = np.random.default_rng(4)
    S2 = szI * szJ
    T = 10.0 + 5.0 * rng.random(N * S2)
    Sa = 35.0 + rng.standard_normal(N * S2)
    tr = Tracers(GRID, OCEAN, T=T.reshape(N, szJ, szI), S=Sa.reshape(N, szJ, szI), Scoef=7.6e-4)
    ru0 = rng.standard_normal((N + 1) * S2)
    rv0 = rng.standard_normal((N + 1) * S2)
    OCEAN.ru_t2.copy_(torch.from_numpy(ru0))
    OCEAN.rv_t2.copy_(torch.from_numpy(rv0))
    tr.pressure_gradient(GRID, OCEAN)
    Hu = 40.0 * rng.standard_normal((N + 1) * S2)
    Hv = 40.0 * rng.standard_normal((N + 1) * S2)
    OCEAN.Huon.copy_(torch.from_numpy(Hu.reshape(OCEAN.Huon.shape)))
    OCEAN.Hvom.copy_(torch.from_numpy(Hv.reshape(OCEAN.Hvom.shape)))
    Hz_old = GRID.Hz.detach().cpu().numpy().ravel().copy()
    Hz_new = Hz_old * (1.0 + 1e-3 * rng.standard_normal(Hz_old.size))
    tr.Hz_old.copy_(torch.from_numpy(Hz_old.reshape(tr.Hz_old.shape)))
    GRID.Hz.copy_(torch.from_numpy(Hz_new.reshape(GRID.Hz.shape)))
    tr.step(GRID, OCEAN, 60.0)
    torch.cuda.synchronize()

    host = GRID.host
    g = O.OGrid(szI, szJ, N, dict(h=cfg["h"], pm=cfg["pm"], pn=cfg["pn"], on_u=host["on_u"], om_v=host["om_v"]))
    BC = -np.ones(S2, np.int32)
    m = O3.OModel3(g, N, host["fomn"], host["sc_r"], host["sc_w"], host["Cs_r"], host["Cs_w"], GRID.hc, 2, BC, BC, BC)
    m.z_r[:] = GRID.z_r.detach().cpu().numpy().ravel()
    m.z_w[:N * S2 + S2] = GRID.z_w.detach().cpu().numpy().ravel()
    L, p = O.lib(), O.ptr
    rho = np.zeros(N * S2)
    L.sfo_rho_eos(m.cp, p(T), p(Sa), *tr.eos, p(rho))
    ru, rv = ru0.copy(), rv0.copy()
    L.sfo_prsgrd(m.cp, p(rho), p(m.z_r), p(m.z_w), p(Hz_old), tr.g, tr.rho0, p(ru), p(rv))
    Tn, Sn = np.zeros_like(T), np.zeros_like(Sa)
    L.sfo_step3d_t(m.cp, p(T), p(Tn), p(Hu), p(Hv), p(Hz_old), p(Hz_new), 60.0)
    L.sfo_step3d_t(m.cp, p(Sa), p(Sn), p(Hu), p(Hv), p(Hz_old), p(Hz_new), 60.0)
    get = lambda t: t.detach().cpu().numpy().ravel()
    assert bitwise_equal(get(tr.rho), rho)
    assert bitwise_equal(get(OCEAN.ru_t2), ru) and bitwise_equal(get(OCEAN.rv_t2), rv)
    assert bitwise_equal(get(tr.T), Tn) and bitwise_equal(get(tr.S), Sn)
    assert np.abs(ru - ru0).max() > 0 and np.abs(Tn - T).max() > 1e-6


def _coupled_gpu(cfg, N, dtype, nsteps, tracers):
    from sailfish_amd.coupled3d import main3d_coupled
    from sailfish_amd.model2d import build_model3d
    ct, GRID, OCEAN, BOUNDARY = build_model3d(cfg, N=N, ndtfast=cfg.get("ndtfast", 6), dt=cfg.get("dt", 12.0),
                                              fix_q5=True, dtype=dtype)
    st = main3d_coupled(ct, GRID, OCEAN, BOUNDARY, nsteps, tracers=tracers)
    return ct, GRID, OCEAN, BOUNDARY, st


def test_coupled_with_tracers_fp64_against_oracle():
    import torch
    from oracle import oracle_coupled as OC
    from sailfish_amd.mod_constants import g as G
    szI, szJ, N, nsteps = 48, 30, 7, 3
    print(f"[case] coupled + tracers FP64 {szI}x{szJ}x{N}", file=sys.stderr, flush=True)
    cfg = _cfg(szI, szJ, N, seed=23)
    ct, GRID, OCEAN, BOUNDARY, st = _coupled_gpu(cfg, N, torch.float64, nsteps, {"Scoef": 7.6e-4})
    torch.cuda.synchronize()
    tr = st.tracers
    host = GRID.host
    # the oracle starts from the same initial tracers (the stratification at the initial depths)
    from sailfish_amd.main3d import initial_depths
    from sailfish_amd.model2d import build_model3d
    from sailfish_amd.baroclinic import stratification
    _ct, G0, O0, _B0 = build_model3d(cfg, N=N, ndtfast=6, dt=12.0, fix_q5=True)
    initial_depths(G0, O0)
    T0 = stratification(G0).ravel()
    o = OC.OCoupled(cfg, N, host, host["sc_r"], host["sc_w"], host["Cs_r"], host["Cs_w"], GRID.hc,
                    BOUNDARY.zetaBC.bcIdxFieldIdx2, BOUNDARY.uvelBC.bcIdxFieldType,
                    BOUNDARY.vvelBC.bcIdxFieldType, rdrg=3e-4, g=G,
                    tracers=dict(T=T0, S=np.full(T0.size, 35.0), eos=tr.eos, rho0=tr.rho0))
    m = o.m
    for k in range(3):
        m.u[k][:] = np.broadcast_to(cfg["u0"], (N + 1, szJ, szI)).ravel()
        m.v[k][:] = np.broadcast_to(cfg["v0"], (N + 1, szJ, szI)).ravel()
    m.AKv[:] = OCEAN.AKv.detach().cpu().numpy().ravel()
    o.s.Zt_avg1[:] = np.asarray(cfg["zeta0"], float).ravel()
    m.set_depth(o.s)
    time = 0.0
    for iic in range(nsteps):
        o.baroclinic_step(iic, 0, ct.nfast, ct.dtfast, ct.dt, ct.weight1, ct.weight2, time)
        time += ct.dt
    get = lambda t: t.detach().cpu().numpy().ravel()
    bad = [n for n in ("u", "v", "ru", "rv") if not bitwise_equal(get(getattr(OCEAN, n)[2]), getattr(m, n)[2])]
    bad += [n for n in ("W", "Huon", "Hvom") if not bitwise_equal(get(getattr(OCEAN, n)), getattr(m, n))]
    bad += [n for n in ("zeta", "ubar", "vbar") if not bitwise_equal(get(getattr(OCEAN, f"{n}_t2")), o.s.v(n, 2))]
    bad += [n for n in ("T", "S") if not bitwise_equal(get(getattr(tr, n)), o.tr[n])]
    assert not bad, bad
    assert np.abs(get(tr.T) - T0).max() > 1e-8        # the tracer has moved


def test_coupled_with_tracers_fp32_against_fp64():
    import torch
    szI, szJ, N = 96, 64, 8
    print(f"[case] coupled + tracers FP32 vs FP64 {szI}x{szJ}x{N}", file=sys.stderr, flush=True)
    cfg = _cfg(szI, szJ, N, seed=41)
    out = {}
    for dt in (torch.float64, torch.float32):
        _ct, _G, O, _B, st = _coupled_gpu(cfg, N, dt, 4, True)
        g = lambda t: t.detach().cpu().numpy().astype(np.float64)
        out[dt] = {"zeta": g(O.zeta_t2), "u": g(O.u_t2), "T": g(st.tracers.T)}
    errs = {k: float(np.abs(out[torch.float32][k] - out[torch.float64][k]).max()
                     / np.abs(out[torch.float64][k]).max()) for k in out[torch.float64]}
    print("[coupled+tracers fp32 error vs fp64] " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()),
          file=sys.stderr, flush=True)
    assert errs["zeta"] <= 1e-4 and errs["u"] <= 1e-4 and errs["T"] <= 1e-5, errs


def test_workload_3d_with_stratification_runs():
    import torch
    cfg = configs.basin3()
    N = cfg["N"]
    print(f"[case] 3D workload + stratification FP32 {cfg['szI']}x{cfg['szJ']}x{N}", file=sys.stderr, flush=True)
    _ct, GRID, OCEAN, _B, st = _coupled_gpu(cfg, N, torch.float32, 2, True)
    T = st.tracers.T.detach().cpu().numpy()
    u = OCEAN.u_t2.detach().cpu().numpy()
    assert np.isfinite(T).all() and np.isfinite(u).all()
    assert T.min() >= 4.0 - 1e-3 and T.max() <= 20.0 + 1e-3
    assert np.abs(u).max() > 1e-4
