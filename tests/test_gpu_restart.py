"""Checkpoint/resume on the GPU: a run that writes a restart, is rebuilt from its
configuration, reads the restart back and continues ends bit for bit where the
uninterrupted run ends (baroclinic main3d in FP64, physics barotropic run in FP32 and
FP64); the history file of a main3d run holds the records main3d.py:202-216 writes."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from sailfish_amd import configs
from tests.helpers2d import bitwise_equal

pytestmark = pytest.mark.gpu


def _cfg3(szI, szJ, N, seed):
    cfg = configs.random_smooth(szI, szJ, seed=seed)
    rng = np.random.default_rng(seed + 1)
    sm = lambda: configs.gaussian_smooth(rng.standard_normal((szJ, szI)), 1.5)
    cfg["u0"] = np.stack([0.05 * sm() for _ in range(N + 1)])
    cfg["v0"] = np.stack([0.05 * sm() for _ in range(N + 1)])
    cfg["dtfast"] = 2.0
    return cfg


def _state(GRID, OCEAN):
    from sailfish_amd.io_netcdf import _state_arrays
    return {n: t.detach().cpu().numpy().copy() for n, t in _state_arrays(GRID, OCEAN)}


def test_main3d_resume_from_restart_is_bitwise(tmp_path):
    import torch
    from sailfish_amd.io_netcdf import HistoryWriter, read_restart, write_restart
    from sailfish_amd.main3d import main3d
    from sailfish_amd.model2d import build_model3d
    from sailfish_amd.ncfile import NCReader

    szI, szJ, N = 40, 22, 6
    print(f"[case] main3d restart {szI}x{szJ}x{N}", file=sys.stderr, flush=True)
    cfg = _cfg3(szI, szJ, N, seed=5)
    # uninterrupted: 4 baroclinic steps of 8 s with a history interval of 8 model seconds
    ct, G, O, B = build_model3d(cfg, N=N, ndtfast=4, dt=8.0, fix_q5=True)
    hw = HistoryWriter(tmp_path / "his.nc", G, interval=8.0)
    main3d(ct, G, O, B, nsteps=4, fused=True, history=hw)
    hw.close()
    torch.cuda.synchronize()
    ref = _state(G, O)
    ref_slots = {n: O.slot_of(n, "t2") for n in ("zeta", "ubar", "u")}
    ref_zeta_t2 = O.zeta_t2.detach().cpu().numpy().copy()

    # 2 steps, restart, fresh model, 2 more steps
    ct, G, O, B = build_model3d(cfg, N=N, ndtfast=4, dt=8.0, fix_q5=True)
    main3d(ct, G, O, B, nsteps=2, fused=True)
    write_restart(tmp_path / "rst.nc", ct, G, O)
    ct2, G2, O2, B2 = build_model3d(cfg, N=N, ndtfast=4, dt=8.0, fix_q5=True)
    read_restart(tmp_path / "rst.nc", ct2, G2, O2)
    main3d(ct2, G2, O2, B2, nsteps=2, fused=True)
    torch.cuda.synchronize()
    got = _state(G2, O2)
    bad = [n for n in ref if not bitwise_equal(ref[n].ravel(), got[n].ravel())]
    assert not bad, bad
    assert {n: O2.slot_of(n, "t2") for n in ref_slots} == ref_slots
    assert ct2.iic == 4 and ct2.time == 32.0

    # the history: a record after the barotropic loop of the first step (t = 0 s), then
    # whenever t reaches the previous due time plus one interval (t = 16 s and 24 s);
    # the last record's zeta is the final zeta(t2) in FP32
    h = NCReader(tmp_path / "his.nc")
    assert list(h["ocean_time"]) == [0.0, 16.0, 24.0]
    assert np.isfinite(h["u"]).all() and h["u"].shape == (3, N, szJ, szI - 1)
    assert np.array_equal(h["zeta"][2], ref_zeta_t2.reshape(szJ, szI).astype(np.float32))


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_physics_resume_from_restart_is_bitwise(tmp_path, dt):
    import torch
    from sailfish_amd.io_netcdf import read_restart, write_restart
    from sailfish_amd.model2d import build_model
    from sailfish_amd.phys2d import PhysStepper, run_phys

    print(f"[case] physics restart {dt}", file=sys.stderr, flush=True)
    tdt = torch.float32 if dt == "f32" else torch.float64
    cfg = configs.basin1(130, 66)
    ct, G, O, B = build_model(cfg, dtype=tdt, ndtfast=40)
    st = PhysStepper(G, O)
    run_phys(ct, G, O, 24, st)
    torch.cuda.synchronize()
    ref = _state(G, O)

    ct, G, O, B = build_model(cfg, dtype=tdt, ndtfast=40)
    st = PhysStepper(G, O)
    run_phys(ct, G, O, 11, st)
    write_restart(tmp_path / "rst.nc", ct, G, O)
    ct2, G2, O2, B2 = build_model(cfg, dtype=tdt, ndtfast=40)
    read_restart(tmp_path / "rst.nc", ct2, G2, O2)
    run_phys(ct2, G2, O2, 13, PhysStepper(G2, O2), start_first=False)
    torch.cuda.synchronize()
    got = _state(G2, O2)
    bad = [n for n in ref if not bitwise_equal(ref[n].ravel(), got[n].ravel())]
    assert not bad, bad
