"""GPU parity of the ensemble launch: every member of a batched run equals its solo
fused run bit for bit (FP64 and FP32), for every schedule (tile height, order, grid)
and for graph replay; FP64 members also equal the CPU oracle."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from sailfish_amd import configs
from tests.helpers2d import bitwise_equal

pytestmark = pytest.mark.gpu

SIZES = [(40, 24), (67, 33), (130, 66), (65, 17), (129, 40)]
FIELDS = ("zeta", "ubar", "vbar", "rzeta", "rubar", "rvbar")


def _members(dtype):
    from sailfish_amd.ensemble import Member
    cfgs = [configs.ensemble_member(a, b, 0.05 + 0.02 * k, 1000 + k) for k, (a, b) in enumerate(SIZES)]
    for c in cfgs:
        c["dtfast"] = 20.0
    return [Member(c, dtype=dtype) for c in cfgs]


def _solo(dtype, nsteps):
    from sailfish_amd.phys2d import run_phys
    ms = _members(dtype)
    for m in ms:
        run_phys(m.ct, m.GRID, m.OCEAN, nsteps, stepper=m.stepper)
    return ms


def _batched(dtype, nsteps, sched, graph=False):
    import torch
    from sailfish_amd.ensemble import PhysEnsemble
    ms = _members(dtype)
    ens = PhysEnsemble(ms, dtype)
    ens.configure(*sched)
    ct = ms[0].ct
    if graph:
        # one Euler substep, then the rest from a captured graph of 3 substeps
        ct.next2DTimeStep()
        ens.substep(True, ct.cff1())
        rest = nsteps - 1
        assert rest % 3 == 0
        w = []
        for _ in range(3):
            ct.next2DTimeStep()
            w.append(ct.cff1())
        ens.capture(w)
        for _ in range(rest // 3):
            ens.replay()
    else:
        for n in range(nsteps):
            ct.next2DTimeStep()
            ens.substep(n == 0, ct.cff1())
    torch.cuda.synchronize()
    ens.sync()
    return ms, ens


def _diff(A, B):
    bad = []
    for k, (a, b) in enumerate(zip(A, B)):
        for name in FIELDS:
            for lev in range(3 if name in ("zeta", "ubar", "vbar") else 2):
                x = getattr(a.OCEAN, f"{name}_t{lev}").cpu().numpy()
                y = getattr(b.OCEAN, f"{name}_t{lev}").cpu().numpy()
                if not bitwise_equal(x, y):
                    bad.append((k, name, lev))
        for name in ("Zt_avg1", "DU_avg1", "DV_avg1"):
            if not bitwise_equal(getattr(a.OCEAN, name).cpu().numpy(), getattr(b.OCEAN, name).cpu().numpy()):
                bad.append((k, name))
    return bad


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("sched", [(8, 1, 0), (4, 0, 0), (16, 1, 7), (8, 0, 5), (4, 1, 3)])
def test_ensemble_equals_solo(dt, sched):
    import torch
    print(f"[case] ensemble {dt} sched={sched}", file=sys.stderr, flush=True)
    dtype = torch.float64 if dt == "f64" else torch.float32
    solo = _solo(dtype, 7)
    bat, ens = _batched(dtype, 7, sched)
    bad = _diff(solo, bat)
    ens.close()
    assert not bad, bad[:10]


class _GraphWeights:
    """CompTimes stand-in for the solo run: the weights the captured graph bakes in
    (substep 1: weight1[0]; then weight1[1..3] repeated)."""

    def __init__(self, ct):
        self.ct, self.n = ct, 0
        self.dtfast = ct.dtfast

    def next2DTimeStep(self):
        self.n += 1

    def cff1(self):
        w = self.ct.weight1
        return float(w[0]) if self.n == 1 else float(w[1 + (self.n - 2) % 3])


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_ensemble_graph_replay_equals_solo(dt):
    import torch
    from sailfish_amd.phys2d import run_phys
    print(f"[case] ensemble graph {dt}", file=sys.stderr, flush=True)
    dtype = torch.float64 if dt == "f64" else torch.float32
    solo = _members(dtype)
    for m in solo:
        run_phys(_GraphWeights(m.ct), m.GRID, m.OCEAN, 10, stepper=m.stepper)
    bat, ens = _batched(dtype, 10, (8, 1, 0), graph=True)
    bad = _diff(solo, bat)
    ens.close()
    assert not bad, bad[:10]


def test_ensemble_fp64_member_matches_oracle():
    import torch
    from oracle import oracle_phys as OP
    print("[case] ensemble fp64 vs oracle", file=sys.stderr, flush=True)
    bat, ens = _batched(torch.float64, 7, (8, 1, 0))
    for m in bat[:3]:
        o = OP.OPhys(m.cfg, m.GRID.host["fomn"], rdrg=m.cfg["rdrg"])
        o.run(7, m.dt, m.ct.weight1)
        for name in ("zeta", "ubar", "vbar"):
            assert bitwise_equal(getattr(m.OCEAN, f"{name}_t2").cpu().numpy().ravel(), o.st.v(name, 2)), name
        assert bitwise_equal(m.OCEAN.Zt_avg1.cpu().numpy(), o.st.Zt_avg1)
    ens.close()


def test_autotuner_restores_state_and_caches(tmp_path, monkeypatch):
    import torch
    from sailfish_amd import autotune
    print("[case] autotune", file=sys.stderr, flush=True)
    monkeypatch.setenv("SF_AUTOTUNE_CACHE", str(tmp_path / "at.json"))
    dtype = torch.float32
    ref = _solo(dtype, 4)
    from sailfish_amd.ensemble import PhysEnsemble
    ms = _members(dtype)
    ens = PhysEnsemble(ms, dtype)
    ct = ms[0].ct
    ct.next2DTimeStep()
    ens.substep(True, ct.cff1())
    rec = autotune.tune(ens, nsub=3, reps=1)
    assert not rec["cached"] and rec["tile_j"] in (4, 8, 16)
    for n in range(3):
        ct.next2DTimeStep()
        ens.substep(False, ct.cff1())
    torch.cuda.synchronize()
    ens.sync()
    assert not _diff(ref, ms)
    rec2 = autotune.tune(ens, nsub=3, reps=1)
    assert rec2["cached"] and rec2["tile_j"] == rec["tile_j"]
    ens.close()
