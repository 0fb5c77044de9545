"""The ensemble and closure workloads at their full sizes.

- 64-member ensemble (configs.ensemble_configs: uneven sizes drawn from 256x128 ..
  1024x512, per-member wind and noise): every member of the batched launch equals its
  solo run bit for bit after four substeps (FP32), and FP64 members equal the oracle.
- Closure workload: the fine grid 2042 x 1018 and the coarse grid 512 x 256 each take
  their physics substeps bit for bit as the oracle (FP64); the whole experiment at its
  specified size (2e6 samples, 5000 coarse steps) selects the model its validation
  scores favour and reports finite errors for both coarse runs."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from sailfish_amd import configs
from tests.helpers2d import bitwise_equal

pytestmark = pytest.mark.gpu

NSUB = 4
FIELDS = ("zeta", "ubar", "vbar", "rzeta", "rubar", "rvbar")


def _members(dtype):
    from sailfish_amd.ensemble import Member
    return [Member(c, dtype=dtype) for c in configs.ensemble_configs()]


def _batched(ms, dtype):
    import torch
    from sailfish_amd.ensemble import PhysEnsemble
    ens = PhysEnsemble(ms, dtype)
    ens.configure(8, 1, 0)
    ct = ms[0].ct
    for n in range(NSUB):
        ct.next2DTimeStep()
        ens.substep(n == 0, ct.cff1())
    torch.cuda.synchronize()
    ens.sync()
    return ens


def test_ensemble_64_members_batched_equals_solo():
    import torch
    from sailfish_amd.phys2d import run_phys
    print("[case] 64-member ensemble FP32, batched vs solo", file=sys.stderr, flush=True)
    solo = _members(torch.float32)
    assert len(solo) == 64 and sum(m.points for m in solo) > 10_000_000
    for m in solo:
        run_phys(m.ct, m.GRID, m.OCEAN, NSUB, stepper=m.stepper)
    torch.cuda.synchronize()
    ref = [{n: getattr(m.OCEAN, f"{n}_t2").cpu().numpy() for n in ("zeta", "ubar", "vbar")} for m in solo]
    del solo
    bat = _members(torch.float32)
    ens = _batched(bat, torch.float32)
    bad = []
    for k, (r, m) in enumerate(zip(ref, bat)):
        for n in ("zeta", "ubar", "vbar"):
            if not bitwise_equal(r[n], getattr(m.OCEAN, f"{n}_t2").cpu().numpy()):
                bad.append((k, n))
        assert np.isfinite(ref[k]["zeta"]).all()
    ens.close()
    assert not bad, bad[:10]


def test_ensemble_64_members_fp64_against_oracle():
    import torch
    from oracle import oracle_phys as OP
    print("[case] 64-member ensemble FP64, members vs oracle", file=sys.stderr, flush=True)
    bat = _members(torch.float64)
    ens = _batched(bat, torch.float64)
    order = np.argsort([m.points for m in bat])
    for k in (int(order[0]), int(order[1]), int(order[-1])):      # two smallest and the largest member
        m = bat[k]
        o = OP.OPhys(m.cfg, m.GRID.host["fomn"], rdrg=m.cfg["rdrg"])
        o.run(NSUB, m.dt, m.ct.weight1)
        for n in ("zeta", "ubar", "vbar"):
            assert bitwise_equal(getattr(m.OCEAN, f"{n}_t2").cpu().numpy().ravel(), o.st.v(n, 2)), (k, n)
        for n in ("Zt_avg1", "DU_avg1", "DV_avg1"):
            assert bitwise_equal(getattr(m.OCEAN, n).cpu().numpy(), getattr(o.st, n)), (k, n)
    ens.close()


@pytest.mark.parametrize("which", ["fine", "coarse"])
def test_closure_grids_fp64_against_oracle(which):
    import torch
    from oracle import oracle_phys as OP
    from sailfish_amd.closure import Model, coarse_config, fine_config
    fcfg = fine_config(512, 256, 4)
    cfg = fcfg if which == "fine" else coarse_config(fcfg)
    if which == "coarse":    # a seeded smooth state, so the coarse grid is not at rest
        rng = np.random.default_rng(5)
        cfg["zeta0"] = configs.gaussian_smooth(0.05 * rng.standard_normal((cfg["szJ"], cfg["szI"])), 4.0)
    print(f"[case] closure {which} grid {cfg['szI']}x{cfg['szJ']} FP64 vs oracle", file=sys.stderr, flush=True)
    m = Model(cfg, torch.float64)
    for _ in range(3):
        m.step()
    torch.cuda.synchronize()
    o = OP.OPhys(cfg, m.GRID.host["fomn"], rdrg=cfg["rdrg"])
    o.run(3, cfg["dtfast"], m.ct.weight1)
    for n in ("zeta", "ubar", "vbar"):
        assert bitwise_equal(getattr(m.OCEAN, f"{n}_t2").cpu().numpy().ravel(), o.st.v(n, 2)), n
    assert np.abs(m.OCEAN.ubar_t2.cpu().numpy()).max() > 0


def test_closure_experiment_at_full_size_selects_by_validation():
    import torch
    from sailfish_amd.closure import run_experiment
    print("[case] closure experiment at full size", file=sys.stderr, flush=True)
    res = run_experiment(dtype=torch.float64, log=lambda *a: print("[closure]", *a, file=sys.stderr, flush=True))
    assert res["samples"] == 2_000_000
    v = res["validation"]
    assert v["steps"] == 5000
    assert v["selected"] == ("closure" if v["score"]["closure"] < v["score"]["plain"] else "plain")
    for row in res["series"]:
        assert all(np.isfinite(row["plain"])) and all(np.isfinite(row["closure"]))
    res["mlp"].close()
