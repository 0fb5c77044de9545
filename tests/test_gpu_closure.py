"""GPU run of the closure pipeline at a small scale: fine runs -> coarse-grained
samples -> BF16-MFMA training -> coarse runs with and without the closure, and the
graph-captured closure loop equal to the stepped one."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("target", ["probe", "step"])
def test_small_closure_experiment_runs_and_learns(target):
    import torch
    from sailfish_amd.closure import run_experiment
    print(f"[case] closure experiment small, {target} targets", file=sys.stderr, flush=True)
    res = run_experiment(coarse_shape=(66, 34), nsteps=120, n_train_runs=2, snaps_per_run=3, snap_spacing=12,
                         nsamples=10_000, epochs=8, batch=1024, lr=2e-3, eval_every=60, dtype=torch.float64,
                         n_val=60, target=target, log=lambda *a: print(*a, file=sys.stderr))
    assert res["samples"] > 5000
    # training-set MSE of the normalised targets: training must bring it below the
    # untrained network's and below predicting the mean (MSE 1)
    assert res["losses"][-1] < min(res["losses"][0], 1.0)
    for row in res["series"]:
        assert all(np.isfinite(row["plain"])) and all(np.isfinite(row["closure"]))
    # the gain of the network over the plain scheme, as the validation run measures it,
    # decides which model is used
    v = res["validation"]
    print("validation", v, "final rmse", res["series"][-1], file=sys.stderr)
    assert v["selected"] == ("closure" if v["score"]["closure"] < v["score"]["plain"] else "plain")
    res["mlp"].close()


@pytest.mark.parametrize("with_closure", [False, True])
def test_graph_loop_equals_stepped_loop(with_closure):
    import torch
    from sailfish_amd.closure import ClosureMLP, GraphLoop, Model, closure_buffers, coarse_config, fine_config
    print(f"[case] closure graph loop closure={with_closure}", file=sys.stderr, flush=True)
    ccfg = coarse_config(fine_config(40, 24))
    mlp = None
    if with_closure:
        mlp = ClosureMLP(seed=3)
        mlp.set_norm(np.zeros(45), np.full(45, 0.1), np.array([0.0, 1e-6, 0.0, 1e-6]))
    a, b = Model(ccfg, torch.float32), Model(ccfg, torch.float32)
    z0 = torch.from_numpy(np.random.default_rng(1).standard_normal(a.shape) * 0.01).cuda().float()
    for m in (a, b):
        m.set_state(z0, torch.zeros_like(z0), torch.zeros_like(z0))
    su, sv = closure_buffers(a)
    a.step(mlp, su, sv)
    b.step(mlp, su, sv)
    loop = GraphLoop(b, 3, mlp)
    for _ in range(4):
        loop.replay()
    torch.cuda.synchronize()
    for r in range(4):
        for k in range(3):
            a.ct.iif = 1 + k          # the weights the graph captured
            a.step(mlp, su, sv)
    torch.cuda.synchronize()
    diffs = [float((x - y).abs().max()) for x, y in zip(a.state(), b.state())]
    print("graph vs stepped max diffs", diffs, file=sys.stderr)
    assert diffs == [0.0, 0.0, 0.0]
    if mlp is not None:
        mlp.close()
