"""GPU parity: the one-to-one FP64 step2d kernels (through the C ABI) against the
CPU oracle, bit for bit, on every time slot of every 2D state array."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from sailfish_amd import configs
from tests.helpers2d import compare_states, oracle_arrays, oracle_model, product_arrays

pytestmark = pytest.mark.gpu

# non-square grids: szI a multiple of szJ (Q5 mask hits whole columns), szI not a
# multiple (mask scattered, admits column-0 points), square, and tall
GRIDS = [(24, 12), (30, 20), (16, 16), (20, 33)]


def _run(cfg, nsteps, mode, lbc=None, fix_q5=False):
    import torch
    from sailfish_amd.main2d import barotropic_loop, main2d
    from sailfish_amd.model2d import build_model
    from oracle import oracle as O

    ct, GRID, OCEAN, BOUNDARY = build_model(cfg, lbc=lbc, ndtfast=4, fix_q5=fix_q5)
    t, g, s, b = oracle_model(cfg, ct.weight1, ct.nfast, lbc, fix_q5=fix_q5)
    if mode == "main2d":
        main2d(ct, GRID, OCEAN, BOUNDARY, nsteps)
        O.main2d(t, g, s, b, nsteps)
    else:
        for _ in range(nsteps):
            barotropic_loop(ct, GRID, OCEAN, BOUNDARY)
            O.barotropic_loop(t, g, s, b)
    torch.cuda.synchronize()
    return product_arrays(OCEAN), oracle_arrays(s)


@pytest.mark.parametrize("szI,szJ", GRIDS)
@pytest.mark.parametrize("mode", ["main2d", "main3d"])
def test_unfused_bitwise(szI, szJ, mode):
    print(f"[case] unfused {mode} {szI}x{szJ}", file=sys.stderr, flush=True)
    cfg = configs.random_smooth(szI, szJ, seed=szI * 100 + szJ)
    prod, orc = _run(cfg, 7 if mode == "main2d" else 2, mode)
    bad = compare_states(prod, orc)
    assert not bad, bad


def test_unfused_gradient_and_clamped_bcs():
    print("[case] unfused mixed BCs", file=sys.stderr, flush=True)
    # vbar's gradient copy reads ubarGradientBCIdx2 (barotropicVelocityBC.py:31), so
    # its list must not be longer than ubar's (E block: M entries >= W block: M-1)
    lbc = {"zeta": ["GRA", "CLO", "GRA", "CLO"], "ubar": ["CLA", "CLO", "GRA", "CLO"],
           "vbar": ["GRA", "CLO", "CLO", "CLO"]}
    cfg = configs.random_smooth(22, 14, seed=7)
    prod, orc = _run(cfg, 6, "main2d", lbc=lbc)
    assert not compare_states(prod, orc)


def test_basin1_reference_mask_blows_up_identically():
    """128x64 with the reference's i%szJ mask: the frozen interior column receives
    the corrector's unscaled AM3 update and the run diverges within a few substeps,
    on the GPU exactly as in the oracle."""
    print("[case] basin1 128x64 reference mask", file=sys.stderr, flush=True)
    cfg = configs.basin1()
    prod, orc = _run(cfg, 40, "main2d")
    assert not compare_states(prod, orc)
    z = prod["zeta"][prod["slots"]["zeta"][2]]
    assert not np.all(np.isfinite(z))


def test_basin1_fixed_mask_200_substeps():
    """Same basin with the column-0/row-0 mask and dt=5 s (inside the scheme's
    stability limit): 200 substeps stay finite and bitwise equal."""
    print("[case] basin1 128x64 fixed mask 200 substeps", file=sys.stderr, flush=True)
    cfg = dict(configs.basin1(), dtfast=5.0)
    prod, orc = _run(cfg, 200, "main2d", fix_q5=True)
    assert not compare_states(prod, orc)
    z = prod["zeta"][prod["slots"]["zeta"][2]]
    assert np.all(np.isfinite(z)) and np.abs(z).max() < 0.2


@pytest.mark.parametrize("szI,szJ", [(24, 12), (20, 33)])
def test_adams_moulton_corr3rd2_intended_bitwise(szI, szJ):
    """The depth-weighted AM3 kernel the reference defines but never launches
    (mod_cppkernels.cpp:213-251) through the C ABI, against the oracle, bit for bit,
    on random fields (points of row 0 read before the array start: 0.0 on both sides)."""
    print(f"[case] AdamsMoultonCorr3rd2 {szI}x{szJ}", file=sys.stderr, flush=True)
    import torch
    from oracle import oracle as O
    from sailfish_amd import mod_operators as ops
    from sailfish_amd.model2d import build_model
    from tests.helpers2d import bitwise_equal

    cfg = configs.random_smooth(szI, szJ, seed=szI + 7 * szJ)
    ct, GRID, OCEAN, _B = build_model(cfg, ndtfast=4)
    _t, g, _s, _b = oracle_model(cfg, ct.weight1, ct.nfast)
    rng = np.random.default_rng(szI * szJ)
    n = szI * szJ
    host = {k: rng.standard_normal(n) for k in ("u", "v", "ru0", "ru1", "ru2", "rv0", "rv1", "rv2")}
    host["z1"] = 0.1 * rng.standard_normal(n)
    host["z2"] = 0.1 * rng.standard_normal(n)
    dev = {k: torch.tensor(a, dtype=torch.float64, device="cuda") for k, a in host.items()}
    h_dev = GRID.h.reshape(-1)
    Dt = 3.0
    ops.AdamsMoultonCorr3rd2_intended(None, None, (Dt, dev["u"], dev["v"], dev["ru0"], dev["ru1"], dev["ru2"],
                                                   dev["rv0"], dev["rv1"], dev["rv2"], h_dev, dev["z1"],
                                                   dev["z2"]), ctx=ops.ctx_of(GRID))
    torch.cuda.synchronize()
    u, v = host["u"].copy(), host["v"].copy()
    P = O.ptr
    O.lib().sfo_AdamsMoultonCorr3rd2(g.cp, Dt, P(u), P(v), P(host["ru0"]), P(host["ru1"]), P(host["ru2"]),
                                     P(host["rv0"]), P(host["rv1"]), P(host["rv2"]), P(g.h), P(host["z1"]),
                                     P(host["z2"]))
    assert bitwise_equal(dev["u"].cpu().numpy(), u)
    assert bitwise_equal(dev["v"].cpu().numpy(), v)
