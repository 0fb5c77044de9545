"""GPU parity of the fused barotropic substep (sf_step2d_substep) against the CPU
oracle's one-kernel-per-reference-kernel sequence.

FP64: bit for bit on every state array the path reads, i.e. everything except the
buffers that no kernel reads (rvbar, the corrector's rubar slot) and the dead AM3
increment of ubar's t0/t1 slots at unmasked points (compared at masked points).
FP32: error against the FP64 oracle within the stated tolerance."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from sailfish_amd import configs
from tests.helpers2d import compare_fused, oracle_arrays, oracle_model, product_arrays

pytestmark = pytest.mark.gpu


def _run(cfg, nsteps, mode, fix_q5, dtype="f64", lbc=None):
    import torch
    from oracle import oracle as O
    from sailfish_amd.main2d import barotropic_loop, main2d
    from sailfish_amd.model2d import build_model

    tdt = torch.float64 if dtype == "f64" else torch.float32
    ct, GRID, OCEAN, BOUNDARY = build_model(cfg, lbc=lbc, ndtfast=4, fix_q5=fix_q5, dtype=tdt)
    t, g, s, b = oracle_model(cfg, ct.weight1, ct.nfast, lbc, fix_q5=fix_q5)
    if mode == "main2d":
        main2d(ct, GRID, OCEAN, BOUNDARY, nsteps, fused=True)
        O.main2d(t, g, s, b, nsteps)
    else:
        for _ in range(nsteps):
            barotropic_loop(ct, GRID, OCEAN, BOUNDARY, fused=True)
            O.barotropic_loop(t, g, s, b)
    torch.cuda.synchronize()
    return product_arrays(OCEAN), oracle_arrays(s)


# (szI, szJ, fix_q5): tile-edge-straddling sizes, partial tiles, masked interior
# columns at tile starts (faithful mask with szI a multiple of szJ)
CASES = [(130, 17, True), (64, 16, True), (67, 29, True), (128, 64, False), (96, 32, False),
         (200, 100, True), (40, 40, False)]


@pytest.mark.parametrize("szI,szJ,fix_q5", CASES)
@pytest.mark.parametrize("mode", ["main2d", "main3d"])
def test_fused_fp64_bitwise(szI, szJ, fix_q5, mode):
    print(f"[case] fused fp64 {mode} {szI}x{szJ} fix_q5={fix_q5}", file=sys.stderr, flush=True)
    cfg = configs.random_smooth(szI, szJ, seed=szI + 7 * szJ)
    prod, orc = _run(cfg, 6 if mode == "main2d" else 2, mode, fix_q5)
    stride = szI if fix_q5 else szJ
    bad = compare_fused(prod, orc, szI, szJ, stride)
    assert not bad, bad


def test_fused_clamped_velocity_boundary():
    print("[case] fused clamped ubar", file=sys.stderr, flush=True)
    lbc = {"ubar": ["CLA", "CLO", "CLO", "CLO"], "zeta": ["GRA", "CLO", "GRA", "CLO"]}
    cfg = configs.random_smooth(70, 20, seed=3)
    prod, orc = _run(cfg, 5, "main2d", True, lbc=lbc)
    assert not compare_fused(prod, orc, 70, 20, 70)


# gradient (copy) velocity boundaries: west / east ubar copies, clamped and closed sides;
# grids with a shifted tile origin (szI % 64 == 1), several tiles, and the reference's
# i%szJ mask (masked interior columns finished by the masked-points kernel)
GRAD_CASES = [
    (70, 20, True, {"ubar": ["GRA", "CLO", "CLO", "CLO"], "zeta": ["GRA", "CLO", "GRA", "CLO"]}),
    (129, 33, True, {"ubar": ["GRA", "CLO", "GRA", "CLO"]}),
    (130, 17, True, {"ubar": ["CLA", "CLO", "GRA", "CLO"], "zeta": ["GRA", "CLO", "GRA", "CLO"]}),
    (128, 64, False, {"ubar": ["GRA", "CLO", "GRA", "CLO"]}),
    (96, 32, False, {"ubar": ["GRA", "CLO", "CLO", "CLO"]}),
]


@pytest.mark.parametrize("szI,szJ,fix_q5,lbc", GRAD_CASES)
@pytest.mark.parametrize("mode", ["main2d", "main3d"])
def test_fused_gradient_velocity_boundary(szI, szJ, fix_q5, lbc, mode):
    print(f"[case] fused gradient ubar {mode} {szI}x{szJ} fix_q5={fix_q5} {lbc}", file=sys.stderr, flush=True)
    from sailfish_amd.step2d_fused import check_maps
    from sailfish_amd.model2d import build_model
    cfg = configs.random_smooth(szI, szJ, seed=szI + 3 * szJ)
    _, GRID, _, BOUNDARY = build_model(cfg, lbc=lbc, ndtfast=4, fix_q5=fix_q5)
    ok, why = check_maps(szI, szJ, szI if fix_q5 else szJ, BOUNDARY)
    assert ok, why
    prod, orc = _run(cfg, 6 if mode == "main2d" else 2, mode, fix_q5, lbc=lbc)
    bad = compare_fused(prod, orc, szI, szJ, szI if fix_q5 else szJ)
    assert not bad, bad


def test_fused_rejects_far_velocity_copy():
    """The reference's vbar gradient copy reads ubar's source list
    (barotropicVelocityBC.py:31): its sources are not neighbours, so the fused
    substep rejects the maps (the caller keeps the one-kernel-per-step path)."""
    import torch
    from sailfish_amd import _lib
    from sailfish_amd.model2d import build_model
    from sailfish_amd.step2d_fused import FusedStepper
    lbc = {"ubar": ["CLA", "CLO", "GRA", "CLO"], "vbar": ["GRA", "CLO", "CLO", "CLO"]}
    cfg = configs.random_smooth(40, 20, seed=3)
    ct, GRID, OCEAN, BOUNDARY = build_model(cfg, lbc=lbc, ndtfast=4, fix_q5=True)
    with pytest.raises(_lib.SailfishError):
        FusedStepper(GRID, BOUNDARY, torch.float64)


def test_fused_fp32_error_vs_fp64():
    """FP32 bulk / FP64 accumulators against the FP64 oracle after 40 substeps on a
    stable configuration.  Tolerance: max |zeta error| <= 1e-5 * max|zeta| + 1e-7 m,
    velocity errors <= 1e-4 relative to the velocity scale."""
    print("[case] fused fp32 vs fp64 oracle", file=sys.stderr, flush=True)
    cfg = configs.random_smooth(160, 80, seed=5)
    prod, orc = _run(cfg, 40, "main2d", True, dtype="f32")
    sl = orc["slots"]
    z_p = prod["zeta"][sl["zeta"][2]]
    z_o = orc["zeta"][sl["zeta"][2]]
    u_p, u_o = prod["ubar"][sl["ubar"][2]], orc["ubar"][sl["ubar"][2]]
    ez = np.max(np.abs(z_p - z_o))
    eu = np.max(np.abs(u_p - u_o))
    assert np.all(np.isfinite(z_o))
    print(f"fp32: max|dzeta|={ez:.3e} (scale {np.abs(z_o).max():.3e}), max|dubar|={eu:.3e} "
          f"(scale {np.abs(u_o).max():.3e})", file=sys.stderr)
    assert ez <= 1e-5 * np.abs(z_o).max() + 1e-7
    assert eu <= 1e-4 * max(np.abs(u_o).max(), 1e-3)
