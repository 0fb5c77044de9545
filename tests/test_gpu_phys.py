"""GPU parity of the physics-mode substep against the CPU oracle: the four-launch
form (sf_phys_substep) and the fused one-launch form (sf_phys_substep_fused, every
tile height) equal the oracle bit for bit in FP64.  In FP32 the fused form (compiled
with fused multiply-adds) agrees with the four-launch form (separately rounded
products) to a few FP32 roundings, and stays within a stated tolerance of FP64."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from sailfish_amd import configs
from tests.helpers2d import bitwise_equal

pytestmark = pytest.mark.gpu

FIELDS = ("zeta", "ubar", "vbar")
RHS = ("rzeta", "rubar", "rvbar")
AVG = ("Zt_avg1", "DU_avg1", "DV_avg1")


def _gpu_run(cfg, nsteps, dtype, fused, tile_j=0):
    import torch
    from sailfish_amd.model2d import build_model
    from sailfish_amd.phys2d import PhysStepper, run_phys

    ct, GRID, OCEAN, _B = build_model(cfg, dtype=dtype, ndtfast=30)
    st = PhysStepper(GRID, OCEAN, fused=fused, tile_j=tile_j)
    run_phys(ct, GRID, OCEAN, nsteps, stepper=st)
    torch.cuda.synchronize()
    return ct, GRID, OCEAN


def _oracle_run(cfg, GRID, ct, nsteps):
    from oracle import oracle_phys as OP
    o = OP.OPhys(cfg, GRID.host["fomn"])
    o.run(nsteps, cfg["dtfast"], ct.weight1)
    return o


def _mismatches(OCEAN, o):
    bad = []
    for name in FIELDS:
        for k in range(3):
            if not bitwise_equal(getattr(OCEAN, f"{name}_t{k}").cpu().numpy().ravel(), o.st.v(name, k)):
                bad.append((name, k))
    for name in RHS:
        for k in range(2):
            if not bitwise_equal(getattr(OCEAN, f"{name}_t{k}").cpu().numpy().ravel(), o.st.v(name, k)):
                bad.append((name, k))
    for name in AVG:
        if not bitwise_equal(getattr(OCEAN, name).cpu().numpy(), getattr(o.st, name)):
            bad.append(name)
    return bad


def _case_cfg(szI, szJ):
    if (szI, szJ) == (128, 64):
        return configs.basin1(), 60
    return configs.random_smooth(szI, szJ, seed=szI * 3 + szJ), 12


# sizes: small, sz % 64 == 1 and sz % TJ == 1 (shifted tile origin), non-square, config-1 basin
SIZES = [(40, 24), (65, 17), (129, 33), (67, 41), (128, 64)]


@pytest.mark.parametrize("szI,szJ", SIZES)
@pytest.mark.parametrize("mode", ["unfused", "fused4", "fused8", "fused16"])
def test_phys_fp64_bitwise(szI, szJ, mode):
    import torch
    print(f"[case] phys fp64 {mode} {szI}x{szJ}", file=sys.stderr, flush=True)
    cfg, n = _case_cfg(szI, szJ)
    fused = mode != "unfused"
    tj = int(mode[5:]) if fused else 0
    ct, GRID, OCEAN = _gpu_run(cfg, n, torch.float64, fused, tj)
    o = _oracle_run(cfg, GRID, ct, n)
    bad = _mismatches(OCEAN, o)
    assert not bad, bad
    assert np.isfinite(OCEAN.zeta_t2.cpu().numpy()).all()


@pytest.mark.parametrize("szI,szJ", [(65, 17), (130, 66)])
@pytest.mark.parametrize("tj", [4, 8, 16])
def test_phys_fp32_fused_matches_unfused(szI, szJ, tj):
    """10 substeps: every field of the fused FP32 form within 2e-6 of the field's
    magnitude of the four-launch FP32 form (the two differ only in the rounding of
    multiply-adds, ~1e-7 relative per operation)."""
    import torch
    print(f"[case] phys fp32 fused{tj} vs unfused {szI}x{szJ}", file=sys.stderr, flush=True)
    cfg = configs.random_smooth(szI, szJ, seed=7)
    _, _, A = _gpu_run(cfg, 10, torch.float32, True, tj)
    _, _, B = _gpu_run(cfg, 10, torch.float32, False)
    bad = []

    def close(x, y, name):
        x, y = x.astype(np.float64), y.astype(np.float64)
        scale = max(np.abs(y).max(), 1e-30)
        err = np.abs(x - y).max() / scale
        print(f"  {name}: max rel diff {err:.2e}", file=sys.stderr)
        return err <= 2e-6

    for name in FIELDS + RHS:
        for k in range(3 if name in FIELDS else 2):
            x = getattr(A, f"{name}_t{k}").cpu().numpy()
            y = getattr(B, f"{name}_t{k}").cpu().numpy()
            if not close(x, y, f"{name}_t{k}"):
                bad.append((name, k))
    for name in AVG:
        if not close(getattr(A, name).cpu().numpy(), getattr(B, name).cpu().numpy(), name):
            bad.append(name)
    assert not bad, bad


def test_phys_fp32_error_vs_fp64():
    """basin1 (128x64 closed basin, 5 km, dt 20 s), 400 substeps: FP32 state with FP64
    averages against the FP64 oracle.  Tolerance: max|dzeta| <= 2e-4 max|zeta|,
    max|dubar| <= 2e-3 max|ubar|, Zt_avg1 within 1e-4 relative."""
    import torch
    print("[case] phys fp32 error", file=sys.stderr, flush=True)
    cfg = configs.basin1()
    ct, GRID, OCEAN = _gpu_run(cfg, 400, torch.float32, True)
    o = _oracle_run(cfg, GRID, ct, 400)
    z = OCEAN.zeta_t2.cpu().double().numpy().ravel()
    zo = o.st.v("zeta", 2)
    u = OCEAN.ubar_t2.cpu().double().numpy().ravel()
    uo = o.st.v("ubar", 2)
    ez = np.abs(z - zo).max() / np.abs(zo).max()
    eu = np.abs(u - uo).max() / np.abs(uo).max()
    ea = np.abs(OCEAN.Zt_avg1.cpu().numpy() - o.st.Zt_avg1).max() / max(np.abs(o.st.Zt_avg1).max(), 1e-30)
    print(f"fp32 phys: rel zeta {ez:.2e}, rel ubar {eu:.2e}, rel Zt_avg1 {ea:.2e}", file=sys.stderr)
    assert ez <= 2e-4 and eu <= 2e-3 and ea <= 1e-4


def test_phys_full_size_conservation_and_fp32():
    """The benchmark workload itself (2048 x 1024 closed basin, basin2 recipe), 30
    substeps: in FP64 the free-surface volume sum(zeta / (pm pn)) over the interior is
    conserved to 1e-12 of sum(|zeta| / (pm pn)) (closed walls, flux-form continuity);
    the FP32 run stays within 1e-4 of max|zeta| and 1e-3 of max|ubar| of the FP64 run;
    both stay finite."""
    import torch
    print("[case] phys full size 2048x1024", file=sys.stderr, flush=True)
    cfg = configs.basin2(2048, 1024)
    area = (1.0 / (cfg["pm"] * cfg["pn"]))[1:-1, 1:-1]

    def volume(z):
        return float((z.reshape(1024, 2048)[1:-1, 1:-1] * area).sum()), float((np.abs(z.reshape(1024, 2048)[1:-1, 1:-1]) * area).sum())

    _, _, O64 = _gpu_run(cfg, 1, torch.float64, True)
    V0, S0 = volume(O64.zeta_t2.cpu().numpy())
    _, _, O64 = _gpu_run(cfg, 30, torch.float64, True)
    _, _, O32 = _gpu_run(cfg, 30, torch.float32, True)
    z64 = O64.zeta_t2.cpu().numpy()
    V1, _ = volume(z64)
    print(f"volume drift {abs(V1 - V0) / S0:.2e}", file=sys.stderr)
    assert abs(V1 - V0) <= 1e-12 * S0
    z32 = O32.zeta_t2.cpu().double().numpy()
    u64, u32 = O64.ubar_t2.cpu().numpy(), O32.ubar_t2.cpu().double().numpy()
    assert np.isfinite(z64).all() and np.isfinite(z32).all()
    ez = np.abs(z32 - z64).max() / np.abs(z64).max()
    eu = np.abs(u32 - u64).max() / np.abs(u64).max()
    print(f"fp32 vs fp64 at full size: zeta {ez:.2e}, ubar {eu:.2e}", file=sys.stderr)
    assert ez <= 1e-4 and eu <= 1e-3
