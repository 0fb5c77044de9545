"""The closure evaluated inside the substep kernel (sf_phys_substep_closure_fused) against
the CPU oracle, bit for bit in FP64: every substep of the GPU run equals the oracle's
substep whose corrector stage takes sf_closure_stress's stresses of the same t1 state
(sfo_phys_substep_cstress; the predictor keeps the wind stress).  sf_closure_stress itself
equals the network forward kernel bit for bit (tests/test_gpu_mlp.py), so this pins the
in-kernel features, network, stress formula and the ghost rows/columns of the corrector.
Grids include shifted tile origins (size % 64 == 1, size % tile height == 1).  FP32: the
fused run stays within a stated bound of the FP64 run."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from oracle import mlp_ref as R
from sailfish_amd import configs
from tests.helpers2d import bitwise_equal

pytestmark = pytest.mark.gpu


def _params(seed=3):
    th = R.init_params(seed)
    p = R.unpack(th)
    rng = np.random.default_rng(seed)
    p["b1"] = 0.1 * rng.standard_normal(64)
    p["b2"] = 0.1 * rng.standard_normal(64)
    p["b3"][:2] = 0.1 * rng.standard_normal(2)
    return R.pack(p).astype(np.float32).astype(np.float64)


def _cfg(shape):
    if shape == "coarse":
        from sailfish_amd.closure import coarse_config, fine_config
        cfg = coarse_config(fine_config(128, 64, 4))
    else:
        cfg = configs.random_smooth(*shape, seed=shape[0] + 7 * shape[1])
    J, I = cfg["szJ"], cfg["szI"]
    rng = np.random.default_rng(J * I)
    cfg["zeta0"] = configs.gaussian_smooth(0.05 * rng.standard_normal((J, I)), 3.0)
    cfg["sustr"] = 1e-4 * (1.0 + configs.gaussian_smooth(rng.standard_normal((J, I)), 3.0))
    cfg["svstr"] = -5e-5 * (1.0 + configs.gaussian_smooth(rng.standard_normal((J, I)), 3.0))
    cfg.setdefault("f", np.full((J, I), 1e-4))
    cfg.setdefault("rdrg", 3e-4)
    return cfg


def _mlp_for(m, seed):
    import torch
    from sailfish_amd.closure import ClosureMLP
    F = m.features()
    mu, sd = F.mean(0), F.std(0)
    sd = torch.where(sd > 0, sd, torch.ones_like(sd))
    mlp = ClosureMLP(params=_params(seed))
    mlp.set_norm(mu.cpu().numpy(), sd.cpu().numpy(), np.array([1e-8, 2e-7, -1e-8, 2e-7]))
    return mlp


@pytest.mark.parametrize("shape,tile_j", [("coarse", 4), ("coarse", 8), ((65, 33), 4), ((67, 41), 8), ((129, 17), 4)])
def test_fused_closure_fp64_equals_oracle_with_corrector_stress(shape, tile_j):
    import torch
    from oracle import oracle_phys as OP
    from sailfish_amd.closure import Model, closure_buffers
    print(f"[case] fused closure fp64 {shape} tile_j={tile_j}", file=sys.stderr, flush=True)
    cfg = _cfg(shape)
    m = Model(cfg, torch.float64)
    m.stepper.tile_j = tile_j
    m.begin()
    mlp = _mlp_for(m, 11)
    o = OP.OPhys(cfg, m.GRID.host["fomn"], rdrg=cfg["rdrg"])
    su, sv = closure_buffers(m)
    st = m.stepper
    w1 = m.ct.weight1
    nsub = 5
    R_max = 0.0
    for n in range(nsub):
        if n > 0:
            m.begin()
        a = st.fields(m.OCEAN)
        mlp.lib.sf_closure_stress(mlp.handle, C.byref(a), st.precision, m.f.data_ptr(), su.data_ptr(),
                                  sv.data_ptr(), None)
        m.advance(mlp)
        torch.cuda.synchronize()
        cs_u, cs_v = su.cpu().numpy().ravel(), sv.cpu().numpy().ravel()
        R_max = max(R_max, float(np.abs(cs_u - o.static["sustr"]).max()))
        o.st.cycleTimes2D()
        k = min(max(n, 0), len(w1) - 1)
        o.substep_cstress(cfg["dtfast"], n == 0, float(w1[k]), cs_u, cs_v)
        for name in ("zeta", "ubar", "vbar"):
            assert bitwise_equal(getattr(m.OCEAN, f"{name}_t2").cpu().numpy().ravel(), o.st.v(name, 2)), (n, name)
        for name in ("rzeta", "rubar", "rvbar"):
            assert bitwise_equal(getattr(m.OCEAN, f"{name}_t1").cpu().numpy().ravel(), o.st.v(name, 1)), (n, name)
        for name in ("Zt_avg1", "DU_avg1", "DV_avg1"):
            assert bitwise_equal(getattr(m.OCEAN, name).cpu().numpy(), getattr(o.st, name)), (n, name)
    # the closure stress is not negligible against the wind stress it adds to
    assert R_max > 1e-6, R_max
    mlp.close()


def test_fused_closure_fp32_close_to_fp64():
    """FP32 state with the fused closure against the FP64 run of the same network: the
    corrected velocities stay within FP32 rounding of the state (relative to the field's
    range) after 20 substeps."""
    import torch
    from sailfish_amd.closure import Model
    print("[case] fused closure fp32 vs fp64", file=sys.stderr, flush=True)
    cfg = _cfg("coarse")
    m0 = Model(cfg, torch.float64)
    m0.begin()
    mlp = _mlp_for(m0, 11)
    del m0
    out = {}
    for dt in (torch.float64, torch.float32):
        m = Model(cfg, dt)
        for _ in range(20):
            m.step(mlp)
        torch.cuda.synchronize()
        out[dt] = [x.double().cpu().numpy() for x in m.state()]
    for name, a, b in zip(("zeta", "ubar", "vbar"), out[torch.float32], out[torch.float64]):
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
        print(f"{name}: max |fp32 - fp64| / max |fp64| = {err:.2e}", file=sys.stderr)
        assert err < 2e-4, (name, err)
    mlp.close()
