"""GPU obc_volcons against the oracle: FP64 sums (area, flux, ubar_xs) within the
reordering bound of an FP64 sum (1e-12 of the summed magnitudes), the corrected
velocities, and the FP32-state form (sums in FP64 of the FP32 values)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import oracle_phys as OP

pytestmark = pytest.mark.gpu


def _fields(szI, szJ, seed):
    rng = np.random.default_rng(seed)
    h = 20 + 30 * rng.random((szJ, szI))
    pm = 1.0 / (1e3 * (1 + 0.2 * rng.random((szJ, szI))))
    pn = 1.0 / (1e3 * (1 + 0.2 * rng.random((szJ, szI))))
    z = 0.1 * rng.standard_normal((szJ, szI))
    u = 0.2 * rng.standard_normal((szJ, szI))
    v = 0.2 * rng.standard_normal((szJ, szI))
    return z, h, pm, pn, u, v


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("flags", [1, 2 | 8, 15])
@pytest.mark.parametrize("szI,szJ", [(23, 17), (700, 300)])
def test_volcons_matches_oracle(dt, flags, szI, szJ):
    import torch
    from sailfish_amd.obc_volcons import VolCons
    print(f"[case] volcons {dt} flags={flags} {szI}x{szJ}", file=sys.stderr, flush=True)
    tdt = torch.float64 if dt == "f64" else torch.float32
    z, h, pm, pn, u, v = _fields(szI, szJ, seed=szI + flags)
    # the oracle sees exactly the values the GPU state holds
    cast = (lambda a: a.astype(np.float32).astype(np.float64)) if dt == "f32" else (lambda a: a)
    z, h, pm, pn, u, v = map(cast, (z, h, pm, pn, u, v))
    dev = [torch.from_numpy(a).to("cuda", tdt).contiguous() for a in (z, h, pm, pn, u, v)]
    vc = VolCons(szI, szJ)
    out = vc(*dev, flags, apply=True).cpu().numpy()
    uo, vo = u.copy(), v.copy()
    area, flux, xs = OP.obc_volcons(szI, szJ, z, h, pm, pn, uo, vo, flags, apply=True)
    mag = area * 0.2 * 4
    assert abs(out[0] - area) <= 1e-12 * area
    assert abs(out[1] - flux) <= 1e-12 * mag
    assert abs(out[2] - xs) <= 1e-12 * mag / area
    ug = dev[4].double().cpu().numpy()
    vg = dev[5].double().cpu().numpy()
    if dt == "f64":
        # corrections identical up to the last-bit difference of ubar_xs
        assert np.abs(ug - uo).max() <= 4e-16 + abs(out[2] - xs)
        assert np.abs(vg - vo).max() <= 4e-16 + abs(out[2] - xs)
    else:
        assert np.abs(ug - uo).max() <= 1e-7 and np.abs(vg - vo).max() <= 1e-7
