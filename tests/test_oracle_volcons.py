"""Pins of the obc_volcons oracle (oracle/sf_oracle_phys.c:sfo_obc_volcons) by closed
forms: uniform inflow through one open segment gives ubar_xs = inflow velocity and
zero flux after the correction; balanced in/outflow gives ubar_xs = 0; a closed basin
has no open area; after the correction the recomputed net flux vanishes for any
state (the constraint the routine enforces)."""
from __future__ import annotations

import numpy as np

from oracle import oracle_phys as OP


def _grid(szI=20, szJ=12, dx=1e3, dy=2e3, depth=50.0):
    h = np.full((szJ, szI), depth)
    pm = np.full((szJ, szI), 1.0 / dx)
    pn = np.full((szJ, szI), 1.0 / dy)
    return h, pm, pn


def test_uniform_west_inflow():
    szI, szJ = 20, 12
    h, pm, pn = _grid(szI, szJ)
    z = np.zeros((szJ, szI))
    u = np.zeros((szJ, szI))
    v = np.zeros((szJ, szI))
    u[:, 1] = 0.3
    area, flux, xs = OP.obc_volcons(szI, szJ, z, h, pm, pn, u, v, 1, apply=True)
    assert abs(area - (szJ - 2) * 50.0 * 2e3) < 1e-6
    assert abs(xs - 0.3) < 1e-15
    assert np.all(np.abs(u[1:-1, 1]) < 1e-15)


def test_balanced_flow_needs_no_correction():
    szI, szJ = 16, 10
    h, pm, pn = _grid(szI, szJ)
    z = np.zeros((szJ, szI))
    u = np.zeros((szJ, szI))
    v = np.zeros((szJ, szI))
    u[:, 1] = 0.2        # in through the west
    u[:, szI - 1] = 0.2  # out through the east
    _a, flux, xs = OP.obc_volcons(szI, szJ, z, h, pm, pn, u, v, 1 | 2, apply=False)
    assert abs(flux) < 1e-9 and abs(xs) < 1e-15


def test_closed_basin_has_no_open_area():
    h, pm, pn = _grid()
    z = np.zeros_like(h)
    u = np.ones_like(h)
    v = np.ones_like(h)
    area, flux, xs = OP.obc_volcons(20, 12, z, h, pm, pn, u, v, 0, apply=True)
    assert area == 0.0 and flux == 0.0 and xs == 0.0
    assert np.all(u == 1.0)


def test_correction_zeroes_net_flux_for_any_state():
    rng = np.random.default_rng(4)
    szI, szJ = 23, 17
    h = 20 + 30 * rng.random((szJ, szI))
    pm = 1.0 / (1e3 * (1 + 0.2 * rng.random((szJ, szI))))
    pn = 1.0 / (1e3 * (1 + 0.2 * rng.random((szJ, szI))))
    z = 0.1 * rng.standard_normal((szJ, szI))
    u = 0.2 * rng.standard_normal((szJ, szI))
    v = 0.2 * rng.standard_normal((szJ, szI))
    area, flux, xs = OP.obc_volcons(szI, szJ, z, h, pm, pn, u, v, 15, apply=True)
    _a2, flux2, xs2 = OP.obc_volcons(szI, szJ, z, h, pm, pn, u, v, 15, apply=False)
    assert abs(flux2) <= 1e-12 * area * max(abs(xs), 1e-3)
    assert abs(xs2) < 1e-13
