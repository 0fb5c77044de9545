"""Pins of the physics-mode oracle (oracle/sf_oracle_phys.c) by the physics the
ROMS LF-AM3 step must reproduce: exact volume conservation in a closed basin,
the gravity-wave seiche period of a flat basin, mirror symmetry, and the inertial
period 2*pi/f."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle_phys as OP
from sailfish_amd import configs
from sailfish_amd.mod_grid import update_metrics


def flat_basin(szI, szJ, dx, H, f0=0.0):
    h = np.full((szJ, szI), H)
    pm = np.full_like(h, 1 / dx)
    return dict(szI=szI, szJ=szJ, h=h, pm=pm, pn=pm.copy(), f=np.full_like(h, f0))


def oracle_for(cfg, rdrg=0.0):
    m = update_metrics(cfg["pm"], cfg["pn"], cfg["f"])
    return OP.OPhys(cfg, m["fomn"], rdrg=rdrg)


def test_volume_conserved_closed_basin():
    cfg = configs.random_smooth(40, 24, seed=2)
    o = oracle_for(cfg, rdrg=3e-4)
    area = (1 / (cfg["pm"] * cfg["pn"]))[1:-1, 1:-1]
    V = lambda: (o.st.v("zeta", 2).reshape(24, 40)[1:-1, 1:-1] * area).sum()
    # the random initial velocities are not zero on the walls: the first substep and
    # the AM3 history of the second carry that wall flux; from then on it is closed
    o.run(2, 2.0, np.zeros(257))
    V0 = V()
    o.run(300, 2.0, np.zeros(257), start_first=False)
    assert abs(V() - V0) <= 1e-13 * abs(cfg["zeta0"]).sum() * area.max()


def test_seiche_period():
    """Flat basin of length Lx = (szI-2)*dx, mode-1 cosine seiche zeta = a*cos(pi*x/Lx):
    period 2*Lx/sqrt(g*H) (C-grid dispersion correction < 0.2% at 64 cells)."""
    szI, szJ, dx, H, g = 66, 4, 1e3, 40.0, 10.0
    cfg = flat_basin(szI, szJ, dx, H)
    Lx = (szI - 2) * dx
    x = (np.arange(szI) - 0.5) * dx
    cfg["zeta0"] = np.tile(1e-3 * np.cos(np.pi * x / Lx), (szJ, 1))
    o = oracle_for(cfg)
    c = np.sqrt(g * H)
    T = 2 * Lx / c
    dt = 0.2 * dx / c
    n = int(round(1.5 * T / dt))
    zs = []
    for k in range(n):
        o.run(1, dt, np.zeros(257), start_first=(k == 0))
        zs.append(o.st.v("zeta", 2).reshape(szJ, szI)[1, 1])
    zs = np.array(zs)
    # first up-crossing after the minimum: zeta(x=0) follows a*cos(2*pi*t/T)
    t = dt * (np.arange(n) + 1)
    imin = int(np.argmin(zs[: int(0.75 * T / dt)]))
    assert abs(t[imin] - T / 2) < 0.01 * T
    imax = imin + int(np.argmax(zs[imin:]))
    assert abs(t[imax] - T) < 0.01 * T


def test_mirror_symmetry_without_rotation():
    szI, szJ = 30, 20
    cfg = flat_basin(szI, szJ, 1e3, 30.0)
    X, Y = np.meshgrid(np.arange(szI) - (szI - 1) / 2, np.arange(szJ))
    cfg["zeta0"] = 0.01 * np.exp(-(X ** 2 + (Y - 9.5) ** 2) / 20.0)
    o = oracle_for(cfg, rdrg=1e-4)
    o.run(60, 3.0, np.zeros(257))
    z = o.st.v("zeta", 2).reshape(szJ, szI)
    # rho points i and szI-1-i mirror about the basin centre (ghost columns 0 and L)
    inner = z[1:-1, 1:-1]
    assert np.allclose(inner, inner[:, ::-1], rtol=1e-10, atol=1e-15)


def test_coriolis_acceleration_first_step():
    """Uniform ubar = U, flat surface, f-plane: one short forward/backward-Euler step
    gives dv/dt = -f*U and du/dt = 0 away from the walls (Coriolis sign and size,
    obsolete/step2d_LF_AM3.h.py:1503-1530)."""
    szI, szJ, dx, H, f0, U, dt = 40, 40, 2e3, 50.0, 1e-4, 0.1, 1.0
    cfg = flat_basin(szI, szJ, dx, H, f0)
    cfg["zeta0"] = np.zeros((szJ, szI))
    cfg["ubar0"] = np.full((szJ, szI), U)
    o = oracle_for(cfg)
    o.run(1, dt, np.zeros(257))
    v = o.st.v("vbar", 2).reshape(szJ, szI)[10:30, 10:30]
    u = o.st.v("ubar", 2).reshape(szJ, szI)[10:30, 10:30]
    assert np.allclose(v, -f0 * U * dt, rtol=1e-6)
    assert np.allclose(u, U, rtol=1e-12)
