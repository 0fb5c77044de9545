"""Assemble a barotropic model (CompTimes, Grid, T_OCEAN, Boundary) from a
configuration dict of host arrays (see configs.py), the way tmp/testProgram.py:36-79
assembles the reference's objects from test.in."""
from __future__ import annotations

import numpy as np
import torch

from . import mod_operators
from .mod_boundary import Boundary
from .mod_comptimes import CompTimes
from .mod_grid import Grid
from .mod_ocean import T_OCEAN
from .set_weights import set_weights


def build_model(cfg: dict, *, device="cuda", dtype=torch.float64, lbc=None, N: int = 1,
                ndtfast: int = 30, with3d: bool = False, fix_q5: bool = False):
    szI, szJ = int(cfg["szI"]), int(cfg["szJ"])
    GRID = Grid(szI - 2, szJ - 2, N, device=device)
    GRID.fix_q5 = bool(fix_q5)
    extra = {k: cfg[k] for k in ("sustr", "svstr") if k in cfg}
    GRID.set_fields(h=cfg["h"], pm=cfg["pm"], pn=cfg["pn"], f=cfg["f"], **extra)
    compTimes = CompTimes(dt=cfg["dtfast"] * ndtfast, ndtfast=ndtfast)
    set_weights(compTimes)
    OCEAN = T_OCEAN(GRID, dtype=dtype, with3d=with3d)
    init_state(OCEAN, cfg)
    BOUNDARY = Boundary(GRID, lbc, device=device)
    mod_operators.initModule(GRID)
    return compTimes, GRID, OCEAN, BOUNDARY


def init_state(OCEAN, cfg: dict):
    """Initial free surface in all three time slots (the slots the first rotated
    substep reads, cf. set_zeta at main3d.py:151); optional initial velocities."""
    dt = OCEAN.dtype
    dev = OCEAN.zeta.device
    z = torch.from_numpy(np.asarray(cfg["zeta0"], dtype=np.float64)).to(dev, dt)
    for k in range(3):
        OCEAN.zeta[k].copy_(z)
    for name in ("ubar", "vbar"):
        if f"{name}0" in cfg:
            a = torch.from_numpy(np.asarray(cfg[f"{name}0"], dtype=np.float64)).to(dev, dt)
            for k in range(3):
                getattr(OCEAN, name)[k].copy_(a)
