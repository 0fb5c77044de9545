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


def build_model3d(cfg: dict, *, N: int, device="cuda", ndtfast: int = 30, dt: float | None = None,
                  lbc=None, fix_q5: bool = True, Vtransform: int = 2, Vstretching: int = 4,
                  theta_s: float = 5.0, theta_b: float = 0.4, Tcline: float = 50.0,
                  dtype=torch.float64):
    """A 3D model as tmp/testProgram.py:36-75 assembles it: grid + s-coordinates,
    initial depths at zeta = 0, AKv = 2e-3 + 8e-3*exp(z_w/150).

    ``dtype`` is the precision of the bulk state (2D and 3D fields, the depths z_w, z_r,
    Hz); torch.float32 gives the FP32 bulk form, whose barotropic averages Zt/DU/DV_avg1
    and grid metrics stay FP64."""
    from .mod_grid import Grid
    from .set_coords import s_coordinates
    szI, szJ = int(cfg["szI"]), int(cfg["szJ"])
    GRID = Grid(szI - 2, szJ - 2, N, device=device, Vtransform=Vtransform, Vstretching=Vstretching,
                theta_s=theta_s, theta_b=theta_b, Tcline=Tcline)
    GRID.fix_q5 = bool(fix_q5)
    extra = {k: cfg[k] for k in ("sustr", "svstr") if k in cfg}
    GRID.set_fields(h=cfg["h"], pm=cfg["pm"], pn=cfg["pn"], f=cfg["f"], **extra)
    sc_r, sc_w, Cs_r, Cs_w = s_coordinates(N, Vstretching, theta_s, theta_b)
    hc = min(GRID.hmin, Tcline) if Vtransform == 1 else Tcline
    GRID.set_vertical(sc_r, sc_w, Cs_r, Cs_w, hc=hc, dtype=dtype)
    dt = float(dt) if dt is not None else cfg["dtfast"] * ndtfast
    compTimes = CompTimes(dt=dt, ndtfast=ndtfast)
    set_weights(compTimes)
    OCEAN = T_OCEAN(GRID, dtype=dtype, with3d=True)
    init_state(OCEAN, cfg)
    # main3d's set_zeta copies Zt_avg1 into zeta t1/t2 before the first barotropic loop
    # (main3d.py:151) and the first set_depth reads it (main3d.py:62-63): the initial free
    # surface starts there too
    OCEAN.Zt_avg1.copy_(torch.from_numpy(np.asarray(cfg["zeta0"], dtype=np.float64).reshape(-1)).to(device))
    BOUNDARY = Boundary(GRID, lbc, device=device)
    mod_operators.initModule(GRID)
    z_w0 = rest_depths(GRID)
    OCEAN.AKv.copy_(torch.from_numpy(2.0e-3 + 8.0e-3 * np.exp(z_w0 / 150.0)).to(device, dtype))
    for name in ("u0", "v0"):
        if name in cfg:
            a = torch.from_numpy(np.asarray(cfg[name], dtype=np.float64)).to(device, dtype)
            getattr(OCEAN, name[0]).copy_(a.expand_as(getattr(OCEAN, name[0])))
    return compTimes, GRID, OCEAN, BOUNDARY


def rest_depths(GRID) -> np.ndarray:
    """z_w at zeta = 0 (host FP64), the reference's set_depth formulas."""
    h = GRID.host["h"]
    N = GRID.N
    sc_w, Cs_w = GRID.host["sc_w"], GRID.host["Cs_w"]
    z = np.zeros((N + 1,) + h.shape)
    hc = GRID.hc
    for k in range(N):
        if GRID.Vtransform == 2:
            z[k] = 0.0 + (0.0 + h) * ((hc * sc_w[k] + Cs_w[k] * h) / (hc + h))
        else:
            z0 = hc * (sc_w[k] - Cs_w[k]) + Cs_w[k] * h
            z[k] = z0 + 0.0 * (1.0 + z0 / h)
    z[N] = 0.0
    return z
