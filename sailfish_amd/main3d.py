"""Baroclinic driver (nonlinear/main3d.py:25-330) without plotting.

Per baroclinic step: set_maxflux, omega, set_zeta, rhs3d, the nfast+1 barotropic
substeps (main3d.py:166-189), the NetCDF history record when one is due
(main3d.py:202-216; ``history`` is an ``io_netcdf.HistoryWriter``), set_depth,
step3d_UV (on copies tmpU/tmpV above 32 levels), setLateralUVBCs, omega, nextTimeStep.  ``fused=True``
runs each barotropic substep as one fused kernel (same state).
"""
from __future__ import annotations

import torch

from . import mod_operators as ops
from .main2d import barotropic_loop
from .rhs import rhs3d


# columns of at most this many levels are solved in registers by step3d_UV, which then
# reads the old velocities before it writes the new ones (SF_STEP3D_REG_LEVELS)
STEP3D_REG_LEVELS = 32


class Main3dState:
    """The driver's scratch: tridiagonal SD/D/RHS and tmpU/tmpV (main3d.py:48-52), in the
    precision of the 3D fields (FP64, or FP32 for the FP32 bulk form).  Up to
    STEP3D_REG_LEVELS levels step3d_UV takes u, v themselves as tmpU/tmpV (no copies)."""

    def __init__(self, GRID, OCEAN):
        z = lambda shape: torch.zeros(shape, dtype=OCEAN.u.dtype, device=GRID.device)
        self.eqSD = z(GRID.shape3DW)
        self.eqD = z(GRID.shape3DW)
        self.eqRHS = z(GRID.shape3DW)
        self.copies = GRID.N > STEP3D_REG_LEVELS
        self.tmpU = z(OCEAN.u_t2.shape) if self.copies else None
        self.tmpV = z(OCEAN.v_t2.shape) if self.copies else None

    def old_velocities(self, OCEAN):
        """tmpU, tmpV for step3d_UV: copies of u, v, or u, v themselves."""
        if not self.copies:
            return OCEAN.u_t2, OCEAN.v_t2
        self.tmpU.copy_(OCEAN.u_t2)
        self.tmpV.copy_(OCEAN.v_t2)
        return self.tmpU, self.tmpV


def initial_depths(GRID, OCEAN):
    """set_depth before the loop (main3d.py:62-63)."""
    ops.set_depth(None, None, (GRID.Vtransform, OCEAN.Zt_avg1, GRID.z_w.reshape(-1), GRID.z_r.reshape(-1),
                               GRID.h.reshape(-1), GRID.hc, GRID.Hz.reshape(-1), GRID.sc_r, GRID.sc_w, GRID.Cs_r,
                               GRID.Cs_w), ctx=ops.ctx_of(GRID))


def baroclinic_step(compTimes, GRID, OCEAN, BOUNDARY, st: Main3dState, fused: bool = False, history=None):
    """One pass of the main3d.py:87-325 loop body."""
    c = ops.ctx_of(GRID)
    dm = BOUNDARY.device_maps()
    BC = dm["bcIdxFieldIdx2"]
    Huon, Hvom, W = OCEAN.Huon.reshape(-1), OCEAN.Hvom.reshape(-1), OCEAN.W.reshape(-1)
    Hz, z_w, z_r = GRID.Hz.reshape(-1), GRID.z_w.reshape(-1), GRID.z_r.reshape(-1)
    ops.set_maxflux(None, None, (OCEAN.u_t2, OCEAN.v_t2, Huon, Hvom, Hz), ctx=c)
    ops.omega(None, None, (W, OCEAN.u_t2, OCEAN.v_t2, Huon, Hvom, z_w, BC), ctx=c)
    ops.set_zeta(None, None, (OCEAN.zeta_t1, OCEAN.zeta_t2, OCEAN.Zt_avg1), ctx=c)
    rhs3d(GRID, OCEAN, BOUNDARY)
    barotropic_loop(compTimes, GRID, OCEAN, BOUNDARY, fused=fused)
    if history is not None:
        history.maybe_write(compTimes, OCEAN)
    initial_depths(GRID, OCEAN)
    tmpU, tmpV = st.old_velocities(OCEAN)
    ops.setVerticalVelEq(None, None, (st.eqSD, st.eqD, st.eqRHS), ctx=c)
    ops.step3d_UV(None, None, (OCEAN.u_t2, OCEAN.v_t2, OCEAN.ru_t2, OCEAN.rv_t2, OCEAN.ubar_t1, OCEAN.vbar_t1,
                               OCEAN.ubar_t2, OCEAN.vbar_t2, Hz, OCEAN.AKv.reshape(-1), z_r, OCEAN.DU_avg1,
                               OCEAN.DV_avg1, tmpU, tmpV, GRID.h.reshape(-1), compTimes.iic,
                               compTimes.ntfirst, 1.0, OCEAN.AK, compTimes.dt), ctx=c)
    ops.setLateralUVBCs(None, None, (compTimes.time, OCEAN.u_t2, OCEAN.v_t2, dm["uFieldIdx2"], dm["vFieldIdx2"],
                                     dm["uFieldType"], dm["vFieldType"]), ctx=c)
    ops.omega(None, None, (W, OCEAN.u_t2, OCEAN.v_t2, Huon, Hvom, z_w, BC), ctx=c)
    compTimes.nextTimeStep()


def main3d(compTimes, GRID, OCEAN, BOUNDARY, nsteps: int | None = None, fused: bool = False, history=None):
    """main3d.py:25-330: initial set_depth, then baroclinic steps until
    isFinalTimeStep (or ``nsteps`` steps).  A run resumed from a restart
    (``io_netcdf.read_restart``) continues exactly where the written run was."""
    st = Main3dState(GRID, OCEAN)
    ops.setVerticalVelEq(None, None, (st.eqSD, st.eqD, st.eqRHS), ctx=ops.ctx_of(GRID))
    initial_depths(GRID, OCEAN)
    n = 0
    while not compTimes.isFinalTimeStep():
        if nsteps is not None and n >= nsteps:
            break
        baroclinic_step(compTimes, GRID, OCEAN, BOUNDARY, st, fused=fused, history=history)
        n += 1
    return st
