"""3D momentum right-hand side (nonlinear/rhs.py:16-93), drop-in signature."""
from __future__ import annotations

from . import mod_operators as ops

CURVGRID = False
UV_COR = True
UV_ADV = True
GPUMUL = 2        # utility/misc.py:10 -- the reference's reduced-thread launches
FUSED = True      # one launch (sf_rhs3d) instead of zero + three kernels; same bits


def rhs3d(GRID, OCEAN, BOUNDARY, fused: bool | None = None):
    """ru, rv := 0; Coriolis; horizontal advection; vertical advection (rhs.py:38-71)."""
    c = ops.ctx_of(GRID)
    BC = BOUNDARY.device_maps()["bcIdxFieldIdx2"]
    S2 = GRID.szI * GRID.szJ
    grsz, bksz = (S2 // 512 + 1,), (512,)
    if FUSED if fused is None else fused:
        ops.rhs3d_fused((GRID.fomn.reshape(-1), OCEAN.u_t2, OCEAN.v_t2, OCEAN.Huon.reshape(-1),
                         OCEAN.Hvom.reshape(-1), OCEAN.W.reshape(-1), GRID.Hz.reshape(-1), OCEAN.ru_t2,
                         OCEAN.rv_t2, BC), launch_threads=grsz[0] * GPUMUL * (bksz[0] // GPUMUL),
                        cor=UV_COR, adv=UV_ADV, ctx=c)
        return
    OCEAN.ru_t2.zero_()
    OCEAN.rv_t2.zero_()
    if UV_COR:
        ops.addCoriolis(grsz, bksz, (GRID.fomn.reshape(-1), OCEAN.u_t2, OCEAN.v_t2, OCEAN.ru_t2, OCEAN.rv_t2,
                                     GRID.Hz.reshape(-1)), ctx=c)
    ops.horizontalAdvection((grsz[0] * GPUMUL,), (bksz[0] // GPUMUL,),
                            (OCEAN.u_t2, OCEAN.v_t2, OCEAN.Huon.reshape(-1), OCEAN.Hvom.reshape(-1),
                             OCEAN.ru_t2, OCEAN.rv_t2, BC), ctx=c)
    if UV_ADV:
        ops.verticalAdvection((grsz[0] * GPUMUL,), (bksz[0] // GPUMUL,),
                              (OCEAN.u_t2, OCEAN.v_t2, OCEAN.W.reshape(-1), OCEAN.ru_t2, OCEAN.rv_t2, BC), ctx=c)
