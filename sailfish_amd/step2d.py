"""Barotropic predictor/corrector (nonlinear/step2d.py:23-127), drop-in signatures.

``step2dPredictor`` / ``step2dCorrector`` issue the same kernel sequence as the
reference, each kernel a one-to-one gfx950 replacement.  ``step2d_substep`` runs a
whole substep (predictor + corrector + both boundary passes) as one fused launch;
it produces the same state bit for bit.
"""
from __future__ import annotations

from . import mod_operators as ops
from .mod_boundary import clamped_value
from .mod_constants import g


def zetabc(zeta, compTimes, BOUNDARY):
    """nonlinear/zetabc.py:5-12 (closed/gradient copy list as one composed map)."""
    ops.apply_bc_map(zeta, BOUNDARY.device_maps()["zeta"])


def barotropicVelocityBC(ubar, vbar, compTimes, BOUNDARY):
    """nonlinear/barotropicVelocityBC.py:7-31."""
    m = BOUNDARY.device_maps()
    val = clamped_value(compTimes.time2D)
    ops.apply_bc_map(ubar, m["ubar"], val)
    ops.apply_bc_map(vbar, m["vbar"], val)


def step2dPredictor(compTimes, GRID, OCEAN, BOUNDARY):
    """step2d.py:23-80"""
    c = ops.ctx_of(GRID)
    h = GRID.h.reshape(-1)
    Dt = compTimes.dtfast
    ops.computeZetaRHS(None, None, (OCEAN.zeta_t1, h, OCEAN.ubar_t1, OCEAN.vbar_t1, OCEAN.DU_avg1,
                                    OCEAN.DV_avg1, OCEAN.Zt_avg1, True, compTimes.weight1,
                                    compTimes.weight2, compTimes.iif, OCEAN.rzeta_t1), ctx=c)
    if compTimes.isFirst2DStep():
        # step2d.py:46-52: forward-Euler start (unreachable from main2d/main3d, whose
        # next2DTimeStep has already advanced iif)
        OCEAN.zeta_t2.copy_(OCEAN.zeta_t1 + Dt * OCEAN.rzeta_t1)
        wGzeta = 0.5
    else:
        ops.computeZetaPred(None, None, (Dt, OCEAN.zeta_t0, OCEAN.zeta_t1, OCEAN.zeta_t2, OCEAN.rzeta_t1), ctx=c)
        wGzeta = 2.0 / 5.0
    zetabc(OCEAN.zeta_t2, compTimes, BOUNDARY)
    ops.computeMomentumRHSPred(None, None, (h, OCEAN.rubar_t1, OCEAN.rvbar_t1, OCEAN.zeta_t0, OCEAN.zeta_t2,
                                            g, wGzeta), ctx=c)
    ops.computeMomentumPred(None, None, (Dt, OCEAN.ubar_t1, OCEAN.ubar_t2, OCEAN.vbar_t1, OCEAN.vbar_t2,
                                         OCEAN.rubar_t1, OCEAN.rvbar_t1, h, OCEAN.zeta_t1, OCEAN.zeta_t2), ctx=c)
    barotropicVelocityBC(OCEAN.ubar_t2, OCEAN.vbar_t2, compTimes, BOUNDARY)


def step2dCorrector(compTimes, GRID, OCEAN, BOUNDARY):
    """step2d.py:84-127"""
    c = ops.ctx_of(GRID)
    h = GRID.h.reshape(-1)
    Dt = compTimes.dtfast
    ops.computeZetaRHS(None, None, (OCEAN.zeta_t2, h, OCEAN.ubar_t2, OCEAN.vbar_t2, OCEAN.DU_avg1,
                                    OCEAN.DV_avg1, OCEAN.Zt_avg1, False, compTimes.weight1,
                                    compTimes.weight2, compTimes.iif, OCEAN.rzeta_t1), ctx=c)
    ops.AdamsMoultonCorr3rd(None, None, (Dt, OCEAN.zeta_t2, OCEAN.rzeta_t0, OCEAN.rzeta_t1, OCEAN.rzeta_t2), ctx=c)
    zetabc(OCEAN.zeta_t2, compTimes, BOUNDARY)
    ops.computeMomentumRHSCorr(None, None, (h, OCEAN.rubar_t2, OCEAN.rvbar_t2, OCEAN.zeta_t0, OCEAN.zeta_t1,
                                            OCEAN.zeta_t2, g), ctx=c)
    # step2d.py:120-123 through the handle bound at mod_operators.py:146
    ops.AdamsMoultonCorr3rd2(None, None, (Dt, OCEAN.ubar_t1, OCEAN.vbar_t1,
                                          OCEAN.rubar_t0, OCEAN.rubar_t1, OCEAN.rubar_t2,
                                          OCEAN.rvbar_t0, OCEAN.rvbar_t1, OCEAN.rvbar_t2,
                                          h, OCEAN.zeta_t1, OCEAN.zeta_t2), ctx=c)
    barotropicVelocityBC(OCEAN.ubar_t2, OCEAN.vbar_t2, compTimes, BOUNDARY)
