"""Barotropic drivers (nonlinear/main2d.py:9-88 and the substep loop of
nonlinear/main3d.py:166-189).

The reference's main2d never ends (its step counter ``iic`` is not advanced,
main2d.py:38-83, mod_comptimes.py:69-71); here the caller gives the number of substeps.
"""
from __future__ import annotations

from .step2d import step2dCorrector, step2dPredictor


def main2d(compTimes, GRID, OCEAN, BOUNDARY, nsteps: int, fused: bool = False):
    """main2d.py:38-83 without plotting: next2DTimeStep, cycleTimes2D, predictor,
    corrector; ``iif`` keeps growing (weight1 index clamped past 257)."""
    from .step2d_fused import step2d_substep
    for _ in range(int(nsteps)):
        compTimes.next2DTimeStep()
        OCEAN.cycleTimes2D()
        if fused:
            step2d_substep(compTimes, GRID, OCEAN, BOUNDARY, zero_vbar=False)
        else:
            step2dPredictor(compTimes, GRID, OCEAN, BOUNDARY)
            step2dCorrector(compTimes, GRID, OCEAN, BOUNDARY)


def barotropic_loop(compTimes, GRID, OCEAN, BOUNDARY, fused: bool = False):
    """The fast-time loop of one baroclinic step (main3d.py:166-189): nfast+1
    substeps, averages zeroed on the first, vbar_t2 zeroed after each corrector
    (main3d.py:189)."""
    from .step2d_fused import step2d_substep
    compTimes.first2DTimeStep()
    for _ in range(compTimes.nfast + 1):
        if compTimes.isFirst2DStep():
            OCEAN.Zt_avg1.zero_()
            OCEAN.DU_avg1.zero_()
            OCEAN.DV_avg1.zero_()
        compTimes.next2DTimeStep()
        OCEAN.cycleTimes2D()
        if fused:
            step2d_substep(compTimes, GRID, OCEAN, BOUNDARY, zero_vbar=True)
        else:
            step2dPredictor(compTimes, GRID, OCEAN, BOUNDARY)
            step2dCorrector(compTimes, GRID, OCEAN, BOUNDARY)
            OCEAN.vbar_t2.zero_()
