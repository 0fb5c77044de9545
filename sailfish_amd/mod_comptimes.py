"""Time counters of the split-explicit scheme (modules/mod_comptimes.py:7-108).

Same attribute names as the reference.  ``weight1/weight2`` are host FP64 arrays
(257 entries): the kernels take ``weight1[iif-1]`` as a scalar argument.
"""
from __future__ import annotations

import numpy as np


class CompTimes:
    def __init__(self, dt: float, ndtfast: int, ntimes: int = 0, ntend: int | None = None):
        self.ntimes = int(ntimes)
        self.dt = float(dt)                     # baroclinic step (s)
        self.ndtfast = int(ndtfast)             # barotropic steps per baroclinic step
        self.dtfast = self.dt / self.ndtfast    # mod_comptimes.py:15
        self.nfast = 0
        self.is2DPredictorStep = False
        self.is2DCorrectorStep = True
        self.indx1 = 0
        self.iic = 0                            # 3D step counter
        self.iif = 0                            # 2D step counter
        self.tdays = 0.0
        self.time = 0.0
        self.time2D = 0.0
        self.ntfirst = 0
        self.ntstart = 0
        # mod_comptimes.py:57 hard-codes ntend=500; callers may set the run length
        self.ntend = 500 if ntend is None else int(ntend)
        self.weight1 = np.zeros(257, dtype=np.float64)
        self.weight2 = np.zeros(257, dtype=np.float64)

    def isInitialTimeStep(self):
        return self.iic == self.ntstart

    def isFinalTimeStep(self):
        return self.iic >= self.ntend + 1

    def isFirst2DStep(self):
        return self.iif == 0

    def first2DTimeStep(self):
        self.iif = 0

    def next2DTimeStep(self):
        self.iif += 1
        self.time2D += self.dtfast

    def nextTimeStep(self):
        self.iic += 1
        self.time += self.dt
        self.time2D = self.time

    def getDeltaTime(self):
        return 2.0 * self.dtfast if self.is2DPredictorStep else self.dtfast

    def cff1(self) -> float:
        """``weight1[iif-1]`` as read by computeZetaRHS (mod_cppkernels.cpp:116).

        The reference reads past the 257-entry array once iif > 257 in main2d (an
        out-of-bounds read); here the index is clamped to the last entry (zero)."""
        k = min(max(self.iif - 1, 0), self.weight1.shape[0] - 1)
        return float(self.weight1[k])
