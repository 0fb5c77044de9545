"""Barotropic time-averaging weights (nonlinear/set_weights.py:45-231).

Host-side restatement of the ROMS power-law filter
F(xi) = xi^Falpha*(1 - xi^Fbeta) - Fgamma*xi (set_weights.py:36-42) followed by the
centroid-shift post-processing and the secondary weights, all in FP64 Python floats
(the reference runs the same scalar loops on CuPy element views).  The two reference
branches that cannot run are kept as errors: ``range(compTimes)`` (set_weights.py:163)
raises in the reference too.
"""
from __future__ import annotations

from math import cos, pi

import numpy as np

Falpha = 2.0
Fbeta = 4.0
Fgamma = 0.284

POWER_LAW = True
COSINE2 = False


def set_weights(compTimes, power_law: bool = POWER_LAW, cosine2: bool = COSINE2):
    """Fill ``compTimes.weight1/weight2`` (257 entries) and ``compTimes.nfast``."""
    w1 = np.zeros(257, dtype=np.float64)
    w2 = np.zeros(257, dtype=np.float64)
    nfast = 0
    ndtfast = int(compTimes.ndtfast)
    gamma = 0.0

    if power_law:
        scale = (Falpha + 1.0) * (Falpha + Fbeta + 1.0) / ((Falpha + 2.0) * (Falpha + Fbeta + 2.0) * ndtfast)
        gamma = Fgamma * max(0.0, 1.0 - 10.0 / ndtfast)
        for _ in range(16):
            nfast = 0
            for i in range(2 * ndtfast):
                cff = scale * (i + 1)
                w1[i] = (cff ** Falpha) - (cff ** (Falpha + Fbeta)) - gamma * cff
                if w1[i] > 0.0:
                    nfast = i + 1
                if nfast > 0 and w1[i] < 0.0:
                    w1[i] = 0.0
            wsum = 0.0
            shift = 0.0
            for i in range(nfast):
                wsum += w1[i]
                shift += (i + 1) * w1[i]
            scale = scale * shift / (wsum * ndtfast)
    elif cosine2:
        cff = pi / ndtfast
        for i in range(2 * ndtfast):
            arg = cff * (i + 1 - ndtfast)
            if 2.0 * abs(arg) < pi:
                w1[i] = 0.0882 + cos(arg) ** 2
                nfast = i + 1

    for _ in range(ndtfast):
        wsum = 0.0
        shift = 0.0
        for i in range(nfast):
            wsum += w1[i]
            shift += (i + 1) * w1[i]
        shift = shift / wsum
        cff = ndtfast - shift
        if cff > 1.0:
            nfast += 1
            for i in range(nfast - 1, 1, -1):      # stops at i=2, as set_weights.py:148
                w1[i] = w1[i - 1]
            w1[0] = 0.0
        elif cff > 0.0:
            wsum = 1.0 - cff
            for i in range(nfast - 1, 1, -1):      # set_weights.py:156
                w1[i] = wsum * w1[i] + cff * w1[i - 1]
            w1[0] = wsum * w1[0]
        elif cff < -1.0:
            raise RuntimeError("set_weights: centroid shift < -1 reaches the reference's "
                               "range(compTimes) branch (set_weights.py:163), which cannot run")
        elif cff < 0.0:
            wsum = 1.0 + cff
            for i in range(nfast - 1):
                w1[i] = wsum * w1[i] - cff * w1[i + 1]
            w1[nfast] = wsum * w1[nfast]

    for j in range(nfast):
        cff = w1[j]
        for i in range(j):
            w2[i] += cff

    wsum = 0.0
    cff = 0.0
    for i in range(nfast):
        wsum += w1[i]
        cff += w2[i]
    wsum = 1.0 / wsum
    cff = 1.0 / cff
    for i in range(nfast):
        w1[i] = wsum * w1[i]
        w2[i] = cff * w2[i]

    compTimes.nfast = nfast
    compTimes.weight1 = w1
    compTimes.weight2 = w2
    compTimes.gamma = gamma
    return w1, w2, nfast


def weight_moments(weight1, weight2, nfast, ndtfast):
    """The reference's self-reported checks (set_weights.py:205-231):
    returns (cff1, cff2, shift, sum_w1, sum_w2); expected (1, >=1.0001, ~1/2, 1, 1)."""
    cff = cff1 = cff2 = wsum = shift = 0.0
    for i in range(nfast):
        cff += weight1[i]
        cff1 += (i + 1) * weight1[i]
        cff2 += (i + 1) * (i + 1) * weight1[i]
        wsum += weight2[i]
        shift += (i + 0.5) * weight2[i]
    return cff1 / ndtfast, cff2 / (ndtfast * ndtfast), shift / ndtfast, cff, wsum
