"""Terrain-following s-coordinates (nonlinear/set_coords.py:10-454), host FP64.

Stretching functions C(s) of ROMS (Song and Haidvogel 1994; Shchepetkin's
Vstretching 2 and 4), the sc_w/sc_r grids sc_w[k] = (k-N)/N, sc_r[k-1] = (k-N-0.5)/N
and the end values sc_w[0] = Cs_w[0] = -1, sc_w[N] = Cs_w[N] = 0.
"""
from __future__ import annotations

import math

import numpy as np


def calcCs1(theta_b, theta_s, sc):
    if theta_s != 0.0:
        cff1 = 1.0 / math.sinh(theta_s)
        cff2 = 0.5 / math.tanh(theta_s * 0.5)
        w2, w1 = 1.0 - theta_b, theta_b
        return w2 * cff1 * math.sinh(theta_s * sc) + w1 * (cff2 * math.tanh(theta_s * (sc + 0.5)) - 0.5)
    return sc


def calcCs2(theta_b, theta_s, sc):
    """set_coords.py:36-82 (Shchepetkin 2005 with bottom blending)."""
    alfa, beta = 1.0, 1.0
    if sc == 0.0:
        return 0.0
    if theta_s > 0.0:
        Csur = (1.0 - math.cosh(theta_s * sc)) / (math.cosh(theta_s) - 1.0)
        if theta_b > 0.0:
            Cbot = math.sinh(theta_b * (sc + 1.0)) / math.sinh(theta_b) - 1.0
            Cweight = (sc + 1.0) ** alfa * (1.0 + (alfa / beta) * (1.0 - (sc + 1.0) ** beta))
            return Cweight * Csur + (1.0 - Cweight) * Cbot
        return Csur
    return sc


def calcCs3(theta_b, theta_s, sc):
    """set_coords.py:84-134 (Geyer): theta_s, theta_b are the surface/bottom exponents."""
    if sc == 0.0:
        return 0.0
    Hscale = 3.0
    Cbot = math.log(math.cosh(Hscale * (sc + 1.0) ** theta_b)) / math.log(math.cosh(Hscale)) - 1.0
    Csur = -math.log(math.cosh(Hscale * abs(sc) ** theta_s)) / math.log(math.cosh(Hscale))
    Cweight = 0.5 * (1.0 - math.tanh(Hscale * (sc + 0.5)))
    return Cweight * Cbot + (1.0 - Cweight) * Csur


def calcCs4(theta_b, theta_s, sc):
    """set_coords.py:136-192"""
    if sc == 0.0:
        return 0.0
    if theta_s > 0.0:
        Csur = (1.0 - math.cosh(theta_s * sc)) / (math.cosh(theta_s) - 1.0)
    else:
        Csur = -sc ** 2
    if theta_b > 0.0:
        return (math.exp(theta_b * Csur) - 1.0) / (1.0 - math.exp(-theta_b))
    return Csur


_CS = {1: calcCs1, 2: calcCs2, 3: calcCs3, 4: calcCs4}


def s_coordinates(N: int, Vstretching: int, theta_s: float, theta_b: float):
    """Returns (sc_r[N], sc_w[N+1], Cs_r[N], Cs_w[N+1]) (set_coords.py:363-411)."""
    if Vstretching not in _CS:
        raise ValueError(f"Vstretching={Vstretching} not supported (1-4; 5 is untested in the reference)")
    Cs = _CS[Vstretching]
    sc_w = np.zeros(N + 1)
    sc_r = np.zeros(N)
    Cs_w = np.zeros(N + 1)
    Cs_r = np.zeros(N)
    sc_w[0] = -1.0
    Cs_w[0] = -1.0
    ds = 1.0 / N
    for k in range(1, N + 1):
        sc_w[k] = ds * (k - N)
        sc_r[k - 1] = ds * (k - N - 0.5)
        Cs_w[k] = Cs(theta_b, theta_s, sc_w[k])
        Cs_r[k - 1] = Cs(theta_b, theta_s, sc_r[k - 1])
    sc_w[N] = 0.0
    Cs_w[N] = 0.0
    return sc_r, sc_w, Cs_r, Cs_w


def set_coords(GRID):
    """Fill GRID.sc_r/sc_w/Cs_r/Cs_w and hc (set_coords.py:345-411)."""
    if GRID.Vtransform == 1:
        hc = min(GRID.hmin, GRID.Tcline)
    else:
        hc = GRID.Tcline
    sc_r, sc_w, Cs_r, Cs_w = s_coordinates(GRID.N, GRID.Vstretching, GRID.theta_s, GRID.theta_b)
    GRID.set_vertical(sc_r, sc_w, Cs_r, Cs_w, hc=hc)
    return GRID
