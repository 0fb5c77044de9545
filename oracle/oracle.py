"""Python side of the CPU oracle: ctypes binding of liboracle.so plus a restatement
of the reference's orchestration (state layout, time-level rotation, BC lists,
predictor/corrector sequence, drivers) on numpy FP64 arrays.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Restated from:
  modules/mod_ocean.py:14-191        state layout and cycleTimes2D/3D
  modules/mod_boundary.py:103-297    buildLBC index sets
  nonlinear/zetabc.py:5-12, nonlinear/barotropicVelocityBC.py:7-31   BC call lists
  nonlinear/step2d.py:23-127         predictor / corrector kernel sequence
  nonlinear/main2d.py:38-83, nonlinear/main3d.py:166-189             drivers
BC lists are applied entry by entry in list order (serial schedule of mod_BC.c).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import math
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

P = C.c_void_p
F64 = C.c_double
I32 = C.c_int32


class sfo_grid(C.Structure):
    _fields_ = [("szI", C.c_int), ("szJ", C.c_int), ("szK", C.c_int),
                ("on_u", P), ("om_v", P), ("pn", P), ("pm", P), ("q5_stride", C.c_int)]


_G = C.POINTER(sfo_grid)
_SIGS2D = {
    "sfo_computeZetaRHS": ([_G, P, P, P, P, P, P, P, C.c_int, F64, P], None),
    "sfo_computeZetaPred": ([_G, F64, P, P, P, P], None),
    "sfo_computeMomentumRHSPred": ([_G, P, P, P, P, P, F64, F64], None),
    "sfo_computeMomentumRHSCorr": ([_G, P, P, P, P, P, P, F64], None),
    "sfo_computeMomentumPred": ([_G, F64, P, P, P, P, P, P, P, P, P], None),
    "sfo_AdamsMoultonCorr3rd": ([_G, F64, P, P, P, P], None),
    "sfo_AdamsMoultonCorr3rd2": ([_G, F64, P, P, P, P, P, P, P, P, P, P, P], None),
    "sfo_setBC": ([P, P, C.c_int, F64], None),
    "sfo_copyBC": ([P, P, P, C.c_int], None),
    "sfo_num_threads": ([], C.c_int),
    "sfo_set_num_threads": ([C.c_int], None),
}


_lib = None
_lib_path = None

# ctypes signatures of every exported oracle function; the 3D and physics drivers add
# theirs at import (declare()), and every load of a library binds all of them
_SIGS: dict = {}


def declare(sigs: dict):
    """Register {name: (argtypes, restype)}; binds them on the loaded library too."""
    _SIGS.update(sigs)
    if _lib is not None:
        _bind(_lib)


def _bind(L):
    for name, (at, rt) in _SIGS.items():
        fn = getattr(L, name)
        fn.argtypes = at
        fn.restype = rt


def _load(path):
    L = C.CDLL(str(path))
    _bind(L)
    return L


def lib():
    """The loaded oracle library: oracle/liboracle.so (built on first use), or the file
    named by the environment variable SF_ORACLE_LIB."""
    global _lib, _lib_path
    if _lib is None:
        env = os.environ.get("SF_ORACLE_LIB")
        path = Path(env) if env else LIB_PATH
        if not env and not path.exists():
            from . import build
            build.build(verbose=False)
        _lib, _lib_path = _load(path), path
    return _lib


@contextlib.contextmanager
def use_library(path):
    """Route every oracle call to another build of the oracle sources (a mutation
    test builds deliberately wrong variants and checks that the pins reject them)."""
    global _lib, _lib_path
    saved = (_lib, _lib_path)
    _lib, _lib_path = _load(path), Path(path)
    try:
        yield _lib
    finally:
        _lib, _lib_path = saved


declare(_SIGS2D)


def ptr(a: np.ndarray) -> int:
    assert a.flags.c_contiguous, "oracle arrays must be C-contiguous"
    return a.ctypes.data


def set_threads(n: int):
    lib().sfo_set_num_threads(int(n))


def num_threads() -> int:
    return lib().sfo_num_threads()


# ---------------------------------------------------------------------------------
# grid / state
# ---------------------------------------------------------------------------------

class OGrid:
    """Host grid: h, pm, pn, f (+ metrics) as flat FP64 arrays."""

    def __init__(self, szI, szJ, N, fields: dict, fix_q5: bool = False):
        self.szI, self.szJ, self.N = int(szI), int(szJ), int(N)
        self.L, self.M = self.szI - 1, self.szJ - 1
        self.S2 = self.szI * self.szJ
        self.f = {k: np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
                  for k, v in fields.items()}
        self.h = self.f["h"]
        self.fix_q5 = bool(fix_q5)
        self.c = sfo_grid(self.szI, self.szJ, self.N, ptr(self.f["on_u"]), ptr(self.f["om_v"]),
                          ptr(self.f["pn"]), ptr(self.f["pm"]), self.szI if fix_q5 else 0)
        self.cp = C.byref(self.c)


class OState:
    """mod_ocean.py:14-191 on numpy: slot arrays + rotating view names."""

    NAMES2D = ("zeta", "ubar", "vbar", "rzeta", "rubar", "rvbar")

    def __init__(self, grid: OGrid):
        S2 = grid.S2
        self.grid = grid
        for n in self.NAMES2D:
            setattr(self, n, np.zeros((3, S2)))
        # slot index aliased by t0, t1, t2 for every rotating field
        self.slot = {n: [0, 1, 2] for n in self.NAMES2D}
        self.Zt_avg1 = np.zeros(S2)
        self.DU_avg1 = np.zeros(S2)
        self.DV_avg1 = np.zeros(S2)

    def v(self, name, level) -> np.ndarray:
        return getattr(self, name)[self.slot[name][level]]

    def cycleTimes2D(self):
        for n in self.NAMES2D:
            s = self.slot[n]
            self.slot[n] = [s[1], s[2], s[0]]

    def copy(self):
        o = OState.__new__(OState)
        o.grid = self.grid
        for n in self.NAMES2D:
            setattr(o, n, getattr(self, n).copy())
        o.slot = {k: list(v) for k, v in self.slot.items()}
        o.Zt_avg1, o.DU_avg1, o.DV_avg1 = self.Zt_avg1.copy(), self.DU_avg1.copy(), self.DV_avg1.copy()
        return o


# ---------------------------------------------------------------------------------
# boundary lists (mod_boundary.py:103-297)
# ---------------------------------------------------------------------------------

BC_CODES = {"CHA": 1 << 1, "CHE": 1 << 2, "CLA": (1 << 3) + 1, "CLO": 1 << 4, "FLA": (1 << 5) + 1,
            "GRA": 1 << 6, "NES": 1 << 8, "RAD": 1 << 11, "RADNUD": (1 << 11) + (1 << 9) + 1,
            "RED": 1 << 12, "SHC": 1 << 13}
bcClamped, bcClosed, bcGradient = 1 << 3, 1 << 4, 1 << 6


def build_lbc(val, i0, j0, L, M):
    """Returns (idx1, idx2, LBC) int32 lists in W, E, N, S block order with the
    N/S blocks placed on row i0 / rows M-1, M (mod_boundary.py:130-163)."""
    W, S, E, N = (BC_CODES[v.upper()] for v in val)
    idx1, idx2, lbc = [], [], []
    row = L + 1
    for j in range(j0, M):
        idx1.append(j * row + i0); idx2.append(j * row + i0 + 1); lbc.append(W)
    for j in range(j0, M):
        idx1.append(j * row + L); idx2.append(j * row + L - 1); lbc.append(E)
    for i in range(i0, L):
        idx1.append(i0 * row + i); idx2.append((i0 + 1) * row + i); lbc.append(N)
    for i in range(i0, L):
        idx1.append(M * row + i); idx2.append((M - 1) * row + i); lbc.append(S)
    return (np.array(idx1, np.int32), np.array(idx2, np.int32), np.array(lbc, np.int32))


class OBoundary:
    def __init__(self, grid: OGrid, lbc: dict | None = None):
        lbc = lbc or {}
        clo = ["CLO"] * 4
        L, M = grid.L, grid.M
        z1, z2, zt = build_lbc(lbc.get("zeta", clo), 0, 0, L, M)
        u1, u2, ut = build_lbc(lbc.get("ubar", clo), 1, 0, L, M)
        v1, v2, vt = build_lbc(lbc.get("vbar", clo), 0, 1, L, M)
        sel = (zt & (bcClosed | bcGradient)) != 0
        self.zeta_copy = (np.ascontiguousarray(z1[sel]), np.ascontiguousarray(z2[sel]))
        self.ubar_closed = np.ascontiguousarray(u1[(ut & bcClosed) != 0])
        self.vbar_closed = np.ascontiguousarray(v1[(vt & bcClosed) != 0])
        self.ubar_clamped = np.ascontiguousarray(u1[(ut & bcClamped) != 0])
        sel = (ut & bcGradient) != 0
        self.ubar_grad = (np.ascontiguousarray(u1[sel]), np.ascontiguousarray(u2[sel]))
        sel = (vt & bcGradient) != 0
        self.vbar_grad1 = np.ascontiguousarray(v1[sel])

    def zetabc(self, zeta):
        a, b = self.zeta_copy
        lib().sfo_copyBC(ptr(zeta), ptr(a), ptr(b), len(a))

    def velocity_bc(self, ubar, vbar, time2D):
        L = lib()
        L.sfo_setBC(ptr(ubar), ptr(self.ubar_closed), len(self.ubar_closed), 0.0)
        L.sfo_setBC(ptr(vbar), ptr(self.vbar_closed), len(self.vbar_closed), 0.0)
        val = 0.05 * math.sin(time2D * 0.0002)
        L.sfo_setBC(ptr(ubar), ptr(self.ubar_clamped), len(self.ubar_clamped), val)
        a, b = self.ubar_grad
        L.sfo_copyBC(ptr(ubar), ptr(a), ptr(b), len(a))
        n = len(self.vbar_grad1)
        assert n <= len(b), "vbar gradient list longer than ubarGradientBCIdx2"
        L.sfo_copyBC(ptr(vbar), ptr(self.vbar_grad1), ptr(b), n)


# ---------------------------------------------------------------------------------
# step2d.py:23-127 and the drivers
# ---------------------------------------------------------------------------------

class OTimes:
    def __init__(self, dtfast, weight1, nfast=0):
        self.dtfast = float(dtfast)
        self.weight1 = np.asarray(weight1, dtype=np.float64)
        self.nfast = int(nfast)
        self.iif = 0
        self.time2D = 0.0

    def cff1(self):
        k = min(max(self.iif - 1, 0), len(self.weight1) - 1)
        return float(self.weight1[k])


G_CONST = 10.0


def step2d_predictor(t: OTimes, g: OGrid, s: OState, b: OBoundary):
    L = lib()
    Dt = t.dtfast
    v = s.v
    L.sfo_computeZetaRHS(g.cp, ptr(v("zeta", 1)), ptr(g.h), ptr(v("ubar", 1)), ptr(v("vbar", 1)),
                         ptr(s.DU_avg1), ptr(s.DV_avg1), ptr(s.Zt_avg1), 1, t.cff1(), ptr(v("rzeta", 1)))
    if t.iif == 0:
        v("zeta", 2)[:] = v("zeta", 1) + Dt * v("rzeta", 1)
        w = 0.5
    else:
        L.sfo_computeZetaPred(g.cp, Dt, ptr(v("zeta", 0)), ptr(v("zeta", 1)), ptr(v("zeta", 2)),
                              ptr(v("rzeta", 1)))
        w = 2.0 / 5.0
    b.zetabc(v("zeta", 2))
    L.sfo_computeMomentumRHSPred(g.cp, ptr(g.h), ptr(v("rubar", 1)), ptr(v("rvbar", 1)), ptr(v("zeta", 0)),
                                 ptr(v("zeta", 2)), G_CONST, w)
    L.sfo_computeMomentumPred(g.cp, Dt, ptr(v("ubar", 1)), ptr(v("ubar", 2)), ptr(v("vbar", 1)),
                              ptr(v("vbar", 2)), ptr(v("rubar", 1)), ptr(v("rvbar", 1)), ptr(g.h),
                              ptr(v("zeta", 1)), ptr(v("zeta", 2)))
    b.velocity_bc(v("ubar", 2), v("vbar", 2), t.time2D)


def step2d_corrector(t: OTimes, g: OGrid, s: OState, b: OBoundary):
    L = lib()
    Dt = t.dtfast
    v = s.v
    L.sfo_computeZetaRHS(g.cp, ptr(v("zeta", 2)), ptr(g.h), ptr(v("ubar", 2)), ptr(v("vbar", 2)),
                         ptr(s.DU_avg1), ptr(s.DV_avg1), ptr(s.Zt_avg1), 0, t.cff1(), ptr(v("rzeta", 1)))
    L.sfo_AdamsMoultonCorr3rd(g.cp, Dt, ptr(v("zeta", 2)), ptr(v("rzeta", 0)), ptr(v("rzeta", 1)),
                              ptr(v("rzeta", 2)))
    b.zetabc(v("zeta", 2))
    L.sfo_computeMomentumRHSCorr(g.cp, ptr(g.h), ptr(v("rubar", 2)), ptr(v("rvbar", 2)), ptr(v("zeta", 0)),
                                 ptr(v("zeta", 1)), ptr(v("zeta", 2)), G_CONST)
    # the 12-argument call of step2d.py:120-123 reaches the 5-parameter
    # AdamsMoultonCorr3rd (mod_operators.py:146): (Dt, ubar_t1, vbar_t1, rubar_t0, rubar_t1)
    L.sfo_AdamsMoultonCorr3rd(g.cp, Dt, ptr(v("ubar", 1)), ptr(v("vbar", 1)), ptr(v("rubar", 0)),
                              ptr(v("rubar", 1)))
    b.velocity_bc(v("ubar", 2), v("vbar", 2), t.time2D)


def main2d(t: OTimes, g: OGrid, s: OState, b: OBoundary, nsteps: int):
    """main2d.py:38-83 for ``nsteps`` substeps."""
    for _ in range(int(nsteps)):
        t.iif += 1
        t.time2D += t.dtfast
        s.cycleTimes2D()
        step2d_predictor(t, g, s, b)
        step2d_corrector(t, g, s, b)


def barotropic_loop(t: OTimes, g: OGrid, s: OState, b: OBoundary):
    """main3d.py:166-189: nfast+1 substeps, averages zeroed first, vbar_t2 := 0."""
    t.iif = 0
    for _ in range(t.nfast + 1):
        if t.iif == 0:
            s.Zt_avg1[:] = 0.0
            s.DU_avg1[:] = 0.0
            s.DV_avg1[:] = 0.0
        t.iif += 1
        t.time2D += t.dtfast
        s.cycleTimes2D()
        step2d_predictor(t, g, s, b)
        step2d_corrector(t, g, s, b)
        s.v("vbar", 2)[:] = 0.0
