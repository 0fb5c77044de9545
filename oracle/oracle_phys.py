"""ctypes binding of the physics-mode oracle (oracle/sf_oracle_phys.c) and a driver
over the same slot rotation as T_OCEAN.  TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import oracle as O

P = C.c_void_p
F64 = C.c_double


class sfo_phys(C.Structure):
    _fields_ = [("szI", C.c_int), ("szJ", C.c_int), ("h", P), ("pm", P), ("pn", P), ("fomn", P),
                ("sustr", P), ("svstr", P), ("g", F64), ("rdrg", F64)]


_PP = C.POINTER(sfo_phys)


class sfo_open(C.Structure):
    _fields_ = [("flags", C.c_int), ("volcons", C.c_int), ("ext_z", P * 4), ("ext_u", P * 4), ("out", P)]


O.declare({
    "sfo_phys_substep_open": ([_PP, F64, C.c_int, F64] + [P] * 22 + [C.POINTER(sfo_open), P], None),
    "sfo_phys_substep": ([_PP, F64, C.c_int, F64] + [P] * 22, None),
    "sfo_phys_substep_cstress": ([_PP, P, P, F64, C.c_int, F64] + [P] * 22, None),
    "sfo_phys_rhs_zeta": ([_PP, P, P, P, C.c_int, C.c_int], F64),
    "sfo_phys_rhs_u": ([_PP, P, P, P, P, C.c_int, C.c_int], F64),
    "sfo_phys_rhs_v": ([_PP, P, P, P, P, C.c_int, C.c_int], F64),
    "sfo_obc_volcons": ([C.c_int, C.c_int] + [P] * 6 + [C.c_int, C.c_int, P], None),
})


def lib():
    return O.lib()


class OPhys:
    """State slots like T_OCEAN (rotating t0/t1/t2 views) plus static fields."""

    def __init__(self, cfg, fomn, rdrg=3e-4, g=10.0):
        self.szI, self.szJ = int(cfg["szI"]), int(cfg["szJ"])
        S2 = self.szI * self.szJ
        flat = lambda a: np.ascontiguousarray(np.asarray(a, float).reshape(-1))
        z = np.zeros(S2)
        self.static = {"h": flat(cfg["h"]), "pm": flat(cfg["pm"]), "pn": flat(cfg["pn"]), "fomn": flat(fomn),
                       "sustr": flat(cfg.get("sustr", z)), "svstr": flat(cfg.get("svstr", z))}
        s = self.static
        self.c = sfo_phys(self.szI, self.szJ, O.ptr(s["h"]), O.ptr(s["pm"]), O.ptr(s["pn"]), O.ptr(s["fomn"]),
                          O.ptr(s["sustr"]), O.ptr(s["svstr"]), g, rdrg)
        self.st = O.OState(O.OGrid(self.szI, self.szJ, 1, dict(h=s["h"], pm=s["pm"], pn=s["pn"],
                                                                 on_u=s["pm"], om_v=s["pm"])))
        for k in range(3):
            self.st.zeta[k][:] = flat(cfg["zeta0"])
            if "ubar0" in cfg:
                self.st.ubar[k][:] = flat(cfg["ubar0"])
            if "vbar0" in cfg:
                self.st.vbar[k][:] = flat(cfg["vbar0"])
        self.scr = [np.zeros(S2) for _ in range(4)]

    def substep(self, dt, first, w1):
        v = self.st.v
        p = O.ptr
        lib().sfo_phys_substep(C.byref(self.c), dt, int(first), w1,
                               p(v("zeta", 0)), p(v("zeta", 1)), p(v("zeta", 2)),
                               p(v("ubar", 0)), p(v("ubar", 1)), p(v("ubar", 2)),
                               p(v("vbar", 0)), p(v("vbar", 1)), p(v("vbar", 2)),
                               p(v("rzeta", 0)), p(v("rzeta", 1)), p(v("rubar", 0)), p(v("rubar", 1)),
                               p(v("rvbar", 0)), p(v("rvbar", 1)),
                               p(self.scr[0]), p(self.scr[1]), p(self.scr[2]), p(self.scr[3]),
                               p(self.st.Zt_avg1), p(self.st.DU_avg1), p(self.st.DV_avg1))

    def substep_cstress(self, dt, first, w1, csustr, csvstr):
        """substep with the corrector stage's surface stress taken from csustr / csvstr."""
        v = self.st.v
        p = O.ptr
        self._cs = (np.ascontiguousarray(csustr, dtype=float), np.ascontiguousarray(csvstr, dtype=float))
        lib().sfo_phys_substep_cstress(C.byref(self.c), p(self._cs[0]), p(self._cs[1]), dt, int(first), w1,
                                       p(v("zeta", 0)), p(v("zeta", 1)), p(v("zeta", 2)),
                                       p(v("ubar", 0)), p(v("ubar", 1)), p(v("ubar", 2)),
                                       p(v("vbar", 0)), p(v("vbar", 1)), p(v("vbar", 2)),
                                       p(v("rzeta", 0)), p(v("rzeta", 1)), p(v("rubar", 0)), p(v("rubar", 1)),
                                       p(v("rvbar", 0)), p(v("rvbar", 1)),
                                       p(self.scr[0]), p(self.scr[1]), p(self.scr[2]), p(self.scr[3]),
                                       p(self.st.Zt_avg1), p(self.st.DU_avg1), p(self.st.DV_avg1))

    def open_bc(self, flags, volcons=False, ext_z=None, ext_u=None):
        """Open edges for substep_open (ext_z / ext_u: 4 per-edge arrays or None)."""
        ob = sfo_open(int(flags), int(bool(volcons)))
        self._ext = []
        for e in range(4):
            for arrs, field in ((ext_z, "ext_z"), (ext_u, "ext_u")):
                a = None if arrs is None else arrs[e]
                if a is not None:
                    a = np.ascontiguousarray(np.asarray(a, float))
                    self._ext.append(a)
                    getattr(ob, field)[e] = O.ptr(a)
        self.ob = ob
        self.volcons_out = np.zeros(3)
        return ob

    def substep_open(self, dt, first, w1):
        v = self.st.v
        p = O.ptr
        lib().sfo_phys_substep_open(C.byref(self.c), dt, int(first), w1,
                                    p(v("zeta", 0)), p(v("zeta", 1)), p(v("zeta", 2)),
                                    p(v("ubar", 0)), p(v("ubar", 1)), p(v("ubar", 2)),
                                    p(v("vbar", 0)), p(v("vbar", 1)), p(v("vbar", 2)),
                                    p(v("rzeta", 0)), p(v("rzeta", 1)), p(v("rubar", 0)), p(v("rubar", 1)),
                                    p(v("rvbar", 0)), p(v("rvbar", 1)),
                                    p(self.scr[0]), p(self.scr[1]), p(self.scr[2]), p(self.scr[3]),
                                    p(self.st.Zt_avg1), p(self.st.DU_avg1), p(self.st.DV_avg1),
                                    C.byref(self.ob), p(self.volcons_out))

    def run_open(self, nsteps, dt, weight1, start_first=True, iif0=0):
        iif = iif0
        for n in range(int(nsteps)):
            iif += 1
            self.st.cycleTimes2D()
            k = min(max(iif - 1, 0), len(weight1) - 1)
            self.substep_open(dt, start_first and n == 0, float(weight1[k]))
        return iif

    def run(self, nsteps, dt, weight1, start_first=True, iif0=0):
        iif = iif0
        for n in range(int(nsteps)):
            iif += 1
            self.st.cycleTimes2D()
            k = min(max(iif - 1, 0), len(weight1) - 1)
            self.substep(dt, start_first and n == 0, float(weight1[k]))
        return iif


def obc_volcons(szI, szJ, z, h, pm, pn, u, v, flags, apply=True):
    """sfo_obc_volcons: returns (area, flux, ubar_xs); corrects u, v in place when apply."""
    L = lib()
    out = np.zeros(3)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (z, h, pm, pn)]
    for a in (u, v):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    L.sfo_obc_volcons(int(szI), int(szJ), *[O.ptr(a) for a in arrs], O.ptr(u), O.ptr(v), int(flags), int(apply),
                      O.ptr(out))
    return float(out[0]), float(out[1]), float(out[2])
