"""Oracle of the baroclinic step: ctypes binding of the 3D kernels in liboracle.so
and a restatement of nonlinear/main3d.py:25-330 and nonlinear/rhs.py:16-93 on numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import oracle as O

P = C.c_void_p
F64 = C.c_double
I32 = C.c_int32


class sfo_grid3(C.Structure):
    _fields_ = [("szI", C.c_int), ("szJ", C.c_int), ("N", C.c_int),
                ("on_u", P), ("om_v", P), ("pn", P), ("pm", P)]


_G3 = C.POINTER(sfo_grid3)
O.declare({
    "sfo_addCoriolis": ([_G3, P, P, P, P, P, P], None),
    "sfo_horizontalAdvection": ([_G3, P, P, P, P, P, P, I32], None),
    "sfo_verticalAdvection": ([_G3, P, P, P, P, P, P], None),
    "sfo_set_maxflux": ([_G3, P, P, P, P, P], None),
    "sfo_omega": ([_G3, P, P, P, P, P], None),
    "sfo_set_depth": ([_G3, C.c_int, P, P, P, P, F64, P, P, P, P, P], None),
    "sfo_set_zeta": ([_G3, P, P, P], None),
    "sfo_setLateralUVBCs": ([_G3, F64, P, P, P, P, P, P], C.c_int),
    "sfo_step3d_UV": ([_G3, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, C.c_int, C.c_int, F64, F64,
                       P, P, P], None),
    "sfo_rho_eos": ([_G3, P, P, F64, F64, F64, F64, F64, P], None),
    "sfo_prsgrd": ([_G3, P, P, P, P, F64, F64, P, P], None),
    "sfo_step3d_t": ([_G3, P, P, P, P, P, P, F64], None),
})


def lib():
    return O.lib()


ptr = O.ptr


class OModel3:
    """Host copy of everything main3d touches (FP64, flat arrays)."""

    def __init__(self, g2: O.OGrid, N, fomn, sc_r, sc_w, Cs_r, Cs_w, hc, Vtransform, BCfield, uType, vType):
        self.g2 = g2
        self.N = N
        S2 = g2.S2
        f = g2.f
        self.c = sfo_grid3(g2.szI, g2.szJ, N, ptr(f["on_u"]), ptr(f["om_v"]), ptr(f["pn"]), ptr(f["pm"]))
        self.cp = C.byref(self.c)
        self.fomn = np.ascontiguousarray(np.asarray(fomn, float).ravel())
        self.sc_r, self.sc_w = np.ascontiguousarray(sc_r, float), np.ascontiguousarray(sc_w, float)
        self.Cs_r, self.Cs_w = np.ascontiguousarray(Cs_r, float), np.ascontiguousarray(Cs_w, float)
        self.hc, self.Vtransform = float(hc), int(Vtransform)
        self.BC = np.ascontiguousarray(np.asarray(BCfield, np.int32).ravel())
        self.uType = np.ascontiguousarray(np.asarray(uType, np.int32).ravel())
        self.vType = np.ascontiguousarray(np.asarray(vType, np.int32).ravel())
        n3 = (N + 1) * S2
        for name in ("u", "v", "ru", "rv"):
            setattr(self, name, np.zeros((3, n3)))
        for name in ("W", "Huon", "Hvom", "AKv", "z_w", "eqSD", "eqD", "eqRHS"):
            setattr(self, name, np.zeros(n3))
        self.z_r = np.zeros(N * S2)
        self.Hz = np.zeros(N * S2)
        self.AK = 1e-5

    def set_depth(self, s: O.OState):
        lib().sfo_set_depth(self.cp, self.Vtransform, ptr(s.Zt_avg1), ptr(self.z_w), ptr(self.z_r),
                            ptr(self.g2.h), self.hc, ptr(self.Hz), ptr(self.sc_r), ptr(self.sc_w), ptr(self.Cs_r),
                            ptr(self.Cs_w))

    def rhs3d(self):
        """rhs.py:38-71 with the reference's launch sizes."""
        L = lib()
        S2 = self.g2.S2
        ru, rv = self.ru[2], self.rv[2]
        ru[:] = 0.0
        rv[:] = 0.0
        L.sfo_addCoriolis(self.cp, ptr(self.fomn), ptr(self.u[2]), ptr(self.v[2]), ptr(ru), ptr(rv), ptr(self.Hz))
        nthreads = (S2 // 512 + 1) * 2 * (512 // 2)
        L.sfo_horizontalAdvection(self.cp, ptr(self.u[2]), ptr(self.v[2]), ptr(self.Huon), ptr(self.Hvom), ptr(ru),
                                  ptr(rv), nthreads)
        L.sfo_verticalAdvection(self.cp, ptr(self.u[2]), ptr(self.v[2]), ptr(self.W), ptr(ru), ptr(rv),
                                ptr(self.BC))

    def baroclinic_step(self, t: O.OTimes, s: O.OState, b: O.OBoundary, iic: int, ntfirst: int, dt: float,
                        time: float):
        """main3d.py:87-325 (no 3D time-level rotation: cycleTimes3D is commented out)."""
        L = lib()
        u2, v2 = self.u[2], self.v[2]
        L.sfo_set_maxflux(self.cp, ptr(u2), ptr(v2), ptr(self.Huon), ptr(self.Hvom), ptr(self.Hz))
        L.sfo_omega(self.cp, ptr(self.W), ptr(self.Huon), ptr(self.Hvom), ptr(self.z_w), ptr(self.BC))
        L.sfo_set_zeta(self.cp, ptr(s.v("zeta", 1)), ptr(s.v("zeta", 2)), ptr(s.Zt_avg1))
        self.rhs3d()
        O.barotropic_loop(t, self.g2, s, b)
        self.set_depth(s)
        tmpU, tmpV = u2.copy(), v2.copy()
        L.sfo_step3d_UV(self.cp, ptr(u2), ptr(v2), ptr(self.ru[2]), ptr(self.rv[2]), ptr(s.v("ubar", 1)),
                        ptr(s.v("vbar", 1)), ptr(s.v("ubar", 2)), ptr(s.v("vbar", 2)), ptr(self.Hz), ptr(self.AKv),
                        ptr(self.z_r), ptr(s.DU_avg1), ptr(s.DV_avg1), ptr(tmpU), ptr(tmpV), iic, ntfirst, 1.0, dt,
                        ptr(self.eqSD), ptr(self.eqD), ptr(self.eqRHS))
        rc = L.sfo_setLateralUVBCs(self.cp, time, ptr(u2), ptr(v2), ptr(self.BC), ptr(self.BC), ptr(self.uType),
                                   ptr(self.vType))
        assert rc == 0, "gradient v boundaries are not supported"
        L.sfo_omega(self.cp, ptr(self.W), ptr(self.Huon), ptr(self.Hvom), ptr(self.z_w), ptr(self.BC))
