"""Oracle of the coupled split-explicit baroclinic step: the 3D kernels of
oracle/sf_oracle3d.c (the reference's rhs3d, set_maxflux, omega, set_depth, step3d_UV
solve) with the 2D/3D coupling of ROMS (sfo_couple_rufrc, sfo_step3d_uv_coupled) and
the physics-mode barotropic substep coupled to them (sfo_phys_substep_coupled).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

One baroclinic step (nonlinear/main3d.py:87-325 with the coupling of
obsolete/step2d_LF_AM3.h.py:778-846, :2377-2590 and nonlinear/step3d_uv.py:365-544):
set_maxflux, omega, set_zeta, rhs3d, rufrc/rvfrc, nfast+1 coupled physics substeps (the
first forward/backward Euler, averages reset on it), set_depth, the coupled
step3d_UV, setLateralUVBCs, omega.  No vbar = 0 after the substeps (the reference's
main3d.py:189 work-around is not part of this form).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import oracle as O
from . import oracle3d as O3
from . import oracle_phys as OP

P = C.c_void_p
F64 = C.c_double


class sfo_couple(C.Structure):
    _fields_ = [("rufrc", P), ("rvfrc", P), ("Fu", P), ("Fv", P), ("Fu1", P), ("Fv1", P), ("Fu2", P), ("Fv2", P),
                ("DU2", P), ("DV2", P), ("w2p", F64), ("w2c", F64), ("ab", C.c_int)]


class sfo_uv_coupled(C.Structure):
    _fields_ = [(n, P) for n in ("u", "v", "ru", "rv", "ru1", "rv1", "ru2", "rv2", "ru_keep", "rv_keep", "ubar_t1",
                                 "vbar_t1", "ubar_t2", "vbar_t2", "Hz_old", "Hz", "Akv", "z_r", "DU_avg1", "DV_avg1",
                                 "DU_avg2", "DV_avg2", "Huon", "Hvom", "sustr", "svstr")] + \
        [("rdrg", F64), ("ab", C.c_int), ("SD", P), ("D", P), ("RHS", P)]


O.declare({
    "sfo_phys_substep_coupled": ([OP._PP, F64, C.c_int, F64] + [P] * 22 + [C.POINTER(sfo_couple)], None),
    "sfo_couple_rufrc": ([O3._G3, P, P, P, P, P, P, F64, P, P], None),
    "sfo_step3d_uv_coupled": ([O3._G3, C.POINTER(sfo_uv_coupled), F64, F64], None),
})


def avg2_weights(weight2, iif: int) -> tuple[float, float]:
    """Weights of DU_avg2 for substep iif (1-based) of a barotropic loop: the fluxes of
    the predictor's state and of the predicted state (step2d_LF_AM3.h.py:778-846;
    ROMS weight(2, k) is weight2[k-1] here, indices past the table read 0)."""
    w = lambda k: float(weight2[k - 1]) if 1 <= k <= len(weight2) else 0.0
    if iif == 1:
        return (-1.0 / 12.0) * w(iif + 1), w(iif)
    return (8.0 / 12.0) * w(iif) - (1.0 / 12.0) * w(iif + 1), (5.0 / 12.0) * w(iif)


def ab_order(n: int) -> int:
    """Adams-Bashforth order index of the forcing extrapolation in baroclinic step n
    (0-based from the start of the run): 0 (F), 1 (3/2, -1/2), 2 (23/12, -16/12, 5/12)."""
    return 0 if n == 0 else (1 if n == 1 else 2)


class OCoupled:
    """Host copies of the coupled model's state (FP64, flat arrays)."""

    def __init__(self, cfg, N, metrics, sc_r, sc_w, Cs_r, Cs_w, hc, BCfield, uType, vType, rdrg=3e-4, g=10.0,
                 tracers=None):
        """tracers: None, or dict(T=..., S=..., eos=(R0, T0, S0, Tcoef, Scoef), rho0=...) for the
        baroclinic pressure gradient (after rhs3d) and the tracer step (after step3d_UV)."""
        self.ph = OP.OPhys(cfg, metrics["fomn"], rdrg=rdrg, g=g)
        self.s = self.ph.st
        fields = dict(h=cfg["h"], pm=cfg["pm"], pn=cfg["pn"], on_u=metrics["on_u"], om_v=metrics["om_v"])
        self.g2 = O.OGrid(cfg["szI"], cfg["szJ"], N, fields)
        self.m = O3.OModel3(self.g2, N, metrics["fomn"], sc_r, sc_w, Cs_r, Cs_w, hc, 2, BCfield, uType, vType)
        S2 = self.g2.S2
        self.N = N
        self.rdrg = float(rdrg)
        self.rufrc, self.rvfrc = np.zeros(S2), np.zeros(S2)
        self.Fu = [np.zeros(S2) for _ in range(3)]
        self.Fv = [np.zeros(S2) for _ in range(3)]
        self.DU2, self.DV2 = np.zeros(S2), np.zeros(S2)
        # right-hand sides of the 3D momentum of this step and the two before (rotated)
        self.ruh = [np.zeros(N * S2) for _ in range(3)]
        self.rvh = [np.zeros(N * S2) for _ in range(3)]
        self.g = float(g)
        self.tr = None
        if tracers is not None:
            n = N * S2
            self.tr = dict(T=np.ascontiguousarray(np.broadcast_to(tracers["T"], (n,)).astype(float)),
                           S=np.ascontiguousarray(np.broadcast_to(tracers["S"], (n,)).astype(float)),
                           eos=tracers["eos"], rho0=float(tracers["rho0"]), rho=np.zeros(n))

    def substep(self, dt, first, w1, cp):
        v = self.s.v
        p = O.ptr
        sc = self.ph.scr
        O.lib().sfo_phys_substep_coupled(C.byref(self.ph.c), dt, int(first), w1,
                                         p(v("zeta", 0)), p(v("zeta", 1)), p(v("zeta", 2)),
                                         p(v("ubar", 0)), p(v("ubar", 1)), p(v("ubar", 2)),
                                         p(v("vbar", 0)), p(v("vbar", 1)), p(v("vbar", 2)),
                                         p(v("rzeta", 0)), p(v("rzeta", 1)), p(v("rubar", 0)), p(v("rubar", 1)),
                                         p(v("rvbar", 0)), p(v("rvbar", 1)),
                                         p(sc[0]), p(sc[1]), p(sc[2]), p(sc[3]),
                                         p(self.s.Zt_avg1), p(self.s.DU_avg1), p(self.s.DV_avg1), C.byref(cp))

    def baroclinic_step(self, iic, ntfirst, nfast, dtfast, dt, weight1, weight2, time):
        L = O.lib()
        m, s, p = self.m, self.s, O.ptr
        u2, v2 = m.u[2], m.v[2]
        L.sfo_set_maxflux(m.cp, p(u2), p(v2), p(m.Huon), p(m.Hvom), p(m.Hz))
        L.sfo_omega(m.cp, p(m.W), p(m.Huon), p(m.Hvom), p(m.z_w), p(m.BC))
        L.sfo_set_zeta(m.cp, p(s.v("zeta", 1)), p(s.v("zeta", 2)), p(s.Zt_avg1))
        m.rhs3d()
        tr = self.tr
        if tr is not None:
            L.sfo_rho_eos(m.cp, p(tr["T"]), p(tr["S"]), *tr["eos"], p(tr["rho"]))
            L.sfo_prsgrd(m.cp, p(tr["rho"]), p(m.z_r), p(m.z_w), p(m.Hz), self.g, tr["rho0"], p(m.ru[2]),
                         p(m.rv[2]))
        st = self.ph.static
        L.sfo_couple_rufrc(m.cp, p(m.ru[2]), p(m.rv[2]), p(u2), p(v2), p(st["sustr"]), p(st["svstr"]), self.rdrg,
                           p(self.rufrc), p(self.rvfrc))
        n = iic - ntfirst
        cur, h1, h2 = n % 3, (n - 1) % 3, (n - 2) % 3
        for iif in range(1, nfast + 2):
            if iif == 1:
                s.Zt_avg1[:] = 0.0
                s.DU_avg1[:] = 0.0
                s.DV_avg1[:] = 0.0
            s.cycleTimes2D()
            k1 = min(max(iif - 1, 0), len(weight1) - 1)
            w2p, w2c = avg2_weights(weight2, iif)
            cp = sfo_couple(p(self.rufrc), p(self.rvfrc), p(self.Fu[cur]), p(self.Fv[cur]), p(self.Fu[h1]),
                            p(self.Fv[h1]), p(self.Fu[h2]), p(self.Fv[h2]), p(self.DU2), p(self.DV2), w2p, w2c,
                            ab_order(n))
            self.substep(dtfast, iif == 1, float(weight1[k1]), cp)
        Hz_old = m.Hz.copy()
        m.set_depth(s)
        st = self.ph.static
        a = sfo_uv_coupled(p(u2), p(v2), p(m.ru[2]), p(m.rv[2]), p(self.ruh[h1]), p(self.rvh[h1]), p(self.ruh[h2]),
                           p(self.rvh[h2]), p(self.ruh[cur]), p(self.rvh[cur]), p(s.v("ubar", 1)), p(s.v("vbar", 1)),
                           p(s.v("ubar", 2)), p(s.v("vbar", 2)), p(Hz_old), p(m.Hz), p(m.AKv), p(m.z_r),
                           p(s.DU_avg1), p(s.DV_avg1), p(self.DU2), p(self.DV2), p(m.Huon), p(m.Hvom),
                           p(st["sustr"]), p(st["svstr"]), self.rdrg, ab_order(n), p(m.eqSD), p(m.eqD), p(m.eqRHS))
        L.sfo_step3d_uv_coupled(m.cp, C.byref(a), 1.0, float(dt))
        if tr is not None:
            for name in ("T", "S"):
                new = np.zeros_like(tr[name])
                L.sfo_step3d_t(m.cp, p(tr[name]), p(new), p(m.Huon), p(m.Hvom), p(Hz_old), p(m.Hz), float(dt))
                tr[name][:] = new
        rc = L.sfo_setLateralUVBCs(m.cp, time, p(u2), p(v2), p(m.BC), p(m.BC), p(m.uType), p(m.vType))
        assert rc == 0, "gradient v boundaries are not supported"
        L.sfo_omega(m.cp, p(m.W), p(m.Huon), p(m.Hvom), p(m.z_w), p(m.BC))
