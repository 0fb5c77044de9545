"""Coupled split-explicit baroclinic step: the reference's 3D kernels (rhs3d,
set_maxflux, omega, set_depth) with the barotropic mode
in physics form (ROMS LF-AM3 with advection, Coriolis, wind, drag) and the 2D/3D
coupling of ROMS that the port keeps only as text and comments:

- the vertical integral of the 3D right-hand side plus surface and bottom stress
  forces the barotropic substeps (nonlinear/rhs.py:76-89, commented in the reference;
  obsolete/step2d_LF_AM3.h.py:2377-2590), with an Adams-Bashforth extrapolation over
  baroclinic steps at the first substep;
- the substeps also accumulate the filtered transport DU_avg2/DV_avg2
  (step2d_LF_AM3.h.py:778-846);
- the 3D momentum step is ROMS's: Hz_new u_new = Hz_old u + dt pm pn AB3(ru) with
  the implicit vertical viscosity, bottom drag and surface stress (the reference's
  step3D_uv solve multiplies ru by Hz a second time and keeps no history of ru, which
  grows by about Hz f dt per step once the coupling moves the 3D flow); it then
  replaces the vertical mean of u, v by the averaged barotropic transport and corrects
  the mass fluxes Huon/Hvom to DU_avg2/DV_avg2 (nonlinear/step3d_uv.py:155-180,
  :365-544), so the vbar_t2 = 0 work-around of nonlinear/main3d.py:189 is not needed.

One baroclinic step (``coupled_step``): set_maxflux, omega, set_zeta, rhs3d, rufrc,
nfast+1 coupled substeps (the first forward/backward Euler, averages reset on it),
set_depth, coupled step3d_UV, setLateralUVBCs, omega.  FP64 or FP32 bulk state (the
averages, forcing and coupling sums stay FP64).  The barotropic substep is the fused
physics tile kernel with the coupling terms (``sf_phys_substep_coupled_fused``; the
four-launch ``sf_phys_substep_coupled`` with ``fused=False``)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from . import mod_operators as ops
from .main3d import initial_depths
from .phys2d import PhysStepper, sf_phys_fields
from .rhs import rhs3d

P = C.c_void_p


class sf_couple_fields(C.Structure):
    _fields_ = [("rufrc", P), ("rvfrc", P), ("Fu", P), ("Fv", P), ("Fu1", P), ("Fv1", P), ("Fu2", P), ("Fv2", P),
                ("DU_avg2", P), ("DV_avg2", P), ("w2p", C.c_double), ("w2c", C.c_double), ("ab", C.c_int32)]


class sf_uv_coupled(C.Structure):
    """Fields of the coupled 3D momentum step (sf_uv_coupled of include/sailfish_hip.h)."""
    _fields_ = [(n, P) for n in ("u", "v", "ru", "rv", "ru1", "rv1", "ru2", "rv2", "ru_keep", "rv_keep", "ubar_t1",
                                 "vbar_t1", "ubar_t2", "vbar_t2", "Hz_old", "Hz", "Akv", "z_r", "DU_avg1", "DV_avg1",
                                 "DU_avg2", "DV_avg2", "Huon", "Hvom", "sustr", "svstr")] + \
        [("rdrg", C.c_double), ("ab", C.c_int32), ("SD", P), ("D", P), ("RHS", P)]


def _bind():
    lib = _lib.load()
    if not getattr(lib, "_sf_coupled_bound", False):
        lib.sf_phys_substep_coupled.argtypes = [C.POINTER(sf_phys_fields), C.POINTER(sf_couple_fields), C.c_int32,
                                                C.c_double, C.c_int32, C.c_double, P]
        lib.sf_phys_substep_coupled.restype = C.c_int32
        lib.sf_phys_substep_coupled_fused.argtypes = [C.POINTER(sf_phys_fields), C.POINTER(sf_couple_fields),
                                                      C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_int32, P]
        lib.sf_phys_substep_coupled_fused.restype = C.c_int32
        lib.sf_couple_rufrc.argtypes = [P, C.c_int32, P, P, P, P, P, P, C.c_double, P, P, P]
        lib.sf_couple_rufrc.restype = C.c_int32
        lib.sf_step3d_uv_coupled.argtypes = [P, C.c_int32, C.POINTER(sf_uv_coupled), C.c_double, C.c_double, P]
        lib.sf_step3d_uv_coupled.restype = C.c_int32
        lib._sf_coupled_bound = True
    return lib


def filtered_transport_weights(weight2, iif: int) -> tuple[float, float]:
    """(w2p, w2c): weights of the predictor-state and predicted-state fluxes in
    DU_avg2 at substep iif (1-based) of a barotropic loop, step2d_LF_AM3.h.py:778-846.
    ROMS's weight(2, k) is weight2[k-1]; entries past the table are 0."""
    def w(k):
        return float(weight2[k - 1]) if 1 <= k <= len(weight2) else 0.0
    if iif == 1:
        return (-1.0 / 12.0) * w(2), w(1)
    return (8.0 / 12.0) * w(iif) - (1.0 / 12.0) * w(iif + 1), (5.0 / 12.0) * w(iif)


def ab_order(n: int) -> int:
    """Adams-Bashforth order index in baroclinic step n (iic - ntfirst): 0 (first step),
    1 (3/2, -1/2), 2 (23/12, -16/12, 5/12)."""
    return 0 if n == 0 else (1 if n == 1 else 2)


class CoupledState:
    """Device fields of the coupling: rufrc/rvfrc, three forcing buffers per component
    (this step and the two before it, rotated per baroclinic step), DU_avg2/DV_avg2,
    three 3D right-hand sides per component for the Adams-Bashforth step of u, v (rotated
    the same way), the thicknesses before set_depth, the FP64 stresses, and the physics
    substep (``fused``: one launch per substep, sf_phys_substep_coupled_fused; else the
    four-launch sf_phys_substep_coupled)."""

    def __init__(self, GRID, OCEAN, rdrg: float = 3e-4, g: float | None = None, tracers=None, fused: bool = True,
                 tile_j: int = 0):
        _lib.require_gpu()
        self.lib = _bind()
        dev, S2 = GRID.device, GRID.szI * GRID.szJ
        z = lambda: torch.zeros(S2, dtype=torch.float64, device=dev)
        self.rufrc, self.rvfrc = z(), z()
        self.Fu = [z() for _ in range(3)]
        self.Fv = [z() for _ in range(3)]
        self.DU_avg2, self.DV_avg2 = z(), z()
        shape3 = (GRID.N,) + GRID.shape2D
        z3 = lambda: torch.zeros(shape3, dtype=OCEAN.u.dtype, device=dev)
        self.ruh = [z3() for _ in range(3)]
        self.rvh = [z3() for _ in range(3)]
        self.Hz_old = torch.zeros_like(GRID.Hz)
        zeros = np.zeros(GRID.shape2D)
        host = GRID.host
        self.sustr = torch.from_numpy(np.ascontiguousarray(host.get("sustr", zeros)).reshape(-1)).to(dev)
        self.svstr = torch.from_numpy(np.ascontiguousarray(host.get("svstr", zeros)).reshape(-1)).to(dev)
        self.rdrg = float(rdrg)
        self.fused = bool(fused)
        self.tile_j = int(tile_j)   # fused substep tile height (0: the library's default)
        kw = {} if g is None else {"g": g}
        self.stepper = PhysStepper(GRID, OCEAN, rdrg=rdrg, fused=False, **kw)
        from .main3d import Main3dState
        self.m3 = Main3dState(GRID, OCEAN)
        # baroclinic.Tracers: pressure gradient of T, S added to the 3D right-hand side and
        # a tracer step per baroclinic step (None: the reference's terms only)
        self.tracers = tracers
        if tracers is not None:
            tracers.Hz_old = self.Hz_old     # the tracer step reads the same old thicknesses
        OCEAN.extra_state = {**getattr(OCEAN, "extra_state", {}), "DU_avg2": self.DU_avg2, "DV_avg2": self.DV_avg2,
                             **{f"Fu{k}": t for k, t in enumerate(self.Fu)},
                             **{f"Fv{k}": t for k, t in enumerate(self.Fv)},
                             **{f"ru_h{k}": t for k, t in enumerate(self.ruh)},
                             **{f"rv_h{k}": t for k, t in enumerate(self.rvh)}}

    def couple_fields(self, n: int, iif: int, weight2) -> sf_couple_fields:
        """Coupling fields of substep iif of baroclinic step n (iic - ntfirst): the
        forcing buffers rotate per baroclinic step, the extrapolation order is that of
        ROMS (first step F, second 3/2 -1/2, then 23/12 -16/12 5/12)."""
        cur, h1, h2 = n % 3, (n - 1) % 3, (n - 2) % 3
        w2p, w2c = filtered_transport_weights(weight2, iif)
        ab = ab_order(n)
        p = lambda t: t.data_ptr()
        return sf_couple_fields(p(self.rufrc), p(self.rvfrc), p(self.Fu[cur]), p(self.Fv[cur]), p(self.Fu[h1]),
                                p(self.Fv[h1]), p(self.Fu[h2]), p(self.Fv[h2]), p(self.DU_avg2), p(self.DV_avg2),
                                w2p, w2c, ab)


def coupled_step(compTimes, GRID, OCEAN, BOUNDARY, st: CoupledState, history=None):
    """One baroclinic step of the coupled form."""
    c = ops.ctx_of(GRID)
    lib = st.lib
    dm = BOUNDARY.device_maps()
    BC = dm["bcIdxFieldIdx2"]
    prec = _lib.SF_FP64 if OCEAN.u.dtype == torch.float64 else _lib.SF_FP32
    stream = _lib.stream_handle()
    Huon, Hvom, W = OCEAN.Huon.reshape(-1), OCEAN.Hvom.reshape(-1), OCEAN.W.reshape(-1)
    Hz, z_w, z_r = GRID.Hz.reshape(-1), GRID.z_w.reshape(-1), GRID.z_r.reshape(-1)
    ops.set_maxflux(None, None, (OCEAN.u_t2, OCEAN.v_t2, Huon, Hvom, Hz), ctx=c)
    ops.omega(None, None, (W, OCEAN.u_t2, OCEAN.v_t2, Huon, Hvom, z_w, BC), ctx=c)
    ops.set_zeta(None, None, (OCEAN.zeta_t1, OCEAN.zeta_t2, OCEAN.Zt_avg1), ctx=c)
    rhs3d(GRID, OCEAN, BOUNDARY)
    if st.tracers is not None:
        st.tracers.pressure_gradient(GRID, OCEAN)
    _lib.check(lib.sf_couple_rufrc(c.handle, prec, OCEAN.ru_t2.data_ptr(), OCEAN.rv_t2.data_ptr(),
                                   OCEAN.u_t2.data_ptr(), OCEAN.v_t2.data_ptr(), st.sustr.data_ptr(),
                                   st.svstr.data_ptr(), st.rdrg, st.rufrc.data_ptr(), st.rvfrc.data_ptr(), stream),
               "sf_couple_rufrc")
    compTimes.first2DTimeStep()
    for _ in range(compTimes.nfast + 1):
        if compTimes.isFirst2DStep():
            OCEAN.Zt_avg1.zero_()
            OCEAN.DU_avg1.zero_()
            OCEAN.DV_avg1.zero_()
        compTimes.next2DTimeStep()
        OCEAN.cycleTimes2D()
        a = st.stepper.fields(OCEAN)
        cp = st.couple_fields(compTimes.iic - compTimes.ntfirst, compTimes.iif, compTimes.weight2)
        if st.fused:
            _lib.check(lib.sf_phys_substep_coupled_fused(C.byref(a), C.byref(cp), st.stepper.precision,
                                                         float(compTimes.dtfast), int(compTimes.iif == 1),
                                                         compTimes.cff1(), st.tile_j, stream),
                       "sf_phys_substep_coupled_fused")
        else:
            _lib.check(lib.sf_phys_substep_coupled(C.byref(a), C.byref(cp), st.stepper.precision,
                                                   float(compTimes.dtfast), int(compTimes.iif == 1),
                                                   compTimes.cff1(), stream), "sf_phys_substep_coupled")
    if history is not None:
        history.maybe_write(compTimes, OCEAN)
    st.Hz_old.copy_(GRID.Hz)
    initial_depths(GRID, OCEAN)
    m3 = st.m3
    n = compTimes.iic - compTimes.ntfirst
    cur, h1, h2 = n % 3, (n - 1) % 3, (n - 2) % 3
    q = lambda t: t.data_ptr()
    a = sf_uv_coupled(q(OCEAN.u_t2), q(OCEAN.v_t2), q(OCEAN.ru_t2), q(OCEAN.rv_t2), q(st.ruh[h1]), q(st.rvh[h1]),
                      q(st.ruh[h2]), q(st.rvh[h2]), q(st.ruh[cur]), q(st.rvh[cur]), q(OCEAN.ubar_t1),
                      q(OCEAN.vbar_t1), q(OCEAN.ubar_t2), q(OCEAN.vbar_t2), q(st.Hz_old), q(Hz), q(OCEAN.AKv), q(z_r),
                      q(OCEAN.DU_avg1), q(OCEAN.DV_avg1), q(st.DU_avg2), q(st.DV_avg2), q(Huon), q(Hvom),
                      q(st.sustr), q(st.svstr), st.rdrg, ab_order(n), q(m3.eqSD), q(m3.eqD), q(m3.eqRHS))
    _lib.check(lib.sf_step3d_uv_coupled(c.handle, prec, C.byref(a), 1.0, float(compTimes.dt), stream),
               "sf_step3d_uv_coupled")
    if st.tracers is not None:
        st.tracers.step(GRID, OCEAN, compTimes.dt)
    ops.setLateralUVBCs(None, None, (compTimes.time, OCEAN.u_t2, OCEAN.v_t2, dm["uFieldIdx2"], dm["vFieldIdx2"],
                                     dm["uFieldType"], dm["vFieldType"]), ctx=c)
    ops.omega(None, None, (W, OCEAN.u_t2, OCEAN.v_t2, Huon, Hvom, z_w, BC), ctx=c)
    compTimes.nextTimeStep()


def main3d_coupled(compTimes, GRID, OCEAN, BOUNDARY, nsteps: int, st: CoupledState | None = None,
                   rdrg: float = 3e-4, history=None, tracers: bool | dict = False, fused: bool = True,
                   tile_j: int = 0) -> CoupledState:
    """``nsteps`` coupled baroclinic steps after the initial set_depth (main3d.py:62-63).
    ``tracers``: True (or a dict of baroclinic.Tracers arguments, e.g. T=...) adds the
    baroclinic pressure gradient and the tracer step; the default temperature is the
    3D workload's stratification 4 + 16 exp(z/500) C at the initial depths."""
    if st is None:
        initial_depths(GRID, OCEAN)
        tr = None
        if tracers:
            from .baroclinic import Tracers, stratification
            kw = dict(tracers) if isinstance(tracers, dict) else {}
            kw.setdefault("T", stratification(GRID))
            tr = Tracers(GRID, OCEAN, **kw)
        st = CoupledState(GRID, OCEAN, rdrg=rdrg, tracers=tr, fused=fused, tile_j=tile_j)
    for _ in range(int(nsteps)):
        coupled_step(compTimes, GRID, OCEAN, BOUNDARY, st, history=history)
    return st
