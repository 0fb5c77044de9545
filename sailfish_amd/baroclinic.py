"""Baroclinic terms of the 3D step: the linear equation of state, the baroclinic
pressure gradient and the tracer step that the reference names but leaves out
(nonlinear/rhs.py:31-32 "prsgrd()" commented out; nonlinear/step3d_t.py; the state
equation of utility/test.in:2130-2160), on the GPU (sf_rho_eos, sf_prsgrd,
sf_step3d_t).

``Tracers`` holds temperature and salinity (N levels each, the bulk precision of the
model) and the density; ``coupled3d.coupled_step`` calls ``pressure_gradient`` after
rhs3d (so the gradient also enters the barotropic forcing) and ``step`` after the
coupled step3d_UV with the thicknesses before (``Hz_old``, kept by the coupled step) and
after set_depth."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from . import mod_operators as ops
from .mod_constants import g as G_CONST, rho0 as RHO0

P = C.c_void_p

# linear equation of state (utility/test.in:449-453)
R0, T0, S0, TCOEF, SCOEF = 1027.0, 14.0, 35.0, 1.7e-4, 0.0


def _bind():
    lib = _lib.load()
    if not getattr(lib, "_sf_baroclinic_bound", False):
        lib.sf_rho_eos.argtypes = [P, C.c_int32, P, P] + [C.c_double] * 5 + [P, P]
        lib.sf_rho_eos.restype = C.c_int32
        lib.sf_prsgrd.argtypes = [P, C.c_int32, P, P, P, P, C.c_double, C.c_double, P, P, P]
        lib.sf_prsgrd.restype = C.c_int32
        lib.sf_step3d_t.argtypes = [P, C.c_int32, P, P, P, P, P, P, C.c_double, P]
        lib.sf_step3d_t.restype = C.c_int32
        lib._sf_baroclinic_bound = True
    return lib


def stratification(GRID, T_surface_excess: float = 16.0, T_deep: float = 4.0, scale: float = 500.0):
    """T = T_deep + T_surface_excess * exp(z_r / scale) at the model's current depths
    (the 3D workload's stratification, 4 + 16 exp(z/500) C), as host FP64."""
    z_r = GRID.z_r.detach().cpu().numpy().astype(np.float64)
    return T_deep + T_surface_excess * np.exp(z_r / scale)


class Tracers:
    """Temperature, salinity and density of a 3D model (N levels, bulk precision)."""

    def __init__(self, GRID, OCEAN, T=None, S=None, R0_=R0, T0_=T0, S0_=S0, Tcoef=TCOEF, Scoef=SCOEF,
                 g: float = G_CONST, rho0: float = RHO0):
        _lib.require_gpu()
        self.lib = _bind()
        dt = OCEAN.u.dtype
        self.precision = _lib.SF_FP64 if dt == torch.float64 else _lib.SF_FP32
        shape = (GRID.N,) + GRID.shape2D
        dev = GRID.device
        mk = lambda a, default: torch.from_numpy(np.broadcast_to(np.asarray(default if a is None else a, np.float64),
                                                                 shape).copy()).to(dev, dt)
        self.T = mk(T, T0_)
        self.S = mk(S, S0_)
        self.T_new = torch.empty_like(self.T)
        self.S_new = torch.empty_like(self.S)
        self.rho = torch.zeros_like(self.T)
        self.Hz_old = torch.zeros_like(GRID.Hz)
        self.eos = (float(R0_), float(T0_), float(S0_), float(Tcoef), float(Scoef))
        self.g, self.rho0 = float(g), float(rho0)
        OCEAN.extra_state = {**getattr(OCEAN, "extra_state", {}), "T": self.T, "S": self.S}

    def density(self, GRID, stream=None):
        c = ops.ctx_of(GRID)
        _lib.check(self.lib.sf_rho_eos(c.handle, self.precision, self.T.data_ptr(), self.S.data_ptr(), *self.eos,
                                       self.rho.data_ptr(), _lib.stream_handle(stream)), "sf_rho_eos")
        return self.rho

    def pressure_gradient(self, GRID, OCEAN, stream=None):
        """ru, rv += the baroclinic pressure gradient of the current T, S and depths."""
        self.density(GRID, stream)
        c = ops.ctx_of(GRID)
        _lib.check(self.lib.sf_prsgrd(c.handle, self.precision, self.rho.data_ptr(), GRID.z_r.data_ptr(),
                                      GRID.z_w.data_ptr(), GRID.Hz.data_ptr(), self.g, self.rho0,
                                      OCEAN.ru_t2.data_ptr(), OCEAN.rv_t2.data_ptr(), _lib.stream_handle(stream)),
                   "sf_prsgrd")

    def step(self, GRID, OCEAN, dt: float, stream=None):
        """Advect T and S by the step's mass fluxes from Hz_old to the current Hz."""
        c = ops.ctx_of(GRID)
        s = _lib.stream_handle(stream)
        for old, new in ((self.T, self.T_new), (self.S, self.S_new)):
            _lib.check(self.lib.sf_step3d_t(c.handle, self.precision, old.data_ptr(), new.data_ptr(),
                                            OCEAN.Huon.data_ptr(), OCEAN.Hvom.data_ptr(), self.Hz_old.data_ptr(),
                                            GRID.Hz.data_ptr(), float(dt), s), "sf_step3d_t")
            old.copy_(new)
