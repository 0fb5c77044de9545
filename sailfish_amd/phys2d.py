"""Physics-mode barotropic model: ROMS LF-AM3 in 2D form with C2 momentum
advection, Coriolis, wind stress and linear bottom drag, on the reference's grid
layout and T_OCEAN arrays (obsolete/step2d_LF_AM3.h.py:917-2780).  One substep =
four gfx950 launches (sf_phys_substep) or one fused launch (sf_phys_substep_fused)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .mod_constants import g as G_CONST

P = C.c_void_p


class sf_phys_fields(C.Structure):
    _fields_ = [
        ("szI", C.c_int32), ("szJ", C.c_int32), ("g", C.c_double), ("rdrg", C.c_double),
        ("h", P), ("pm", P), ("pn", P), ("fomn", P), ("sustr", P), ("svstr", P),
        ("zeta", P * 3), ("ubar", P * 3), ("vbar", P * 3),
        ("rzeta", P * 3), ("rubar", P * 3), ("rvbar", P * 3),
        ("scratch", P * 4),
        ("Zt_avg1", P), ("DU_avg1", P), ("DV_avg1", P),
    ]


class sf_open_bc(C.Structure):
    _fields_ = [("flags", C.c_int32), ("volcons", C.c_int32), ("zeta_ext", P * 4), ("vel_ext", P * 4)]


def _bind():
    lib = _lib.load()
    lib.sf_phys_substep_open.argtypes = [C.POINTER(sf_phys_fields), C.POINTER(sf_open_bc), C.c_int32, C.c_double,
                                         C.c_int32, C.c_double, P, P, P]
    lib.sf_phys_substep_open.restype = C.c_int32
    lib.sf_phys_substep.argtypes = [C.POINTER(sf_phys_fields), C.c_int32, C.c_double, C.c_int32, C.c_double, P]
    lib.sf_phys_substep.restype = C.c_int32
    lib.sf_phys_substep_fused.argtypes = [C.POINTER(sf_phys_fields), C.c_int32, C.c_double, C.c_int32, C.c_double,
                                          C.c_int32, P]
    lib.sf_phys_substep_fused.restype = C.c_int32
    lib.sf_phys_fused_default_tile.argtypes = [C.c_int32, C.c_int32, C.c_int32]
    lib.sf_phys_fused_default_tile.restype = C.c_int32
    return lib


class PhysStepper:
    """Static fields in the state precision plus the predicted-state scratch.

    ``fused``: one launch per substep (tile height ``tile_j``, 0 = library default)
    instead of the four-launch form.

    ``open_edges``: SF_OBC_* bits of the open edges (Chapman free surface, Flather normal
    velocity; the four-launch form), ``volcons``: the volume constraint on them after
    each substep, ``zeta_ext`` / ``vel_ext``: four per-edge host arrays (or None) of the
    external free surface and normal velocity."""

    def __init__(self, GRID, OCEAN, rdrg: float = 3e-4, g: float = G_CONST, fused: bool = True,
                 tile_j: int = 0, open_edges: int = 0, volcons: bool = False, zeta_ext=None, vel_ext=None):
        _lib.require_gpu()
        self.lib = _bind()
        dt = OCEAN.dtype
        if dt not in (torch.float64, torch.float32):
            raise _lib.SailfishError(f"unsupported precision {dt}")
        self.dtype = dt
        self.precision = _lib.SF_FP64 if dt == torch.float64 else _lib.SF_FP32
        dev = GRID.device
        host = GRID.host
        S2 = GRID.szI * GRID.szJ
        get = lambda k: torch.from_numpy(np.ascontiguousarray(host[k]).reshape(-1)).to(dev, dt)
        zeros = np.zeros(GRID.shape2D)
        self.static = {
            "h": get("h"), "pm": get("pm"), "pn": get("pn"), "fomn": get("fomn"),
            "sustr": torch.from_numpy(host.get("sustr", zeros).reshape(-1).copy()).to(dev, dt),
            "svstr": torch.from_numpy(host.get("svstr", zeros).reshape(-1).copy()).to(dev, dt),
        }
        self.szI, self.szJ = GRID.szI, GRID.szJ
        self.g, self.rdrg = float(g), float(rdrg)
        self.open_edges = int(open_edges)
        self.fused, self.tile_j = bool(fused) and not self.open_edges, int(tile_j)
        # predicted-state scratch of the four-launch form (the fused form keeps it in LDS)
        self.scratch = [] if self.fused else [torch.zeros(S2, dtype=dt, device=dev) for _ in range(4)]
        if self.open_edges:
            ob = sf_open_bc(self.open_edges, int(bool(volcons)))
            self._ext = []
            for e in range(4):
                for arrs, field in ((zeta_ext, "zeta_ext"), (vel_ext, "vel_ext")):
                    a = None if arrs is None else arrs[e]
                    if a is not None:
                        t = torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64))).to(dev)
                        self._ext.append(t)
                        getattr(ob, field)[e] = t.data_ptr()
            self.open_bc = ob
            n = int(self.lib.sf_obc_volcons_work_size(self.szI, self.szJ))
            self.vc_work = torch.zeros(n, dtype=torch.float64, device=dev)
            self.vc_out = torch.zeros(3, dtype=torch.float64, device=dev)

    def fields(self, OCEAN) -> sf_phys_fields:
        a = sf_phys_fields()
        a.szI, a.szJ, a.g, a.rdrg = self.szI, self.szJ, self.g, self.rdrg
        s = self.static
        a.h, a.pm, a.pn, a.fomn, a.sustr, a.svstr = (s[k].data_ptr() for k in
                                                     ("h", "pm", "pn", "fomn", "sustr", "svstr"))
        for name in ("zeta", "ubar", "vbar", "rzeta", "rubar", "rvbar"):
            arr = getattr(a, name)
            for k in range(3):
                arr[k] = getattr(OCEAN, f"{name}_t{k}").data_ptr()
        for k, buf in enumerate(self.scratch):
            a.scratch[k] = buf.data_ptr()
        a.Zt_avg1, a.DU_avg1, a.DV_avg1 = OCEAN.Zt_avg1.data_ptr(), OCEAN.DU_avg1.data_ptr(), OCEAN.DV_avg1.data_ptr()
        return a

    def step(self, compTimes, OCEAN, first: bool, stream=None):
        a = self.fields(OCEAN)
        if self.open_edges:
            rc = self.lib.sf_phys_substep_open(C.byref(a), C.byref(self.open_bc), self.precision,
                                               float(compTimes.dtfast), int(bool(first)), compTimes.cff1(),
                                               self.vc_work.data_ptr(), self.vc_out.data_ptr(),
                                               _lib.stream_handle(stream))
            _lib.check(rc, "sf_phys_substep_open")
            return
        if self.fused:
            rc = self.lib.sf_phys_substep_fused(C.byref(a), self.precision, float(compTimes.dtfast),
                                                int(bool(first)), compTimes.cff1(), self.tile_j,
                                                _lib.stream_handle(stream))
            _lib.check(rc, "sf_phys_substep_fused")
            return
        rc = self.lib.sf_phys_substep(C.byref(a), self.precision, float(compTimes.dtfast), int(bool(first)),
                                      compTimes.cff1(), _lib.stream_handle(stream))
        _lib.check(rc, "sf_phys_substep")


def run_phys(compTimes, GRID, OCEAN, nsteps: int, stepper: PhysStepper | None = None, start_first: bool = True):
    """``nsteps`` physics substeps (next2DTimeStep, cycleTimes2D, substep); the
    first substep of the run is forward/backward Euler."""
    st = stepper or PhysStepper(GRID, OCEAN)
    for n in range(int(nsteps)):
        compTimes.next2DTimeStep()
        OCEAN.cycleTimes2D()
        st.step(compTimes, OCEAN, first=(start_first and n == 0))
    return st
