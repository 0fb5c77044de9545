"""Fused barotropic substep: ``step2dPredictor`` + ``step2dCorrector``
(nonlinear/step2d.py:23-127) in one tiled gfx950 kernel (``sf_step2d_substep``).

Same state in, same state out as the one-kernel-per-reference-kernel sequence, with
two documented exceptions that no kernel of the path ever reads: the corrector's
momentum RHS slots (``rubar_t2``/``rvbar_t2``) and ``rvbar_t1`` are not written, and
the AM3 update that step2d.py:120-123 applies to ``ubar_t1`` is kept only at the
points computeMomentumPred masks (everywhere else the next predictor overwrites it
before it is read).  Works for FP64 and FP32 ``T_OCEAN`` states (FP32 keeps the
time-averaging accumulators in FP64).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from . import mod_operators as ops
from .mod_boundary import clamped_value

_I32P = C.POINTER(C.c_int32)
SF_SUBSTEP_ZERO_VBAR = 1


class sf_step2d_fields(C.Structure):
    _fields_ = [
        ("h", C.c_void_p), ("on_u", C.c_void_p), ("om_v", C.c_void_p), ("pm", C.c_void_p), ("pn", C.c_void_p),
        ("zeta", C.c_void_p * 3), ("ubar", C.c_void_p * 3), ("vbar", C.c_void_p * 3),
        ("rzeta", C.c_void_p * 3), ("rubar", C.c_void_p * 3), ("rvbar", C.c_void_p * 3),
        ("Zt_avg1", C.c_void_p), ("DU_avg1", C.c_void_p), ("DV_avg1", C.c_void_p),
    ]


def _bind():
    lib = _lib.load()
    lib.sf_ctx_set_bc_maps.argtypes = [C.c_void_p, _I32P, _I32P, C.c_int32, _I32P, _I32P, C.c_int32,
                                       _I32P, _I32P, C.c_int32]
    lib.sf_ctx_set_bc_maps.restype = C.c_int32
    lib.sf_step2d_substep.argtypes = [C.c_void_p, C.POINTER(sf_step2d_fields), C.c_int32, C.c_double,
                                      C.c_double, C.c_double, C.c_int32, C.c_void_p]
    lib.sf_step2d_substep.restype = C.c_int32
    lib.sf_step2d_substep_parts.argtypes = [C.c_void_p, C.POINTER(sf_step2d_fields), C.c_int32, C.c_double,
                                            C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_void_p]
    lib.sf_step2d_substep_parts.restype = C.c_int32
    lib.sf_step2d_tile_shape.argtypes = [_I32P, _I32P]
    lib.sf_step2d_tile_shape.restype = C.c_int32
    return lib


def tile_shape():
    lib = _bind()
    a, b = C.c_int32(), C.c_int32()
    _lib.check(lib.sf_step2d_tile_shape(C.byref(a), C.byref(b)), "sf_step2d_tile_shape")
    return a.value, b.value


def check_maps(szI: int, szJ: int, q5_stride: int, BOUNDARY) -> tuple[bool, str]:
    """Host-only check (no GPU) that the fused substep accepts these boundary maps."""
    lib = _lib.load()
    arrs = [_i32(x) for x in (*BOUNDARY.zeta_map, *BOUNDARY.ubar_map2, *BOUNDARY.vbar_map2)]
    p = [a[0].ctypes.data for a in arrs]
    rc = lib.sf_step2d_check_maps(szI, szJ, q5_stride, p[0], p[1], len(arrs[0][0]), p[2], p[3],
                                  len(arrs[2][0]), p[4], p[5], len(arrs[4][0]))
    return rc == _lib.SF_OK, (_lib.last_error() if rc != _lib.SF_OK else "")


def _i32(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    return a, a.ctypes.data_as(_I32P)


class FusedStepper:
    """Holds the boundary maps in the context and the grid fields in the state's
    precision; ``step(compTimes, OCEAN)`` runs one substep on the current stream."""

    def __init__(self, GRID, BOUNDARY, dtype=torch.float64):
        lib = _bind()
        self.ctx = ops.ctx_of(GRID)
        self.dtype = dtype
        self.precision = _lib.SF_FP64 if dtype == torch.float64 else _lib.SF_FP32
        if dtype not in (torch.float64, torch.float32):
            raise _lib.SailfishError(f"unsupported state precision {dtype}")
        # velocity maps: predictor + corrector passes composed
        zd, zs = BOUNDARY.zeta_map
        ud, us = BOUNDARY.ubar_map2
        vd, vs = BOUNDARY.vbar_map2
        keep = [_i32(x) for x in (zd, zs, ud, us, vd, vs)]
        rc = lib.sf_ctx_set_bc_maps(self.ctx.handle, keep[0][1], keep[1][1], len(zd), keep[2][1], keep[3][1],
                                    len(ud), keep[4][1], keep[5][1], len(vd))
        _lib.check(rc, "sf_ctx_set_bc_maps")
        cast = (lambda t: t.reshape(-1).contiguous()) if dtype == torch.float64 else \
            (lambda t: t.reshape(-1).to(torch.float32).contiguous())
        self.grid_fields = {k: cast(getattr(GRID, k)) for k in ("h", "on_u", "om_v", "pm", "pn")}
        self.lib = lib

    def fields(self, OCEAN) -> sf_step2d_fields:
        if OCEAN.dtype != self.dtype:
            raise _lib.SailfishError(f"state is {OCEAN.dtype}, stepper built for {self.dtype}")
        a = sf_step2d_fields()
        g = self.grid_fields
        a.h, a.on_u, a.om_v, a.pm, a.pn = (g[k].data_ptr() for k in ("h", "on_u", "om_v", "pm", "pn"))
        for name in ("zeta", "ubar", "vbar", "rzeta", "rubar", "rvbar"):
            arr = getattr(a, name)
            for k in range(3):
                arr[k] = getattr(OCEAN, f"{name}_t{k}").data_ptr()
        a.Zt_avg1 = OCEAN.Zt_avg1.data_ptr()
        a.DU_avg1 = OCEAN.DU_avg1.data_ptr()
        a.DV_avg1 = OCEAN.DV_avg1.data_ptr()
        return a

    def step(self, compTimes, OCEAN, zero_vbar: bool = False, stream=None, parts: int = 3):
        a = self._cached_fields(OCEAN)
        flags = SF_SUBSTEP_ZERO_VBAR if zero_vbar else 0
        rc = self.lib.sf_step2d_substep_parts(self.ctx.handle, C.byref(a), self.precision,
                                              float(compTimes.dtfast), compTimes.cff1(),
                                              clamped_value(compTimes.time2D), flags, parts,
                                              _lib.stream_handle(stream))
        _lib.check(rc, "sf_step2d_substep")

    def _cached_fields(self, OCEAN):
        key = (id(OCEAN), OCEAN.zeta_t0.data_ptr(), OCEAN.ubar_t0.data_ptr(), OCEAN.rubar_t0.data_ptr())
        cache = self.__dict__.setdefault("_fcache", {})
        a = cache.get(key)
        if a is None:
            if len(cache) > 12:
                cache.clear()
            a = cache[key] = self.fields(OCEAN)
        return a


def step2d_substep(compTimes, GRID, OCEAN, BOUNDARY, zero_vbar: bool = False):
    """One fused barotropic substep (after the caller's next2DTimeStep/cycleTimes2D)."""
    st = getattr(GRID, "_sf_fused", None)
    if st is None or st.dtype != OCEAN.dtype:
        st = FusedStepper(GRID, BOUNDARY, OCEAN.dtype)
        GRID._sf_fused = st
    st.step(compTimes, OCEAN, zero_vbar=zero_vbar)
