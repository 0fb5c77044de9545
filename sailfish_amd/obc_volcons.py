"""Open-boundary volume conservation (nonlinear/obc_volcons.py:1-72): FP64 sums of the
open segments' cross-section area and net inflow of a state, the uniform correction
ubar_xs = flux / area and its application to the boundary velocities
(conserve_mass), on the GPU (sf_obc_volcons)."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib

WEST, EAST, SOUTH, NORTH = 1, 2, 4, 8
P = C.c_void_p


def _bind(lib):
    if not getattr(lib, "_sf_volcons_bound", False):
        lib.sf_obc_volcons_work_size.argtypes = [C.c_int32, C.c_int32]
        lib.sf_obc_volcons_work_size.restype = C.c_int32
        lib.sf_obc_volcons.argtypes = [P] * 6 + [C.c_int32] * 5 + [P, P, P]
        lib.sf_obc_volcons.restype = C.c_int32
        lib._sf_volcons_bound = True
    return lib


class VolCons:
    """Device work space and result of obc_volcons for one grid size."""

    def __init__(self, szI: int, szJ: int, device="cuda"):
        _lib.require_gpu()
        self.lib = _bind(_lib.load())
        self.szI, self.szJ = int(szI), int(szJ)
        n = int(self.lib.sf_obc_volcons_work_size(self.szI, self.szJ))
        self.work = torch.zeros(n, dtype=torch.float64, device=device)
        self.out = torch.zeros(3, dtype=torch.float64, device=device)

    def __call__(self, zeta, h, pm, pn, ubar, vbar, flags: int, apply: bool = True, stream=None):
        """(area, flux, ubar_xs) stay on the device in ``self.out``; the state arrays
        (one precision: FP64 or FP32) are corrected in place when ``apply``."""
        prec = _lib.SF_FP64 if zeta.dtype == torch.float64 else _lib.SF_FP32
        for t in (h, pm, pn, ubar, vbar):
            if t.dtype != zeta.dtype:
                raise _lib.SailfishError("obc_volcons: all fields must share one precision")
        rc = self.lib.sf_obc_volcons(zeta.data_ptr(), h.data_ptr(), pm.data_ptr(), pn.data_ptr(), ubar.data_ptr(),
                                     vbar.data_ptr(), self.szI, self.szJ, prec, int(flags), int(bool(apply)),
                                     self.work.data_ptr(), self.out.data_ptr(), _lib.stream_handle(stream))
        _lib.check(rc, "sf_obc_volcons")
        return self.out
