"""Ocean state and time-level rotation (modules/mod_ocean.py:14-191).

Same array names, shapes and flat views as the reference: ``zeta/ubar/vbar`` and
their RHS arrays are (3, M+1, L+1); ``u/v/ru/rv`` are (3, N+1, M+1, L+1);
``W/Huon/Hvom/AKv`` are (N+1, M+1, L+1); the ``*_avg1`` couplings are flat (M+1)*(L+1).
``_t0/_t1/_t2`` are flat views of the time slots that ``cycleTimes2D/3D`` rotate
(t0 <- t1, t1 <- t2, t2 <- old t0) without copying.

Precision: ``dtype`` (torch.float64 or torch.float32) is the type of every state
array except the barotropic time-averaging accumulators ``Zt_avg1``, ``DU_avg1``
and ``DV_avg1``, which stay FP64.
"""
from __future__ import annotations

import torch

_ROT2D = ("zeta", "ubar", "vbar", "rzeta", "rubar", "rvbar")
_ROT3D = ("u", "v", "ru", "rv")


class T_OCEAN:
    def __init__(self, GRID, dtype=torch.float64, with3d: bool = True):
        N, M, L = GRID.N, GRID.M, GRID.L
        dev = GRID.device
        self.dtype = dtype
        z2 = lambda *s: torch.zeros(*s, dtype=dtype, device=dev)
        for name in _ROT2D:
            setattr(self, name, z2((3, M + 1, L + 1)))
        self._views2d()
        f64 = dict(dtype=torch.float64, device=dev)
        self.Zt_avg1 = torch.zeros((M + 1) * (L + 1), **f64)
        self.U_avg1 = torch.zeros((M + 1) * (L + 1), **f64)
        self.V_avg1 = torch.zeros((M + 1) * (L + 1), **f64)
        self.DU_avg1 = torch.zeros((M + 1) * (L + 1), **f64)
        self.DV_avg1 = torch.zeros((M + 1) * (L + 1), **f64)
        self.with3d = with3d
        if with3d:
            for name in _ROT3D:
                setattr(self, name, z2((3, N + 1, M + 1, L + 1)))
            self.W = z2((N + 1, M + 1, L + 1))
            self.Huon = z2((N + 1, M + 1, L + 1))
            self.Hvom = z2((N + 1, M + 1, L + 1))
            self.AKv = 0.001 + z2((N + 1, M + 1, L + 1))   # mod_ocean.py:91
            self.AK = 0.00001
            self._views3d()

    def _views2d(self):
        for name in _ROT2D:
            arr = getattr(self, name)
            for k in range(3):
                setattr(self, f"{name}_t{k}", arr[k].reshape(-1))

    def _views3d(self):
        for name in _ROT3D:
            arr = getattr(self, name)
            for k in range(3):
                setattr(self, f"{name}_t{k}", arr[k].reshape(-1))

    @staticmethod
    def _rotate(obj, name):
        t0 = getattr(obj, f"{name}_t0")
        setattr(obj, f"{name}_t0", getattr(obj, f"{name}_t1"))
        setattr(obj, f"{name}_t1", getattr(obj, f"{name}_t2"))
        setattr(obj, f"{name}_t2", t0)

    def cycleTimes2D(self):
        """mod_ocean.py:120-155"""
        for name in _ROT2D:
            self._rotate(self, name)

    def cycleTimes3D(self):
        """mod_ocean.py:157-191"""
        for name in _ROT3D:
            self._rotate(self, name)

    def slot_of(self, name: str, level: str) -> int:
        """Index (0..2) of the array slot currently aliased by ``{name}_{level}``."""
        view = getattr(self, f"{name}_{level}")
        base = getattr(self, name)
        stride = base[0].numel() * base.element_size()
        return (view.data_ptr() - base.data_ptr()) // stride
