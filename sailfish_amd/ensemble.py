"""Ensembles of physics-mode basins advanced together: one persistent launch per
substep over all members (sf_ens_substep), optionally replayed from a HIP graph
(sf_ens_capture / sf_ens_replay), with the schedule picked by the autotuner in
``autotune.py``.  Members keep the T_OCEAN layout of a single basin; the library
owns the time-slot rotation and the host views are brought up to date by ``sync``.

``looped_substeps`` is the comparison: the same members advanced one after the
other with one sf_phys_substep_fused launch each."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .model2d import build_model
from .phys2d import PhysStepper, sf_phys_fields

P = C.c_void_p


def _bind(lib):
    if getattr(lib, "_sf_ens_bound", False):
        return lib
    lib.sf_ens_create.argtypes = [C.POINTER(sf_phys_fields), C.POINTER(C.c_double), C.c_int32, C.c_int32,
                                  C.POINTER(P)]
    lib.sf_ens_destroy.argtypes = [P]
    lib.sf_ens_configure.argtypes = [P, C.c_int32, C.c_int32, C.c_int32]
    lib.sf_ens_num_tiles.argtypes = [P]
    lib.sf_ens_rotation.argtypes = [P]
    lib.sf_ens_substep.argtypes = [P, C.c_int32, C.c_double, P]
    lib.sf_ens_capture.argtypes = [P, C.c_int32, C.POINTER(C.c_double), P]
    lib.sf_ens_replay.argtypes = [P, P]
    for f in ("sf_ens_create", "sf_ens_destroy", "sf_ens_configure", "sf_ens_num_tiles", "sf_ens_rotation",
              "sf_ens_substep", "sf_ens_capture", "sf_ens_replay"):
        getattr(lib, f).restype = C.c_int32
    lib._sf_ens_bound = True
    return lib


class Member:
    """One basin: its CompTimes, GRID, OCEAN and the physics stepper holding its
    static fields in the state precision."""

    def __init__(self, cfg: dict, dtype=torch.float32, ndtfast: int = 30):
        self.cfg = cfg
        self.ct, self.GRID, self.OCEAN, _B = build_model(cfg, dtype=dtype, ndtfast=ndtfast)
        self.stepper = PhysStepper(self.GRID, self.OCEAN, rdrg=cfg.get("rdrg", 3e-4), fused=True)
        self.dt = float(cfg["dtfast"])

    @property
    def points(self) -> int:
        return self.GRID.szI * self.GRID.szJ


class PhysEnsemble:
    def __init__(self, members: list, precision_dtype=torch.float32):
        _lib.require_gpu()
        self.lib = _bind(_lib.load())
        self.members = members
        self.dtype = precision_dtype
        self.precision = _lib.SF_FP64 if precision_dtype == torch.float64 else _lib.SF_FP32
        n = len(members)
        arr = (sf_phys_fields * n)(*[m.stepper.fields(m.OCEAN) for m in members])
        dts = (C.c_double * n)(*[m.dt for m in members])
        h = P()
        _lib.check(self.lib.sf_ens_create(arr, dts, n, self.precision, C.byref(h)), "sf_ens_create")
        self.handle = h
        self.pending = 0          # library rotations not yet applied to the host views
        self.schedule = (8, 1, 0)

    def close(self):
        if self.handle:
            self.lib.sf_ens_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def points(self) -> int:
        return sum(m.points for m in self.members)

    def num_tiles(self) -> int:
        return int(self.lib.sf_ens_num_tiles(self.handle))

    def configure(self, tile_j: int, order: int, grid: int):
        _lib.check(self.lib.sf_ens_configure(self.handle, int(tile_j), int(order), int(grid)), "sf_ens_configure")
        self.schedule = (int(tile_j), int(order), int(grid))

    def substep(self, first: bool, cff1: float, stream=None):
        _lib.check(self.lib.sf_ens_substep(self.handle, int(bool(first)), float(cff1), _lib.stream_handle(stream)),
                   "sf_ens_substep")
        self.pending += 1

    def capture(self, cff1_list, stream=None):
        n = len(cff1_list)
        w = (C.c_double * n)(*[float(x) for x in cff1_list])
        _lib.check(self.lib.sf_ens_capture(self.handle, n, w, _lib.stream_handle(stream)), "sf_ens_capture")
        self.graph_nsub = n

    def replay(self, stream=None):
        _lib.check(self.lib.sf_ens_replay(self.handle, _lib.stream_handle(stream)), "sf_ens_replay")
        self.pending += self.graph_nsub

    def sync(self):
        """Apply the library's pending slot rotations to every member's views."""
        for _ in range(self.pending % 3):
            for m in self.members:
                m.OCEAN.cycleTimes2D()
        self.pending = 0


class LoopedMembers:
    """The comparison: each member advanced with its own sf_phys_substep_fused launch
    (argument structs prepared once per time rotation)."""

    def __init__(self, members: list, tile_j: int = 0):
        self.members = members
        self.lib = members[0].stepper.lib
        self.tile_j = int(tile_j)
        self.structs = []
        for m in members:
            rots = []
            for _ in range(3):
                m.OCEAN.cycleTimes2D()
                rots.append(m.stepper.fields(m.OCEAN))
            self.structs.append(rots)       # rots[r]: fields after r+1 rotations (mod 3)
        self.rot = 0
        self.pending = 0

    def substep(self, first: bool, cff1: float, stream=None):
        s = _lib.stream_handle(stream)
        self.rot = (self.rot + 1) % 3
        fn = self.lib.sf_phys_substep_fused
        for m, rots in zip(self.members, self.structs):
            rc = fn(C.byref(rots[(self.rot - 1) % 3]), m.stepper.precision, m.dt, int(bool(first)), float(cff1),
                    self.tile_j, s)
            if rc:
                _lib.check(rc, "sf_phys_substep_fused")
        self.pending += 1

    def sync(self):
        for _ in range(self.pending % 3):
            for m in self.members:
                m.OCEAN.cycleTimes2D()
        self.pending = 0
