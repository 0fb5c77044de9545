"""Lateral boundary conditions (modules/mod_boundary.py:28-297, modules/mod_BC.c).

``Boundary`` rebuilds the reference's index sets with the same rules as
``buildLBC`` (mod_boundary.py:103-200), including its use of ``i0`` as the *row* of
the "N"/"S" blocks (mod_boundary.py:149-163).  The ordered setBC/copyBC lists that
``zetabc`` (nonlinear/zetabc.py:5-12) and ``barotropicVelocityBC``
(nonlinear/barotropicVelocityBC.py:7-31) launch are composed once, on the host, into
one map per field: every destination gets either the pre-BC value of one source
point or a constant.  Composition runs the lists in list order, which fixes one
legal schedule of the reference's racy parallel copies (a source that is also a
destination).  The GPU applies a map in one launch (sf_bc_apply_map).
"""
from __future__ import annotations

import math

import numpy as np

bcNone = 0
bcAcquire = 1 << 0
bcChapman_explicit = 1 << 1
bcChapman_implicit = 1 << 2
bcClamped = 1 << 3
bcClosed = 1 << 4
bcFlather = 1 << 5
bcGradient = 1 << 6
bcMixed = 1 << 7
bcNested = 1 << 8
bcNudging = 1 << 9
bcPeriodic = 1 << 10
bcRadiation = 1 << 11
bcReduced = 1 << 12
bcShchepetkin = 1 << 13

dictLBC = {
    "CHA": bcChapman_explicit,
    "CHE": bcChapman_implicit,
    "CLA": bcClamped + bcAcquire,
    "CLO": bcClosed,
    "FLA": bcFlather + bcAcquire,
    "GRA": bcGradient,
    "NES": bcNested,
    "RAD": bcRadiation,
    "RADNUD": bcRadiation + bcNudging + bcAcquire,
    "RED": bcReduced,
    "SHC": bcShchepetkin,
}

iW, iS, iE, iN = (0, 1, 2, 3)

# map source codes understood by sf_bc_apply_map
SRC_ZERO = -1
SRC_CLAMP = -2


def decodeLBC(val: str) -> int:
    try:
        return dictLBC[val.upper()]
    except KeyError:
        raise ValueError(f'LBC type "{val}" is not recognized') from None


def checkLBC(bc, comp) -> bool:
    return (bc & comp) != 0


class BCIdx:
    def __init__(self, bcIdx, bcIdx1, bcIdx2, bcIdxFieldIdx2, bcIdxFieldType, LBC):
        self.bcIdx = bcIdx
        self.bcIdx1 = bcIdx1
        self.bcIdx2 = bcIdx2
        self.bcIdxFieldIdx2 = bcIdxFieldIdx2
        self.bcIdxFieldType = bcIdxFieldType
        self.LBC = LBC

    def unpack(self):
        return self.bcIdx, self.bcIdx1, self.bcIdx2, self.LBC


def buildLBC(val, i0: int, j0: int, L: int, M: int) -> BCIdx:
    """Index sets of one field, in the reference's W, E, "N", "S" block order
    (mod_boundary.py:103-200).  All arrays are host int32."""
    shp = (M + 1, L + 1)
    bcW, bcE = decodeLBC(val[iW]), decodeLBC(val[iE])
    bcN, bcS = decodeLBC(val[iN]), decodeLBC(val[iS])
    iRange = np.arange(i0, L)
    jRange = np.arange(j0, M)
    rv = lambda r, c: np.ravel_multi_index(np.broadcast_arrays(r, c), shp).ravel()
    blocks = [
        (rv(jRange, i0), rv(jRange, i0), rv(jRange, i0 + 1), bcW, M - j0),
        (rv(jRange, L - 1), rv(jRange, L), rv(jRange, L - 1), bcE, M - j0),
        (rv(i0, iRange), rv(i0, iRange), rv(i0 + 1, iRange), bcN, L - i0),
        (rv(M - 1, iRange), rv(M, iRange), rv(M - 1, iRange), bcS, L - i0),
    ]
    bcIdx = np.concatenate([b[0] for b in blocks]).astype(np.int32)
    bcIdx1 = np.concatenate([b[1] for b in blocks]).astype(np.int32)
    bcIdx2 = np.concatenate([b[2] for b in blocks]).astype(np.int32)
    LBC = np.concatenate([np.full(b[4], b[3]) for b in blocks]).astype(np.int32)
    fieldIdx2 = -np.ones(shp, np.int32)
    fieldType = -np.ones(shp, np.int32)
    # numpy scatter: the last duplicate wins
    fieldIdx2.ravel()[bcIdx1] = bcIdx2
    fieldType.ravel()[bcIdx1] = LBC
    return BCIdx(bcIdx, bcIdx1, bcIdx2, fieldIdx2, fieldType, LBC)


def compose_bc_ops(ops, n_points: int):
    """Compose an ordered list of BC operations into (dst, src) int32 arrays.

    ``ops`` items: ("set", idx, code) with code SRC_ZERO/SRC_CLAMP, or
    ("copy", idx1, idx2).  Operations run in list order and, inside one copy
    list, entry by entry.  Identity entries are dropped; dst is sorted."""
    origin: dict[int, int] = {}
    for op in ops:
        if op[0] == "set":
            for p in np.asarray(op[1]).tolist():
                origin[int(p)] = int(op[2])
        elif op[0] == "copy":
            idx1 = np.asarray(op[1]).tolist()
            idx2 = np.asarray(op[2]).tolist()
            if len(idx2) < len(idx1):
                raise ValueError("copy list reads past its source index array "
                                 "(undefined in the reference)")
            for d, s in zip(idx1, idx2):
                origin[int(d)] = origin.get(int(s), int(s))
        else:
            raise ValueError(op[0])
    for p in origin:
        if not (0 <= p < n_points):
            raise ValueError(f"boundary index {p} outside the grid")
    pairs = sorted((d, s) for d, s in origin.items() if d != s)
    dst = np.array([d for d, _ in pairs], dtype=np.int32)
    src = np.array([s for _, s in pairs], dtype=np.int32)
    return dst, src


def apply_map_numpy(var: np.ndarray, dst, src, clamp_val=0.0):
    """Host application of a composed map (used by tests and setup code)."""
    pre = var.copy()
    vals = np.where(src >= 0, pre[np.maximum(src, 0)], np.where(src == SRC_CLAMP, clamp_val, 0.0))
    var[dst] = vals
    return var


class Boundary:
    """Boundary index sets and composed maps for one grid.

    ``lbc`` maps field name ('zeta', 'ubar', 'vbar', 'u', 'v') to four LBC keywords
    in the reference's W, S, E, N order (e.g. ``['CLO'] * 4``)."""

    def __init__(self, GRID, lbc: dict | None = None, device=None):
        lbc = lbc or {}
        clo = ["CLO"] * 4
        L, M = GRID.L, GRID.M
        self.GRID = GRID
        self.n_points = (M + 1) * (L + 1)
        self.zetaBC = buildLBC(lbc.get("zeta", clo), 0, 0, L, M)
        self.ubarBC = buildLBC(lbc.get("ubar", clo), 1, 0, L, M)
        self.vbarBC = buildLBC(lbc.get("vbar", clo), 0, 1, L, M)
        self.uvelBC = buildLBC(lbc.get("u", clo), 1, 0, L, M)
        self.vvelBC = buildLBC(lbc.get("v", clo), 0, 1, L, M)

        # Flather/Shchepetkin momentum BCs mark the zeta entry of the same cell as
        # "acquire" (mod_boundary.py:259-267); the lookup is by cell index.
        first = {}
        for k, c in enumerate(self.zetaBC.bcIdx.tolist()):
            first.setdefault(c, k)
        for bcset in (self.ubarBC, self.vbarBC):
            for idx, _, _, bc in zip(*bcset.unpack()):
                if checkLBC(bc, bcFlather) or checkLBC(bc, bcShchepetkin):
                    if int(idx) not in first:
                        raise ValueError(f"Index {idx} not in the zeta BC list")
                    self.zetaBC.LBC[first[int(idx)]] |= bcAcquire

        z, ub, vb = self.zetaBC, self.ubarBC, self.vbarBC
        sel = ((z.LBC & bcClosed) | (z.LBC & bcGradient)) != 0
        self.zetaClosedOrGradientBCIdx1 = z.bcIdx1[sel]
        self.zetaClosedOrGradientBCIdx2 = z.bcIdx2[sel]
        self.zetaClampedBCIdx1 = z.bcIdx1[(z.LBC & bcClamped) != 0]
        self.ubarClosedBCIdx1 = ub.bcIdx1[(ub.LBC & bcClosed) != 0]
        self.vbarClosedBCIdx1 = vb.bcIdx1[(vb.LBC & bcClosed) != 0]
        self.ubarClampedBCIdx1 = ub.bcIdx1[(ub.LBC & bcClamped) != 0]
        self.vbarClampedBCIdx1 = vb.bcIdx1[(vb.LBC & bcClamped) != 0]
        sel = (ub.LBC & bcGradient) != 0
        self.ubarGradientBCIdx1, self.ubarGradientBCIdx2 = ub.bcIdx1[sel], ub.bcIdx2[sel]
        sel = (vb.LBC & bcGradient) != 0
        self.vbarGradientBCIdx1, self.vbarGradientBCIdx2 = vb.bcIdx1[sel], vb.bcIdx2[sel]

        # zetabc.py:13 -- one copy list
        self.zeta_ops = [("copy", self.zetaClosedOrGradientBCIdx1, self.zetaClosedOrGradientBCIdx2)]
        # barotropicVelocityBC.py:11-31 -- per-field order; vbar's gradient copy reads
        # ubarGradientBCIdx2 (barotropicVelocityBC.py:31)
        self.ubar_ops = [
            ("set", self.ubarClosedBCIdx1, SRC_ZERO),
            ("set", self.ubarClampedBCIdx1, SRC_CLAMP),
            ("copy", self.ubarGradientBCIdx1, self.ubarGradientBCIdx2),
        ]
        self.vbar_ops = [
            ("set", self.vbarClosedBCIdx1, SRC_ZERO),
            ("copy", self.vbarGradientBCIdx1, self.ubarGradientBCIdx2),
        ]
        n = self.n_points
        self.zeta_map = compose_bc_ops(self.zeta_ops, n)
        self.ubar_map = compose_bc_ops(self.ubar_ops, n)
        self.vbar_map = compose_bc_ops(self.vbar_ops, n)
        # the predictor and the corrector both apply barotropicVelocityBC to the same
        # arrays with nothing in between: their composition is what a substep does
        self.ubar_map2 = compose_bc_ops(self.ubar_ops + self.ubar_ops, n)
        self.vbar_map2 = compose_bc_ops(self.vbar_ops + self.vbar_ops, n)
        self._dev = None
        self.device = device if device is not None else getattr(GRID, "device", "cuda")

    # -- device copies ------------------------------------------------------------
    def device_maps(self):
        if self._dev is None:
            import torch
            to = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(self.device)
            self._dev = {
                "zeta": tuple(map(to, self.zeta_map)),
                "ubar": tuple(map(to, self.ubar_map)),
                "vbar": tuple(map(to, self.vbar_map)),
                "bcIdxFieldIdx2": to(self.zetaBC.bcIdxFieldIdx2.ravel()),
                "uFieldIdx2": to(self.uvelBC.bcIdxFieldIdx2.ravel()),
                "vFieldIdx2": to(self.vvelBC.bcIdxFieldIdx2.ravel()),
                "uFieldType": to(self.uvelBC.bcIdxFieldType.ravel()),
                "vFieldType": to(self.vvelBC.bcIdxFieldType.ravel()),
            }
        return self._dev


def clamped_value(time2D: float) -> float:
    """The placeholder clamped value of barotropicVelocityBC.py:21-22."""
    return 0.05 * math.sin(time2D * 0.0002)
