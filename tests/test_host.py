"""CPU tests of the host-side logic and of the C-ABI library's exports (no GPU)."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from sailfish_amd import configs
from sailfish_amd.mod_boundary import (SRC_CLAMP, SRC_ZERO, apply_map_numpy, buildLBC,
                                       compose_bc_ops)

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        names |= set(re.findall(r"^\s*(?:int|const char \*|void)\s*\*?\s*(sf_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from sailfish_amd import _lib
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 10
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the ctypes layer binds only declared names
    assert set(_lib.exported_symbols()) <= set(syms)


def test_version_without_gpu():
    from sailfish_amd import _lib
    assert _lib.load().sf_version() >= 100


def test_build_lbc_hand_enumerated():
    """buildLBC on L=5, M=4 (6x5 rho points), mod_boundary.py:103-200, written out by
    hand: W block col i0, E block col L, "N" block on row i0, "S" block on row M."""
    z = buildLBC(["CLO"] * 4, 0, 0, 5, 4)
    # row stride 6
    W1 = [0, 6, 12, 18]; W2 = [1, 7, 13, 19]
    E1 = [5, 11, 17, 23]; E2 = [4, 10, 16, 22]
    N1 = [0, 1, 2, 3, 4]; N2 = [6, 7, 8, 9, 10]
    S1 = [24, 25, 26, 27, 28]; S2 = [18, 19, 20, 21, 22]
    assert z.bcIdx1.tolist() == W1 + E1 + N1 + S1
    assert z.bcIdx2.tolist() == W2 + E2 + N2 + S2
    u = buildLBC(["CLO"] * 4, 1, 0, 5, 4)   # ubar: i0=1
    assert u.bcIdx1.tolist() == [1, 7, 13, 19] + E1 + [7, 8, 9, 10] + [25, 26, 27, 28]
    v = buildLBC(["CLO"] * 4, 0, 1, 5, 4)   # vbar: j0=1
    assert v.bcIdx1.tolist() == [6, 12, 18] + [11, 17, 23] + N1 + S1


def _sequential(var, ops, clamp):
    var = var.copy()
    for op in ops:
        if op[0] == "set":
            for p in op[1]:
                var[p] = 0.0 if op[2] == SRC_ZERO else clamp
        else:
            for d, s in zip(op[1], op[2]):
                var[d] = var[s]
    return var


@pytest.mark.parametrize("lbc", [["CLO"] * 4, ["GRA", "CLO", "GRA", "GRA"], ["CLA", "GRA", "CLO", "GRA"]])
def test_composed_map_equals_sequential(lbc):
    rng = np.random.default_rng(0)
    L, M = 9, 6
    b = buildLBC(lbc, 0, 0, L, M)
    ops = [("copy", b.bcIdx1, b.bcIdx2), ("set", b.bcIdx1[:3], SRC_ZERO), ("copy", b.bcIdx1[::-1], b.bcIdx2[::-1]),
           ("set", b.bcIdx1[5:9], SRC_CLAMP)]
    var = rng.standard_normal((M + 1) * (L + 1))
    dst, src = compose_bc_ops(ops, var.size)
    got = apply_map_numpy(var.copy(), dst, src, clamp_val=0.7)
    want = _sequential(var, ops, 0.7)
    assert np.array_equal(got, want)
    assert len(set(dst.tolist())) == len(dst)


def test_ocean_rotation_cpu():
    import torch
    from sailfish_amd.mod_grid import Grid
    from sailfish_amd.mod_ocean import T_OCEAN
    cfg = configs.random_smooth(8, 6)
    g = Grid(6, 4, 2, device="cpu").set_fields(h=cfg["h"], pm=cfg["pm"], pn=cfg["pn"], f=cfg["f"])
    o = T_OCEAN(g, dtype=torch.float64)
    assert [o.slot_of("zeta", f"t{k}") for k in range(3)] == [0, 1, 2]
    o.cycleTimes2D()
    assert [o.slot_of("zeta", f"t{k}") for k in range(3)] == [1, 2, 0]
    o.zeta_t2.fill_(3.0)
    assert float(o.zeta[0].sum()) == 3.0 * 48
    o.cycleTimes3D()
    assert [o.slot_of("u", f"t{k}") for k in range(3)] == [1, 2, 0]


class _G:
    device = "cpu"

    def __init__(self, szI, szJ):
        self.L, self.M = szI - 1, szJ - 1


FUSED_SIZES = [(130, 17), (64, 16), (67, 29), (128, 64), (96, 32), (200, 100), (40, 40), (70, 20),
               (2048, 1024), (256, 128), (384, 256), (512, 256), (768, 384), (1024, 512), (65, 9),
               (129, 33), (33, 129)]


@pytest.mark.parametrize("szI,szJ", FUSED_SIZES)
@pytest.mark.parametrize("lbc", [None, {"ubar": ["CLA", "CLO", "CLO", "CLO"], "zeta": ["GRA", "CLO", "GRA", "CLO"]},
                                 {"ubar": ["GRA", "CLO", "GRA", "CLO"], "zeta": ["GRA", "CLO", "GRA", "CLO"]}])
def test_fused_maps_accepted(szI, szJ, lbc):
    from sailfish_amd.mod_boundary import Boundary
    from sailfish_amd.step2d_fused import check_maps
    b = Boundary(_G(szI, szJ), lbc)
    ok, msg = check_maps(szI, szJ, szI, b)
    assert ok, msg


def test_fused_maps_rejections():
    from types import SimpleNamespace
    from sailfish_amd.mod_boundary import Boundary
    from sailfish_amd.step2d_fused import check_maps
    # the reference's vbar gradient copy reads ubar's source list: far sources
    b = Boundary(_G(40, 20), {"ubar": ["CLA", "CLO", "GRA", "CLO"], "vbar": ["GRA", "CLO", "CLO", "CLO"]})
    ok, msg = check_maps(40, 20, 40, b)
    assert not ok and "velocity" in msg
    # hand-made ubar copies on a 140 x 20 grid (64-wide tiles from i = 0): a neighbour
    # copy inside one tile region is accepted; a source that is itself a destination, a
    # source two points away, and a copy across a tile's left edge are not
    z = (np.zeros(0, np.int32), np.zeros(0, np.int32))
    v = z
    f = lambda i, j: j * 140 + i

    def chk(pairs):
        d = np.array([p[0] for p in pairs], np.int32)
        s_ = np.array([p[1] for p in pairs], np.int32)
        return check_maps(140, 20, 140, SimpleNamespace(zeta_map=z, ubar_map2=(d, s_), vbar_map2=v))

    assert chk([(f(139, 5), f(138, 5))])[0]
    assert chk([(f(1, 5), f(2, 5)), (f(139, 5), f(138, 5))])[0]
    ok, msg = chk([(f(1, 5), f(2, 5)), (f(2, 5), f(3, 5))])       # f(2,5) is both
    assert not ok and "neighbour" in msg
    ok, msg = chk([(f(1, 5), f(3, 5))])                            # two points away
    assert not ok and "neighbour" in msg
    ok, msg = chk([(f(64, 1), f(63, 1))])   # tile 1's region [64, 128] holds 64, not 63
    assert not ok and "outside the region" in msg
    b = Boundary(_G(30, 20), None)
    ok, msg = check_maps(30, 20, 20, b)     # i%szJ mask with 30 % 20 != 0
    assert not ok and "column-0" in msg
    ok, msg = check_maps(40, 20, 20, Boundary(_G(40, 20), None))
    assert ok, msg


def test_c_abi_rejects_bad_arguments_without_a_gpu():
    """Argument checks of the C ABI return SF_ERR_ARG with a message before any HIP
    call: null handles and fields, unknown precision, unsupported tile heights, grids
    too small, empty ensembles."""
    from sailfish_amd import _lib
    from sailfish_amd.phys2d import sf_phys_fields
    lib = _lib.load()
    lib.sf_last_error_string.restype = C.c_char_p
    P = C.c_void_p
    err = lambda: lib.sf_last_error_string().decode()

    f = lib.sf_phys_substep_fused
    f.argtypes = [C.POINTER(sf_phys_fields), C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_int32, P]
    f.restype = C.c_int32
    assert f(None, 0, 1.0, 0, 0.5, 8, None) == -1 and "null" in err()
    a = sf_phys_fields()
    a.szI, a.szJ = 2, 64
    assert f(C.byref(a), 0, 1.0, 0, 0.5, 8, None) == -1 and "too small" in err()
    # every pointer set to a dummy non-null address: the tile and precision checks
    # reject the call before anything is launched or dereferenced
    a.szI, a.szJ = 128, 64
    for name, ct in sf_phys_fields._fields_:
        if ct is C.c_void_p:
            setattr(a, name, 0x1000)
        elif getattr(ct, "_type_", None) is C.c_void_p:
            fld = getattr(a, name)
            for k in range(len(fld)):
                fld[k] = 0x1000
    assert f(C.byref(a), _lib.SF_FP64, 1.0, 0, 0.5, 5, None) == -1 and "tile height" in err()
    assert f(C.byref(a), 7, 1.0, 0, 0.5, 8, None) == -1 and "precision" in err()

    e = lib.sf_ens_create
    e.argtypes = [C.POINTER(sf_phys_fields), C.POINTER(C.c_double), C.c_int32, C.c_int32, C.POINTER(P)]
    e.restype = C.c_int32
    out = P()
    dt = (C.c_double * 1)(1.0)
    assert e(C.byref(a), dt, 0, 0, C.byref(out)) == -1
    assert e(C.byref(a), dt, 1, 9, C.byref(out)) == -1 and "precision" in err()
    lib.sf_ens_configure.argtypes = [P, C.c_int32, C.c_int32, C.c_int32]
    assert lib.sf_ens_configure(None, 8, 0, 0) == -1 and "null" in err()

    lib.sf_mlp_create.argtypes = [C.c_int32, C.POINTER(P)]
    assert lib.sf_mlp_create(0, C.byref(out)) == -1
    lib.sf_closure_stress.argtypes = [P, P, C.c_int32, P, P, P, P]
    assert lib.sf_closure_stress(None, None, 0, None, None, None, None) == -1 and "null" in err()
    lib.sf_rhs3d.argtypes = [P] * 11 + [C.c_int64, C.c_int32, C.c_int32, P]
    assert lib.sf_rhs3d(*([None] * 11), 0, 1, 1, None) == -1 and "null context" in err()


def test_filtered_transport_weights_sum_to_one():
    """DU_avg2 weights of the coupled substeps (step2d_LF_AM3.h.py:778-846): over the
    nfast+1 substeps of a barotropic loop they sum to the filter's total weight, 1."""
    from sailfish_amd.coupled3d import filtered_transport_weights
    from sailfish_amd.mod_comptimes import CompTimes
    from sailfish_amd.set_weights import set_weights
    for ndtfast in (4, 6, 30, 40):
        ct = CompTimes(dt=300.0, ndtfast=ndtfast)
        set_weights(ct)
        tot = sum(sum(filtered_transport_weights(ct.weight2, k)) for k in range(1, ct.nfast + 2))
        assert abs(tot - 1.0) <= 1e-12
        w2p, w2c = filtered_transport_weights(ct.weight2, 1)
        assert abs(w2p + ct.weight2[1] / 12.0) <= 1e-17 and w2c == ct.weight2[0]
