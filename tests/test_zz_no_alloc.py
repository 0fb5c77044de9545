"""sf_obc_volcons_noalloc (static device work space, no allocation) against
sf_obc_volcons (caller work buffer), both called through the C ABI: two calls in a row
on one device, the second on a smaller grid, each bitwise equal in the sums (area,
flux, ubar_xs) and in the corrected boundary velocities.  Also the size limit of the
static work space."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = C.c_void_p


def _lib():
    from sailfish_amd import _lib as L
    lib = L.load()
    lib.sf_obc_volcons_work_size.argtypes = [C.c_int32, C.c_int32]
    lib.sf_obc_volcons_work_size.restype = C.c_int32
    lib.sf_obc_volcons.argtypes = [P] * 6 + [C.c_int32] * 5 + [P, P, P]
    lib.sf_obc_volcons.restype = C.c_int32
    lib.sf_obc_volcons_noalloc.argtypes = [P] * 6 + [C.c_int32] * 5 + [P, P]
    lib.sf_obc_volcons_noalloc.restype = C.c_int32
    lib.sf_last_error_string.restype = C.c_char_p
    return lib


def _fields(szI, szJ, seed, dtype):
    import torch
    rng = np.random.default_rng(seed)
    arrs = (0.1 * rng.standard_normal((szJ, szI)),            # zeta
            20 + 30 * rng.random((szJ, szI)),                  # h
            1.0 / (1e3 * (1 + 0.2 * rng.random((szJ, szI)))),  # pm
            1.0 / (1e3 * (1 + 0.2 * rng.random((szJ, szI)))),  # pn
            0.2 * rng.standard_normal((szJ, szI)),             # ubar
            0.2 * rng.standard_normal((szJ, szI)))             # vbar
    return [torch.from_numpy(a).to("cuda", dtype).contiguous() for a in arrs]


def _run(lib, fields, szI, szJ, prec, flags, noalloc):
    import torch
    out = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda")
    ptrs = [t.data_ptr() for t in fields]
    stream = torch.cuda.current_stream().cuda_stream
    if noalloc:
        rc = lib.sf_obc_volcons_noalloc(*ptrs, szI, szJ, prec, flags, 1, out.data_ptr(), stream)
    else:
        work = torch.zeros(lib.sf_obc_volcons_work_size(szI, szJ), dtype=torch.float64, device="cuda")
        rc = lib.sf_obc_volcons(*ptrs, szI, szJ, prec, flags, 1, work.data_ptr(), out.data_ptr(), stream)
    assert rc == 0, lib.sf_last_error_string()
    torch.cuda.synchronize()
    return out.cpu().numpy(), fields[4].cpu().numpy(), fields[5].cpu().numpy()


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_noalloc_matches_work_buffer_form_two_calls(dt):
    import torch
    lib = _lib()
    tdt, prec = (torch.float64, 0) if dt == "f64" else (torch.float32, 1)
    # the first call covers the most boundary points (many partial blocks), the second
    # fewer, so it runs over partials the first call left in the static array
    for k, (szI, szJ, flags) in enumerate([(3000, 1200, 15), (300, 130, 2 | 8)]):
        print(f"[case] noalloc {dt} call {k}: {szI}x{szJ} flags={flags}", file=sys.stderr, flush=True)
        ref = _run(lib, _fields(szI, szJ, 40 + k, tdt), szI, szJ, prec, flags, noalloc=False)
        got = _run(lib, _fields(szI, szJ, 40 + k, tdt), szI, szJ, prec, flags, noalloc=True)
        assert np.isfinite(ref[0]).all() and ref[0][0] > 0
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)


def test_noalloc_rejects_grids_beyond_the_static_work_space():
    import torch
    lib = _lib()
    fields = _fields(8, 8, 1, torch.float64)
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    ptrs = [t.data_ptr() for t in fields]
    # 2*(200000-2) + 2*(100000-2) boundary points exceed the 4096-double work space; the
    # size check comes before any device access, so the small buffers are never read
    assert lib.sf_obc_volcons_work_size(200000, 100000) > 4096
    rc = lib.sf_obc_volcons_noalloc(*ptrs, 200000, 100000, 0, 15, 1, out.data_ptr(), None)
    assert rc == -1
    assert b"static maximum" in lib.sf_last_error_string()
    # the largest grid that fits is accepted by the check (square grid: 4*(n-2) points)
    assert lib.sf_obc_volcons_work_size(131010, 131010) <= 4096
