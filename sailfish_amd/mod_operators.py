"""Kernel handles of the hot path (modules/mod_operators.py:31-185).

``initModule(GRID)`` replaces the reference's ``initModule`` + three ``initialize``
launches: it creates an ``sf_ctx`` (sizes + metric pointers) in libsfhip.so.
The raw handles keep the reference's ``kernel(grid, block, args)`` calling form; the
launch geometry arguments are accepted and ignored (each C-ABI entry point picks
its own), and every call goes to a hand-written gfx950 kernel on the current torch
stream.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import sf_grid_desc

G = None
CTX = None
shp = None
sz = None
bksz = (512,)     # reference launch constants (utility/misc.py:17), informative only
grsz = None


def _ptr(t) -> int:
    if isinstance(t, torch.Tensor):
        if not t.is_cuda:
            raise _lib.SailfishError("tensor argument is not on the GPU")
        return t.data_ptr()
    raise TypeError(f"expected a torch GPU tensor, got {type(t)}")


def _f64(t):
    if t.dtype != torch.float64:
        raise _lib.SailfishError(f"FP64 kernel called with {t.dtype} array")
    return _ptr(t)


def _bulk(name: str, fields) -> tuple[str, list]:
    """Entry point for the precision of the 3D bulk fields: ``name`` when all are FP64,
    ``name + "_f32"`` when all are FP32 (the FP32 bulk form); mixed precisions raise."""
    dts = {t.dtype for t in fields}
    if dts == {torch.float64}:
        return name, [_ptr(t) for t in fields]
    if dts == {torch.float32}:
        return name + "_f32", [_ptr(t) for t in fields]
    raise _lib.SailfishError(f"{name}: 3D fields must all be FP64 or all FP32, got {sorted(map(str, dts))}")


class Context:
    """Owns one sf_ctx (freed with the object)."""

    def __init__(self, GRID):
        _lib.require_gpu()
        self.GRID = GRID
        self._keep = (GRID.on_u, GRID.om_v, GRID.pn, GRID.pm)
        flags = _lib.SF_FLAG_FIX_Q5_MASK if getattr(GRID, "fix_q5", False) else 0
        d = sf_grid_desc(GRID.szI, GRID.szJ, GRID.N, flags,
                         _f64(GRID.on_u), _f64(GRID.om_v), _f64(GRID.pn), _f64(GRID.pm))
        h = C.c_void_p()
        _lib.check(_lib.load().sf_ctx_create(C.byref(d), C.byref(h)), "sf_ctx_create")
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.load().sf_ctx_destroy(h)
            except Exception:
                pass
            self.handle = None


def initModule(GRID):
    """mod_operators.py:31-45"""
    global G, CTX, shp, sz, grsz
    G = GRID
    shp = tuple(GRID.on_u.shape)
    sz = GRID.on_u.numel()
    grsz = (sz // bksz[0] + 1,)
    CTX = Context(GRID)
    GRID._sf_ctx = CTX
    return CTX


def ctx_of(GRID=None):
    if GRID is not None and getattr(GRID, "_sf_ctx", None) is not None:
        return GRID._sf_ctx
    if CTX is None:
        raise _lib.SailfishError("initModule(GRID) has not been called")
    return CTX


def _stream():
    return _lib.stream_handle()


def _cff1(weight1, iif) -> float:
    w = weight1.detach().cpu().numpy() if isinstance(weight1, torch.Tensor) else np.asarray(weight1)
    k = min(max(int(iif) - 1, 0), w.shape[0] - 1)
    return float(w[k])


# ---- raw handles (reference argument tuples) ---------------------------------------

def computeZetaRHS(grid, block, args, ctx=None):
    """mod_cppkernels.cpp:80-128; args = (zeta, h, ubar, vbar, DU_avg1, DV_avg1,
    Zt_avg1, isPredictor, weight1, weight2, iif, res)."""
    zeta, h, ubar, vbar, DU, DV, Zt, isPred, w1, _w2, iif, res = args
    c = ctx or ctx_of()
    _lib.call("sf_compute_zeta_rhs", c.handle, _f64(zeta), _f64(h), _f64(ubar), _f64(vbar),
              _f64(DU), _f64(DV), _f64(Zt), int(bool(isPred)), _cff1(w1, iif), _f64(res), _stream())


def computeZetaPred(grid, block, args, ctx=None):
    """mod_cppkernels.cpp:133-150; args = (Dt, zeta_t0, zeta_t1, zeta_t2, rzeta_t1)."""
    Dt, z0, z1, z2, rz1 = args
    c = ctx or ctx_of()
    _lib.call("sf_compute_zeta_pred", c.handle, float(Dt), _f64(z0), _f64(z1), _f64(z2), _f64(rz1), _stream())


def computeMomentumRHSPred(grid, block, args, ctx=None):
    """mod_cppkernels.cpp:12-42; args = (h, rhs_ubar, rhs_vbar, zeta_t0, zeta_t2, g, weight)."""
    h, ru, rv, z0, z2, gc, w = args
    c = ctx or ctx_of()
    _lib.call("sf_compute_momentum_rhs_pred", c.handle, _f64(h), _f64(ru), _f64(rv), _f64(z0), _f64(z2),
              float(gc), float(w), _stream())


def computeMomentumRHSCorr(grid, block, args, ctx=None):
    """mod_cppkernels.cpp:45-77; args = (h, rhs_ubar, rhs_vbar, zeta_t0, zeta_t1, zeta_t2, g)."""
    h, ru, rv, z0, z1, z2, gc = args
    c = ctx or ctx_of()
    _lib.call("sf_compute_momentum_rhs_corr", c.handle, _f64(h), _f64(ru), _f64(rv), _f64(z0), _f64(z1),
              _f64(z2), float(gc), _stream())


def computeMomentumPred(grid, block, args, ctx=None):
    """mod_cppkernels.cpp:155-186; args = (Dt, u_t1, u_t2, v_t1, v_t2, rhsu, rhsv, h,
    zeta_t1, zeta_t2)."""
    Dt, u1, u2, v1, v2, ru, rv, h, z1, z2 = args
    c = ctx or ctx_of()
    _lib.call("sf_compute_momentum_pred", c.handle, float(Dt), _f64(u1), _f64(u2), _f64(v1), _f64(v2),
              _f64(ru), _f64(rv), _f64(h), _f64(z1), _f64(z2), _stream())


def AdamsMoultonCorr3rd(grid, block, args, ctx=None):
    """mod_cppkernels.cpp:190-209; args = (Dt, v_t2, rhs_t0, rhs_t1, rhs_t2, ...).

    Arguments past the fifth are ignored, exactly as the reference kernel ignores the
    12-argument tuple that step2d.py:120-123 passes through the mis-bound handle
    (mod_operators.py:146)."""
    Dt, v2, r0, r1, r2 = args[:5]
    c = ctx or ctx_of()
    _lib.call("sf_adams_moulton_corr3rd", c.handle, float(Dt), _f64(v2), _f64(r0), _f64(r1), _f64(r2),
              _stream())


# mod_operators.py:146 binds this name to the 5-argument kernel
AdamsMoultonCorr3rd2 = AdamsMoultonCorr3rd


def AdamsMoultonCorr3rd2_intended(grid, block, args, ctx=None):
    """The 12-argument kernel of mod_cppkernels.cpp:213-251 (never launched by the
    reference); args = (Dt, u_t2, v_t2, rhsu_t0..t2, rhsv_t0..t2, h, zeta_t1, zeta_t2)."""
    Dt, u2, v2, ru0, ru1, ru2, rv0, rv1, rv2, h, z1, z2 = args
    c = ctx or ctx_of()
    _lib.call("sf_adams_moulton_corr3rd2", c.handle, float(Dt), _f64(u2), _f64(v2), _f64(ru0), _f64(ru1),
              _f64(ru2), _f64(rv0), _f64(rv1), _f64(rv2), _f64(h), _f64(z1), _f64(z2), _stream())


def apply_bc_map(var, dmap, clamp_val: float = 0.0):
    dst, src = dmap
    n = int(dst.numel())
    if n == 0:
        return
    _lib.call("sf_bc_apply_map", _f64(var), _ptr(dst), _ptr(src), n, float(clamp_val), _stream())


def setBC(grid, block, args):
    """mod_BC.c:6-13; args = (var, idx, size, val)."""
    var, idx, n, val = args
    n = int(n)
    if n == 0:
        return
    _lib.call("sf_bc_set", _f64(var), _ptr(idx), n, float(val), _stream())


# ---- 3D handles (nonlinear/rhs_kernels.cpp, nonlinear/step3D_uv.kernels.cpp) -------

def _launch_threads(grid, block, n_default):
    """Thread count of a reference launch ``kernel(grid, block, args)``."""
    try:
        return int(grid[0]) * int(block[0])
    except (TypeError, IndexError):
        return int(n_default)


def addCoriolis(grid, block, args, ctx=None):
    """rhs_kernels.cpp:6-35; args = (fomn, u, v, ru, rv, Hz)."""
    fomn, u, v, ru, rv, Hz = args
    c = ctx or ctx_of()
    _lib.call("sf_add_coriolis", c.handle, _f64(fomn), _f64(u), _f64(v), _f64(ru), _f64(rv), _f64(Hz), _stream())


def horizontalAdvection(grid, block, args, ctx=None):
    """rhs_kernels.cpp:70-132; args = (u, v, Huon, Hvom, ru, rv, BC)."""
    u, v, Hu, Hv, ru, rv, _BC = args
    c = ctx or ctx_of()
    S2 = c.GRID.szI * c.GRID.szJ
    nt = _launch_threads(grid, block, (S2 // 512 + 1) * 512)
    _lib.call("sf_horizontal_advection", c.handle, _f64(u), _f64(v), _f64(Hu), _f64(Hv), _f64(ru), _f64(rv),
              nt, _stream())


def rhs3d_fused(args, launch_threads: int, cor: bool = True, adv: bool = True, ctx=None):
    """nonlinear/rhs.py:38-71 as one launch: ru, rv := 0, addCoriolis (cor),
    horizontalAdvection with the reference's launch of ``launch_threads`` threads,
    verticalAdvection (adv).  args = (fomn, u, v, Huon, Hvom, W, Hz, ru, rv, BC); the 3D
    fields all FP64 (bitwise the reference's arithmetic) or all FP32 (FP32 bulk form),
    fomn FP64 in both."""
    fomn, u, v, Hu, Hv, W, Hz, ru, rv, BC = args
    c = ctx or ctx_of()
    name, p = _bulk("sf_rhs3d", (u, v, Hu, Hv, W, Hz, ru, rv))
    _lib.call(name, c.handle, _f64(fomn), *p[:6], _ptr(BC), *p[6:], int(launch_threads), int(bool(cor)),
              int(bool(adv)), _stream())


def verticalAdvection(grid, block, args, ctx=None):
    """rhs_kernels.cpp:135-197; args = (u, v, W, ru, rv, BC)."""
    u, v, W, ru, rv, BC = args
    c = ctx or ctx_of()
    _lib.call("sf_vertical_advection", c.handle, _f64(u), _f64(v), _f64(W), _f64(ru), _f64(rv), _ptr(BC),
              _stream())


def set_maxflux(grid, block, args, ctx=None):
    """step3D_uv.kernels.cpp:372-405; args = (u, v, Huon, Hvom, Hz), all FP64 or all FP32."""
    u, v, Hu, Hv, Hz = args
    c = ctx or ctx_of()
    name, p = _bulk("sf_set_maxflux", (u, v, Hu, Hv, Hz))
    _lib.call(name, c.handle, *p, _stream())


def omega(grid, block, args, ctx=None):
    """step3D_uv.kernels.cpp:62-165; args = (W, u, v, Huon, Hvom, z_w, BC); W, Huon, Hvom,
    z_w all FP64 or all FP32 (column sums in FP64 either way)."""
    W, _u, _v, Hu, Hv, z_w, BC = args
    c = ctx or ctx_of()
    name, p = _bulk("sf_omega", (W, Hu, Hv, z_w))
    _lib.call(name, c.handle, *p, _ptr(BC), _stream())


def set_depth(grid, block, args, ctx=None):
    """step3D_uv.kernels.cpp:225-291; args = (Vtransform, Zt_avg1, z_w, z_r, h, hc, Hz,
    sc_r, sc_w, Cs_r, Cs_w); z_w, z_r, Hz all FP64 or all FP32 (FP32 bulk form)."""
    Vt, Zt, z_w, z_r, h, hc, Hz, sc_r, sc_w, Cs_r, Cs_w = args
    c = ctx or ctx_of()
    name, (pw, pr, ph) = _bulk("sf_set_depth", (z_w, z_r, Hz))
    _lib.call(name, c.handle, int(Vt), _f64(Zt), pw, pr, _f64(h), float(hc), ph,
              _f64(sc_r), _f64(sc_w), _f64(Cs_r), _f64(Cs_w), _stream())


def set_zeta(grid, block, args, ctx=None):
    """step3D_uv.kernels.cpp:294-310; args = (zeta_t1, zeta_t2, Zt_avg1); zeta FP64 or
    FP32, Zt_avg1 FP64."""
    z1, z2, Zt = args
    c = ctx or ctx_of()
    name, p = _bulk("sf_set_zeta", (z1, z2))
    _lib.call(name, c.handle, *p, _f64(Zt), _stream())


_lateral_checked: dict = {}


def setLateralUVBCs(grid, block, args, ctx=None):
    """step3D_uv.kernels.cpp:314-368; args = (t, u, v, uIdx2, vIdx2, uType, vType)."""
    t, u, v, _ui, _vi, ut, vt = args
    c = ctx or ctx_of()
    # the boundary types never change: check them once per array (no host copy, and so
    # no synchronisation, on later calls, e.g. inside a HIP graph capture)
    key = (vt.data_ptr(), vt.numel())
    if key not in _lateral_checked:
        vt_host = vt.detach().cpu().numpy()
        _lateral_checked[key] = bool(((vt_host > 0) & ((vt_host & (1 << 4)) == 0) & ((vt_host & (1 << 3)) == 0)
                                      & ((vt_host & (1 << 6)) != 0)).any())
    if _lateral_checked[key]:
        raise _lib.SailfishError("setLateralUVBCs: gradient v boundaries index the 2D index field at 3D "
                                 "offsets in the reference (step3D_uv.kernels.cpp:363); not supported")
    name, p = _bulk("sf_set_lateral_uv_bcs", (u, v))
    _lib.call(name, c.handle, float(t), *p, _ptr(ut), _ptr(vt), _stream())


_velEq = {}


def setVerticalVelEq(grid, block, args, ctx=None):
    """step3D_uv.kernels.cpp:21-27: remember the tridiagonal scratch (SD, D, RHS)."""
    SD, D, RHS = args
    c = ctx or ctx_of()
    _velEq[id(c)] = (SD, D, RHS)


def step3d_UV(grid, block, args, ctx=None):
    """step3D_uv.kernels.cpp:646-707; args = (u_t2, v_t2, ru_t2, rv_t2, ubar_t1, vbar_t1,
    ubar_t2, vbar_t2, Hz, AKv, z_r, DU_avg1, DV_avg1, tmpU, tmpV, h, iic, ntfirst,
    lambda, AK, dt) as passed by main3d.py:304-306."""
    (u, v, ru, rv, ub1, vb1, ub2, vb2, Hz, AKv, z_r, DU, DV, tmpU, tmpV, _h, iic, ntfirst, lam, _AK, dt) = args
    c = ctx or ctx_of()
    if id(c) not in _velEq:
        raise _lib.SailfishError("setVerticalVelEq has not been called for this grid")
    SD, D, RHS = _velEq[id(c)]
    if u.dtype == torch.float32 and SD.dtype != torch.float32:
        # FP32 bulk form: 3D fields, ubar/vbar and the scratch in FP32, DU/DV_avg1 in FP64;
        # FP64 scratch registered by setVerticalVelEq is replaced by FP32 scratch of its size
        SD, D, RHS = _velEq32(c, SD)
    name, p = _bulk("sf_step3d_uv", (u, v, ru, rv, ub1, vb1, ub2, vb2, Hz, AKv, z_r))
    q = [_ptr(x) for x in (tmpU, tmpV)] if name.endswith("_f32") else [_f64(tmpU), _f64(tmpV)]
    if name.endswith("_f32") and {tmpU.dtype, tmpV.dtype} != {torch.float32}:
        raise _lib.SailfishError("step3d_UV: tmpU/tmpV must be FP32 with FP32 fields")
    _lib.call(name, c.handle, *p, _f64(DU), _f64(DV), *q, int(iic), int(ntfirst), float(lam), float(dt),
              _ptr(SD), _ptr(D), _ptr(RHS), _stream())


_velEq32_cache: dict = {}


def _velEq32(c, like):
    """FP32 scratch of the FP32 step3d_UV (same shape and device as the FP64 eqSD/eqD/eqRHS)."""
    key = id(c)
    if (key not in _velEq32_cache or _velEq32_cache[key][0].numel() != like.numel()
            or _velEq32_cache[key][0].device != like.device):
        _velEq32_cache[key] = tuple(torch.zeros(like.numel(), dtype=torch.float32, device=like.device)
                                    for _ in range(3))
    return _velEq32_cache[key]
