"""ctypes binding of ``libsfhip.so`` (the C ABI declared in ``include/sailfish_hip.h``).

The product path has no CPU fallback: if the library is missing or a GPU is not
available, every entry point raises.  Build the library with
``python -m sailfish_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsfhip.so"

SF_OK = 0
SF_FP64 = 0
SF_FP32 = 1
SF_FLAG_FIX_Q5_MASK = 1

_lock = threading.Lock()
_lib = None

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
F64 = C.c_double


class SailfishError(RuntimeError):
    """A C-ABI entry point returned an error code."""


class sf_grid_desc(C.Structure):
    _fields_ = [
        ("szI", I32), ("szJ", I32), ("szK", I32), ("flags", I32),
        ("on_u", P), ("om_v", P), ("pn", P), ("pm", P),
    ]


# name -> argtypes (restype is always int32 unless listed in _RESTYPES)
_SIGS = {
    "sf_ctx_create": [C.POINTER(sf_grid_desc), C.POINTER(P)],
    "sf_ctx_destroy": [P],
    "sf_last_error_string": [],
    "sf_version": [],
    "sf_compute_zeta_rhs": [P, P, P, P, P, P, P, P, I32, F64, P, P],
    "sf_compute_zeta_pred": [P, F64, P, P, P, P, P],
    "sf_compute_momentum_rhs_pred": [P, P, P, P, P, P, F64, F64, P],
    "sf_compute_momentum_rhs_corr": [P, P, P, P, P, P, P, F64, P],
    "sf_compute_momentum_pred": [P, F64, P, P, P, P, P, P, P, P, P, P],
    "sf_adams_moulton_corr3rd": [P, F64, P, P, P, P, P],
    "sf_adams_moulton_corr3rd2": [P, F64, P, P, P, P, P, P, P, P, P, P, P, P],
    "sf_bc_set": [P, P, I32, F64, P],
    "sf_bc_apply_map": [P, P, P, I32, F64, P],
    "sf_step2d_check_maps": [I32, I32, I32, P, P, I32, P, P, I32, P, P, I32],
    "sf_add_coriolis": [P, P, P, P, P, P, P, P],
    "sf_horizontal_advection": [P, P, P, P, P, P, P, I64, P],
    "sf_vertical_advection": [P, P, P, P, P, P, P, P],
    "sf_rhs3d": [P, P, P, P, P, P, P, P, P, P, P, I64, I32, I32, P],
    "sf_set_maxflux": [P, P, P, P, P, P, P],
    "sf_omega": [P, P, P, P, P, P, P],
    "sf_rhs3d_f32": [P, P, P, P, P, P, P, P, P, P, P, I64, I32, I32, P],
    "sf_set_maxflux_f32": [P, P, P, P, P, P, P],
    "sf_omega_f32": [P, P, P, P, P, P, P],
    "sf_step3d_uv_f32": [P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, I32, I32, F64, F64, P, P, P, P],
    "sf_set_depth": [P, I32, P, P, P, P, F64, P, P, P, P, P, P],
    "sf_set_zeta": [P, P, P, P, P],
    "sf_set_lateral_uv_bcs": [P, F64, P, P, P, P, P],
    "sf_set_depth_f32": [P, I32, P, P, P, P, F64, P, P, P, P, P, P],
    "sf_set_zeta_f32": [P, P, P, P, P],
    "sf_set_lateral_uv_bcs_f32": [P, F64, P, P, P, P, P],
    "sf_step3d_uv": [P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, I32, I32, F64, F64, P, P, P, P],
}
_RESTYPES = {"sf_last_error_string": C.c_char_p}


def exported_symbols() -> list[str]:
    """Names the ctypes layer binds (every one must be exported by the library)."""
    return list(_SIGS)


def load(path: Path | None = None):
    """Load the shared library (no GPU work).  Raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path else LIB_PATH
        if not p.exists():
            raise SailfishError(
                f"{p} is missing: build it with `python -m sailfish_amd.build` "
                "(the HIP path has no CPU fallback)")
        lib = C.CDLL(str(p))
        for name, argtypes in _SIGS.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue
            fn.argtypes = argtypes
            fn.restype = _RESTYPES.get(name, I32)
        if path is None:
            _lib = lib
        return lib


def last_error() -> str:
    msg = load().sf_last_error_string()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> None:
    if rc != SF_OK:
        raise SailfishError(f"{what or 'sailfish call'} failed ({rc}): {last_error()}")


def call(name: str, *args) -> None:
    rc = getattr(load(), name)(*args)
    check(rc, name)


_gpu_checked = False


def require_gpu() -> None:
    """Fail loudly when no HIP device is visible (the product path never runs on CPU)."""
    global _gpu_checked
    if _gpu_checked:
        return
    import torch
    if not torch.cuda.is_available():
        raise SailfishError("no HIP device visible: the Sailfish MI355X path requires a GPU")
    load()
    _gpu_checked = True


def stream_handle(stream=None) -> int:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)
