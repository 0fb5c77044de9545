"""Build the gfx950 HIP library ``libsfhip.so`` in-tree.

``python -m sailfish_amd.build`` (or ``__graft_entry__.build()``) compiles every
``csrc/*.hip`` with hipcc for gfx950 and links them into
``sailfish_amd/libsfhip.so``.  Object files go to ``_build/`` at the repository root.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build" / "sfhip"
LIB = PKG / "libsfhip.so"

ARCH = os.environ.get("SF_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-munsafe-fp-atomics",
    "-Wall",
    "-Wno-unused-function",
    # no packed-FP32 instructions (v_pk_*_f32) anywhere in the library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]


def _hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libsfhip.so")
    return exe


# Matrix-core products of the closure network's inference write their results to
# ordinary VGPRs (no accumulation-register copies before the ReLU packs) in the units
# that evaluate the network inside other kernels.
UNIT_FLAGS = {
    name: ["-mllvm", "-amdgpu-mfma-vgpr-form"]
    for name in ("phys2d_fused.hip", "phys2d_fused_f32.hip", "closure_infer.hip")
}


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [_hipcc(), *HIPCC_FLAGS, *UNIT_FLAGS.get(src.name, []), "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"[sailfish_amd.build] built {LIB} from {len(sources)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
