"""Compile the oracle's C restatement into ``oracle/liboracle.so`` (gcc, FP64,
-ffp-contract=off so that no multiply-add is fused; OpenMP for the CPU baseline).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference's own kernel bodies are not built here: they need CUDA built-ins
(``__global__``, ``threadIdx``...) that this image does not provide, and a source
patch for dangling references, so ``oracle/_ref`` stays empty.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
SOURCES = sorted(HERE.glob("sf_oracle*.c"))
CFLAGS = ["-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-fopenmp", "-Wall", "-Wno-unused-function"]


def build(verbose: bool = True) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", *CFLAGS, *map(str, SOURCES), "-o", str(tmp), "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"[oracle.build] built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
