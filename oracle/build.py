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


def build(verbose: bool = True, sources=None, out=None, opt=None, extra=()) -> Path:
    """Compile ``sources`` (default: oracle/sf_oracle*.c) into ``out`` (default:
    oracle/liboracle.so).  ``opt`` replaces the optimisation flag and ``extra`` adds
    flags: the mutation test builds its variants at -O1 to save time, the CPU-baseline
    leg of bench.py builds with -O3 -march=native on the machine it times.
    -ffp-contract=off and IEEE semantics hold in every variant, so all of them compute
    the same bits."""
    sources = list(sources) if sources is not None else SOURCES
    out = Path(out) if out is not None else LIB
    flags = [opt if (f == "-O2" and opt) else f for f in CFLAGS] + list(extra)
    tmp = out.with_suffix(".so.tmp")
    cmd = ["gcc", *flags, *map(str, sources), "-o", str(tmp), "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"[oracle.build] built {out}", file=sys.stderr)
    return out


if __name__ == "__main__":
    build()
