#!/usr/bin/env python3
"""Benchmark of the barotropic LF-AM3 substep (step2dPredictor + step2dCorrector of
nonlinear/step2d.py) on one MI355X.

Workload (default): closed basin of 2048 x 1024 rho points, dx = 2 km, seeded
bathymetry 200..4000 m, beta plane, seeded free-surface noise, dt = 5 s; FP64 state,
fused substep kernel, reference discretisation (sf_step2d_substep), with the
column-0/row-0 mask variant of computeMomentumPred (SF_FLAG_FIX_Q5_MASK) so that the
non-square grid evolves like the square case.  A "step" is one barotropic substep
over the whole grid; inputs are resident in HBM before timing starts.

Prints one JSON line.  ``roofline`` is the dominant kernel (the tiled substep
kernel): algorithmic bytes per launch over its mean launch time, measured with HIP
events on the launch stream.  ``cpu_baseline`` times the CPU oracle (C restatement
of the reference kernels, OpenMP) on a bounded prefix of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "cell-steps/s + % HBM roofline; closure-coarse vs fine run at equal zeta RMSE"

# Algorithmic HBM bytes per rho point of the tiled substep kernel: every field read
# or written once (boundary-only reads omitted).
#   reads : zeta_t1, zeta_t0, h, ubar_t1, vbar_t1, on_u, om_v, pm, pn,
#           rzeta_t0, rzeta_t2                                  (11 bulk)
#           Zt_avg1, DU_avg1, DV_avg1                           (3 FP64)
#   writes: zeta_t2, ubar_t2, vbar_t2, rzeta_t1, rubar_t1       (5 bulk)
#           Zt_avg1, DU_avg1, DV_avg1                           (3 FP64)
BULK_FIELDS = 16
F64_FIELDS = 6


def bytes_per_point(bulk_bytes: int) -> int:
    return BULK_FIELDS * bulk_bytes + F64_FIELDS * 8


def run_gpu(args):
    import torch
    from sailfish_amd import configs
    from sailfish_amd.model2d import build_model
    from sailfish_amd.step2d_fused import FusedStepper

    dtype = torch.float64 if args.dtype == "f64" else torch.float32
    cfg = configs.basin2(args.szI, args.szJ, dx=args.dx)
    ct, GRID, OCEAN, BOUNDARY = build_model(cfg, dtype=dtype, ndtfast=30, fix_q5=True)
    st = FusedStepper(GRID, BOUNDARY, dtype)
    S2 = GRID.szI * GRID.szJ

    def substep(parts=3):
        ct.next2DTimeStep()
        OCEAN.cycleTimes2D()
        st.step(ct, OCEAN, parts=parts)

    for _ in range(args.warmup):
        substep()
    torch.cuda.synchronize()

    # timed region: K whole substeps
    t0 = time.perf_counter()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(args.steps):
        substep()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev0.elapsed_time(ev1)
    ms_per_step = ms_total / args.steps

    # dominant kernel: HIP events around each tiled launch (the masked-points
    # launch follows outside the events), same stream as the launches
    nk = max(10, min(args.steps, 100))
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(nk)]
    for a, b in evs:
        ct.next2DTimeStep()
        OCEAN.cycleTimes2D()
        a.record()
        st.step(ct, OCEAN, parts=1)
        b.record()
        st.step(ct, OCEAN, parts=2)
    torch.cuda.synchronize()
    kern_ms = sum(a.elapsed_time(b) for a, b in evs) / nk

    z = OCEAN.zeta_t2.detach().float().cpu()
    finite = bool(torch.isfinite(z).all())
    bulk = 8 if args.dtype == "f64" else 4
    bpp = bytes_per_point(bulk)
    traffic = bpp * S2
    achieved = traffic / (kern_ms * 1e-3) / 1e9
    return dict(cfg=cfg, S2=S2, ms_per_step=ms_per_step, wall=wall, kern_ms=kern_ms, bpp=bpp,
                achieved=achieved, finite=finite, nk=nk)


def run_cpu_baseline(args, budget_s: float):
    """Oracle (one C function per reference kernel, OpenMP) on the same grid for a
    bounded number of substeps."""
    import numpy as np
    from oracle import oracle as O
    from sailfish_amd import configs
    from sailfish_amd.mod_comptimes import CompTimes
    from sailfish_amd.mod_grid import update_metrics
    from sailfish_amd.set_weights import set_weights

    cfg = configs.basin2(args.szI, args.szJ, dx=args.dx)
    threads = int(os.environ.get("SF_CPU_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    O.set_threads(threads)
    ct = CompTimes(cfg["dtfast"] * 30, 30)
    set_weights(ct)
    m = update_metrics(cfg["pm"], cfg["pn"], cfg["f"])
    g = O.OGrid(cfg["szI"], cfg["szJ"], 1, dict(h=cfg["h"], pm=cfg["pm"], pn=cfg["pn"], on_u=m["on_u"],
                                                   om_v=m["om_v"]), fix_q5=True)
    s = O.OState(g)
    for k in range(3):
        s.zeta[k][:] = cfg["zeta0"].ravel()
    b = O.OBoundary(g)
    t = O.OTimes(cfg["dtfast"], ct.weight1, ct.nfast)
    O.main2d(t, g, s, b, 1)    # warm-up
    n = 0
    t0 = time.perf_counter()
    while True:
        O.main2d(t, g, s, b, 1)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 1000:
            break
    S2 = cfg["szI"] * cfg["szJ"]
    return dict(value=S2 * n / el, unit="cell-steps/s", cores=threads, kind="port",
                sample=f"{n} substeps of the {cfg['szI']}x{cfg['szJ']} basin (oracle C restatement of "
                       f"the reference kernels, FP64, {threads} OpenMP threads, {el:.1f} s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    ap.add_argument("--szI", type=int, default=2048)
    ap.add_argument("--szJ", type=int, default=1024)
    ap.add_argument("--dx", type=float, default=2e3)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the barotropic path would shard as row slabs with halo exchange; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    r = run_gpu(args)
    value = r["S2"] / (r["ms_per_step"] * 1e-3)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {
            "workload": f"barotropic LF-AM3 substep (step2d predictor+corrector, fused), closed basin "
                        f"{args.szI}x{args.szJ} rho points, dx={args.dx/1e3:g} km, dt=5 s, "
                        f"{'FP64' if args.dtype == 'f64' else 'FP32 bulk + FP64 averages'}",
            "grid": [args.szI, args.szJ],
            "kernel": "sf_step2d_substep",
            "mask": "SF_FLAG_FIX_Q5_MASK",
            "finite_after_run": r["finite"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": r["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": r["achieved"] / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_step2d_fused",
            "kernel_ms": r["kern_ms"],
            "bytes_per_point": r["bpp"],
            "launches_timed": r["nk"],
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = run_cpu_baseline(args, args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
