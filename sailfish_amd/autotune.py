"""Schedule autotuner of the ensemble launch.

For an ensemble (its member shapes and precision) it times every candidate
schedule -- tile height (4, 8, 16), tile order (0: strided over blocks, 1:
contiguous chunk per block with XCD banding) and grid size (one block per tile,
or a persistent grid of 2/4/8 blocks per CU) -- on a scratch copy of the members'
state, picks the fastest and caches the choice in a JSON file keyed by device,
precision and the sorted member shapes (``SF_AUTOTUNE_CACHE`` or
``~/.cache/sailfish_amd/autotune.json``)."""
from __future__ import annotations

import hashlib
import json
import os
from pathlib import Path

import torch

N_CU = 256


def cache_path() -> Path:
    p = os.environ.get("SF_AUTOTUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "sailfish_amd" / "autotune.json"


def _key(ens) -> str:
    shapes = sorted((m.GRID.szI, m.GRID.szJ) for m in ens.members)
    h = hashlib.sha1(json.dumps(shapes).encode()).hexdigest()[:16]
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    return f"{dev}|{'f64' if ens.dtype == torch.float64 else 'f32'}|{len(shapes)}|{h}"


def candidates(ntiles_by_tj: dict):
    out = []
    for tj in (4, 8, 16):
        nt = ntiles_by_tj[tj]
        grids = [0] + [g for g in (2 * N_CU, 4 * N_CU, 8 * N_CU) if g < nt]
        for order in (0, 1):
            for g in grids:
                out.append((tj, order, g))
    return out


def _time(ens, nsub: int, reps: int) -> float:
    best = float("inf")
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _k in range(nsub):
            ens.substep(False, 0.0)
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b) / nsub)
    return best


def tune(ens, nsub: int = 6, reps: int = 3, use_cache: bool = True, verbose: bool = False) -> dict:
    """Pick and apply the fastest schedule for ``ens``.  The timing runs advance a
    scratch copy of the state (the members' arrays are restored afterwards)."""
    if nsub % 3:
        raise ValueError("nsub must be a multiple of 3 (the slot rotation must come back)")
    key = _key(ens)
    path = cache_path()
    cache = {}
    if use_cache and path.exists():
        try:
            cache = json.loads(path.read_text())
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        c = cache[key]
        ens.configure(c["tile_j"], c["order"], c["grid"])
        return dict(c, cached=True)

    # save the state the timing runs overwrite
    names = ("zeta", "ubar", "vbar", "rzeta", "rubar", "rvbar")
    saved = [{n: getattr(m.OCEAN, n).clone() for n in names} |
             {n: getattr(m.OCEAN, n).clone() for n in ("Zt_avg1", "DU_avg1", "DV_avg1")} for m in ens.members]
    pending = ens.pending
    ntiles = {}
    for tj in (4, 8, 16):
        ens.configure(tj, 1, 0)
        ntiles[tj] = ens.num_tiles()
    results = []
    for cand in candidates(ntiles):
        ens.configure(*cand)
        for _k in range(3):                # warm-up (code load); multiples of 3 keep the rotation
            ens.substep(False, 0.0)
        torch.cuda.synchronize()
        t = _time(ens, nsub, reps)
        results.append((t, cand))
        if verbose:
            print(f"autotune {cand}: {t * 1e3:.1f} us/substep")
    results.sort()
    t_best, best = results[0]
    for m, sv in zip(ens.members, saved):
        for n, v in sv.items():
            getattr(m.OCEAN, n).copy_(v)
    torch.cuda.synchronize()
    # every candidate ran a multiple of 3 substeps: the library's slot rotation is back
    # where it was, and the restored slot arrays are the state before tuning
    ens.pending = pending
    ens.configure(*best)
    rec = dict(tile_j=best[0], order=best[1], grid=best[2], ms_per_substep=t_best,
               tried=[dict(tile_j=c[0], order=c[1], grid=c[2], ms=t) for t, c in results])
    if use_cache:
        cache[key] = {k: rec[k] for k in ("tile_j", "order", "grid", "ms_per_substep")}
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            path.write_text(json.dumps(cache, indent=1))
        except OSError:
            pass
    return dict(rec, cached=False)
