"""Shared helpers: build the oracle state from the same configuration dict the
product model is built from, and compare states."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from sailfish_amd.mod_grid import update_metrics


def oracle_model(cfg, weight1, nfast, lbc=None, fix_q5=False):
    m = update_metrics(np.asarray(cfg["pm"], float), np.asarray(cfg["pn"], float), np.asarray(cfg["f"], float))
    fields = dict(h=cfg["h"], pm=cfg["pm"], pn=cfg["pn"], on_u=m["on_u"], om_v=m["om_v"])
    g = O.OGrid(cfg["szI"], cfg["szJ"], 1, fields, fix_q5=fix_q5)
    s = O.OState(g)
    z = np.asarray(cfg["zeta0"], float).reshape(-1)
    for k in range(3):
        s.zeta[k][:] = z
        if "ubar0" in cfg:
            s.ubar[k][:] = np.asarray(cfg["ubar0"], float).reshape(-1)
        if "vbar0" in cfg:
            s.vbar[k][:] = np.asarray(cfg["vbar0"], float).reshape(-1)
    b = O.OBoundary(g, lbc)
    t = O.OTimes(cfg["dtfast"], weight1, nfast)
    return t, g, s, b


STATE2D = ("zeta", "ubar", "vbar", "rzeta", "rubar", "rvbar")


def product_arrays(OCEAN):
    """All 2D state arrays of a T_OCEAN as host numpy, keyed like the oracle."""
    out = {}
    for n in STATE2D:
        out[n] = getattr(OCEAN, n).detach().cpu().numpy().reshape(3, -1)
    for n in ("Zt_avg1", "DU_avg1", "DV_avg1"):
        out[n] = getattr(OCEAN, n).detach().cpu().numpy()
    out["slots"] = {n: [OCEAN.slot_of(n, f"t{k}") for k in range(3)] for n in STATE2D}
    return out


def oracle_arrays(s):
    out = {n: getattr(s, n) for n in STATE2D}
    for n in ("Zt_avg1", "DU_avg1", "DV_avg1"):
        out[n] = getattr(s, n)
    out["slots"] = {n: list(s.slot[n]) for n in STATE2D}
    return out


def bitwise_equal(a, b):
    """Bit-for-bit equality, except that any NaN equals any NaN: the payload of a
    generated NaN differs between x86 (0xfff8...) and AMD GPUs (0x7ff8...)."""
    a = np.ascontiguousarray(a); b = np.ascontiguousarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    u = {8: np.uint64, 4: np.uint32, 2: np.uint16}[a.dtype.itemsize]
    return np.array_equal(a[~na].view(u), b[~nb].view(u))


def compare_states(prod, orc, names=STATE2D + ("Zt_avg1", "DU_avg1", "DV_avg1"), skip=()):
    """Returns list of (name, slot, n_mismatch, max_abs_diff) for non-bitwise fields."""
    bad = []
    assert prod["slots"] == orc["slots"], (prod["slots"], orc["slots"])
    for n in names:
        pa, oa = np.asarray(prod[n]), np.asarray(orc[n])
        if pa.ndim == 1:
            pa, oa = pa[None], oa[None]
        for k in range(pa.shape[0]):
            if (n, k) in skip:
                continue
            x, y = pa[k], oa[k]
            if not bitwise_equal(x, y):
                with np.errstate(invalid="ignore"):
                    d = np.abs(x - y)
                    nmis = int(np.sum((x.view(np.uint64) != y.view(np.uint64)) & ~(np.isnan(x) & np.isnan(y))))
                bad.append((n, k, nmis, float(np.nanmax(d)) if np.isfinite(d).any() else float("nan")))
    return bad


def masked_points(S2, stride):
    f = np.arange(S2)
    return (f % stride == 0) | (f // stride == 0)


def compare_fused(prod, orc, szI, szJ, stride):
    """Mismatches of the fused substep's state against the oracle on every array the path
    reads: all of zeta, vbar, rzeta, the FP64 averages, ubar t2 and rubar, and ubar t0/t1
    at the masked points (stride: the computeMomentumPred mask stride).  The fused kernel
    does not write what no kernel reads (rvbar, the dead AM3 increment of ubar t0/t1 at
    unmasked points)."""
    bad = []
    assert prod["slots"] == orc["slots"]
    sl = prod["slots"]
    masked = masked_points(szI * szJ, stride)
    for n in ("zeta", "vbar", "rzeta"):
        for k in range(3):
            if not bitwise_equal(prod[n][k], orc[n][k]):
                bad.append((n, k))
    for n in ("Zt_avg1", "DU_avg1", "DV_avg1"):
        if not bitwise_equal(prod[n], orc[n]):
            bad.append((n, -1))
    t2 = sl["ubar"][2]
    if not bitwise_equal(prod["ubar"][t2], orc["ubar"][t2]):
        bad.append(("ubar", "t2"))
    for lev in (0, 1):
        s = sl["ubar"][lev]
        if not bitwise_equal(prod["ubar"][s][masked], orc["ubar"][s][masked]):
            bad.append(("ubar", f"t{lev}@masked"))
    for lev in (0, 1):
        s = sl["rubar"][lev]
        if not bitwise_equal(prod["rubar"][s], orc["rubar"][s]):
            bad.append(("rubar", f"t{lev}"))
    return bad
