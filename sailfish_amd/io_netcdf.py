"""ROMS-style NetCDF history output and restart files (checkpoint/resume).

History: the file layout of nonlinear/adhocNetcdfOutput.py:3-158 (dimensions
xi_rho/xi_u/xi_v, eta_rho/eta_u/eta_v, s_rho, s_w, N and the unlimited ocean_time;
grid variables, s-coordinate curves, and zeta/ubar/vbar/u/v records in FP32 with the
same attributes), written every 30 model minutes with the record slicing of
nonlinear/main3d.py:202-216.  The reference writes NETCDF4 through the netCDF4
package; this image has no HDF5/netCDF4 library, so the file is NetCDF-3 with 64-bit
offsets (``ncfile.NCWriter``), which every netCDF reader opens with the same
dimension, variable and attribute names.  Each record is complete in the file as soon
as it is written (the record count in the header is updated per record).

Restart: every array of the split-explicit state in its own precision (the three time
slots of each rotated field, the right-hand-side histories, the FP64 barotropic
averages, the 3D diagnostics W/Huon/Hvom, AKv and the depths z_w/z_r/Hz), the slot
each time-level view points at, and the CompTimes counters.  Reading a restart into a
model built from the same configuration puts it in exactly the state it had when the
file was written, so a resumed run continues bit for bit.

All device data passes through host memory: the files are written from CPU copies
(one field at a time) and read back into the existing device tensors in place.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from .mod_ocean import _ROT2D, _ROT3D
from .ncfile import NCReader, NCWriter

HISTORY_INTERVAL = 30 * 60.0     # nonlinear/main3d.py:202 (every 30 model minutes)
TIME_UNITS = "seconds since 2001-01-01 00:00:00"


def _host(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu").numpy()


def grid_coordinates(GRID) -> dict:
    """x/y of rho, u and v points from the metrics: rho point i at (i - 1/2) dx from
    the west wall (the u point of column 1), u point i on the face between rho i-1 and
    i, v point j on the face between rho j-1 and j (the flat C-grid of the kernels)."""
    host = GRID.host
    eta_rho, xi_rho = GRID.M + 1, GRID.L + 1
    dx, dy = 1.0 / host["pm"], 1.0 / host["pn"]
    x_face = np.concatenate([np.zeros((eta_rho, 1)), np.cumsum(dx, axis=1)[:, :-1]], axis=1) - dx[:, :1]
    y_face = np.concatenate([np.zeros((1, xi_rho)), np.cumsum(dy, axis=0)[:-1, :]], axis=0) - dy[:1, :]
    x_rho, y_rho = x_face + 0.5 * dx, y_face + 0.5 * dy
    return {"x_rho": x_rho, "y_rho": y_rho, "x_u": x_face[:, :-1], "y_u": y_rho[:, :-1],
            "x_v": x_rho[:-1, :], "y_v": y_face[:-1, :],
            "xl": float(np.sum(dx[1, 1:-1])), "el": float(np.sum(dy[1:-1, 1]))}


def create_roms_history(path, GRID) -> NCWriter:
    """Create the history file (nonlinear/adhocNetcdfOutput.py:3-158) for a grid of
    xi_rho = L+1, eta_rho = M+1 and N levels, and write its grid variables.  Returns
    the open writer."""
    xi_rho, eta_rho, s_rho = GRID.L + 1, GRID.M + 1, GRID.N
    nc = NCWriter(path)
    nc.dim("ocean_time", None)
    for name, size in (("xi_rho", xi_rho), ("xi_u", xi_rho - 1), ("xi_v", xi_rho), ("eta_rho", eta_rho),
                       ("eta_u", eta_rho), ("eta_v", eta_rho - 1), ("s_rho", s_rho), ("s_w", s_rho + 1),
                       ("N", s_rho)):
        nc.dim(name, size)
    nc.var("ntimes", "i4", (), long_name="number of long time-steps")
    nc.var("xl", "f8", (), long_name="domain length in the XI-direction", units="meter")
    nc.var("el", "f8", (), long_name="domain length in the ETA-direction", units="meter")
    for name, dim, where, cs in (("s_rho", "s_rho", "RHO", "Cs_r"), ("s_w", "s_w", "W", "Cs_w")):
        nc.var(name, "f8", (dim,), long_name=f"S-coordinate at {where}-points", valid_min=-1.0, valid_max=0.0,
               positive="up", standard_name="ocean_s_coordinate_g2",
               formula_terms=f"s: {name} C: {cs} eta: zeta depth: h depth_c: hc", field=f"{name}, scalar")
    for name, dim, where in (("Cs_r", "s_rho", "RHO"), ("Cs_w", "s_w", "W")):
        nc.var(name, "f8", (dim,), long_name=f"S-coordinate stretching curves at {where}-points",
               valid_min=-1.0, valid_max=0.0, field=f"{name}, scalar")
    nc.var("h", "f8", ("eta_rho", "xi_rho"), long_name="bathymetry at RHO-points", units="meter", grid="grid",
           location="face", coordinates="x_rho y_rho", field="bath, scalar")
    for pt, dims in (("rho", ("eta_rho", "xi_rho")), ("u", ("eta_u", "xi_u")), ("v", ("eta_v", "xi_v"))):
        for ax in ("x", "y"):
            nc.var(f"{ax}_{pt}", "f8", dims, long_name=f"{ax}-locations of {pt.upper()}-points", units="meter",
                   field=f"{ax}_{pt}, scalar")
    nc.var("ocean_time", "f8", ("ocean_time",), long_name="time since initialization", units=TIME_UNITS,
           calendar="proleptic_gregorian", field="time, scalar, series")
    nc.var("zeta", "f4", ("ocean_time", "eta_rho", "xi_rho"), long_name="free-surface", units="meter",
           time="ocean_time", grid="grid", location="face", coordinates="x_rho y_rho ocean_time",
           field="free-surface, scalar, series")
    nc.var("ubar", "f4", ("ocean_time", "eta_u", "xi_u"), long_name="vertically integrated u-momentum component",
           units="meter second-1", time="ocean_time", grid="grid", location="edge1",
           coordinates="x_u y_u ocean_time", field="ubar-velocity, scalar, series")
    nc.var("vbar", "f4", ("ocean_time", "eta_v", "xi_v"), long_name="vertically integrated v-momentum component",
           units="meter second-1", time="ocean_time", grid="grid", location="edge2",
           coordinates="x_v y_v ocean_time", field="vbar-velocity, scalar, series")
    nc.var("u", "f4", ("ocean_time", "s_rho", "eta_u", "xi_u"), long_name="u-momentum component",
           units="meter second-1", time="ocean_time", grid="grid", location="edge1",
           coordinates="x_u y_u s_rho ocean_time", field="u-velocity, scalar, series")
    nc.var("v", "f4", ("ocean_time", "s_rho", "eta_v", "xi_v"), long_name="v-momentum component",
           units="meter second-1", time="ocean_time", grid="grid", location="edge2",
           coordinates="x_v y_v s_rho ocean_time", field="v-velocity, scalar, series")
    nc.begin()
    xy = grid_coordinates(GRID)
    for k in ("x_rho", "y_rho", "x_u", "y_u", "x_v", "y_v", "xl", "el"):
        nc.put(k, xy[k])
    nc.put("h", GRID.host["h"])
    nc.put("ntimes", 0)
    host = GRID.host
    if "Cs_r" in host:
        nc.put("Cs_r", host["Cs_r"])
        nc.put("Cs_w", host["Cs_w"])
        nc.put("s_rho", host["sc_r"])
        nc.put("s_w", host["sc_w"])
    return nc


class HistoryWriter:
    """History records every ``interval`` model seconds (nonlinear/main3d.py:202-216):
    the first call writes, then one record each time the model time reaches the next
    multiple of the interval after the first.  zeta at t2 on all rho points; ubar and u
    without their last xi column, vbar and v without their last eta row, u and v
    without the extra top level of the (N+1)-level arrays (the reference's slicing)."""

    def __init__(self, path, GRID, interval: float = HISTORY_INTERVAL):
        self.path = Path(path)
        self.GRID = GRID
        self.interval = float(interval)
        self.nc = create_roms_history(self.path, GRID)
        self.previous = None
        self.records = 0

    def due(self, time: float) -> bool:
        return self.previous is None or time >= self.previous + self.interval - 1e-4

    def maybe_write(self, compTimes, OCEAN) -> bool:
        if not self.due(compTimes.time):
            return False
        self.previous = (0.0 if self.previous is None else self.previous) + self.interval
        self.write(compTimes.time, OCEAN)
        return True

    def write(self, time: float, OCEAN):
        G = self.GRID
        N, M, L = G.N, G.M, G.L
        k = self.records
        nc = self.nc
        nc.put_record("zeta", k, _host(OCEAN.zeta_t2).reshape(M + 1, L + 1))
        nc.put_record("ubar", k, _host(OCEAN.ubar_t2).reshape(M + 1, L + 1)[:, :-1])
        nc.put_record("vbar", k, _host(OCEAN.vbar_t2).reshape(M + 1, L + 1)[:-1, :])
        if getattr(OCEAN, "with3d", False):
            nc.put_record("u", k, _host(OCEAN.u_t2).reshape(N + 1, M + 1, L + 1)[:-1, :, :-1])
            nc.put_record("v", k, _host(OCEAN.v_t2).reshape(N + 1, M + 1, L + 1)[:-1, :-1, :])
        else:
            nc.put_record("u", k, np.zeros((N, M + 1, L)))
            nc.put_record("v", k, np.zeros((N, M, L + 1)))
        # ocean_time last: the record count the header holds covers a whole record
        nc.put_record("ocean_time", k, float(time))
        nc.put("ntimes", k + 1)
        nc.flush()
        self.records += 1

    def close(self):
        self.nc.close()


# ---------------------------------------------------------------------------------- restart

_COUNTERS_F = ("dt", "dtfast", "time", "time2D", "tdays")
_COUNTERS_I = ("iic", "iif", "nfast", "ntfirst", "ntstart", "ntend", "ndtfast", "indx1", "ntimes",
               "is2DPredictorStep", "is2DCorrectorStep")
_PLAIN_2D = ("Zt_avg1", "DU_avg1", "DV_avg1", "U_avg1", "V_avg1")
_PLAIN_3D = ("W", "Huon", "Hvom", "AKv")
_GRID_3D = ("z_w", "z_r", "Hz")
_NC_TYPE = {torch.float64: "f8", torch.float32: "f4"}


def _state_arrays(GRID, OCEAN):
    """(name, tensor) of every array a restart holds: the state of T_OCEAN, the depths
    of the grid and any further arrays a model registers in ``OCEAN.extra_state``."""
    out = [(n, getattr(OCEAN, n)) for n in _ROT2D]
    out += [(n, getattr(OCEAN, n)) for n in _PLAIN_2D if hasattr(OCEAN, n)]
    if getattr(OCEAN, "with3d", False):
        out += [(n, getattr(OCEAN, n)) for n in _ROT3D]
        out += [(n, getattr(OCEAN, n)) for n in _PLAIN_3D]
        out += [(n, getattr(GRID, n)) for n in _GRID_3D if hasattr(GRID, n)]
    out += sorted(getattr(OCEAN, "extra_state", {}).items())
    return out


def write_restart(path, compTimes, GRID, OCEAN):
    """Write every state array, the slot of each time-level view and the time counters."""
    with3d = int(bool(getattr(OCEAN, "with3d", False)))
    nc = NCWriter(path)
    nc.gatts.update(title="sailfish_amd restart", szI=GRID.szI, szJ=GRID.szJ, N=GRID.N, with3d=with3d)
    for name in _COUNTERS_F:
        nc.gatts[f"ct_{name}"] = float(getattr(compTimes, name))
    for name in _COUNTERS_I:
        nc.gatts[f"ct_{name}"] = int(getattr(compTimes, name))
    nc.gatts["ct_weight1"] = np.asarray(compTimes.weight1, dtype=np.float64)
    nc.gatts["ct_weight2"] = np.asarray(compTimes.weight2, dtype=np.float64)
    # the slot each _t0/_t1/_t2 view aliases (cycleTimes2D/3D rotate the views)
    names = list(_ROT2D) + (list(_ROT3D) if with3d else [])
    slots = np.array([[OCEAN.slot_of(n, f"t{k}") for k in range(3)] for n in names], dtype=np.int32)
    nc.gatts["rotated_names"] = ",".join(names)
    nc.dim("rotated", len(names))
    nc.dim("level", 3)
    nc.var("slots", "i4", ("rotated", "level"))
    arrays = _state_arrays(GRID, OCEAN)
    for name, t in arrays:
        dim = f"n_{t.numel()}"
        if all(d != dim for d, _ in nc.dims):
            nc.dim(dim, t.numel())
        nc.var(name, _NC_TYPE[t.dtype], (dim,))
    nc.begin()
    nc.put("slots", slots)
    for name, t in arrays:
        nc.put(name, _host(t.reshape(-1)))
    nc.close()


def read_restart(path, compTimes, GRID, OCEAN):
    """Restore a restart into a model built from the same configuration (same grid,
    precision and 3D flag): device arrays are overwritten in place, the time-level views
    re-pointed to the recorded slots and the counters restored."""
    nc = NCReader(path)
    a = nc.attrs
    shape = tuple(int(a[k][0]) for k in ("szI", "szJ", "N"))
    if shape != (GRID.szI, GRID.szJ, GRID.N):
        raise ValueError(f"restart grid {shape} != model {(GRID.szI, GRID.szJ, GRID.N)}")
    if int(a["with3d"][0]) != int(bool(getattr(OCEAN, "with3d", False))):
        raise ValueError("restart and model differ in the 3D state")
    for name, t in _state_arrays(GRID, OCEAN):
        if name not in nc.vars:
            raise ValueError(f"restart has no {name}")
        data = nc[name]
        if data.size != t.numel() or data.dtype.itemsize != t.element_size():
            raise ValueError(f"restart field {name}: {data.dtype}[{data.size}] does not match the model")
        t.reshape(-1).copy_(torch.from_numpy(np.ascontiguousarray(data)))
    slots = nc["slots"]
    for name, row in zip(a["rotated_names"].split(","), slots):
        base = getattr(OCEAN, name)
        for k in range(3):
            setattr(OCEAN, f"{name}_t{k}", base[int(row[k])].reshape(-1))
    for name in _COUNTERS_F:
        setattr(compTimes, name, float(a[f"ct_{name}"][0]))
    for name in _COUNTERS_I:
        v = int(a[f"ct_{name}"][0])
        setattr(compTimes, name, bool(v) if isinstance(getattr(compTimes, name), bool) else v)
    compTimes.weight1 = np.array(a["ct_weight1"], dtype=np.float64)
    compTimes.weight2 = np.array(a["ct_weight2"], dtype=np.float64)
