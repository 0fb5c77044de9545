"""Grid arrays and metrics (modules/mod_grid.py:88-242).

Attribute names, shapes and the flat (M+1, L+1) layout with xi fastest follow the
reference.  Setup arithmetic (metrics, depths at rest) runs once on the host in
numpy FP64; the arrays are then placed on the GPU as torch tensors.
"""
from __future__ import annotations

import numpy as np


def update_metrics(pm: np.ndarray, pn: np.ndarray, f: np.ndarray) -> dict:
    """Restatement of Grid.updateMetrics (mod_grid.py:197-242) on numpy arrays.

    Note the reference's staggering: on_u/om_u average along the first (eta) axis and
    leave row 0 at zero (:222-223); on_v/om_v average along xi and leave column 0 at
    zero (:231-232)."""
    shp = pm.shape
    out = {}
    out["om_r"] = 1.0 / pm
    out["on_r"] = 1.0 / pn
    out["omn"] = 1.0 / (pm * pn)
    out["fomn"] = f * out["omn"]
    out["pnom_r"] = pn / pm
    out["pmon_r"] = pm / pn
    out["pmon_u"] = (pm[:-1, :] + pm[1:, :]) / (pn[:-1, :] + pn[1:, :])
    out["pnom_u"] = (pn[:-1, :] + pn[1:, :]) / (pm[:-1, :] + pm[1:, :])
    om_u = np.zeros(shp); on_u = np.zeros(shp)
    om_u[1:, :] = 2.0 / (pm[:-1, :] + pm[1:, :])
    on_u[1:, :] = 2.0 / (pn[:-1, :] + pn[1:, :])
    out["om_u"], out["on_u"] = om_u, on_u
    out["pmon_v"] = (pm[:, :-1] + pm[:, 1:]) / (pn[:, :-1] + pn[:, 1:])
    out["pnom_v"] = (pn[:, :-1] + pn[:, 1:]) / (pm[:, :-1] + pm[:, 1:])
    om_v = np.zeros(shp); on_v = np.zeros(shp)
    om_v[:, 1:] = 2.0 / (pm[:, :-1] + pm[:, 1:])
    on_v[:, 1:] = 2.0 / (pn[:, :-1] + pn[:, 1:])
    out["om_v"], out["on_v"] = om_v, on_v
    s4m = pm[:-1, :-1] + pm[:-1, 1:] + pm[1:, :-1] + pm[1:, 1:]
    s4n = pn[:-1, :-1] + pn[:-1, 1:] + pn[1:, :-1] + pn[1:, 1:]
    out["pmon_p"] = s4m / s4n
    out["pnom_p"] = s4n / s4m
    out["om_p"] = 4.0 / s4m
    out["on_p"] = 4.0 / s4n
    return out


class Grid:
    """Grid of ``szI = L+1`` by ``szJ = M+1`` rho points and ``N`` s-levels.

    ``Lm``/``Mm`` are the interior counts as in mod_grid.py:93-97 (L = Lm+1, M = Mm+1).
    Host arrays (numpy) are kept in ``self.host``; device tensors carry the reference's
    attribute names (``h``, ``pm``, ``pn``, ``on_u``, ``om_v``, ``fomn`` ...)."""

    def __init__(self, Lm: int, Mm: int, N: int = 1, *, device="cuda",
                 Vtransform: int = 2, Vstretching: int = 4, theta_s: float = 5.0,
                 theta_b: float = 0.4, Tcline: float = 50.0):
        self.Lm, self.Mm = int(Lm), int(Mm)
        self.L, self.M = self.Lm + 1, self.Mm + 1
        self.shape2D = (self.M + 1, self.L + 1)
        self.N = int(N)
        self.shape3D = (self.N, self.M + 1, self.L + 1)
        self.shape3DW = (self.N + 1, self.M + 1, self.L + 1)
        self.Vtransform, self.Vstretching = int(Vtransform), int(Vstretching)
        self.theta_s, self.theta_b = float(theta_s), float(theta_b)
        self.Tcline = float(Tcline)
        self.hc = self.Tcline
        self.device = device
        self.host: dict[str, np.ndarray] = {}
        # False: computeMomentumPred masks with i%szJ as the reference does
        # (mod_cppkernels.cpp:162); True: with i%szI (column 0 and row 0 only)
        self.fix_q5 = False

    @property
    def szI(self):
        return self.L + 1

    @property
    def szJ(self):
        return self.M + 1

    def set_fields(self, *, h, pm, pn, f, **extra):
        """Set bathymetry, metrics and Coriolis (host numpy FP64 arrays of shape2D),
        compute the derived metrics and upload everything to the device."""
        import torch
        arrs = {"h": h, "pm": pm, "pn": pn, "f": f, **extra}
        for k, v in arrs.items():
            v = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
            if v.shape != self.shape2D:
                raise ValueError(f"{k}: shape {v.shape} != {self.shape2D}")
            self.host[k] = v
        self.host.update(update_metrics(self.host["pm"], self.host["pn"], self.host["f"]))
        self.hmin = float(self.host["h"].min())
        self.hmax = float(self.host["h"].max())
        for k, v in self.host.items():
            setattr(self, k, torch.from_numpy(np.ascontiguousarray(v)).to(self.device))
        return self

    def set_vertical(self, sc_r, sc_w, Cs_r, Cs_w, hc=None, dtype=None):
        """s-coordinate curves (FP64) and the depth arrays z_r, z_w, Hz in ``dtype``
        (FP64 by default; FP32 for the FP32 bulk form)."""
        import torch
        dtype = torch.float64 if dtype is None else dtype
        if hc is not None:
            self.hc = float(hc)
        for k, v in (("sc_r", sc_r), ("sc_w", sc_w), ("Cs_r", Cs_r), ("Cs_w", Cs_w)):
            v = np.asarray(v, dtype=np.float64)
            self.host[k] = v
            setattr(self, k, torch.from_numpy(v.copy()).to(self.device))
        self.z_r = torch.zeros((self.N,) + self.shape2D, dtype=dtype, device=self.device)
        self.z_w = torch.zeros((self.N + 1,) + self.shape2D, dtype=dtype, device=self.device)
        self.Hz = torch.zeros(self.shape3D, dtype=dtype, device=self.device)
        return self
