"""Synthetic basin configurations (host numpy FP64) for the benchmark workloads.

Sizes "A x B rho points" mean the full arrays szI = A (xi), szJ = B, i.e. Lm = A-2,
Mm = B-2 (mod_grid.py:93-100).  Rho points sit at x_r = (i-0.5)*dx, y_r = (j-0.5)*dy
so that the walls (u/v points 1 and L/M) are at 0 and Lx = Lm*dx.  Random fields
use numpy's PCG64 generator with the seeds named below; smoothing is a separable
Gaussian (reflect padding, truncated at 4 sigma) written here.

  basin1  closed 128x64 basin, dx=5 km, ridge bathymetry 100..1000 m, f-plane,
          tau_x = 0.1 sin(pi y/Ly), Gaussian zeta bump 0.1 m / 50 km, dt=20 s
  basin2  closed 2048x1024 basin, dx=2 km, beta plane f=1e-4+2e-11 y, bathymetry
          from white noise (seed 2) smoothed sigma=20 cells mapped to 200..4000 m,
          double-gyre tau_x = 0.1 cos(2 pi y/Ly), zeta noise N(0, 1 cm) (seed 3)
          smoothed sigma=5 cells, dt=5 s
  ensemble_configs
          members of the basin1 recipe (dx=5 km, dt=20 s) with seeded sizes from
          {256x128, 384x256, 512x256, 768x384, 1024x512} (seed 4), wind amplitude
          U[0.05, 0.15] N/m^2 (seed 5) and their own free-surface noise (seed 1000+k,
          1 cm smoothed sigma=5 cells) on top of the 0.1 m Gaussian bump
"""
from __future__ import annotations

import numpy as np

RHO0 = 1025.0


def gaussian_smooth(a: np.ndarray, sigma: float) -> np.ndarray:
    """Separable Gaussian filter, reflect padding, truncated at 4 sigma."""
    if sigma <= 0:
        return a.copy()
    r = int(4.0 * sigma + 0.5)
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    k /= k.sum()
    out = a.astype(np.float64)
    for axis in range(out.ndim):
        pad = [(0, 0)] * out.ndim
        pad[axis] = (r, r)
        p = np.pad(out, pad, mode="symmetric")
        p = np.moveaxis(p, axis, -1)
        n = out.shape[axis]
        acc = np.zeros(p.shape[:-1] + (n,))
        for t in range(2 * r + 1):
            acc += k[t] * p[..., t:t + n]
        out = np.moveaxis(acc, -1, axis)
    return np.ascontiguousarray(out)


def _coords(szI, szJ, dx, dy):
    ii = np.arange(szI, dtype=np.float64)
    jj = np.arange(szJ, dtype=np.float64)
    x = (ii - 0.5) * dx
    y = (jj - 0.5) * dy
    X, Y = np.meshgrid(x, y)
    Lx = (szI - 2) * dx
    Ly = (szJ - 2) * dy
    return X, Y, Lx, Ly


def basin1(szI: int = 128, szJ: int = 64, dx: float = 5e3) -> dict:
    X, Y, Lx, Ly = _coords(szI, szJ, dx, dx)
    h = 100.0 + 900.0 * (1.0 - np.exp(-((X - Lx / 2) ** 2) / (2.0 * (Lx / 6.0) ** 2)))
    f = np.full_like(h, 1e-4)
    r2 = (X - Lx / 2) ** 2 + (Y - Ly / 2) ** 2
    zeta0 = 0.1 * np.exp(-r2 / (2.0 * 50e3 ** 2))
    return dict(name="basin1", szI=szI, szJ=szJ, dx=dx, dy=dx, h=h, f=f,
                pm=np.full_like(h, 1.0 / dx), pn=np.full_like(h, 1.0 / dx),
                sustr=0.1 * np.sin(np.pi * Y / Ly) / RHO0, svstr=np.zeros_like(h),
                zeta0=zeta0, dtfast=20.0, nsteps=2000, rdrg=3e-4)


def basin2(szI: int = 2048, szJ: int = 1024, dx: float = 2e3, seed_h: int = 2,
           seed_z: int = 3) -> dict:
    X, Y, Lx, Ly = _coords(szI, szJ, dx, dx)
    rng = np.random.Generator(np.random.PCG64(seed_h))
    noise = gaussian_smooth(rng.standard_normal((szJ, szI)), min(20.0, max(1.0, szI / 16)))
    lo, hi = noise.min(), noise.max()
    h = 200.0 + (noise - lo) / (hi - lo) * (4000.0 - 200.0)
    f = 1e-4 + 2e-11 * Y
    rng = np.random.Generator(np.random.PCG64(seed_z))
    zeta0 = gaussian_smooth(0.01 * rng.standard_normal((szJ, szI)), min(5.0, max(1.0, szI / 64)))
    return dict(name="basin2", szI=szI, szJ=szJ, dx=dx, dy=dx, h=h, f=f,
                pm=np.full_like(h, 1.0 / dx), pn=np.full_like(h, 1.0 / dx),
                sustr=0.1 * np.cos(2 * np.pi * Y / Ly) / RHO0, svstr=np.zeros_like(h),
                zeta0=zeta0, dtfast=5.0, nsteps=1000, rdrg=3e-4)


def basin2_analytic(szI: int = 4096, szJ: int = 4096, dx: float = 2e3) -> dict:
    """The basin2 recipe (beta plane, cos wind, bathymetry in [200, 4000] m, small
    free-surface bumps) with analytic fields instead of smoothed noise, so grids far
    larger than the 256 MB last-level cache set up in seconds (throughput runs)."""
    X, Y, Lx, Ly = _coords(szI, szJ, dx, dx)
    s = np.sin(6 * np.pi * X / Lx) * np.cos(4 * np.pi * Y / Ly)
    h = 200.0 + (0.5 + 0.5 * s) * (4000.0 - 200.0)
    f = 1e-4 + 2e-11 * Y
    zeta0 = 0.01 * np.sin(10 * np.pi * X / Lx) * np.sin(8 * np.pi * Y / Ly)
    return dict(name="basin2_analytic", szI=szI, szJ=szJ, dx=dx, dy=dx, h=h, f=f,
                pm=np.full_like(h, 1.0 / dx), pn=np.full_like(h, 1.0 / dx),
                sustr=0.1 * np.cos(2 * np.pi * Y / Ly) / RHO0, svstr=np.zeros_like(h),
                zeta0=zeta0, dtfast=5.0, nsteps=1000, rdrg=3e-4)


def random_smooth(szI: int, szJ: int, seed: int = 0) -> dict:
    """Small non-square test basin with smooth random bathymetry, metrics that vary in
    space and a random free-surface and velocity state (for parity tests)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sm = lambda s: gaussian_smooth(rng.standard_normal((szJ, szI)), s)
    dx = 1e3 * (1.0 + 0.1 * sm(2.0))
    dy = 1e3 * (1.0 + 0.1 * sm(2.0))
    h = 50.0 + 40.0 * np.abs(sm(1.5))
    return dict(name=f"random{szI}x{szJ}", szI=szI, szJ=szJ, h=h,
                pm=1.0 / dx, pn=1.0 / dy, f=1e-4 + 1e-5 * sm(3.0),
                zeta0=0.05 * sm(1.0), ubar0=0.1 * sm(1.0), vbar0=0.1 * sm(1.0),
                sustr=1e-4 * sm(4.0), svstr=1e-4 * sm(4.0), dtfast=2.0, nsteps=10, rdrg=3e-4)


ENSEMBLE_SIZES = ((256, 128), (384, 256), (512, 256), (768, 384), (1024, 512))


def ensemble_member(szI: int, szJ: int, amp: float, seed: int, dx: float = 5e3) -> dict:
    """One ensemble member: the basin1 recipe at any size with wind amplitude ``amp``
    and its own seeded free-surface noise added to the Gaussian bump."""
    cfg = basin1(szI, szJ, dx=dx)
    _X, Y, _Lx, Ly = _coords(szI, szJ, dx, dx)
    cfg["sustr"] = amp * np.sin(np.pi * Y / Ly) / RHO0
    rng = np.random.Generator(np.random.PCG64(seed))
    cfg["zeta0"] = cfg["zeta0"] + gaussian_smooth(0.01 * rng.standard_normal((szJ, szI)), 5.0)
    cfg["name"] = f"member{seed}_{szI}x{szJ}"
    return cfg


def ensemble_configs(n: int = 64, sizes=ENSEMBLE_SIZES, seed_size: int = 4, seed_amp: int = 5) -> list:
    """The 64-member ensemble workload (uneven sizes, per-member wind and noise)."""
    rs = np.random.Generator(np.random.PCG64(seed_size))
    ra = np.random.Generator(np.random.PCG64(seed_amp))
    pick = rs.integers(0, len(sizes), size=n)
    amps = ra.uniform(0.05, 0.15, size=n)
    return [ensemble_member(sizes[p][0], sizes[p][1], float(a), 1000 + k) for k, (p, a) in enumerate(zip(pick, amps))]


def basin3(szI: int = 1024, szJ: int = 512, N: int = 32, dx: float = 5e3) -> dict:
    """The 3D mode-split workload: the basin1 recipe (depth, f-plane, wind, Gaussian
    free-surface bump, fluid at rest) on 1024 x 512 rho points with N = 32 s-levels
    (Vtransform 2, Vstretching 4, theta_s 5, theta_b 0.4, hc 50 m), 40 barotropic
    substeps per baroclinic step of 300 s (dtfast 7.5 s).  The temperature profile
    T = 4 + 16 exp(z/500) with S = 35 and a linear equation of state only matters with
    a baroclinic pressure gradient (see the ``baroclinic`` module)."""
    cfg = basin1(szI, szJ, dx=dx)
    cfg.update(name=f"basin3_{szI}x{szJ}x{N}", N=N, ndtfast=40, dt=300.0, dtfast=7.5, nsteps=200,
               theta_s=5.0, theta_b=0.4, hc=50.0)
    return cfg
