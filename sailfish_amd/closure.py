"""Learned subgrid momentum closure for the coarse physics-mode model.

Pipeline (all state on the GPU; the network runs on the project's BF16-MFMA kernels
in csrc/mlp.hip):

1. fine runs of the basin1 recipe at 4x the coarse resolution (``fine_config``);
2. ``coarse_grain``: 4x4 block means of zeta, face means of ubar/vbar and h, then
   the closed-basin boundary rules of the coarse grid;
3. samples (``make_samples``): features = 3x3 stencils of (ubar, vbar, zeta, h, f) of
   the coarse-grained state, as centre value and neighbour - centre differences
   (sf_closure_features); targets = subgrid momentum tendency CG(T_fine(B)) -
   T_coarse(CG(B)), each instantaneous tendency measured by a tiny forward/backward
   Euler probe substep;
4. training with Adam on normalised features/targets (sf_mlp_train_step);
5. coarse runs with the closure: every substep evaluates the network inside the substep
   kernel (sf_phys_substep_closure_fused) and adds the momentum stress R * D to the wind
   stress of the corrector stage; the plain coarse scheme is the same run without it.
   sf_closure_stress computes the same stresses in a separate launch (the tests compare
   the two paths).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib, configs
from .model2d import build_model
from .phys2d import PhysStepper, sf_phys_fields

P = C.c_void_p
NIN, IN, NOUT = 45, 48, 2


def _bind(lib):
    if getattr(lib, "_sf_mlp_bound", False):
        return lib
    f = C.c_float
    lib.sf_mlp_num_params.restype = C.c_int32
    sig = {
        "sf_mlp_create": [C.c_int32, C.POINTER(P)],
        "sf_mlp_destroy": [P],
        "sf_mlp_set_params": [P, C.POINTER(f)],
        "sf_mlp_get_params": [P, C.POINTER(f)],
        "sf_mlp_get_grad": [P, C.POINTER(f), C.POINTER(f)],
        "sf_mlp_set_norm": [P, C.POINTER(f), C.POINTER(f), C.POINTER(f)],
        "sf_mlp_train_step": [P, P, P, P, C.c_int32, C.c_int32, C.c_int32, f, f, f, f, C.c_int32, P],
        "sf_mlp_forward": [P, P, C.c_int32, P, P],
        "sf_closure_features": [C.POINTER(sf_phys_fields), C.c_int32, P, P, P],
        "sf_closure_stress": [P, C.POINTER(sf_phys_fields), C.c_int32, P, P, P, P],
        "sf_phys_substep_closure_fused": [C.POINTER(sf_phys_fields), P, P, C.c_int32, C.c_double, C.c_int32,
                                          C.c_double, C.c_int32, P],
    }
    for name, args in sig.items():
        getattr(lib, name).argtypes = args
        getattr(lib, name).restype = C.c_int32
    lib._sf_mlp_bound = True
    return lib


def to_bf16_bits(x: torch.Tensor) -> torch.Tensor:
    """FP32 -> BF16 (round to nearest even) as int16 bit patterns."""
    return x.to(torch.float32).to(torch.bfloat16).view(torch.int16).contiguous()


def _fp(a: np.ndarray):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


class ClosureMLP:
    """Device network: FP32 master parameters, Adam moments, BF16 copy, normalisation."""

    def __init__(self, max_blocks: int = 1024, params: np.ndarray | None = None, seed: int = 7):
        _lib.require_gpu()
        self.lib = _bind(_lib.load())
        self.nparam = int(self.lib.sf_mlp_num_params())
        h = P()
        _lib.check(self.lib.sf_mlp_create(int(max_blocks), C.byref(h)), "sf_mlp_create")
        self.handle = h
        self.max_blocks = int(max_blocks)
        if params is None:
            params = init_params(seed)
        self.set_params(params)
        self.mu = np.zeros(NIN, np.float32)
        self.sd = np.ones(NIN, np.float32)
        self.ynorm = np.array([0, 1, 0, 1], np.float32)
        self.step_count = 0

    def close(self):
        if self.handle:
            self.lib.sf_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: np.ndarray):
        a, p = _fp(params)
        assert a.size == self.nparam
        _lib.check(self.lib.sf_mlp_set_params(self.handle, p), "sf_mlp_set_params")
        self.step_count = 0

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.nparam, np.float32)
        _lib.check(self.lib.sf_mlp_get_params(self.handle, out.ctypes.data_as(C.POINTER(C.c_float))),
                   "sf_mlp_get_params")
        return out

    def get_grad(self):
        g = np.zeros(self.nparam, np.float32)
        loss = np.zeros(1, np.float32)
        _lib.check(self.lib.sf_mlp_get_grad(self.handle, g.ctypes.data_as(C.POINTER(C.c_float)),
                                            loss.ctypes.data_as(C.POINTER(C.c_float))), "sf_mlp_get_grad")
        return g, float(loss[0])

    def set_norm(self, mu, sd, ynorm):
        self.mu, self.sd, self.ynorm = (np.asarray(mu, np.float32), np.asarray(sd, np.float32),
                                        np.asarray(ynorm, np.float32))
        a, pa = _fp(self.mu)
        b, pb = _fp(self.sd)
        c, pc = _fp(self.ynorm)
        _lib.check(self.lib.sf_mlp_set_norm(self.handle, pa, pb, pc), "sf_mlp_set_norm")

    def train_step(self, Xb: torch.Tensor, Y: torch.Tensor, idx: torch.Tensor, adam: bool = True, lr: float = 1e-3,
                   beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, ntile: int | None = None,
                   stream=None):
        """One step on the samples idx (a multiple of 64 * ntile).  ntile None: the number of
        64-sample tiles per workgroup that keeps the grid at <= 256 workgroups (one per CU),
        so large batches run as few, long workgroups and their gradient partials stay few."""
        n = idx.numel()
        if ntile is None:
            ntile = default_ntile(n)
        assert n % (64 * ntile) == 0
        nblocks = n // (64 * ntile)
        if adam:
            self.step_count += 1
        rc = self.lib.sf_mlp_train_step(self.handle, Xb.data_ptr(), Y.data_ptr(), idx.data_ptr(), nblocks, ntile,
                                        int(adam), lr, beta1, beta2, eps, max(self.step_count, 1),
                                        _lib.stream_handle(stream))
        _lib.check(rc, "sf_mlp_train_step")

    def forward(self, Xb: torch.Tensor) -> torch.Tensor:
        n = Xb.shape[0]
        Y = torch.empty((n, NOUT), dtype=torch.float32, device=Xb.device)
        _lib.check(self.lib.sf_mlp_forward(self.handle, Xb.data_ptr(), n, Y.data_ptr(), _lib.stream_handle()),
                   "sf_mlp_forward")
        return Y


def init_params(seed: int = 7) -> np.ndarray:
    """He-style initial parameters in the kernel layout (padding zero)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W1 = np.zeros((IN, 64))
    W1[:NIN] = rng.standard_normal((NIN, 64)) * math.sqrt(2.0 / NIN)
    W2 = rng.standard_normal((64, 64)) * math.sqrt(2.0 / 64)
    W3 = np.zeros((64, 16))
    W3[:, :NOUT] = rng.standard_normal((64, NOUT)) * math.sqrt(1.0 / 64)
    return np.concatenate([W1.ravel(), np.zeros(64), W2.ravel(), np.zeros(64), W3.ravel(), np.zeros(16)])


# ------------------------------------------------------------------------------------
# grids and coarse-graining
# ------------------------------------------------------------------------------------
def fine_config(coarse_szI: int = 512, coarse_szJ: int = 256, ratio: int = 4, dx_coarse: float = 1250.0,
                amp: float = 0.1, bump: float = 0.1, center=(0.5, 0.5), dt_coarse: float = 5.0) -> dict:
    """The basin1 recipe at ``ratio`` times the coarse resolution: the fine interior is
    exactly ``ratio`` x the coarse interior (szI_f = ratio*(szI_c-2) + 2)."""
    szI = ratio * (coarse_szI - 2) + 2
    szJ = ratio * (coarse_szJ - 2) + 2
    dx = dx_coarse / ratio
    cfg = configs.basin1(szI, szJ, dx=dx)
    X, Y, Lx, Ly = configs._coords(szI, szJ, dx, dx)
    cfg["sustr"] = amp * np.sin(np.pi * Y / Ly) / configs.RHO0
    r2 = (X - center[0] * Lx) ** 2 + (Y - center[1] * Ly) ** 2
    cfg["zeta0"] = bump * np.exp(-r2 / (2.0 * 50e3 ** 2))
    cfg["dtfast"] = dt_coarse / ratio
    cfg["ratio"] = ratio
    return cfg


def _closed_bc(z, u, v):
    """Closed-basin boundary values (free-slip ghosts, zero normal flow), in place,
    in the order of the physics oracle's boundary routine."""
    z[1:-1, 0] = z[1:-1, 1]
    z[1:-1, -1] = z[1:-1, -2]
    z[0, :] = z[1, :]
    z[-1, :] = z[-2, :]
    u[:, 0] = 0.0
    u[:, 1] = 0.0
    u[:, -1] = 0.0
    u[0, 2:-1] = u[1, 2:-1]
    u[-1, 2:-1] = u[-2, 2:-1]
    v[0, :] = 0.0
    v[1, :] = 0.0
    v[-1, :] = 0.0
    v[2:-1, 0] = v[2:-1, 1]
    v[2:-1, -1] = v[2:-1, -2]


def coarse_grain(z, u, v, r: int = 4):
    """Fine (szJ_f, szI_f) fields -> coarse ((szJ_f-2)/r + 2, (szI_f-2)/r + 2):
    zeta by r x r block means of the interior cells, ubar by means over the r fine u
    points of each coarse west face, vbar over each coarse south face."""
    Jf, If = z.shape
    Jc, Ic = (Jf - 2) // r + 2, (If - 2) // r + 2
    zc = torch.zeros((Jc, Ic), dtype=z.dtype, device=z.device)
    uc = torch.zeros_like(zc)
    vc = torch.zeros_like(zc)
    zi = z[1:-1, 1:-1].reshape(Jc - 2, r, Ic - 2, r)
    zc[1:-1, 1:-1] = zi.mean(dim=(1, 3))
    # u faces: fine column 1 + r*(I-1), rows of the coarse cell; coarse columns 1 .. Ic-1
    uf = u[1:-1, 1::r][:, :Ic - 1]                       # (Jf-2, Ic-1)
    uc[1:-1, 1:] = uf.reshape(Jc - 2, r, Ic - 1).mean(dim=1)
    vf = v[1::r, 1:-1][:Jc - 1, :]                       # (Jc-1, If-2)
    vc[1:, 1:-1] = vf.reshape(Jc - 1, Ic - 2, r).mean(dim=2)
    _closed_bc(zc, uc, vc)
    return zc, uc, vc


def coarse_config(fcfg: dict) -> dict:
    """Coarse grid of a fine configuration: block-mean bathymetry, face-mean wind."""
    r = fcfg["ratio"]
    hf = torch.from_numpy(fcfg["h"])
    Jf, If = hf.shape
    Jc, Ic = (Jf - 2) // r + 2, (If - 2) // r + 2
    hc = torch.zeros((Jc, Ic), dtype=torch.float64)
    hc[1:-1, 1:-1] = hf[1:-1, 1:-1].reshape(Jc - 2, r, Ic - 2, r).mean(dim=(1, 3))
    hc[1:-1, 0] = hc[1:-1, 1]
    hc[1:-1, -1] = hc[1:-1, -2]
    hc[0, :] = hc[1, :]
    hc[-1, :] = hc[-2, :]
    sf = torch.from_numpy(fcfg["sustr"])
    sc = torch.zeros_like(hc)
    sc[1:-1, 1:] = sf[1:-1, 1::r][:, :Ic - 1].reshape(Jc - 2, r, Ic - 1).mean(dim=1)
    sc[0, :] = sc[1, :]
    sc[-1, :] = sc[-2, :]
    dx = fcfg["dx"] * r
    cfg = dict(name="coarse", szI=Ic, szJ=Jc, dx=dx, dy=dx, h=hc.numpy(), f=np.full((Jc, Ic), float(fcfg["f"].flat[0])),
               pm=np.full((Jc, Ic), 1.0 / dx), pn=np.full((Jc, Ic), 1.0 / dx), sustr=sc.numpy(),
               svstr=np.zeros((Jc, Ic)), dtfast=fcfg["dtfast"] * r, rdrg=fcfg["rdrg"])
    cfg["zeta0"] = np.zeros((Jc, Ic))
    return cfg


# ------------------------------------------------------------------------------------
# models
# ------------------------------------------------------------------------------------
class Model:
    """A physics-mode basin: CompTimes, GRID, OCEAN, fused stepper, Coriolis field."""

    def __init__(self, cfg: dict, dtype=torch.float64):
        self.cfg = cfg
        self.ct, self.GRID, self.OCEAN, _B = build_model(cfg, dtype=dtype, ndtfast=30)
        self.stepper = PhysStepper(self.GRID, self.OCEAN, rdrg=cfg.get("rdrg", 3e-4), fused=True)
        self.dtype = dtype
        self.shape = (self.GRID.szJ, self.GRID.szI)
        self.f = torch.from_numpy(np.ascontiguousarray(cfg["f"], dtype=np.float64).ravel()).to("cuda", dtype)
        self.n = 0

    def set_state(self, z, u, v):
        """All three time slots := (z, u, v); rhs history and averages cleared; the next
        substep is an Euler start."""
        for k in range(3):
            self.OCEAN.zeta[k].copy_(z.reshape(self.shape))
            self.OCEAN.ubar[k].copy_(u.reshape(self.shape))
            self.OCEAN.vbar[k].copy_(v.reshape(self.shape))
        for name in ("rzeta", "rubar", "rvbar"):
            getattr(self.OCEAN, name).zero_()
        self.n = 0

    def state(self):
        O = self.OCEAN
        return (O.zeta_t2.reshape(self.shape), O.ubar_t2.reshape(self.shape), O.vbar_t2.reshape(self.shape))

    def step(self, closure: "ClosureMLP | None" = None, su=None, sv=None):
        self.begin()
        self.advance(closure, su, sv)

    def begin(self):
        """Start a substep: time counter and slot rotation (t1 is now the current state)."""
        self.ct.next2DTimeStep()
        self.OCEAN.cycleTimes2D()

    def force(self, prev, cur):
        """Overwrite the levels t0 (n-1) and t1 (n) with given states (z, u, v)."""
        O = self.OCEAN
        for lev, st in ((0, prev), (1, cur)):
            for name, x in zip(("zeta", "ubar", "vbar"), st):
                getattr(O, f"{name}_t{lev}").copy_(x.reshape(-1))

    def advance(self, closure: "ClosureMLP | None" = None, su=None, sv=None):
        """One substep; with ``closure`` the network runs inside the substep kernel and its
        stress enters the corrector stage (su, sv are not used: kept for callers that pass
        the buffers of closure_buffers)."""
        st = self.stepper
        a = st.fields(self.OCEAN)
        if closure is not None:
            rc = closure.lib.sf_phys_substep_closure_fused(
                C.byref(a), closure.handle, self.f.data_ptr(), st.precision, float(self.ct.dtfast), int(self.n == 0),
                self.ct.cff1(), st.tile_j, _lib.stream_handle())
            _lib.check(rc, "sf_phys_substep_closure_fused")
        else:
            rc = st.lib.sf_phys_substep_fused(C.byref(a), st.precision, float(self.ct.dtfast), int(self.n == 0),
                                              self.ct.cff1(), st.tile_j, _lib.stream_handle())
            _lib.check(rc, "sf_phys_substep_fused")
        self.n += 1

    def features(self) -> torch.Tensor:
        """Raw features [S2, 45] of the current t1 state (call after cycleTimes2D)."""
        st = self.stepper
        S2 = self.shape[0] * self.shape[1]
        out = torch.empty((S2, NIN), dtype=torch.float32, device="cuda")
        lib = _bind(st.lib)
        a = st.fields(self.OCEAN)
        _lib.check(lib.sf_closure_features(C.byref(a), st.precision, self.f.data_ptr(), out.data_ptr(),
                                           _lib.stream_handle()), "sf_closure_features")
        return out


def closure_buffers(model: Model):
    S2 = model.shape[0] * model.shape[1]
    return (torch.empty(S2, dtype=model.dtype, device="cuda"), torch.empty(S2, dtype=model.dtype, device="cuda"))


def sample_mask(shape) -> torch.Tensor:
    """Points where both the u and the v target exist: 2 <= i <= L-1, 2 <= j <= M-1."""
    J, I = shape
    m = torch.zeros(shape, dtype=torch.bool, device="cuda")
    m[2:J - 1, 2:I - 1] = True
    return m.reshape(-1)


def make_samples(fine: Model, fine_probe: Model, probe: Model, nsnap: int, spacing: int, r: int = 4):
    """Samples of the subgrid momentum tendency along a fine run.

    At nsnap times ``spacing`` coarse steps apart, with B the fine state:
    target = CG(T_fine(B)) - T_coarse(CG(B)), where T is a model's instantaneous
    momentum tendency, measured as one forward/backward Euler substep of a tiny length
    (the probe models' substep) divided by that length.  The fine trajectory is not
    differenced, so the fine scheme's time-stepping error stays out of the target.
    Targets: ubar and vbar components (m/s^2); features: 3x3 stencils of CG(B) (centre
    value and neighbour - centre differences)."""
    feats, targs = [], []
    mask = sample_mask(probe.shape)
    dtp = probe.ct.dtfast
    assert abs(fine_probe.ct.dtfast - dtp) < 1e-15
    for s in range(nsnap):
        for _ in range(spacing * r):
            fine.step()
        Bf = [x.clone() for x in fine.state()]
        fine_probe.set_state(*Bf)
        fine_probe.step()
        Tf = [(x - y) / dtp for x, y in zip(fine_probe.state(), Bf)]
        _tz, Tu, Tv = coarse_grain(*Tf, r)
        B = coarse_grain(*Bf, r)
        probe.set_state(*B)
        F = probe.features()
        probe.step()
        _zp, up, vp = probe.state()
        Ru = (Tu - (up - B[1]) / dtp).reshape(-1)
        Rv = (Tv - (vp - B[2]) / dtp).reshape(-1)
        feats.append(F[mask])
        targs.append(torch.stack([Ru[mask], Rv[mask]], dim=1).float())
    return torch.cat(feats), torch.cat(targs)


def make_samples_step(fine: Model, coarse: Model, nsnap: int, spacing: int, r: int = 4):
    """One-step-error samples along a fine run (the correction a coarse step needs).

    At nsnap times ``spacing`` coarse steps apart, with C(k) the coarse-grained fine state
    k coarse steps on: the coarse model takes three steps whose two input levels are
    forced to (C(k-1), C(k)), k = 1, 2, 3 (an Euler start, then two LF-AM3 steps whose
    right-hand-side history comes from the forced states); X is the result of the last.
    Target = (C(4) - X) / dt for ubar and vbar (m/s^2): the extra momentum tendency that
    would have made that coarse step land on the coarse-grained fine state.  Features:
    3x3 stencils of C(3), as in make_samples."""
    feats, targs = [], []
    mask = sample_mask(coarse.shape)
    dtc = coarse.ct.dtfast
    for s in range(nsnap):
        for _ in range((spacing - 4) * r):
            fine.step()
        C = []
        for q in range(5):
            if q:
                for _ in range(r):
                    fine.step()
            C.append([x.clone() for x in coarse_grain(*fine.state(), r)])
        coarse.set_state(*C[0])
        for k in (1, 2, 3):
            coarse.begin()
            coarse.force(C[k - 1], C[k])
            if k == 3:
                F = coarse.features()
            coarse.advance()
        _z, uX, vX = coarse.state()
        Ru = ((C[4][1] - uX) / dtc).reshape(-1)
        Rv = ((C[4][2] - vX) / dtc).reshape(-1)
        feats.append(F[mask])
        targs.append(torch.stack([Ru[mask], Rv[mask]], dim=1).float())
    return torch.cat(feats), torch.cat(targs)


def linear_r2(F: torch.Tensor, Y: torch.Tensor, n: int = 200_000) -> list:
    """Fraction of the target variance a least-squares linear map of the features
    explains (diagnostic of how predictable the targets are from the features)."""
    k = min(n, F.shape[0])
    Fk = F[:k].double()
    sd = Fk.std(0)
    keep = sd > 1e-5 * Fk.abs().mean(0) + 1e-30   # constant features (up to FP32 rounding) dropped
    Fk = (Fk[:, keep] - Fk[:, keep].mean(0)) / sd[keep]
    A = torch.cat([Fk, torch.ones((k, 1), dtype=torch.float64, device=F.device)], 1).cpu()
    Yk = Y[:k].double().cpu()
    # normal equations solved through the pseudo-inverse of A^T A (eigenvalues below
    # 1e-12 of the largest dropped): an in-sample fit with an intercept, so R^2 >= 0
    ev, V = torch.linalg.eigh(A.T @ A)
    inv = torch.where(ev > 1e-12 * ev.max(), 1.0 / ev, torch.zeros_like(ev))
    sol = V @ (inv[:, None] * (V.T @ (A.T @ Yk)))
    res = Yk - A @ sol
    return [float(1.0 - res[:, c].var() / Yk[:, c].var()) for c in range(Yk.shape[1])]


def normalise(F: torch.Tensor, Y: torch.Tensor):
    """Per-feature mean and spread; a feature constant up to FP32 rounding (spread below
    1e-5 of its magnitude, e.g. the x differences of a beta-plane f) gets spread 1, so
    its rounding noise is not amplified into an input."""
    mu = F.mean(0)
    sd = F.std(0)
    sd = torch.where(sd > 1e-5 * F.abs().mean(0) + 1e-30, sd, torch.ones_like(sd))
    ym = Y.mean(0)
    ys = Y.std(0)
    ys = torch.where(ys > 0, ys, torch.ones_like(ys))
    return mu, sd, ym, ys


def recip_sd(sd) -> np.ndarray:
    """1 / sd rounded to FP32 (the scale sf_mlp_set_norm stores for the inference kernel)."""
    sd = sd.detach().cpu().numpy() if isinstance(sd, torch.Tensor) else np.asarray(sd)
    return np.float32(1.0) / sd.astype(np.float32)


def pack_inputs(F: torch.Tensor, mu, sd) -> torch.Tensor:
    """Normalised BF16 inputs [n, 48] (padding columns zero): (F - mu) * (1 / sd), the
    arithmetic of the closure inference kernel."""
    n = F.shape[0]
    Xn = torch.zeros((n, IN), dtype=torch.float32, device=F.device)
    mu = torch.as_tensor(mu, dtype=torch.float32, device=F.device)
    Xn[:, :NIN] = (F - mu) * torch.from_numpy(recip_sd(sd)).to(F.device)
    return to_bf16_bits(Xn)


def default_ntile(n: int, blocks: int = 256) -> int:
    """Tiles of 64 samples per training workgroup for a batch of n samples (n a multiple of
    64): the largest divisor of n / 64 that is <= ceil(n / (64 * blocks))."""
    assert n % 64 == 0 and n > 0
    tiles = n // 64
    t = max(1, -(-tiles // blocks))
    while tiles % t:
        t -= 1
    return t


def train(mlp: ClosureMLP, Xb: torch.Tensor, Yn: torch.Tensor, epochs: int, batch: int = 4096, lr: float = 1e-3,
          seed: int = 6, lr_final: float | None = None, noise: float = 0.0):
    """Adam over shuffled mini-batches (seeded); the learning rate follows a cosine from
    ``lr`` to ``lr_final`` over the epochs (constant when lr_final is None).  With
    ``noise`` > 0 every epoch trains on the normalised inputs plus Gaussian noise of
    that standard deviation (input-noise regularisation, which keeps the network's
    response smooth for states off the training trajectories).  Returns the
    training-set MSE (clean inputs) before training and after every epoch."""
    n = Xb.shape[0]
    nb = n // batch
    g = torch.Generator(device="cpu").manual_seed(seed)
    gn = torch.Generator(device=Xb.device).manual_seed(seed + 1)
    losses = [dataset_mse(mlp, Xb, Yn)]
    for ep in range(epochs):
        lr_ep = lr if lr_final is None else lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * ep / epochs))
        Xe = Xb
        if noise > 0.0:
            Xn = Xb.view(torch.bfloat16).float()
            Xn[:, :NIN] += noise * torch.randn((n, NIN), generator=gn, device=Xb.device)
            Xe = to_bf16_bits(Xn)
            del Xn
        perm = torch.randperm(n, generator=g).to(torch.int32).cuda()
        for b in range(nb):
            idx = perm[b * batch:(b + 1) * batch]
            mlp.train_step(Xe, Yn, idx, adam=True, lr=lr_ep)
        losses.append(dataset_mse(mlp, Xb, Yn))
    return losses


def dataset_mse(mlp: ClosureMLP, Xb: torch.Tensor, Yn: torch.Tensor) -> float:
    """Mean squared error of the network over a whole (normalised) data set."""
    return float(((mlp.forward(Xb) - Yn) ** 2).mean())


def rmse(a: torch.Tensor, b: torch.Tensor) -> float:
    return float(torch.sqrt(torch.mean((a[1:-1, 1:-1].double() - b[1:-1, 1:-1].double()) ** 2)))


class GraphLoop:
    """``nsub`` (a multiple of 3) substeps of a Model -- with the closure stress kernel
    before each when ``closure`` is given -- captured once in a HIP graph and replayed.
    The substeps' average weights are those of the captured steps."""

    def __init__(self, model: Model, nsub: int = 3, closure: ClosureMLP | None = None):
        assert nsub % 3 == 0 and model.n > 0
        self.model, self.nsub = model, nsub
        self.su, self.sv = closure_buffers(model) if closure is not None else (None, None)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            for _ in range(nsub):
                model.step(closure, self.su, self.sv)
        # capture does not run the kernels; the host rotation is back after 3k substeps

    def replay(self, times: int = 1):
        for _ in range(times):
            self.graph.replay()


def fine_coarse_ics(fine: Model, r: int):
    z, u, v = (x.reshape(fine.shape) for x in (fine.OCEAN.zeta_t1, fine.OCEAN.ubar_t1, fine.OCEAN.vbar_t1))
    return coarse_grain(z, u, v, r)


def run_experiment(coarse_shape=(512, 256), ratio: int = 4, nsteps: int = 5000, n_train_runs: int = 4,
                   snaps_per_run: int = 4, snap_spacing: int = 250, nsamples: int = 2_000_000, probe_dt: float = 0.01,
                   epochs: int = 300, batch: int = 131072, lr: float = 4e-3, eval_every: int = 500, seed: int = 6,
                   n_val: int | None = None, gains=(1.0, 0.5, 0.25, 0.1), lr_final: float | None = 1e-5,
                   noise: float = 1.0, target: str = "probe",
                   dtype=torch.float64, log=None) -> dict:
    """Train the closure on fine runs with seeded initial bumps and winds, then run the
    basin1 configuration (centred bump, wind 0.1 N/m^2) fine, coarse plain and coarse
    with closure for ``nsteps`` coarse steps, and report the RMSE of zeta, ubar and vbar
    of both coarse runs against the coarse-grained fine run."""
    import time as _time
    log = log or (lambda *a: None)
    t_start = _time.perf_counter()
    rng = np.random.Generator(np.random.PCG64(seed))
    Fs, Ys = [], []
    for run in range(n_train_runs):
        center = (float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.3, 0.7)))
        amp, bump = float(rng.uniform(0.05, 0.15)), float(rng.uniform(0.05, 0.15))
        fcfg = fine_config(*coarse_shape, ratio, amp=amp, bump=bump, center=center)
        fine = Model(fcfg, dtype)
        if target == "step":
            coarse = Model(coarse_config(fcfg), dtype)
            F, Y = make_samples_step(fine, coarse, snaps_per_run, snap_spacing, ratio)
            del coarse
        else:
            fpcfg = dict(fcfg, dtfast=probe_dt)
            fine_probe = Model(fpcfg, dtype)
            pcfg = coarse_config(fcfg)
            pcfg["dtfast"] = probe_dt
            probe = Model(pcfg, dtype)
            F, Y = make_samples(fine, fine_probe, probe, snaps_per_run, snap_spacing, ratio)
            del fine_probe, probe
        Fs.append(F)
        Ys.append(Y)
        log(f"training run {run}: {F.shape[0]} samples")
        del fine
    F = torch.cat(Fs)
    Y = torch.cat(Ys)
    g = torch.Generator(device="cpu").manual_seed(seed)
    sel = torch.randperm(F.shape[0], generator=g)[:min(nsamples, F.shape[0])].cuda()
    F, Y = F[sel], Y[sel]
    r2 = linear_r2(F, Y)
    log(f"linear least-squares R^2 of the targets: {r2}")
    mu, sd, ym, ys = normalise(F, Y)
    Xb = pack_inputs(F, mu, sd)
    Yn = ((Y - ym) / ys).contiguous()
    mlp = ClosureMLP(max_blocks=batch // 64, seed=seed)
    mlp.set_norm(mu.cpu().numpy(), sd.cpu().numpy(), [float(ym[0]), float(ys[0]), float(ym[1]), float(ys[1])])
    t0 = _time.perf_counter()
    losses = train(mlp, Xb, Yn, epochs, batch, lr, seed, lr_final, noise)
    torch.cuda.synchronize()
    t_train = _time.perf_counter() - t0
    log(f"training: {epochs} epochs, loss {losses[0]:.4f} -> {losses[-1]:.4f} ({t_train:.1f} s)")
    # held-out skill of the network on the last 5 % of the samples is not separated here:
    # the evaluation below is on a configuration none of the training runs used

    # validation on a held-out fine run: the closure is selected only if it lowers the
    # normalised RMSE there; otherwise the plain coarse scheme is the model (fallback)
    vcenter = (float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.3, 0.7)))
    vamp, vbump = float(rng.uniform(0.05, 0.15)), float(rng.uniform(0.05, 0.15))
    vcfg = fine_config(*coarse_shape, ratio, amp=vamp, bump=vbump, center=vcenter)
    n_val = nsteps if n_val is None else n_val
    # the closure's output is also scaled by a gain chosen on the validation run (a
    # network that fits the tendency a priori can still drift when coupled)
    base = [float(ym[0]), float(ys[0]), float(ym[1]), float(ys[1])]
    params = mlp.get_params()
    variants = {}
    for gain in gains:
        variants[gain] = ClosureMLP(max_blocks=1024, params=params)
        variants[gain].set_norm(mu.cpu().numpy(), sd.cpu().numpy(), [gain * x for x in base])
    vseries = compare_runs(vcfg, None, ratio, n_val, max(n_val // 10, 1), dtype, variants=variants)
    gain_scores = {g: validation_score(vseries, g) for g in gains}
    plain_score = validation_score(vseries, "plain")
    for g in gains:
        log(f"validation gain {g}: plain {plain_score:.3e}, closure {gain_scores[g]:.3e}")
    for v in variants.values():
        v.close()
    # a run that diverged (non-finite score) ranks last
    best = min(gain_scores, key=lambda g: gain_scores[g] if math.isfinite(gain_scores[g]) else math.inf)
    score = dict(plain=plain_score, closure=gain_scores[best])
    selected = "closure" if math.isfinite(score["closure"]) and score["closure"] < score["plain"] else "plain"
    mlp.set_norm(mu.cpu().numpy(), sd.cpu().numpy(), [best * x for x in base])
    log(f"validation ({n_val} steps, mean over time): plain {score['plain']:.3e}, closure (gain {best}) {score['closure']:.3e}"
        f" -> {selected}")

    series = compare_runs(fine_config(*coarse_shape, ratio), mlp, ratio, nsteps, eval_every, dtype, log)
    fcfg = fine_config(*coarse_shape, ratio)
    ccfg = coarse_config(fcfg)
    return dict(losses=losses, train_s=t_train, series=series, samples=int(F.shape[0]), mlp=mlp, linear_r2=r2,
                target_rms=[float(Y[:, 0].pow(2).mean().sqrt()), float(Y[:, 1].pow(2).mean().sqrt())],
                validation=dict(score=score, selected=selected, steps=n_val, gain=best,
                                gain_scores={str(k): v for k, v in gain_scores.items()}),
                fine_shape=[fcfg["szI"], fcfg["szJ"]], coarse_shape=[ccfg["szI"], ccfg["szJ"]],
                dt_fine=fcfg["dtfast"], dt_coarse=ccfg["dtfast"], total_s=_time.perf_counter() - t_start)


def validation_score(series: list, key="plain") -> float:
    """Mean over the checkpoints of a compare_runs series of the RMSE of zeta, ubar and
    vbar, each divided by that field's signal maximum, summed over the three fields.
    ``key`` is "plain", "closure" or the name of one of the variants."""
    def err(r):
        return r[key] if key in ("plain", "closure") else r["variants"][key]
    return sum(err(r)[c] / max(r["signal"][c], 1e-30) for r in series for c in range(3)) / len(series)


def compare_runs(fcfg: dict, mlp: ClosureMLP | None, ratio: int, nsteps: int, every: int, dtype, log=None,
                 variants: dict | None = None) -> list:
    """Fine run vs coarse plain vs coarse with closure from CG(fine initial state);
    RMSE of zeta, ubar, vbar against CG(fine) every ``every`` coarse steps.
    ``variants`` (name -> network) adds one more coarse run per network beside them,
    reported under row["variants"][name]."""
    log = log or (lambda *a: None)
    ccfg = coarse_config(fcfg)
    fine = Model(fcfg, dtype)
    plain = Model(ccfg, dtype)
    zc, uc, vc = fine_coarse_ics(fine, ratio)
    plain.set_state(zc, uc, vc)
    runs = {}
    for name, net in ([("closure", mlp)] if mlp is not None else []) + list((variants or {}).items()):
        m = Model(ccfg, dtype)
        m.set_state(zc, uc, vc)
        runs[name] = (m, net, *closure_buffers(m))
    series = []
    done = 0
    while done < nsteps:
        n = min(every, nsteps - done)
        for _ in range(n * ratio):
            fine.step()
        for _ in range(n):
            plain.step()
            for m, net, su, sv in runs.values():
                m.step(net, su, sv)
        done += n
        zf, uf, vf = coarse_grain(*fine.state(), ratio)
        zp, up, vp = plain.state()
        row = dict(step=done, plain=[rmse(zp, zf), rmse(up, uf), rmse(vp, vf)],
                   signal=[float(zf.abs().max()), float(uf.abs().max()), float(vf.abs().max())])
        errs = {}
        for name, (m, *_rest) in runs.items():
            zq, uq, vq = m.state()
            errs[name] = [rmse(zq, zf), rmse(uq, uf), rmse(vq, vf)]
        if mlp is not None:
            row["closure"] = errs.pop("closure")
        if variants:
            row["variants"] = errs
        series.append(row)
        log(f"step {done}: rmse plain {row['plain']} closure {row.get('closure')}")
    return series


def speed(mlp: ClosureMLP, coarse_shape=(512, 256), ratio: int = 4, nsub: int = 300, dtype=torch.float32,
          training: bool = True) -> dict:
    """GPU time per substep of the fine run and of the coarse run with the closure
    (graph replays), the closure's share of the coarse substep, the wall-time ratio for the
    same simulated period and (``training``) the training-step rate."""
    fcfg = fine_config(*coarse_shape, ratio)
    ccfg = coarse_config(fcfg)
    out = {}
    for name, cfg, cl in (("fine", fcfg, None), ("coarse_closure", ccfg, mlp), ("coarse_plain", ccfg, None)):
        m = Model(cfg, dtype)
        m.step()                           # Euler start outside the graph
        loop = GraphLoop(m, 3, cl)
        loop.replay(5)
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        loop.replay(nsub // 3)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / (nsub // 3 * 3)
        pts = cfg["szI"] * cfg["szJ"]
        out[name] = dict(ms_per_substep=ms, points=pts, cell_steps_per_s=pts / (ms * 1e-3), dt=cfg["dtfast"])
    # wall time to simulate one hour
    for k in out:
        out[k]["s_per_sim_hour"] = 3600.0 / out[k]["dt"] * out[k]["ms_per_substep"] * 1e-3
    out["speedup_closure_vs_fine"] = out["fine"]["s_per_sim_hour"] / out["coarse_closure"]["s_per_sim_hour"]
    out["closure_adds"] = out["coarse_closure"]["ms_per_substep"] / out["coarse_plain"]["ms_per_substep"] - 1.0
    if training:
        out["training"] = train_rate(mlp.get_params())
    return out


MLP_FLOP_PER_SAMPLE = 43008   # forward 16 384 + backward data 10 240 + weight gradients 16 384 (padded shapes)


def train_rate(params: np.ndarray, batch: int = 131072, nstep: int = 50, n: int = 1 << 20) -> dict:
    """Time per Adam step of the training kernels (forward + backward + block sums, then the
    reduction + Adam) at `batch` random samples from n synthetic rows, on a copy of the
    network; matrix-core FLOP rate from MLP_FLOP_PER_SAMPLE."""
    g = torch.Generator(device="cpu").manual_seed(11)
    X = torch.zeros(n, IN)
    X[:, :NIN] = torch.randn(n, NIN, generator=g)
    Xb = to_bf16_bits(X.cuda())
    Y = torch.randn(n, NOUT, generator=g).cuda().contiguous()
    idx = torch.randperm(n, generator=g)[:batch].to(torch.int32).cuda()
    mlp = ClosureMLP(max_blocks=batch // 64, params=params)
    for _ in range(5):
        mlp.train_step(Xb, Y, idx, adam=True, lr=1e-4)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(nstep):
        mlp.train_step(Xb, Y, idx, adam=True, lr=1e-4)
    b.record()
    torch.cuda.synchronize()
    mlp.close()
    us = a.elapsed_time(b) * 1e3 / nstep
    tflops = batch * MLP_FLOP_PER_SAMPLE / (us * 1e-6) / 1e12
    return dict(batch=batch, ntile=default_ntile(batch), us_per_step=us, samples_per_us=batch / us,
                tflop_s=tflops, frac_bf16_peak=tflops / 2500.0)
