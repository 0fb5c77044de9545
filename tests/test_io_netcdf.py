"""NetCDF history output and restart files on the CPU (host tensors, no kernels): the
history layout of nonlinear/adhocNetcdfOutput.py:3-158 and the record slicing and
30-minute cadence of nonlinear/main3d.py:202-216; a restart restores every array, the
rotation of the time-level views and the counters exactly."""
from __future__ import annotations

import numpy as np
import pytest
import torch
from scipy.io import netcdf_file

from sailfish_amd import configs
from sailfish_amd.io_netcdf import HistoryWriter, read_restart, write_restart
from sailfish_amd.mod_comptimes import CompTimes
from sailfish_amd.mod_grid import Grid
from sailfish_amd.mod_ocean import T_OCEAN
from sailfish_amd.set_coords import s_coordinates
from sailfish_amd.set_weights import set_weights


def _model(szI=14, szJ=9, N=5, dtype=torch.float64, seed=0, with3d=True):
    cfg = configs.random_smooth(szI, szJ, seed=3)
    G = Grid(szI - 2, szJ - 2, N, device="cpu")
    G.set_fields(h=cfg["h"], pm=cfg["pm"], pn=cfg["pn"], f=cfg["f"])
    G.set_vertical(*s_coordinates(N, 4, 5.0, 0.4), hc=50.0)
    ct = CompTimes(dt=60.0, ndtfast=12)
    set_weights(ct)
    O = T_OCEAN(G, dtype=dtype, with3d=with3d)
    g = torch.Generator().manual_seed(seed)
    for name in ("zeta", "ubar", "vbar", "rzeta", "rubar", "rvbar") + (
            ("u", "v", "ru", "rv", "W", "Huon", "Hvom", "AKv") if with3d else ()):
        t = getattr(O, name)
        t.copy_(torch.randn(t.shape, generator=g, dtype=torch.float64).to(dtype))
    for name in ("Zt_avg1", "DU_avg1", "DV_avg1"):
        getattr(O, name).copy_(torch.randn(getattr(O, name).shape, generator=g, dtype=torch.float64))
    for name in ("z_w", "z_r", "Hz"):
        t = getattr(G, name)
        t.copy_(torch.randn(t.shape, generator=g, dtype=torch.float64))
    return ct, G, O


@pytest.mark.parametrize("dtype,with3d", [(torch.float64, True), (torch.float32, True), (torch.float32, False)])
def test_restart_round_trip_restores_state_views_and_counters(tmp_path, dtype, with3d):
    ct, G, O = _model(dtype=dtype, with3d=with3d)
    # put the views in a rotated position, as after a few substeps / baroclinic steps
    for _ in range(2):
        O.cycleTimes2D()
    if with3d:
        O.cycleTimes3D()
    ct.iic, ct.iif, ct.time, ct.time2D, ct.ntfirst = 7, 13, 420.0, 417.5, 1
    path = tmp_path / "rst.nc"
    write_restart(path, ct, G, O)

    ct2, G2, O2 = _model(dtype=dtype, with3d=with3d, seed=99)
    read_restart(path, ct2, G2, O2)
    names = ["zeta", "ubar", "vbar", "rzeta", "rubar", "rvbar", "Zt_avg1", "DU_avg1", "DV_avg1"]
    if with3d:
        names += ["u", "v", "ru", "rv", "W", "Huon", "Hvom", "AKv"]
    for n in names:
        a, b = getattr(O, n), getattr(O2, n)
        assert a.dtype == b.dtype and torch.equal(a, b), n
    for n in ("zeta", "ubar", "rvbar") + (("u", "rv") if with3d else ()):
        for k in range(3):
            assert O2.slot_of(n, f"t{k}") == O.slot_of(n, f"t{k}")
            assert torch.equal(getattr(O2, f"{n}_t{k}"), getattr(O, f"{n}_t{k}"))
    if with3d:
        for n in ("z_w", "z_r", "Hz"):
            assert torch.equal(getattr(G, n), getattr(G2, n))
    for n in ("iic", "iif", "time", "time2D", "ntfirst", "dt", "dtfast", "nfast", "ndtfast"):
        assert getattr(ct2, n) == getattr(ct, n), n
    assert np.array_equal(ct2.weight1, ct.weight1) and np.array_equal(ct2.weight2, ct.weight2)


def test_restart_rejects_another_grid(tmp_path):
    ct, G, O = _model()
    write_restart(tmp_path / "r.nc", ct, G, O)
    ct2, G2, O2 = _model(szI=15)
    with pytest.raises(ValueError, match="grid"):
        read_restart(tmp_path / "r.nc", ct2, G2, O2)
    ct3, G3, O3 = _model(dtype=torch.float32)
    with pytest.raises(ValueError, match="does not match"):
        read_restart(tmp_path / "r.nc", ct3, G3, O3)


def test_history_layout_slicing_and_cadence(tmp_path):
    ct, G, O = _model(szI=12, szJ=8, N=4)
    path = tmp_path / "his.nc"
    hw = HistoryWriter(path, G)
    written, snaps = [], []
    for step in range(1, 100):      # dt = 60 s: records at t = 60 s, then every 30 min from 3600 s
        ct.time = 60.0 * step
        O.zeta_t2.add_(0.01)
        if hw.maybe_write(ct, O):
            written.append(ct.time)
            snaps.append([getattr(O, f"{n}_t2").clone() for n in ("zeta", "ubar", "vbar", "u", "v")])
    hw.close()
    assert written == [60.0, 3600.0, 5400.0]
    M, L, N = G.M, G.L, G.N
    with netcdf_file(str(path), "r", mmap=False) as nc:
        d = {k: v for k, v in nc.dimensions.items()}
        assert (d["xi_rho"], d["xi_u"], d["xi_v"]) == (L + 1, L, L + 1)
        assert (d["eta_rho"], d["eta_u"], d["eta_v"]) == (M + 1, M + 1, M)
        assert (d["s_rho"], d["s_w"], d["N"], d["ocean_time"]) == (N, N + 1, N, None)
        v = nc.variables
        assert v["zeta"].dimensions == ("ocean_time", "eta_rho", "xi_rho")
        assert v["u"].dimensions == ("ocean_time", "s_rho", "eta_u", "xi_u")
        assert v["v"].dimensions == ("ocean_time", "s_rho", "eta_v", "xi_v")
        assert v["zeta"].units == b"meter" and v["ocean_time"].units == b"seconds since 2001-01-01 00:00:00"
        assert v["u"].long_name == b"u-momentum component"
        assert np.array_equal(v["ocean_time"][:], written)
        assert int(v["ntimes"].getValue()) == 3
        assert v["ocean_time"].shape == (3,)
        for k, (z, ub, vb, u, vv) in enumerate(snaps):
            assert np.array_equal(v["zeta"][k], z.numpy().reshape(M + 1, L + 1).astype(np.float32))
            assert np.array_equal(v["ubar"][k], ub.numpy().reshape(M + 1, L + 1)[:, :-1].astype(np.float32))
            assert np.array_equal(v["vbar"][k], vb.numpy().reshape(M + 1, L + 1)[:-1, :].astype(np.float32))
            assert np.array_equal(v["u"][k], u.numpy().reshape(N + 1, M + 1, L + 1)[:-1, :, :-1].astype(np.float32))
            assert np.array_equal(v["v"][k], vv.numpy().reshape(N + 1, M + 1, L + 1)[:-1, :-1, :].astype(np.float32))
        # grid variables: rho points at (i - 1/2) dx from the west wall, u points on the faces
        dx = 1.0 / G.host["pm"]
        assert np.allclose(v["x_u"][:][:, 2] - v["x_u"][:][:, 1], dx[:, 1])
        assert np.allclose(v["x_rho"][:][:, 1] - v["x_u"][:][:, 1], 0.5 * dx[:, 1])
        assert np.array_equal(v["h"][:], G.host["h"])
        assert np.array_equal(v["Cs_r"][:], G.host["Cs_r"])
