"""Host checks of the closure data path: the coarse-graining operator against its
definition (block means / face means, closed-basin boundary values), conservation
of the free-surface volume, and the fine/coarse grid correspondence."""
from __future__ import annotations

import numpy as np
import torch

from sailfish_amd.closure import coarse_config, coarse_grain, fine_config


def test_fine_grid_is_exact_refinement():
    f = fine_config(66, 34, ratio=4)
    assert (f["szI"] - 2) == 4 * (66 - 2) and (f["szJ"] - 2) == 4 * (34 - 2)
    assert abs(f["dx"] * 4 - 1250.0) < 1e-12
    c = coarse_config(f)
    assert (c["szI"], c["szJ"]) == (66, 34)
    assert abs(c["dx"] - 1250.0) < 1e-12


def test_coarse_grain_of_prolonged_fields_is_identity():
    rng = np.random.default_rng(1)
    Jc, Ic, r = 10, 14, 4
    zc = rng.standard_normal((Jc, Ic))
    uc = rng.standard_normal((Jc, Ic))
    vc = rng.standard_normal((Jc, Ic))
    Jf, If = r * (Jc - 2) + 2, r * (Ic - 2) + 2
    zf = np.zeros((Jf, If))
    uf = np.zeros((Jf, If))
    vf = np.zeros((Jf, If))
    for J in range(1, Jc - 1):
        for I in range(1, Ic - 1):
            zf[1 + r * (J - 1):1 + r * J, 1 + r * (I - 1):1 + r * I] = zc[J, I]
    for J in range(1, Jc - 1):
        for I in range(1, Ic):
            uf[1 + r * (J - 1):1 + r * J, 1 + r * (I - 1)] = uc[J, I]      # west face of coarse cell I
    for J in range(1, Jc):
        for I in range(1, Ic - 1):
            vf[1 + r * (J - 1), 1 + r * (I - 1):1 + r * I] = vc[J, I]      # south face of coarse cell J
    z2, u2, v2 = coarse_grain(torch.from_numpy(zf), torch.from_numpy(uf), torch.from_numpy(vf), r)
    z2, u2, v2 = z2.numpy(), u2.numpy(), v2.numpy()
    assert np.allclose(z2[1:-1, 1:-1], zc[1:-1, 1:-1], rtol=0, atol=1e-15)
    # u interior faces 2..L-1 (walls 1 and L are zero after the boundary rule)
    assert np.allclose(u2[1:-1, 2:-1], uc[1:-1, 2:-1], rtol=0, atol=1e-15)
    assert np.all(u2[:, [0, 1, -1]] == 0.0)
    assert np.allclose(v2[2:-1, 1:-1], vc[2:-1, 1:-1], rtol=0, atol=1e-15)
    assert np.all(v2[[0, 1, -1], :] == 0.0)
    # zero-gradient ghosts of zeta
    assert np.array_equal(z2[1:-1, 0], z2[1:-1, 1]) and np.array_equal(z2[0, :], z2[1, :])


def test_block_mean_conserves_volume():
    rng = np.random.default_rng(2)
    r = 4
    zf = torch.from_numpy(rng.standard_normal((r * 8 + 2, r * 12 + 2)))
    u = torch.zeros_like(zf)
    zc, _, _ = coarse_grain(zf, u.clone(), u.clone(), r)
    # uniform cells: fine area = coarse area / r^2
    assert abs(float(zc[1:-1, 1:-1].sum()) * r * r - float(zf[1:-1, 1:-1].sum())) < 1e-9


def test_default_ntile_keeps_grid_within_one_workgroup_per_cu():
    from sailfish_amd.closure import default_ntile
    for n in (64, 4096, 16384, 64 * 255, 64 * 256, 64 * 257, 131072, 64 * 3 * 1000, 2_000_000 // 64 * 64):
        t = default_ntile(n)
        tiles = n // 64
        assert tiles % t == 0 and t >= 1
        # the largest divisor of the tile count at or below ceil(tiles / 256)
        cap = -(-tiles // 256)
        assert t <= cap and all(tiles % d for d in range(t + 1, cap + 1))
    assert default_ntile(131072) == 8 and default_ntile(4096) == 1
