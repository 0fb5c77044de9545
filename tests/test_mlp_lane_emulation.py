"""CPU emulation, lane by lane, of the closure network's training kernel algorithm
(csrc/mlp.hip, csrc/mlp_net.h): the packed weight layout, the operand layouts of the
16x16x16 / 16x16x32 matrix-core products, the layer chaining in registers, the identity
products that move operands to the sample-row layout, and the weight-gradient sums.  The
products are emulated exactly in FP64 (the hardware's FP32 sums differ only in rounding),
so with the BF16 roundings at the kernel's points the emulated gradient of one wave's
16-sample groups must equal the FP64 restatement of the BF16 network (oracle/mlp_ref.py,
grad_bf16) to FP64 rounding.  This pins the layout algebra independently of the GPU."""
from __future__ import annotations

import numpy as np

from oracle import mlp_ref as R

bf = R.bf16_round
LANES = np.arange(64)
LR, LQ = LANES % 16, LANES // 16


def c32(q, j):
    """K index of element j of lane group q in a 16x16x32 operand (the kernel's order)."""
    return 4 * q + j if j < 4 else 16 + 4 * q + j - 4


def mma(a, b, c, K):
    """D = A B + C for one 16x16xK product: lane (r, q) element j of A holds A[r][k] and of
    B holds B[k][r], with the same k for both (k = 4q + j for K = 16, c32(q, j) for K = 32 --
    any order shared by A and B gives the same sum); C and D in the result layout, lane
    (r, q) element t = row 4q + t of column r."""
    kmap = (lambda q, j: 4 * q + j) if K == 16 else c32
    A = np.zeros((16, K))
    B = np.zeros((K, 16))
    C = np.zeros((16, 16))
    for lane in range(64):
        r, q = lane % 16, lane // 16
        for j in range(a.shape[1]):
            A[r, kmap(q, j)] = a[lane, j]
            B[kmap(q, j), r] = b[lane, j]
        for t in range(4):
            C[4 * q + t, r] = c[lane, t]
    D = A @ B + C
    return np.array([[D[4 * (lane // 16) + t, lane % 16] for t in range(4)] for lane in range(64)])


def wpk_slot(p):
    """Slot of parameter p in the packed forward operand layout (the kernel's wpk_slot)."""
    H, OUT = R.H, R.OUT
    off_b1 = R.IN * H
    off_w2 = off_b1 + H
    off_b2 = off_w2 + H * H
    off_w3 = off_b2 + H
    off_b3 = off_w3 + H * OUT

    def k32(r, kk):
        return (r + 16 * ((kk & 15) >> 2)) * 8 + (kk & 3) + ((kk >> 4) << 2)

    if p < off_b1:
        k, n = divmod(p, H)
        tn, r = n >> 4, n & 15
        if k < 32:
            return tn * 768 + k32(r, k)
        kk = k - 32
        return tn * 768 + 512 + (r + 16 * (kk >> 2)) * 4 + (kk & 3)
    if off_w2 <= p < off_b2:
        k, n = divmod(p - off_w2, H)
        return 3072 + ((n >> 4) * 2 + (k >> 5)) * 512 + k32(n & 15, k & 31)
    if off_w3 <= p < off_b3:
        k, n = divmod(p - off_w3, OUT)
        return 7168 + (k >> 5) * 512 + k32(n, k & 31)
    return -1


def emulate_wave(theta, X, Y, scale, swap_input_halves=False):
    """One wave of k_mlp_fwd_bwd over 16-sample groups X[g] (16 x 48), Y[g] (16 x 2).
    swap_input_halves plants a layout mistake: the input operand of layer 1 built with its
    two 4-element halves exchanged, so its K order no longer matches the weights'."""
    p = R.unpack(theta)
    wbits = bf(theta)                                     # BF16 copy, parameter order
    wpk = np.zeros(8192)
    for i in range(theta.size):
        s = wpk_slot(i)
        if s >= 0:
            wpk[s] = wbits[i]
    assert np.count_nonzero([wpk_slot(i) >= 0 for i in range(theta.size)]) == 8192
    tile = lambda off, E: wpk[off:off + 64 * E].reshape(64, E)
    w1a = [tile(tn * 768, 8) for tn in range(4)]
    w1b = [tile(tn * 768 + 512, 4) for tn in range(4)]
    w2 = [tile(3072 + e * 512, 8) for e in range(8)]
    w3 = [tile(7168 + kb * 512, 8) for kb in range(2)]
    rows = lambda b, base: np.array([[b[base + 4 * q + t] for t in range(4)] for q in LQ])
    b1 = [rows(p["b1"], 16 * tn) for tn in range(4)]
    b2 = [rows(p["b2"], 16 * tn) for tn in range(4)]
    b3 = rows(p["b3"], 0)
    W2b, W3b = bf(p["W2"]), bf(p["W3"])
    w3r = [np.array([[W3b[16 * tm + r, 4 * q + j] for j in range(4)] for r, q in zip(LR, LQ)]) for tm in range(4)]
    w2r = [[np.array([[W2b[16 * tm + r, 32 * kb + c32(q, j)] for j in range(8)] for r, q in zip(LR, LQ)])
            for kb in range(2)] for tm in range(4)]
    I = np.array([[1.0 if 4 * q + j == r else 0.0 for j in range(4)] for r, q in zip(LR, LQ)])
    z4 = np.zeros((64, 4))
    pack4 = bf
    relu4 = lambda z: np.maximum(bf(z), 0.0)
    cat8 = lambda a, b: np.concatenate([a, b], 1)
    mask = lambda d, a: np.where(a > 0, d, 0.0)
    tr = lambda v: pack4(mma(v, I, z4, 16))
    gW1 = [z4.copy() for _ in range(12)]
    gW2 = [z4.copy() for _ in range(16)]
    gW3 = [z4.copy() for _ in range(4)]
    gb1 = [z4.copy() for _ in range(4)]
    gb2 = [z4.copy() for _ in range(4)]
    gb3 = np.zeros(2)
    for Xg, Yg in zip(X, Y):
        xa = np.array([[Xg[r, c32(q, j)] for j in range(8)] for r, q in zip(LR, LQ)])
        if swap_input_halves:
            xa = np.concatenate([xa[:, 4:], xa[:, :4]], 1)
        xb = np.array([[Xg[r, 32 + 4 * q + j] for j in range(4)] for r, q in zip(LR, LQ)])
        # forward16
        a1 = [relu4(mma(w1b[tn], xb, mma(w1a[tn], xa, b1[tn], 32), 16)) for tn in range(4)]
        h0, h1 = cat8(a1[0], a1[1]), cat8(a1[2], a1[3])
        a2 = [relu4(mma(w2[2 * tn + 1], h1, mma(w2[2 * tn], h0, b2[tn], 32), 32)) for tn in range(4)]
        z3 = mma(w3[1], cat8(a2[2], a2[3]), mma(w3[0], cat8(a2[0], a2[1]), b3, 32), 32)
        # output error on lanes q = 0
        e0 = np.where(LQ == 0, z3[:, 0] - Yg[LR, 0], 0.0)
        e1 = np.where(LQ == 0, z3[:, 1] - Yg[LR, 1], 0.0)
        d0, d1 = scale * e0, scale * e1
        gb3 += [d0.sum(), d1.sum()]
        dz3 = np.stack([bf(d0), bf(d1), np.zeros(64), np.zeros(64)], 1)
        dz2 = []
        for tm in range(4):
            d = mask(mma(w3r[tm], dz3, z4, 16), a2[tm])
            gb2[tm] += d
            dz2.append(pack4(d))
        g0, g1 = cat8(dz2[0], dz2[1]), cat8(dz2[2], dz2[3])
        dz1 = []
        for tm in range(4):
            d = mask(mma(w2r[tm][1], g1, mma(w2r[tm][0], g0, z4, 32), 32), a1[tm])
            gb1[tm] += d
            dz1.append(pack4(d))
        t3 = tr(dz3)
        for tm in range(4):
            gW3[tm] = mma(tr(a2[tm]), t3, gW3[tm], 16)
        td = [tr(v) for v in dz2]
        for tm in range(4):
            ta = tr(a1[tm])
            for tn in range(4):
                gW2[tm * 4 + tn] = mma(ta, td[tn], gW2[tm * 4 + tn], 16)
        td = [tr(v) for v in dz1]
        for tk, xk in enumerate((xa[:, :4], xa[:, 4:], xb)):
            tx = tr(xk)
            for tn in range(4):
                gW1[tk * 4 + tn] = mma(tx, td[tn], gW1[tk * 4 + tn], 16)
    # the kernel's block-sum layouts
    G = {k: np.zeros(shape) for k, shape in R.SHAPES}
    for lane in range(64):
        r, q = lane % 16, lane // 16
        for t in range(4):
            for tm in range(4):
                for tn in range(4):
                    G["W2"][16 * tm + 4 * q + t, 16 * tn + r] = gW2[tm * 4 + tn][lane, t]
                G["W3"][16 * tm + 4 * q + t, r] = gW3[tm][lane, t]
                G["b1"][16 * tm + 4 * q + t] += gb1[tm][lane, t]
                G["b2"][16 * tm + 4 * q + t] += gb2[tm][lane, t]
            for tk in range(3):
                for tn in range(4):
                    G["W1"][16 * tk + 4 * q + t, 16 * tn + r] = gW1[tk * 4 + tn][lane, t]
    G["b3"][:2] = gb3
    return G


def test_lane_emulation_of_the_training_kernel_equals_the_bf16_gradient():
    rng = np.random.default_rng(5)
    p = R.unpack(R.init_params(5))
    p["b1"] = 0.1 * rng.standard_normal(64)
    p["b2"] = 0.1 * rng.standard_normal(64)
    p["b3"][:2] = 0.1 * rng.standard_normal(2)
    theta = R.pack(p).astype(np.float32).astype(np.float64)
    X = np.zeros((32, R.IN))
    X[:, :R.NIN] = rng.standard_normal((32, R.NIN))
    X = bf(X.astype(np.float32))
    Y = rng.standard_normal((32, R.NOUT)).astype(np.float32).astype(np.float64)
    scale = 2.0 / (32 * R.NOUT)
    G = emulate_wave(theta, [X[:16], X[16:]], [Y[:16], Y[16:]], scale)
    E = R.unpack(R.grad_bf16(theta, X, Y)[0])
    for k in E:
        err = np.linalg.norm(G[k] - E[k]) / max(np.linalg.norm(E[k]), 1e-300)
        assert err < 1e-12, (k, err)
    # a planted layout mistake fails the same comparison
    B = emulate_wave(theta, [X[:16], X[16:]], [Y[:16], Y[16:]], scale, swap_input_halves=True)
    assert np.linalg.norm(B["W2"] - E["W2"]) / np.linalg.norm(E["W2"]) > 1e-3


def test_packed_layout_is_a_bijection_onto_the_weight_slots():
    slots = [wpk_slot(i) for i in range(R.NPARAM)]
    used = sorted(s for s in slots if s >= 0)
    assert used == list(range(8192))
    # biases have no slot
    assert all(wpk_slot(i) < 0 for i in range(R.IN * R.H, R.IN * R.H + R.H))
