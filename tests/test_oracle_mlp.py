"""Pins of the oracle's contact MLP (the checker of the device's training kernels).

The exact FP64 network is pinned by a hand-computed single-path network (layout
of inputs, outputs, weights and biases; the SiLU) and its gradient by central
finite differences; the BF16-emulating mode is the exact one up to BF16 rounding.
"""
import math

import numpy as np

from oracle import pyoracle as po

NP = 10448
K = [16, 64, 64, 64]
N = [64, 64, 64, 16]


def offw(l):
    return [0, 1088, 5248, 9408][l]


def offb(l):
    return offw(l) + K[l] * N[l]


def silu(z):
    return z / (1.0 + math.exp(-z))


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, dtype=np.float32)
    fan = [12, 64, 64, 64]
    for l in range(4):
        a = scale * math.sqrt(6.0 / (fan[l] + (3 if l == 3 else 64)))
        w = rng.uniform(-a, a, size=(K[l], N[l])).astype(np.float32)
        if l == 0:
            w[12:, :] = 0
        if l == 3:
            w[:, 3:] = 0
        p[offw(l):offw(l) + K[l] * N[l]] = w.ravel()
        b = rng.uniform(-0.1, 0.1, size=N[l]).astype(np.float32)
        if l == 3:
            b[3:] = 0
        p[offb(l):offb(l) + N[l]] = b
    return p


def random_norm(seed):
    rng = np.random.default_rng(seed)
    z = np.zeros(27, dtype=np.float32)
    z[:12] = rng.uniform(-0.5, 0.5, 12)
    z[12:24] = rng.uniform(0.5, 2.0, 12)
    z[24:27] = rng.uniform(0.5, 2.0, 3)
    return z


def test_single_path_network_closed_form():
    """Input feature 4 -> unit 7 -> unit 3 -> unit 60 -> output 1, everything else zero
    except the biases checked below: y1 = s * (w4 h3 + b4_1) with the SiLU chain."""
    p = np.zeros(NP, dtype=np.float32)
    w = [0.75, -1.25, 0.5, 2.0]
    b = [0.125, 0.25, -0.375, 0.0625]
    p[offw(0) + 4 * 64 + 7] = w[0]
    p[offb(0) + 7] = b[0]
    p[offw(1) + 7 * 64 + 3] = w[1]
    p[offb(1) + 3] = b[1]
    p[offw(2) + 3 * 64 + 60] = w[2]
    p[offb(2) + 60] = b[2]
    p[offw(3) + 60 * 16 + 1] = w[3]
    p[offb(3) + 1] = b[3]
    p[offb(3) + 0] = 0.5
    norm = np.zeros(27, dtype=np.float32)
    norm[12:24] = 1.0
    norm[4] = 0.25        # mean of feature 4
    norm[12 + 4] = 2.0    # std of feature 4
    norm[24:27] = [3.0, 0.5, 7.0]
    x = np.zeros((1, 12), dtype=np.float32)
    x[0, 4] = 1.75
    x[0, 5] = 100.0       # not connected
    h = (1.75 - 0.25) / 2.0
    h = silu(w[0] * h + b[0])
    # units other than the path carry silu(0) = 0 except through their biases (all zero)
    h = silu(w[1] * h + b[1])
    h = silu(w[2] * h + b[2])
    y1 = 0.5 * (w[3] * h + b[3])
    y = po.mlp_forward(p, norm, x, bf16=False)
    assert abs(y[0, 1] - y1) < 1e-12
    assert abs(y[0, 0] - 3.0 * 0.5) < 1e-12
    assert y[0, 2] == 0.0


def test_gradient_matches_central_differences():
    p = random_params(1)
    norm = random_norm(2)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(40, 12)).astype(np.float32)
    t = rng.normal(size=(40, 3)).astype(np.float32)
    g, loss = po.mlp_gradient(p, norm, x, t, bf16=False)
    assert loss > 0
    # every layer's weights and biases, plus padded entries that must stay zero
    idx = list(rng.choice(offb(0), 40, replace=False))
    for l in range(1, 4):
        idx += list(offw(l) + rng.choice(K[l] * N[l], 40, replace=False))
    for l in range(4):
        idx += [offb(l) + j for j in range(min(N[l], 3))]
    worst = 0.0
    for i in idx:
        row, col = None, None
        for l in range(4):
            if offw(l) <= i < offb(l):
                row, col = divmod(i - offw(l), N[l])
                if (l == 0 and row >= 12) or (l == 3 and col >= 3):
                    assert g[i] == 0.0
                    row = None
                break
        pp = p.copy()
        h = np.float32(2.0 ** -12)
        pp[i] = p[i] + h
        up = pp[i]
        lp = po.mlp_gradient(pp, norm, x, t, bf16=False)[1]
        pp[i] = p[i] - h
        dn = pp[i]
        lm = po.mlp_gradient(pp, norm, x, t, bf16=False)[1]
        fd = (lp - lm) / (float(up) - float(dn))
        worst = max(worst, abs(fd - g[i]) / (abs(g).max()))
    assert worst < 1e-5, worst


def test_bf16_mode_is_the_exact_network_up_to_rounding():
    p = random_params(4)
    norm = random_norm(5)
    x = np.random.default_rng(6).normal(size=(64, 12)).astype(np.float32)
    t = np.random.default_rng(7).normal(size=(64, 3)).astype(np.float32)
    y0 = po.mlp_forward(p, norm, x, bf16=False)
    y1 = po.mlp_forward(p, norm, x, bf16=True)
    err = np.abs(y1 - y0).max() / np.abs(y0).max()
    assert 1e-5 < err < 3e-2, err
    g0, _ = po.mlp_gradient(p, norm, x, t, bf16=False)
    g1, _ = po.mlp_gradient(p, norm, x, t, bf16=True)
    rel = np.linalg.norm(g1 - g0) / np.linalg.norm(g0)
    assert rel < 5e-2, rel
