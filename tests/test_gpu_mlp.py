"""GPU: the contact MLP's MFMA kernels (inference, training gradient, Adam) against the
oracle's FP64 network with the same BF16 rounding points.

Tolerances: FP32 accumulation and the hardware exponential against FP64 move a
hidden activation across a BF16 rounding boundary now and then (one BF16 ulp,
2^-8 relative); outputs agree within 2e-3 of their largest magnitude and
gradients within 1e-2 in norm, layer by layer.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

from .test_oracle_mlp import K, N, NP, offb, offw, random_norm, random_params

pytestmark = pytest.mark.gpu


def layer_slices():
    for l in range(4):
        yield "W%d" % (l + 1), slice(offw(l), offb(l))
        yield "b%d" % (l + 1), slice(offb(l), offb(l) + N[l])


@pytest.mark.parametrize("rows", [1, 17, 1000])
def test_forward_matches_oracle(gpu, rows):
    p = random_params(11)
    norm = random_norm(12)
    mlp = gpu.ContactMLP(seed=0)
    mlp.set_params(p, norm)
    x = np.random.default_rng(rows).normal(size=(rows, 12)).astype(np.float32)
    y = mlp.forward(x)
    yo = po.mlp_forward(p, norm, x, bf16=True)
    err = np.abs(y - yo).max() / np.abs(yo).max()
    print("rows", rows, "forward max error / max |y|:", err)
    assert err < 2e-3
    # parameters round-trip unchanged
    p2, n2 = mlp.params()
    assert np.array_equal(p2, p) and np.array_equal(n2, norm)


def test_gradient_matches_oracle(gpu):
    p = random_params(21)
    norm = random_norm(22)
    rng = np.random.default_rng(23)
    rows = 200  # 12 full tiles and a ragged one
    x = rng.normal(size=(rows, 12)).astype(np.float32)
    t = rng.normal(size=(rows, 3)).astype(np.float32)
    mlp = gpu.ContactMLP(seed=0)
    mlp.set_params(p, norm)
    g, loss = mlp.gradient(x, t)
    go, lo = po.mlp_gradient(p, norm, x, t, bf16=True)
    assert abs(loss - lo) / lo < 1e-4, (loss, lo)
    for name, sl in layer_slices():
        rel = np.linalg.norm(g[sl] - go[sl]) / np.linalg.norm(go[sl])
        print(name, "gradient relative error:", rel)
        assert rel < 1e-2, (name, rel)
    # padding of the first and last layers stays exactly zero
    w1 = g[offw(0):offb(0)].reshape(K[0], N[0])
    w4 = g[offw(3):offb(3)].reshape(K[3], N[3])
    assert not w1[12:].any() and not w4[:, 3:].any() and not g[offb(3) + 3:offb(3) + 16].any()
    # and the BF16 network's gradient is the exact network's up to rounding
    ge, _ = po.mlp_gradient(p, norm, x, t, bf16=False)
    cos = g.astype(np.float64) @ ge / (np.linalg.norm(g) * np.linalg.norm(ge))
    assert cos > 0.999, cos


def test_training_fits_a_smooth_map(gpu):
    """Adam on the device fits a smooth synthetic map; the oracle's loss at the trained
    parameters agrees with the device's last epoch."""
    rng = np.random.default_rng(31)
    rows = 16384
    x = rng.normal(size=(rows, 12)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2],
                  np.tanh(x[:, 3] - x[:, 4]),
                  0.3 * x[:, 5] ** 2 - x[:, 11]], axis=1).astype(np.float32)
    norm = np.zeros(27, dtype=np.float32)
    norm[12:24] = 1.0
    norm[24:27] = t.std(axis=0)
    mlp = gpu.ContactMLP(seed=5)
    mlp.set_params(None, norm)
    p0, _ = mlp.params()
    _, l0 = po.mlp_gradient(p0, norm, x[:2048], t[:2048], bf16=True)
    losses = mlp.train(x, t, epochs=40, batch=256, lr=3e-3, seed=7)
    print("loss: initial %.4g, epochs:" % l0, losses[::8], losses[-1])
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.05 * l0
    p1, _ = mlp.params()
    _, l1 = po.mlp_gradient(p1, norm, x[:2048], t[:2048], bf16=True)
    assert l1 < 0.05 * l0
    assert abs(l1 - losses[-1]) < 0.5 * losses[-1] + 1e-3
