"""Pins of the warm-start MLP oracle (oracle/mlp_oracle.py): the analytic
gradient against central finite differences of the FP64 network, the bf16
rounding against hand-worked values, and the effect of the emulated roundings
bounded."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _net(seed=3, n=48):
    rng = np.random.default_rng(seed)
    th = M.init_params(seed).astype(np.float64)
    th[M.O_B1:M.O_W2] = rng.normal(0, 0.3, M.H)
    th[M.O_B2:M.O_W3] = rng.normal(0, 0.3, M.H)
    th[M.O_B3:M.O_W4] = rng.normal(0, 0.3, M.H)
    th[M.O_B4] = 0.2
    x = rng.normal(0, 1, (n, M.NIN))
    t = rng.normal(0, 1, n)
    w = rng.uniform(0.5, 2.0, n)
    return th, x, t, w


def test_bf16_rounding_known_values():
    one = 1.0
    ulp = 2.0 ** -7          # bf16 spacing at 1
    vals = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 2 + 2 ** -20, -2.5, 3.0e-39, 65504.0])
    want = np.array([one, one, one + 2 * ulp, one + ulp, -2.5, M.bf16(3.0e-39)[()], 65536.0])
    got = M.bf16(vals)
    # ties to even: 1 + ulp/2 -> 1, 1 + 3ulp/2 -> 1 + 2ulp; above the tie -> up
    np.testing.assert_array_equal(got[:5], want[:5])
    assert got[6] == 65536.0
    # idempotent, and exact on values with 8 significant bits
    np.testing.assert_array_equal(M.bf16(got), got)


@pytest.mark.parametrize("block", ["W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4"])
def test_gradient_matches_finite_differences(block):
    th, x, t, w = _net()
    g, _ = M.gradient(th, x, t, w, emulate=False)
    lo = {"W1": M.O_W1, "b1": M.O_B1, "W2": M.O_W2, "b2": M.O_B2, "W3": M.O_W3, "b3": M.O_B3, "w4": M.O_W4,
          "b4": M.O_B4}[block]
    hi = {"W1": M.O_B1, "b1": M.O_W2, "W2": M.O_B2, "b2": M.O_W3, "W3": M.O_B3, "b3": M.O_W4, "w4": M.O_B4,
          "b4": M.NP}[block]
    rng = np.random.default_rng(hash(block) % 2**32)
    idx = rng.choice(np.arange(lo, hi), size=min(12, hi - lo), replace=False)
    h = 1e-5
    for i in idx:
        tp = th.copy(); tp[i] += h
        tm = th.copy(); tm[i] -= h
        fd = (M.loss(tp, x, t, w) - M.loss(tm, x, t, w)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (block, i, fd, g[i])


def test_loss_consistent_with_forward():
    th, x, t, w = _net(5)
    _, loss = M.gradient(th, x, t, w, emulate=False)
    y, _ = M.forward(th, x, emulate=False)
    assert loss == pytest.approx(float((w * (y - t) ** 2).sum() / w.sum()), rel=1e-12)


def test_emulated_rounding_is_a_small_perturbation():
    """The bf16 roundings change the gradient by a few parts in a thousand,
    not more: the GPU comparison tolerance is set from this."""
    th, x, t, w = _net(7, n=256)
    ge, le = M.gradient(th, x, t, w, emulate=True)
    gx, lx = M.gradient(th, x, t, w, emulate=False)
    rel = np.linalg.norm(ge - gx) / np.linalg.norm(gx)
    assert 1e-5 < rel < 5e-2
    assert abs(le - lx) < 5e-2 * lx


def test_gelu_closed_forms():
    z = np.array([0.0, 1.0, -1.0, 10.0, -10.0])
    from math import erf, sqrt, pi, exp
    want = [0.0, 0.5 * (1 + erf(1 / sqrt(2))), -0.5 * (1 - erf(1 / sqrt(2))), 10.0, 0.0]
    np.testing.assert_allclose(M.gelu(z), want, atol=1e-15)
    # derivative at 0 is 1/2; at large |z| it tends to the step function
    np.testing.assert_allclose(M.gelu_grad(np.array([0.0, 10.0, -10.0])), [0.5, 1.0, 0.0], atol=1e-15)
    # and it is the derivative of gelu
    zz = np.linspace(-4, 4, 81)
    fd = (M.gelu(zz + 1e-6) - M.gelu(zz - 1e-6)) / 2e-6
    np.testing.assert_allclose(M.gelu_grad(zz), fd, atol=1e-8)
