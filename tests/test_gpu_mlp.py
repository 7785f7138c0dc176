"""Warm-start MLP kernels on the GPU against the CPU oracle
(oracle/mlp_oracle.py): initialisation, forward pass, full-batch gradient and
Adam training, all through the C ABI of include/rewl_mlp.h."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from rewl_simulator_amd import abi
from rewl_simulator_amd.mlp import MLP

pytestmark = pytest.mark.gpu


def _data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 4)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.3 * x[:, 3] ** 2).astype(np.float32)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return x, t, w


def _perturbed(seed):
    rng = np.random.default_rng(seed)
    th = M.init_params(seed).astype(np.float64)
    th[M.O_B1:M.O_W2] = rng.normal(0, 0.3, M.H)
    th[M.O_B2:M.O_W3] = rng.normal(0, 0.3, M.H)
    th[M.O_B3:M.O_W4] = rng.normal(0, 0.3, M.H)
    th[M.O_B4] = 0.2
    return th.astype(np.float32)


def test_init_matches_oracle(gpu):
    with MLP(seed=12345) as m:
        np.testing.assert_array_equal(m.params, M.init_params(12345))


@pytest.mark.parametrize("n", [1, 31, 32, 33, 1000, 20000])
def test_forward_matches_oracle(gpu, n):
    x, _, _ = _data(n, n)
    th = _perturbed(4)
    with MLP() as m:
        m.params = th
        y = m.forward_raw(x)
    yo, _ = M.forward(th, x, emulate=True)
    err = np.abs(y - yo)
    scale = max(1.0, np.abs(yo).max())
    print(f"forward n={n}: max |GPU - oracle| {err.max():.3e}, median {np.median(err):.3e} (scale {scale:.3f})",
          flush=True)
    # FP32 accumulation order and rare one-ulp bf16 flips of activations
    assert err.max() <= 1e-3 * scale, err.max()
    assert np.median(err) <= 1e-4 * scale


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("n", [32, 777, 40000])
def test_gradient_matches_oracle(gpu, n, weighted):
    x, t, w = _data(100 + n, n)
    th = _perturbed(9)
    with MLP() as m:
        m.params = th
        g, loss = m.gradient_raw(x, t, w if weighted else None)
    go, lo = M.gradient(th, x, t, w if weighted else None, emulate=True)
    assert loss == pytest.approx(lo, rel=1e-4)
    blocks = [(M.O_W1, M.O_B1), (M.O_B1, M.O_W2), (M.O_W2, M.O_B2), (M.O_B2, M.O_W3), (M.O_W3, M.O_B3),
              (M.O_B3, M.O_W4), (M.O_W4, M.O_B4), (M.O_B4, M.NP)]
    rels = []
    for lo_, hi_ in blocks:
        a, b = g[lo_:hi_].astype(np.float64), go[lo_:hi_]
        rels.append(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))
    print(f"gradient n={n} weighted={weighted}: relative error per block " + " ".join(f"{r:.2e}" for r in rels),
          flush=True)
    # every parameter block within 1e-3 relative of the bf16-emulating oracle
    assert max(rels) < 1e-3, rels


def test_gradient_is_deterministic(gpu):
    x, t, w = _data(3, 5000)
    with MLP() as m:
        g1, l1 = m.gradient_raw(x, t, w)
        g2, l2 = m.gradient_raw(x, t, w)
    np.testing.assert_array_equal(g1, g2)
    assert l1 == l2


def test_training_fits_and_is_deterministic(gpu):
    x, t, w = _data(21, 20000)
    runs = []
    for _ in range(2):
        with MLP(seed=5) as m:
            hist = m.train_raw(x, t, w, steps=300, batch=4096, lr=3e-3, lr_final=3e-5, seed=11)
            _, full = m.gradient_raw(x, t, w)
            runs.append((hist, full, m.params))
    (h1, f1, p1), (h2, f2, p2) = runs
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(h1, h2)
    initial = float(np.average(t ** 2, weights=w))
    assert f1 < 0.05 * initial, (f1, initial)
    assert h1[-20:].mean() < 0.1 * h1[:5].mean()


def test_bad_arguments_are_rejected(gpu):
    with MLP() as m:
        with pytest.raises(abi.REWLError):
            m.params = np.zeros(10, np.float32)
        bad = np.zeros(M.NP, np.float32)
        bad[3] = np.nan
        with pytest.raises(abi.REWLError):
            m.params = bad
        with pytest.raises(abi.REWLError):
            m.train_raw(np.zeros((4, 4)), np.zeros(4), steps=1, batch=0)
