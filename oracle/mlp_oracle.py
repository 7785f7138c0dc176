"""CPU oracle of the warm-start MLP (include/rewl_mlp.h), FP64 numpy.

Test infrastructure only: the GPU kernels are checked against it.  The
reference has no network (every walker starts from flat ln g,
wang_landau.hpp:35-40); this module restates the network of rewl_mlp.h:

    z1 = W1 x + b1,          a1 = r(gelu(z1))
    z2 = r(W2) a1 + b2,      a2 = r(gelu(z2))
    z3 = r(W3) a2 + b3,      a3 = gelu(z3)
    y  = r(w4) . a3 + b4     (layer 1 also uses r(W1))

with r = round to bf16 (nearest even) when ``emulate`` is set and the
identity otherwise, and the loss sum_i w_i (y_i - t_i)^2 / sum_i w_i.  The
backward pass rounds dL/dz3 and dL/dz2 with r before the products that the
GPU runs on bf16 operands.  With ``emulate=False`` it is the exact gradient of
the FP64 network, which tests pin against central finite differences.
"""
import numpy as np
from scipy.special import erf

NIN, H = 4, 128
O_W1 = 0
O_B1 = O_W1 + H * NIN
O_W2 = O_B1 + H
O_B2 = O_W2 + H * H
O_W3 = O_B2 + H
O_B3 = O_W3 + H * H
O_W4 = O_B3 + H
O_B4 = O_W4 + H
NP = O_B4 + 1


def bf16(a):
    """Round float values to bf16 (nearest even) and return them as float64."""
    f = np.asarray(a, dtype=np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(theta):
    th = np.asarray(theta, dtype=np.float64)
    assert th.size == NP
    return dict(W1=th[O_W1:O_B1].reshape(H, NIN), b1=th[O_B1:O_W2], W2=th[O_W2:O_B2].reshape(H, H),
                b2=th[O_B2:O_W3], W3=th[O_W3:O_B3].reshape(H, H), b3=th[O_B3:O_W4], w4=th[O_W4:O_B4],
                b4=th[O_B4])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(theta, x, emulate=True):
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    p = unpack(theta)
    x = np.asarray(x, dtype=np.float64).reshape(-1, NIN)
    # float32 inputs, as the GPU reads them
    if emulate:
        x = x.astype(np.float32).astype(np.float64)
    W1, W2, W3, w4 = r(p["W1"]), r(p["W2"]), r(p["W3"]), r(p["w4"])
    z1 = x @ W1.T + p["b1"]
    a1 = r(gelu(z1))
    z2 = a1 @ W2.T + p["b2"]
    a2 = r(gelu(z2))
    z3 = a2 @ W3.T + p["b3"]
    a3 = gelu(z3)
    y = a3 @ w4 + p["b4"]
    cache = dict(x=x, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, a3=a3, W1=W1, W2=W2, W3=W3, w4=w4)
    return y, cache


def gradient(theta, x, t, w=None, emulate=True):
    """(grad over the parameter vector, loss) of the weighted MSE."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    y, c = forward(theta, x, emulate)
    t = np.asarray(t, dtype=np.float64)
    w = np.ones_like(t) if w is None else np.asarray(w, dtype=np.float64)
    ws = w.sum()
    d = y - t
    loss = float((w * d * d).sum() / ws)
    dy = 2.0 * w * d / ws
    g = np.zeros(NP)
    g[O_W4:O_B4] = dy @ c["a3"]
    g[O_B4] = dy.sum()
    dz3 = dy[:, None] * c["w4"][None, :] * gelu_grad(c["z3"])
    g[O_B3:O_W4] = dz3.sum(0)
    dz3r = r(dz3)
    g[O_W3:O_B3] = (dz3r.T @ c["a2"]).ravel()
    dz2 = (dz3r @ c["W3"]) * gelu_grad(c["z2"])
    g[O_B2:O_W3] = dz2.sum(0)
    dz2r = r(dz2)
    g[O_W2:O_B2] = (dz2r.T @ c["a1"]).ravel()
    dz1 = (dz2r @ c["W2"]) * gelu_grad(c["z1"])
    g[O_B1:O_W2] = dz1.sum(0)
    g[O_W1:O_B1] = (dz1.T @ c["x"]).ravel()
    return g, loss


def loss(theta, x, t, w=None, emulate=False):
    y, _ = forward(theta, x, emulate)
    t = np.asarray(t, dtype=np.float64)
    w = np.ones_like(t) if w is None else np.asarray(w, dtype=np.float64)
    return float((w * (y - t) ** 2).sum() / w.sum())


def init_params(seed):
    """He-uniform initialisation of rewl_mlp_create: one splitmix64 stream,
    u = (z >> 11) 2^-53, value (2u - 1) sqrt(6 / fan_in) rounded to float32;
    W1, W2, W3 (fan_in 4, 128, 128), then w4 (fan_in 512); biases zero."""
    M = (1 << 64) - 1
    st = [seed & M]

    def uni():
        st[0] = (st[0] + 0x9E3779B97F4A7C15) & M
        z = st[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        return (z >> 11) * 2.0 ** -53

    th = np.zeros(NP, dtype=np.float32)

    def fill(off, count, fan_in):
        a = np.sqrt(6.0 / fan_in)
        th[off:off + count] = [(2.0 * uni() - 1.0) * a for _ in range(count)]

    fill(O_W1, H * NIN, NIN)
    fill(O_W2, H * H, H)
    fill(O_W3, H * H, H)
    fill(O_W4, H, 4 * H)
    return th
