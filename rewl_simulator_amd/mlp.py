"""Warm-start MLP on the GPU (include/rewl_mlp.h) and the warm start itself.

The network maps features (E/N, 1/L, K, sigma) to ln g(E)/N.  ``MLP`` wraps
the C ABI (training and inference run in the library's HIP kernels);
``features`` builds the network input from a system description and
``warm_start`` seeds an engine's windows from a prediction.
"""
import ctypes as C

import numpy as np

from . import abi

NUM_PARAMS = abi.MLP_NUM_PARAMS


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    def __init__(self, seed=1):
        self.lib = abi.load_engine()
        h = C.c_void_p()
        abi.check(self.lib.rewl_mlp_create(C.c_uint64(seed), C.byref(h)))
        self.h = h
        # feature standardisation and target shift, applied on the host
        self.x_shift = np.zeros(4, np.float64)
        self.x_scale = np.ones(4, np.float64)
        self.t_shift = 0.0

    def close(self):
        if self.h:
            self.lib.rewl_mlp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        th = np.empty(NUM_PARAMS, np.float32)
        abi.check(self.lib.rewl_mlp_get_params(self.h, _fptr(th), NUM_PARAMS))
        return th

    @params.setter
    def params(self, th):
        th = np.ascontiguousarray(th, dtype=np.float32)
        abi.check(self.lib.rewl_mlp_set_params(self.h, _fptr(th), th.size))

    def last_ms(self):
        v = C.c_double()
        abi.check(self.lib.rewl_mlp_last_ms(self.h, C.byref(v)))
        return v.value

    # raw network (no standardisation)
    def forward_raw(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 4)
        y = np.empty(len(x), np.float32)
        abi.check(self.lib.rewl_mlp_forward(self.h, _fptr(x), len(x), _fptr(y)))
        return y

    def gradient_raw(self, x, t, w=None):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 4)
        t = np.ascontiguousarray(t, dtype=np.float32)
        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, dtype=np.float32)
            wp = _fptr(w)
        g = np.empty(NUM_PARAMS, np.float32)
        loss = C.c_double()
        abi.check(self.lib.rewl_mlp_gradient(self.h, _fptr(x), _fptr(t), wp, len(x), _fptr(g), C.byref(loss)))
        return g, loss.value

    def train_raw(self, x, t, w=None, steps=1000, batch=8192, lr=2e-3, lr_final=2e-5, beta1=0.9, beta2=0.999,
                  eps=1e-8, seed=7, warmup=50):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 4)
        t = np.ascontiguousarray(t, dtype=np.float32)
        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, dtype=np.float32)
            wp = _fptr(w)
        p = abi.MlpTrainParams(steps=steps, batch=batch, lr=lr, lr_final=lr_final, beta1=beta1, beta2=beta2,
                               eps=eps, seed=seed, warmup=warmup)
        hist = np.empty(max(steps, 1), np.float32)
        abi.check(self.lib.rewl_mlp_train(self.h, _fptr(x), _fptr(t), wp, len(x), C.byref(p), _fptr(hist)))
        return hist[:steps]

    # standardised interface
    def fit_standardisation(self, x, t=None):
        x = np.asarray(x, np.float64).reshape(-1, 4)
        self.x_shift = x.mean(0)
        sd = x.std(0)
        self.x_scale = np.where(sd > 0, sd, 1.0)
        if t is not None:
            self.t_shift = float(np.mean(t))

    def _std(self, x):
        return (np.asarray(x, np.float64).reshape(-1, 4) - self.x_shift) / self.x_scale

    def predict(self, x):
        return self.forward_raw(self._std(x)).astype(np.float64) + self.t_shift

    def train(self, x, t, w=None, **kw):
        return self.train_raw(self._std(x), np.asarray(t, np.float64) - self.t_shift, w, **kw)

    def save(self, path):
        np.savez(path, params=self.params, x_shift=self.x_shift, x_scale=self.x_scale, t_shift=self.t_shift)

    def load(self, path):
        d = np.load(path)
        self.params = d["params"]
        self.x_shift, self.x_scale, self.t_shift = d["x_shift"], d["x_scale"], float(d["t_shift"])


def features(E, N, L, K, sigma):
    """Network input rows (E/N, 1/L, K, sigma) for energies E of one system."""
    E = np.asarray(E, np.float64)
    out = np.empty((E.size, 4))
    out[:, 0] = E / N
    out[:, 1] = 1.0 / L
    out[:, 2] = K
    out[:, 3] = sigma
    return out


def window_bin_energies(window):
    """Lower edge of every bin of a window (the energies the stitched ln g is
    reported at)."""
    return window["e_min"] + window["bin_size"] * np.arange(window["num_bins"])


def warm_start(eng, mlp, sigma_of, lnf):
    """Seed every walker's ln g from the network's prediction at its window's
    bin energies (lower edges, as the stitched curves the network is trained
    on report them), and start every window at ln f = lnf: a few host calls
    (rewl_mlp_forward, rewl_set_logdos, rewl_set_window_lnf) per window.

    ``sigma_of(system_index)`` gives the disorder strength of each system."""
    for k, win in enumerate(eng.windows):
        s = eng.systems[win["system"]]
        L = s["L"]
        N = L * L
        Eb = window_bin_energies(win)
        lng = N * mlp.predict(features(Eb, N, L, s["K"], sigma_of(win["system"])))
        lng = lng - lng[0]
        R = s["replicas_per_window"]
        for r in range(R):
            eng.set_logdos(win["first_walker"] + r, lng)
        eng.set_window_lnf(k, lnf)
