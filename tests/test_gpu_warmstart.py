"""The learned warm start end to end at a small size: curves produced by the
engine, the MLP trained by the library's kernels, and held-out disorder run
from the network's ln g at ln f = 1e-3 against plain flat-init WL from ln f = 1.
Sweep counts are deterministic (Philox streams, fixed seeds), so the gain is
asserted on them; wall times are printed."""
import numpy as np
import pytest

from rewl_simulator_amd import warmstart as W

pytestmark = pytest.mark.gpu


def test_warm_start_cuts_sweeps_on_heldout_disorder(gpu):
    train_specs = W.training_specs(48, sizes=(12, 16), sigmas=(0.0, 0.5, 1.0))
    held = W.heldout_specs(8, L=16, sigma=0.5)
    curves = W.generate(train_specs, batch=48)
    assert all(c["visited"].sum() > 100 for c in curves)
    mlp, hist = W.train(curves, steps=1500, batch=8192)
    assert hist[-50:].mean() < 0.05 * hist[:10].mean()
    res = W.evaluate(mlp, held, max_checks=20000)
    print(res, flush=True)
    plain, warm = res["plain"], res["warm"]
    assert plain["converged"] and warm["converged"]
    # fewer sweeps to ln f < 1e-6, at a comparable ln g error
    assert warm["sweeps"] * 1.5 <= plain["sweeps"], (warm["sweeps"], plain["sweeps"])
    assert warm["lng_error_max"] <= 1.5 * plain["lng_error_max"] + 0.05
