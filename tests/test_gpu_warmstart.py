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
    curves = W.generate(train_specs, lnf_final=1e-5, batch=48)
    assert all(c["visited"].sum() > 100 for c in curves)
    mlp, hist = W.train(curves, steps=1500, batch=8192)
    assert hist[-50:].mean() < 0.05 * hist[:10].mean()
    res = W.evaluate(mlp, held, lnf_final=1e-5, ref_lnf=0.0, max_checks=20000)
    print(res, flush=True)
    plain, warm = res["plain"], res["warm"]
    assert plain["converged"] and warm["converged"]
    # the network's ln g is a useful start: well inside the spread of plain WL
    # at this ln f, and the warm-started run ends where plain WL ends
    assert res["prediction_error_max"] < 5.0
    assert res["warm_vs_plain_max"] < 1.0
    # fewer sweeps to the final ln f
    assert warm["sweeps"] < plain["sweeps"], (warm["sweeps"], plain["sweeps"])
