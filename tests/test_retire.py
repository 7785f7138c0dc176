"""Retirement of finished systems (rewl_params.retire) in the oracle: a system
of a batch ends in the state of its own single-system run, bit for bit, and
the batch's other systems are unaffected by its leaving."""
import numpy as np

from oracle.oracle import Oracle
from rewl_simulator_amd import configs


def run_to_end(ora, cap=5000):
    checks = 0
    while True:
        for _ in range(2):
            ora.sweeps(20)
            ora.exchange()
        done = ora.check()
        checks += 1
        if done or checks >= cap:
            return checks


def test_retired_system_equals_its_own_run():
    systems, params = configs.config4(realisations=4, sizes=(4, 6, 8, 6), bins=64, windows=3, replicas=2,
                                      lnf_final=2e-2)
    params = dict(params, sweeps_per_check=40, sweeps_per_exchange=20, flatness=0.5)
    assert params["retire"] == 1
    batch = Oracle(systems, params)
    batch.adjust_to_range()
    run_to_end(batch)
    per = 3 * 2
    iters = []
    for si, s in enumerate(systems):
        alone = Oracle([s], params)
        alone.adjust_to_range()
        run_to_end(alone)
        for j in range(per):
            g = si * per + j
            assert alone.state(j)["E"] == batch.state(g)["E"], (si, j)
            np.testing.assert_array_equal(alone.spins(j), batch.spins(g))
            np.testing.assert_array_equal(alone.logdos(j).view(np.uint64), batch.logdos(g).view(np.uint64))
            np.testing.assert_array_equal(alone.visits(j), batch.visits(g))
        iters.append(alone.window_state(0, 0)["iteration"])
        alone.close()
    # the systems finished at different times, so retirement took effect
    assert len(set(iters)) > 1
    batch.close()


def test_without_retirement_finished_systems_keep_sweeping():
    systems, params = configs.config4(realisations=4, sizes=(4, 6, 8, 6), bins=64, windows=3, replicas=2,
                                      lnf_final=2e-2)
    params = dict(params, sweeps_per_check=40, sweeps_per_exchange=20, flatness=0.5, retire=0)
    batch = Oracle(systems, params)
    batch.adjust_to_range()
    run_to_end(batch)
    alone = Oracle([systems[0]], params)
    alone.adjust_to_range()
    run_to_end(alone)
    differ = any(alone.state(j)["E"] != batch.state(j)["E"] or
                 not np.array_equal(alone.logdos(j), batch.logdos(j)) for j in range(6))
    assert differ
    batch.close(); alone.close()
