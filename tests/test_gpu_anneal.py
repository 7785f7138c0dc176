"""Simulated annealing on the GPU (rewl_anneal) against the oracle's
restatement of Simulated_Annealer::simulate_annealing: every chain's lowest
energy, final energy, sweep count and final spins bit for bit, for clean and
disordered Ising and Ashkin-Teller lattices, with a partly filled second
workgroup of chains."""
import numpy as np
import pytest

from oracle import oracle as O
from rewl_simulator_amd import abi, configs
from rewl_simulator_amd.engine import anneal, anneal_params

pytestmark = pytest.mark.gpu


def _system(model, field, L=6):
    N = L * L
    h = configs.gaussian_strain(L, 1.5, seed=21) if field else None
    if model == abi.ISING2D:
        return configs.ising_system(L, -4.0 * N, 4.0 * N, 1.0, 1, 1, field=h)
    return configs.at_system(L, -8.0 * N, 6.0 * N, 1.0, 1, 1, field=h)


@pytest.mark.parametrize("field", [False, True])
@pytest.mark.parametrize("model", [abi.ISING2D, abi.ASHKIN_TELLER2D])
def test_anneal_matches_oracle(gpu, model, field):
    print(f"case: anneal model={model} field={field}", flush=True)
    systems = [_system(model, field), _system(model, field, L=4)]
    p = anneal_params(initial_temperature=8.0, final_temperature=0.2, num_temperatures=10, block_size=30,
                      max_blocks=6, chains=70)
    g = anneal(systems, p)
    o = O.anneal(systems, p)
    for k in ("min_energy", "final_energy", "sweeps", "spins"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
    assert np.all(g["min_energy"] <= g["final_energy"])


@pytest.mark.parametrize("model", [abi.ISING2D, abi.ASHKIN_TELLER2D])
def test_anneal_start_spins(gpu, model):
    """At T = 1e300 every flip is accepted and two sweeps return the start
    spins, which are the Philox draws of randomize_dofs."""
    s = _system(model, False, L=5)
    p = anneal_params(initial_temperature=1e300, final_temperature=1e300, num_temperatures=1, block_size=1,
                      max_blocks=1, chains=3)
    g = anneal([s], p)
    ndof = 25 * (1 if model == abi.ISING2D else 2)
    key = [s["seed"] & 0xFFFFFFFF, s["seed"] >> 32]
    want = [1 if O.philox4x32_10([j >> 2, 0, 5, c], key)[j & 3] < 2 ** 31 else -1
            for c in range(3) for j in range(ndof)]
    np.testing.assert_array_equal(g["spins"], want)


def test_anneal_rejects_bad_parameters(gpu):
    s = _system(abi.ISING2D, False)
    for bad in (dict(chains=0), dict(max_blocks=0), dict(block_size=0), dict(final_temperature=0.0)):
        with pytest.raises(abi.REWLError):
            anneal([s], anneal_params(**bad))


def test_anneal_mixed_models_in_one_call(gpu):
    """Ising and Ashkin-Teller systems in one call (one launch per model,
    each skipping the other's workgroups), with and without fields: every
    chain as the oracle has it."""
    systems = [_system(abi.ASHKIN_TELLER2D, True, L=4), _system(abi.ISING2D, False, L=6),
               _system(abi.ISING2D, True, L=5), _system(abi.ASHKIN_TELLER2D, False, L=3)]
    p = anneal_params(initial_temperature=6.0, final_temperature=0.3, num_temperatures=6, block_size=20,
                      max_blocks=4, chains=5)
    g = anneal(systems, p)
    o = O.anneal(systems, p)
    for k in ("min_energy", "final_energy", "sweeps", "spins"):
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)
