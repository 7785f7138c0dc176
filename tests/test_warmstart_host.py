"""Host logic of the warm-start pipeline (no GPU): the visited-bin join of
window curves, normalisation, and the network's input features."""
import numpy as np
import pytest

from rewl_simulator_amd import warmstart as W
from rewl_simulator_amd.mlp import features


def _windows_of(curve, spans, shifts, holes):
    """Cut a known curve into overlapping windows, each carrying its own
    additive constant (as WL ln g does) and unvisited bins."""
    out = []
    for (a, b), c in zip(spans, shifts):
        lk = curve[a:b] + c
        vk = np.ones(b - a, bool)
        vk[[h - a for h in holes if a <= h < b]] = False
        lk = np.where(vk, lk, -1e4)   # an unvisited bin's ln g is no value at all
        out.append((a, lk, vk))
    return out


@pytest.mark.parametrize("holes", [[], [5, 6, 40, 41, 42, 77], list(range(1, 120, 3))])
def test_stitch_recovers_curve_up_to_a_constant(holes):
    x = np.linspace(0, 1, 120)
    curve = 300 * x * (1 - 0.3 * x) + 5 * np.sin(7 * x)
    spans = [(0, 40), (25, 70), (55, 100), (85, 120)]
    shifts = [0.0, -812.25, 33.5, 1e3]
    E, lng, vis = W.stitch_visited(_windows_of(curve, spans, shifts, holes), -10.0, 0.5)
    expect = np.ones(120, bool)
    expect[holes] = False
    np.testing.assert_array_equal(vis, expect)
    d = (lng - curve)[vis]
    assert np.max(np.abs(d - d[0])) < 1e-9
    assert np.all(np.isnan(lng[~vis]))
    np.testing.assert_allclose(E, -10.0 + 0.5 * np.arange(120))


def test_stitch_refuses_disjoint_windows():
    curve = np.arange(20.0)
    ws = [(0, curve[:10], np.ones(10, bool)), (10, curve[10:], np.ones(10, bool))]
    with pytest.raises(ValueError):
        W.stitch_visited(ws, 0.0, 1.0)


def test_normalise_counts_states_per_site():
    N = 9
    lng = np.array([0.0, 2.0, np.nan, 5.0, 1.0])
    vis = ~np.isnan(lng)
    out = W.normalise(lng, vis, N)
    total = np.log(np.exp(out[vis]).sum())
    assert total == pytest.approx(N * np.log(W.STATES_PER_SITE), rel=1e-12)
    np.testing.assert_allclose(np.diff(out[[0, 1, 3, 4]]), np.diff(lng[[0, 1, 3, 4]]))


def test_features_layout():
    f = features(np.array([-64.0, 0.0]), 64, 8, 0.5, 0.25)
    np.testing.assert_array_equal(f, [[-1.0, 0.125, 0.5, 0.25], [0.0, 0.125, 0.5, 0.25]])


def test_rfat_systems_half_range():
    s = W.rfat_systems([(8, 0.5, 3)], bins=512)[0]
    assert s["e_max"] == 0.0
    assert s["bin_size"] == pytest.approx(-s["e_min"] / 512)
    assert s["num_windows"] == 8 and s["replicas_per_window"] == 4


def test_record_stages_first_time_at_or_below_each_power_of_two():
    reach = np.full((4, 21), -1, dtype=np.int64)
    # ln f = 1, 2^-1 exactly, 1e-3 (between 2^-10 and 2^-9), below 1e-6
    W.record_stages(reach, [1.0, 0.5, 1e-3, 5e-7], 250)
    assert reach[0].tolist() == [250] + [-1] * 20
    assert reach[1].tolist() == [250, 250] + [-1] * 19
    assert reach[2].tolist() == [250] * 10 + [-1] * 11      # 2^-9 = 1.95e-3 >= 1e-3 > 2^-10
    assert reach[3].tolist() == [250] * 21                  # clipped at 2^-20 < 1e-6
    # later checks fill only the stages newly reached; earlier entries stay
    W.record_stages(reach, [0.25, 0.5, 2.0 ** -10, 5e-7], 500)
    assert reach[0, :3].tolist() == [250, 500, 500] and reach[0, 3] == -1
    assert reach[1].tolist() == [250, 250] + [-1] * 19
    assert reach[2, 9] == 250 and reach[2, 10] == 500 and reach[2, 11] == -1


def test_ising_training_layout_and_features():
    systems = W.ising_systems((16, 48))
    assert [s["L"] for s in systems] == [16, 48]
    assert [s["num_windows"] for s in systems] == [8, 24]
    for s in systems:
        N = s["L"] ** 2
        assert s["bin_size"] == 8.0 and s["e_min"] == -2.0 * N and s["e_max"] == 2.0 * N + 8.0
        assert s["K"] == 0.0
    # clean-Ising curves carry K = 0 into the network's inputs; the AT curves
    # (no "K" key) keep the Ashkin-Teller coupling
    E = np.array([-512.0, -504.0])
    x, t = W.dataset([dict(L=16, K=0.0, sigma=0.0, E=E, lng=np.array([0.7, 6.2]), visited=np.array([True, True])),
                      dict(L=16, sigma=0.5, E=E, lng=np.array([1.0, 2.0]), visited=np.array([True, False]))])
    assert x.shape == (3, 4) and (x[:2, 2] == 0.0).all() and x[2, 2] == W.K_AT
    assert np.allclose(t, [0.7 / 256, 6.2 / 256, 1.0 / 256])


def test_config5_layout_and_bin_width():
    specs, systems, params = W.config5(n=3, bins=1024)
    assert [s[0] for s in specs] == [48, 48, 48] and all(s[1] == 0.5 for s in specs)
    for s in systems:
        assert s["num_windows"] == 8 and s["replicas_per_window"] == 4 and s["e_max"] == 0.0
        # 1024 bins over [E(all-up), 0) are narrower than one flipped spin (12 +- 2 h_i)
        assert 11.0 < s["bin_size"] < 11.5
    _, wide, _ = W.config5(n=1)
    assert 22.0 < wide[0]["bin_size"] < 23.0     # the stated 512 bins: two excitations wide
    assert params["lnf_final"] == 1e-6
