"""Clips around binary pursuit, CPU side: the oracle's get_clips and unit
means against known answers, and the host logic of the post-pursuit output
(unit -> channel assignment, per-channel output lists)."""
import numpy as np
import pytest

from spikesorting_fullpursuit_amd.full_binary_pursuit import assign_units_to_channels, segment_output


def _ramp(C, K):
    # V[c][k] = 1000 c + k: every clip sample is known in closed form
    return (1000.0 * np.arange(C)[:, None] + np.arange(K)[None, :]).astype(np.float32)


def test_oracle_get_clips_known_answers(oracle_mod):
    C, K = 5, 400
    V = _ramp(C, K)
    pd = {'sampling_rate': 40000, 'n_samples': K, 'v_dtype': np.float32}
    cw = [-7.5e-4, 7.5e-4]   # window [-30, 31)
    ev = np.array([10, 29, 30, 100, 200, 368, 369, 390], dtype=np.int64)
    nb = [3, 0, 4]
    clips, valid = oracle_mod.get_clips(pd, V, nb, ev, cw)
    # trimmed from the front while idx - 30 < 0, from the back while idx + 31 >= K
    np.testing.assert_array_equal(valid, [False, False, True, True, True, True, False, False])
    assert clips.shape == (4, 3 * 61)
    for row, k in enumerate([30, 100, 200, 368]):
        for j, c in enumerate(nb):
            expect = 1000.0 * c + np.arange(k - 30, k + 31)
            np.testing.assert_array_equal(clips[row, j * 61:(j + 1) * 61], expect)
    # no events: an empty (0, width) array; every event invalid: None
    e, v = oracle_mod.get_clips(pd, V, nb, np.zeros(0, dtype=np.int64), cw)
    assert e.shape == (0, 3 * 61) and v.size == 0
    n, v = oracle_mod.get_clips(pd, V, nb, np.array([1, 2, 3]), cw)
    assert n is None and not v.any()
    n, v = oracle_mod.get_clips(pd, V, nb, np.array([380, 390]), cw)
    assert n is None and not v.any()


def test_oracle_unit_means_are_numpys(oracle_mod):
    clips = np.array([[1, 2], [3, 4], [10, 20], [5, 7]], dtype=np.float32)
    m = oracle_mod.unit_mean_clips(clips, np.array([0, 0, 2, 0]), 3)
    np.testing.assert_array_equal(m[0], np.float32([3.0, np.float32(13.0) / np.float32(3.0)]))
    assert np.all(np.isnan(m[1]))
    np.testing.assert_array_equal(m[2], [10, 20])


def _unit_case(seed, n_units=6, n_chans=4, N=9, n=300):
    rng = np.random.default_rng(seed)
    labels = rng.integers(0, n_units, n)
    home = rng.integers(0, n_chans, n_units)
    clips = rng.standard_normal((n, n_chans * N)).astype(np.float32)
    for u in range(n_units):   # a unit is large on its home channel
        sel = labels == u
        clips[np.ix_(sel, np.arange(home[u] * N, (home[u] + 1) * N))] *= 8 + u
    thr = rng.uniform(3.0, 5.0, n_chans)
    return clips, labels, thr, home


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_unit_channel_assignment_matches_oracle(oracle_mod, seed):
    clips, labels, thr, home = _unit_case(seed)
    means = oracle_mod.unit_mean_clips(clips, labels, 6)
    got = assign_units_to_channels(labels, means, 4, 9, thr, 4.0)
    ref = oracle_mod.unit_channels(clips, labels, 4, 9, thr, 4.0)
    assert got == ref
    for c, units in got.items():
        for u in units:
            assert home[u] == c


def test_segment_output_lists():
    crossings = np.array([100, 40, 70, 10, 55], dtype=np.int64)
    labels = np.array([1, 0, 1, 2, 0], dtype=np.int64)
    bp_bool = np.ones(5, dtype=bool)
    N = 3
    clips = np.arange(5 * 2 * N, dtype=np.float32).reshape(5, 2 * N)
    items = [{'channel': 0, 'ID': 7, 'index_window': [1000, 2000], 'neighbors': np.array([0, 1])},
             {'channel': 1, 'ID': 8, 'index_window': [1000, 2000], 'neighbors': np.array([1])}]
    assign = {0: [1, 0], 1: []}
    out = segment_output(crossings, labels, bp_bool, clips, assign, items, 2, N, [0, N], [N, 2 * N], np.float32,
                         output_clips=True)
    ev, lab, cl, bb, ident = out[0]
    np.testing.assert_array_equal(ev, [1100, 1070, 1040, 1055])   # unit 1, then unit 0; onset added
    np.testing.assert_array_equal(lab, [1, 1, 0, 0])
    np.testing.assert_array_equal(cl, clips[[0, 2, 1, 4]])
    assert bb.all() and ident == 7
    assert out[1] == [[], [], [], [], 8]
