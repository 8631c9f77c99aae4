"""The segment driver's binary-pursuit clip width (full_binary_pursuit.py:
14-115), CPU side: the oracle's restatement against answers fixed by
construction.  Voltage +-1 alternating sample to sample plus a narrow
template at events of both parities: the mean clip is exactly 0 outside the
template (as many +1 as -1 at every offset), below 5 % of the median absolute
voltage (1), so both ends are trimmed to the original width; with the
alternation removed on one channel the ends stay."""
import numpy as np

from spikesorting_fullpursuit_amd.full_binary_pursuit import UnitSummary


def _voltage(C, K, events, flat_channel=None):
    V = np.where(np.arange(K) % 2 == 0, 1.0, -1.0).astype(np.float32)[None, :].repeat(C, axis=0)
    if flat_channel is not None:
        V[flat_channel] = 1.0
    shape = np.zeros(11, dtype=np.float32)
    shape[:] = -6.0 * np.hanning(13)[1:-1]
    for e in events:
        V[:, e - 5:e + 6] += shape
    return V


def _sort_info(C, factor):
    return {'n_channels': C, 'n_samples_per_chan': 61, 'sampling_rate': 40000, 'clip_width': [-7.5e-4, 7.5e-4],
            'max_binary_pursuit_clip_width_factor': factor}


def test_clip_width_trims_to_the_original_width(oracle_mod):
    C, K = 2, 40000
    events = 1000 + 201 * np.arange(100)            # 201 is odd: half the events on each parity
    V = _voltage(C, K, events)
    np.random.seed(0)
    width, starts, stops = oracle_mod.binary_pursuit_clip_width([events], V, _sort_info(C, 2.0))
    # wide window [-60, 61): the front stops at index 30, the back at 90 (never inside [-30, 31))
    np.testing.assert_allclose(width, [-7.5e-4, 7.5e-4], rtol=0, atol=1e-18)
    # the reference's column ranges: start 60 - 30 - 30 = 0, stride last - first = 60, width 61
    np.testing.assert_array_equal(starts, [0, 60])
    np.testing.assert_array_equal(stops, [61, 121])


def test_clip_width_keeps_ends_that_are_not_flat(oracle_mod):
    C, K = 2, 40000
    events = 1000 + 201 * np.arange(100)
    V = _voltage(C, K, events, flat_channel=1)      # channel 1: mean clip 1.0 everywhere off the template
    np.random.seed(0)
    width, starts, stops = oracle_mod.binary_pursuit_clip_width([events], V, _sort_info(C, 2.0))
    np.testing.assert_allclose(width, [-1.5e-3, 1.5e-3], rtol=0, atol=1e-18)
    np.testing.assert_array_equal(starts, [30, 150])    # 60 - 30 - 0, stride 120
    np.testing.assert_array_equal(stops, [91, 211])


def test_clip_width_factor_at_most_one(oracle_mod):
    V = _voltage(3, 5000, [1000])
    width, starts, stops = oracle_mod.binary_pursuit_clip_width([np.array([1000])], V, _sort_info(3, 1.0))
    np.testing.assert_allclose(width, [-7.5e-4, 7.5e-4])
    np.testing.assert_array_equal(starts, [0, 61, 122])
    np.testing.assert_array_equal(stops, [61, 122, 183])


def test_unit_summary_one_unit_per_cluster():
    seg_data = [[np.array([50, 10, 30]), np.array([2, 1, 2]), None, None, 7], [[], [], [], [], 8],
                [np.array([5]), np.array([0]), None, None, 9]]
    s = UnitSummary(seg_data, [], {}, np.float32)
    assert [(u['ID'], int(u['label'])) for u in s.summaries] == [(7, 1), (7, 2), (9, 0)]
    np.testing.assert_array_equal(s.summaries[1]['spike_indices'], [30, 50])
    s.set_bp_templates(['a', 'b', 'c'])
    assert [u['bp_template'] for u in s.summaries] == ['a', 'b', 'c']
