"""The segment driver full_binary_pursuit() (full_binary_pursuit.py:117-438)
on the GPU against the oracle's restatement of the same steps, on seeded
segments whose clustering output is the true spike trains plus a cluster of
random times (a noise unit): binary-pursuit clip width, templates, residual
noise covariance, sum-of-two removal, separability, noise-template removal,
binary pursuit at n_max_shift_inds = N - 1 and the per-channel output.  The
float64 statistics agree to rounding (BLAS order); every decision and the
output lists are equal."""
import numpy as np
import pytest

from spikesorting_fullpursuit_amd import synthetic
from spikesorting_fullpursuit_amd.full_binary_pursuit import UnitSummary, full_binary_pursuit

pytestmark = pytest.mark.gpu


def _segment(seed, factor, n_cov, duration_s=1.0):
    rec = synthetic.config("4ch-60s", seed=seed, duration_s=duration_s)
    V = rec['voltage']
    C, K = V.shape
    tm = rec['templates'].reshape(-1, C, 61)
    home = np.argmax(np.max(np.abs(tm), axis=2), axis=1)
    rng = np.random.default_rng(seed)
    items, results = [], {}
    for c in range(C):
        item_id = 10 + c
        items.append({'channel': c, 'ID': item_id, 'seg_number': 0, 'n_samples': K, 'index_window': [0, K],
                      'neighbors': np.array([c]), 'thresholds': np.full(C, 4.0)})
        ev = [rec['events'][t] for t in range(tm.shape[0]) if home[t] == c]
        lab = [np.full(rec['events'][t].size, t) for t in range(tm.shape[0]) if home[t] == c]
        if c == 0:   # a cluster of random times: a noise unit
            ev.append(np.sort(rng.choice(np.arange(200, K - 200), 60, replace=False)))
            lab.append(np.full(60, 99))
        if ev:
            ev, lab = np.concatenate(ev), np.concatenate(lab)
            order = np.argsort(ev, kind='stable')
            results[item_id] = (ev[order].astype(np.int64), lab[order].astype(np.int64))
    si = {'n_channels': C, 'sampling_rate': 40000, 'clip_width': list(rec['sort_info']['clip_width']),
          'n_samples_per_chan': 61, 'max_binary_pursuit_clip_width_factor': factor, 'n_cov_samples': n_cov,
          'sigma': 4.0, 'sigma_bp_noise': 3.719, 'sigma_bp_CI': 12.0, 'n_samples': K,
          'n_threshold_crossings': np.full(C, 300), 'do_overlap_recheck': True, 'get_adjusted_clips': False,
          'output_separability_metrics': False, 'verbose': False}
    data = {'segment_voltages': [(V.tobytes(), V.shape)], 'results_dict': results}
    return rec, items, data, si


@pytest.mark.parametrize("seed,factor,n_cov", [(3, 1.0, 100000), (4, 1.6, 500), (8, 2.0, 100000)])
def test_segment_driver_matches_oracle(oracle_mod, seed, factor, n_cov):
    rec, items, data, si = _segment(seed, factor, n_cov)
    V = rec['voltage']
    K = V.shape[1]
    seg_data = [[data['results_dict'][w['ID']][0], data['results_dict'][w['ID']][1], None, None, w['ID']]
                for w in items if w['ID'] in data['results_dict']]
    units = [u['spike_indices'] for u in UnitSummary(seg_data, items, si, np.float32).summaries]
    print("case segment driver", seed, factor, n_cov, len(units), "units", flush=True)
    np.random.seed(1000 + seed)
    want, want_sm, want_width = oracle_mod.full_binary_pursuit_units(V, np.float32, dict(si), units, items,
                                                                     chunk_len=10 * K)
    np.random.seed(1000 + seed)
    si_run = dict(si)
    got = full_binary_pursuit(items, data, 0, si_run, np.float32, 0.5, 1.5e-3)
    # sort_info is restored, the width used is recorded
    assert si_run['clip_width'] == si['clip_width'] and si_run['n_samples_per_chan'] == 61
    np.testing.assert_allclose(si_run['binary_pursuit_clip_width'], [min(want_width[0], 0), max(want_width[1], 0)],
                               rtol=0, atol=0)
    for w in items:   # work items got their neighbourhoods back
        assert np.array_equal(w['neighbors'], [w['channel']])
    assert len(got) == len(want) == V.shape[0]
    n_events = 0
    for g, r in zip(got, want):
        assert g[4] == r[4]
        if len(r[0]) == 0:
            assert g[:4] == [[], [], [], []]
            continue
        np.testing.assert_array_equal(g[0], r[0])
        np.testing.assert_array_equal(g[1], r[1])
        np.testing.assert_array_equal(g[3], r[3])
        n_events += len(r[0])
    assert n_events > 0
    # the noise cluster never reaches binary pursuit as its own unit
    assert want_sm['templates'].shape[0] <= len(units)


def test_binary_pursuit_clip_width_trims_flat_ends(oracle_mod):
    """Spikes every 200 samples of a template that is zero outside its middle
    21 samples: the mean clip over the widened window is noise averaged over
    ~5000 clips at both ends, mostly below 5 % of the median absolute
    voltage, so the width is trimmed toward the original one (never inside
    it), identically on the GPU path and the oracle."""
    from spikesorting_fullpursuit_amd.full_binary_pursuit import get_binary_pursuit_clip_width
    rng = np.random.default_rng(21)
    C, K, fs, N = 4, 1000000, 40000, 61
    V = rng.standard_normal((C, K)).astype(np.float32)
    shape = np.zeros(N)
    shape[20:41] = -8 * np.hanning(21)
    ev = np.arange(300, K - 300, 200, dtype=np.int64)
    for c, g in enumerate([1.0, 0.6, 0.3, 0.1]):
        for e in ev:
            V[c, e - 30:e + 31] += (g * shape).astype(np.float32)
    items = [{'channel': c, 'ID': c, 'seg_number': 0, 'n_samples': K, 'index_window': [0, K],
              'neighbors': np.array([c]), 'thresholds': np.full(C, 4.0)} for c in range(C)]
    data = {'results_dict': {0: (ev, np.zeros(ev.size, dtype=np.int64))}}
    si = {'n_channels': C, 'n_samples_per_chan': N, 'sampling_rate': fs, 'clip_width': [-7.5e-4, 7.5e-4],
          'max_binary_pursuit_clip_width_factor': 2.0}
    print("case clip width", flush=True)
    np.random.seed(5)
    w, starts, stops = oracle_mod.binary_pursuit_clip_width([ev], V, si)
    np.random.seed(5)
    gw, gstarts, gstops = get_binary_pursuit_clip_width(items, {'sampling_rate': fs, 'n_samples': K,
                                                                'v_dtype': np.float32}, V, data, si)
    assert gw == w
    np.testing.assert_array_equal(gstarts, starts)
    np.testing.assert_array_equal(gstops, stops)
    assert 7.5e-4 <= w[1] < 1.5e-3   # trimmed, but not inside the original width
    np.testing.assert_array_equal(stops - starts, N)
