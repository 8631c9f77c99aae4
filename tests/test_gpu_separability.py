"""Threshold inputs of binary pursuit on the GPU (fbp_channel_covariance,
fbp_template_variances and the separability-metrics mirror) against the
oracle.  The reference evaluates both with numpy/BLAS, whose summation order
is not fixed, so parity is to float64 rounding: relative 1e-12."""
import numpy as np
import pytest

from spikesorting_fullpursuit_amd import _lib, synthetic
from spikesorting_fullpursuit_amd.neuron_separability import (clip_noise_covariance, compute_separability_metrics,
                                                              set_bp_threshold)

pytestmark = pytest.mark.gpu

RTOL = 1e-12


@pytest.mark.parametrize("C,N,n,n_rows", [(4, 61, 3000, 1000), (16, 61, 5000, None), (2, 128, 700, 4096),
                                          (3, 7, 1, None)])
def test_channel_covariance_matches_numpy(oracle_mod, C, N, n, n_rows):
    rng = np.random.default_rng(C * 1000 + N)
    clips = (rng.standard_normal((n, C * N)) * 50 + 3).astype(np.float32)
    rows = None if n_rows is None else rng.integers(0, n, (C, n_rows))
    print("case covariance", C, N, n, n_rows, flush=True)
    got = _lib.channel_covariance(clips, C, N, rows)
    want = oracle_mod.channel_covariance(clips, C, N, rows)
    scale = np.max(np.abs(want)) if want.size else 1.0
    np.testing.assert_allclose(got, want, rtol=RTOL, atol=RTOL * scale)
    for c in range(C):   # exactly symmetric, and the same bits on a second call
        np.testing.assert_array_equal(got[c], got[c].T)
    np.testing.assert_array_equal(got.view(np.uint64), _lib.channel_covariance(clips, C, N, rows).view(np.uint64))


def test_template_variances_and_metrics_match_oracle(oracle_mod):
    rec = synthetic.config("16ch-300s", seed=3, duration_s=0.2)
    tm = rec['separability_metrics']['templates'].astype(np.float64)
    T, CN = tm.shape
    C, N = 16, CN // 16
    rng = np.random.default_rng(11)
    covs = []
    for _ in range(C):
        a = rng.standard_normal((N, N)) * 0.1
        covs.append(a @ a.T + np.eye(N))
    si = {'n_channels': C, 'n_samples_per_chan': N, 'n_samples': 12000, 'n_threshold_crossings': np.arange(C) * 7,
          'sigma_bp_noise': 3.719, 'sigma_bp_CI': 6.0}
    print("case separability", T, C, N, flush=True)
    got = compute_separability_metrics(list(tm), covs, si)
    want = oracle_mod.separability_metrics(list(tm), covs, si)
    np.testing.assert_allclose(got['neuron_variances'], want['neuron_variances'], rtol=RTOL)
    for key in ('neuron_lower_thresholds', 'neuron_lower_CI'):
        np.testing.assert_allclose(got[key], want[key], rtol=RTOL, atol=RTOL * np.max(np.abs(want[key])))
    np.testing.assert_allclose(got['contamination'], want['contamination'], rtol=1e-9, atol=1e-300)
    for key in ('template_SS', 'peak_channel', 'channel_p_noise'):
        np.testing.assert_array_equal(got[key], want[key])
    assert got['templates'].shape == (T, CN) and got['template_SS_by_chan'].shape == (T, C)
    # identity covariance: the variance is the template's sum of squares
    ident = compute_separability_metrics(list(tm), [np.eye(N)] * C, si)
    np.testing.assert_allclose(ident['neuron_variances'], np.sum(tm ** 2, axis=1), rtol=1e-13)
    before = got['neuron_lower_thresholds'].copy()
    thr = set_bp_threshold(got)['neuron_lower_thresholds']
    np.testing.assert_array_equal(thr, np.maximum(np.maximum(before, 0), got['neuron_lower_CI']))


def test_clip_noise_covariance_draws_like_the_reference(oracle_mod):
    rng = np.random.default_rng(21)
    C, N, n = 5, 61, 4000
    clips = rng.standard_normal((n, C * N)).astype(np.float32)
    print("case noise covariance", C, N, n, flush=True)
    got = clip_noise_covariance(clips, C, N, 1500, rng=np.random.RandomState(9))
    draw = np.random.RandomState(9)
    rows = np.stack([draw.randint(0, n, 1500) for _ in range(C)])
    want = oracle_mod.channel_covariance(clips, C, N, rows)
    for c in range(C):
        np.testing.assert_allclose(got[c], want[c], rtol=RTOL, atol=RTOL * np.max(np.abs(want[c])))
    every = clip_noise_covariance(clips, C, N, n + 1)
    np.testing.assert_allclose(np.stack(every), oracle_mod.channel_covariance(clips, C, N, None), rtol=RTOL,
                               atol=RTOL * 2)
