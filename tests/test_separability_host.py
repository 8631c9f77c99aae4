"""Threshold inputs of binary pursuit, CPU side: the oracle's channel
covariance and separability metrics against the definitions (brute force,
closed forms), the host threshold rule, and argument checks of the C-ABI
entry points (rejected before any GPU call)."""
import numpy as np
import pytest
from scipy.stats import norm

from spikesorting_fullpursuit_amd import _lib
from spikesorting_fullpursuit_amd.neuron_separability import set_bp_threshold


def _brute_cov(x):
    # sum_r (x_ri - m_i)(x_rj - m_j) / n with explicit loops (ddof = 0)
    n, N = x.shape
    m = [sum(float(x[r, j]) for r in range(n)) / n for j in range(N)]
    out = np.zeros((N, N))
    for i in range(N):
        for j in range(N):
            out[i, j] = sum((float(x[r, i]) - m[i]) * (float(x[r, j]) - m[j]) for r in range(n)) / n
    return out


def test_oracle_channel_covariance_is_the_ddof0_definition(oracle_mod):
    rng = np.random.default_rng(3)
    C, N, n = 3, 4, 9
    clips = rng.standard_normal((n, C * N)).astype(np.float32)
    rows = rng.integers(0, n, (C, 6))
    got = oracle_mod.channel_covariance(clips, C, N, rows)
    for c in range(C):
        want = _brute_cov(clips[rows[c], c * N:(c + 1) * N])
        np.testing.assert_allclose(got[c], want, rtol=1e-12, atol=1e-14)
    every = oracle_mod.channel_covariance(clips, C, N, None)
    np.testing.assert_allclose(every[1], _brute_cov(clips[:, N:2 * N]), rtol=1e-12, atol=1e-14)


def _sort_info(C, N, sigma_noise=3.719, sigma_ci=np.inf):
    return {'n_channels': C, 'n_samples_per_chan': N, 'n_samples': 1000,
            'n_threshold_crossings': np.arange(C) * 10, 'sigma_bp_noise': sigma_noise, 'sigma_bp_CI': sigma_ci}


def test_oracle_separability_identity_covariance_closed_form(oracle_mod):
    # white noise: cov_c = I, so var = SS, thr = -SS/2 + s*sqrt(SS) and the
    # contamination is the standard normal tail at s
    rng = np.random.default_rng(4)
    C, N, T = 3, 5, 4
    tm = rng.standard_normal((T, C * N))
    si = _sort_info(C, N)
    sm = oracle_mod.separability_metrics(list(tm), [np.eye(N)] * C, si)
    ss = np.sum(tm ** 2, axis=1)
    np.testing.assert_allclose(sm['neuron_variances'], ss, rtol=1e-14)
    np.testing.assert_allclose(sm['neuron_lower_thresholds'], -0.5 * ss + 3.719 * np.sqrt(ss), rtol=1e-13)
    np.testing.assert_allclose(sm['contamination'], norm.sf(3.719), rtol=1e-9)
    assert np.all(sm['neuron_lower_CI'] == -np.inf)
    np.testing.assert_array_equal(sm['channel_p_noise'], (1000 - np.arange(C) * 10) / 1000)


def test_oracle_separability_brute_force_quadratic_forms(oracle_mod):
    rng = np.random.default_rng(5)
    C, N, T = 2, 4, 3
    tm = rng.standard_normal((T, C * N))
    covs = []
    for _ in range(C):
        a = rng.standard_normal((N, N))
        covs.append(a @ a.T + N * np.eye(N))
    sm = oracle_mod.separability_metrics(list(tm), covs, _sort_info(C, N, 1.645, 2.0))
    for n in range(T):
        v = sum(tm[n, c * N + i] * covs[c][i, j] * tm[n, c * N + j]
                for c in range(C) for i in range(N) for j in range(N))
        assert sm['neuron_variances'][n] == pytest.approx(v, rel=1e-12)
        e = 0.5 * np.sum(tm[n] ** 2)
        assert sm['neuron_lower_CI'][n] == pytest.approx(e - 2.0 * np.sqrt(v), rel=1e-12)
        assert sm['peak_channel'][n] == np.argmax(np.abs(tm[n])) // N


def test_peak_channel_known_answer(oracle_mod):
    C, N = 3, 4
    tm = np.zeros((2, C * N))
    tm[0, 2 * N + 1] = -5.0   # channel 2
    tm[0, 0] = 4.0
    tm[1, N + 3] = 1.0        # channel 1
    sm = oracle_mod.separability_metrics(list(tm), [np.eye(N)] * C, _sort_info(C, N))
    np.testing.assert_array_equal(sm['peak_channel'], [2, 1])


def test_set_bp_threshold_rule():
    sm = {'neuron_lower_thresholds': np.array([-1.0, 2.0, 3.0, -4.0]),
          'neuron_lower_CI': np.array([-2.0, 5.0, 1.0, 0.5])}
    set_bp_threshold(sm)
    np.testing.assert_array_equal(sm['neuron_lower_thresholds'], [0.0, 5.0, 3.0, 0.5])


def test_covariance_entry_points_reject_bad_arguments():
    clips = np.zeros((4, 2 * 129), dtype=np.float32)
    with pytest.raises(ValueError, match="samples per channel"):
        _lib.channel_covariance(clips, 2, 129)
    clips = np.zeros((4, 2 * 3), dtype=np.float32)
    with pytest.raises(ValueError, match="out of range"):
        _lib.channel_covariance(clips, 2, 3, rows=np.array([[0, 1], [2, 4]]))
    with pytest.raises(ValueError, match="no clip rows"):
        _lib.channel_covariance(np.zeros((0, 6), dtype=np.float32), 2, 3)
    with pytest.raises(ValueError, match="samples per channel"):
        _lib.template_variances(np.zeros((1, 2 * 129)), [np.eye(129)] * 2)
