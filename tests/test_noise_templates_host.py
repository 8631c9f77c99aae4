"""Noise-template pruning before binary pursuit, CPU side
(neuron_separability.py:155-250, consolidate.py:9-71): the oracle's
restatement pinned by constructions whose answers are known without it
(a template against a shifted copy of itself, disjoint channel supports, the
white-noise closed form of the threshold), and the host halves of the
product (the noisy rule, deletion, shift and truncation, the C-ABI argument
checks)."""
import numpy as np
import pytest
from scipy.stats import norm

from spikesorting_fullpursuit_amd import _lib
from spikesorting_fullpursuit_amd.consolidate import alignment_lags, best_lag, shift_truncate
from spikesorting_fullpursuit_amd.neuron_separability import delete_noise_units, find_noisy_templates


def _shifted_copy(t, C, d):
    """Every channel of t moved by d samples: out[c][j] = t[c][j + d] (0 outside)."""
    N = t.size // C
    a = t.reshape(C, N)
    out = np.zeros_like(a)
    if d >= 0:
        out[:, :N - d] = a[:, d:]
    else:
        out[:, -d:] = a[:, :N + d]
    return out.reshape(-1)


@pytest.mark.parametrize("C,N,d", [(4, 61, 3), (4, 61, -5), (2, 31, 7), (3, 16, -1), (1, 40, 12)])
def test_oracle_align_recovers_a_known_shift(oracle_mod, C, N, d):
    rng = np.random.default_rng(C * 100 + N + d)
    t1 = rng.standard_normal(C * N)
    t2 = _shifted_copy(t1, C, d)
    s1, s2, lag, width = oracle_mod.optimal_align_templates(t1, t2, C)
    assert lag == d and width == N - abs(d)
    # after alignment the two templates are the same samples
    np.testing.assert_array_equal(s1, s2)
    # absolute alignment finds the same lag for the negated copy; plain does not
    _, _, lag_abs, _ = oracle_mod.optimal_align_templates(t1, -t2, C, align_abs=True)
    _, _, lag_neg, _ = oracle_mod.optimal_align_templates(t1, -t2, C, align_abs=False)
    assert lag_abs == d and lag_neg != d
    # a search window that excludes the lag finds another one inside it
    m = abs(d) - 1
    _, _, lag_in, _ = oracle_mod.optimal_align_templates(t1, t2, C, max_shift=m)
    assert abs(lag_in) <= m
    # zero padding: full length, the aligned samples in place, zeros elsewhere
    p1, p2, lag_p, n_p = oracle_mod.optimal_align_templates(t1, t2, C, zero_pad=True)
    assert lag_p == d and n_p == N and p1.shape == (C * N,)
    cols = np.arange(N)
    inside = (cols >= d) if d >= 0 else (cols < N + d)
    for c in range(C):
        np.testing.assert_array_equal(p1[c * N:(c + 1) * N][~inside], 0)
        np.testing.assert_array_equal(p1[c * N:(c + 1) * N][inside], s1[c * width:(c + 1) * width])
        np.testing.assert_array_equal(p2[c * N:(c + 1) * N][inside], s2[c * width:(c + 1) * width])


def test_oracle_align_zero_lag_and_errors(oracle_mod):
    t = np.arange(12.0)
    a, b, lag, n = oracle_mod.optimal_align_templates(t, t, 3)
    assert lag == 0 and a is t and b is t and n == 4
    with pytest.raises(ValueError):
        oracle_mod.optimal_align_templates(t, t[:-1], 3)
    with pytest.raises(ValueError):
        oracle_mod.optimal_align_templates(t, t, 5)


@pytest.mark.parametrize("C,N,d,zero_pad", [(4, 61, 3, False), (4, 61, -5, True), (2, 31, 7, True),
                                            (3, 16, -1, False)])
def test_host_shift_truncate_equals_oracle(oracle_mod, C, N, d, zero_pad):
    rng = np.random.default_rng(7 + N)
    t1 = rng.standard_normal(C * N)
    t2 = _shifted_copy(t1, C, d) + 0.01 * rng.standard_normal(C * N)
    w1, w2, lag, n = oracle_mod.optimal_align_templates(t1, t2, C, zero_pad=zero_pad)
    assert lag == d
    g1, g2, gn = shift_truncate(t1, t2, C, d, zero_pad)
    np.testing.assert_array_equal(g1, w1)
    np.testing.assert_array_equal(g2, w2)
    assert gn == n


def test_host_lag_helpers():
    assert alignment_lags(244, None) == (-243, 243)
    assert alignment_lags(244, 14) == (-14, 14)
    assert alignment_lags(10, 50) == (-9, 9)
    with pytest.raises(ValueError):
        alignment_lags(244, -1)
    xc = np.array([[1.0, -3.0, 3.0, 2.0], [0.5, 0.5, -0.5, 0.0]])
    np.testing.assert_array_equal(best_lag(xc, -2, True), [-1, -2])    # first of equal maxima
    np.testing.assert_array_equal(best_lag(xc, -2, False), [0, -2])


def _white_metrics(tm, sigma=3.719):
    # identity noise covariance: var = SS, thr = -SS/2 + sigma sqrt(SS)
    ss = np.sum(tm ** 2, axis=1)
    return {'templates': tm, 'template_SS': ss, 'neuron_variances': ss.copy(),
            'neuron_lower_thresholds': -0.5 * ss + sigma * np.sqrt(ss),
            'neuron_lower_CI': 0.5 * ss - 6.0 * np.sqrt(ss)}


def test_noisy_rule_closed_form(oracle_mod):
    # noisy iff 2(-SS/2 + s sqrt(SS)) > SS/2, i.e. sqrt(SS) < 4 s / 3
    s = 3.719
    edge = 4 * s / 3
    tm = np.zeros((4, 8))
    tm[:, 0] = [edge * 0.9, edge * 1.1, edge * 0.5, edge * 3]
    sm = _white_metrics(tm, s)
    want = np.array([True, False, True, False])
    np.testing.assert_array_equal(oracle_mod.find_noisy_templates(sm), want)
    np.testing.assert_array_equal(find_noisy_templates(sm, {}), want)


def test_oracle_noise_check_closed_forms(oracle_mod):
    """Three noisy templates against one good template, white noise:
    * 0.9 x the good template: E = 0.9 SS - SS/2 lies above the good
      template's threshold, so the good unit would take its spikes: kept;
    * on a channel the good template does not touch: E = -SS/2, and with the
      threshold raised to 0 the probability is below sf(sigma): deleted;
    * the all-zero good template has zero likelihood variance and is
      skipped for every noisy template."""
    s = 3.719
    C, N = 2, 20
    rng = np.random.default_rng(5)
    good = np.concatenate([rng.standard_normal(N) * 4, np.zeros(N)])
    tm = np.stack([good, 0.9 * good, np.concatenate([np.zeros(N), rng.standard_normal(N)]),
                   np.concatenate([np.zeros(N), np.ones(N) * 0.01]), np.zeros(C * N)])
    sm = _white_metrics(tm, s)
    sm['neuron_lower_thresholds'] = np.maximum(sm['neuron_lower_thresholds'], 0)
    si = {'n_channels': C, 'n_samples_per_chan': N, 'sigma_bp_noise': s}
    noisy = np.array([False, True, True, True, False])
    # closed form for the first noisy template (aligned at lag 0: a scaled copy)
    ss = sm['template_SS'][0]
    p = norm.sf(sm['neuron_lower_thresholds'][0], 0.9 * ss - 0.5 * ss, np.sqrt(ss))
    assert p > norm.sf(s)
    _, flags = oracle_mod.check_noise_templates(sm, si, noisy)
    np.testing.assert_array_equal(flags, [False, False, True, True, False])


def test_oracle_template_pair_closed_form(oracle_mod):
    C, N = 2, 10
    rng = np.random.default_rng(9)
    t = rng.standard_normal(C * N)
    si = {'n_channels': C, 'n_samples_per_chan': N}
    eye = [np.eye(N)] * C
    assert oracle_mod.check_template_pair(t, t, eye, si) == 1.0           # no variance of the difference
    ss = float(np.dot(t, t))
    np.testing.assert_allclose(oracle_mod.check_template_pair(t, np.zeros_like(t), eye, si),
                               norm.cdf(-0.5 * np.sqrt(ss)), rtol=1e-12)


def test_delete_noise_units_matches_oracle(oracle_mod):
    rng = np.random.default_rng(2)
    T, C, N = 6, 3, 5

    def metrics():
        return {'templates': rng.standard_normal((T, C * N)), 'template_SS_by_chan': rng.standard_normal((T, C)),
                'template_SS': np.arange(T) * 1.0, 'neuron_variances': np.arange(T) * 2.0,
                'neuron_lower_thresholds': np.arange(T) * 3.0, 'neuron_lower_CI': np.arange(T) * 4.0,
                'contamination': np.arange(T) * 0.1, 'peak_channel': np.arange(T) % C,
                'channel_covariance_mats': [np.eye(N)] * C, 'channel_p_noise': np.ones(C),
                'bp_n_samples_per_chan': N, 'template_covariance_mats': [np.full((2, 2), k) for k in range(T)]}
    a = metrics()
    b = {k: (v.copy() if isinstance(v, np.ndarray) else (list(v) if isinstance(v, list) else v)) for k, v in a.items()}
    flags = np.array([True, False, False, True, False, True])
    got = delete_noise_units(a, flags, verbose=False)
    want = oracle_mod.delete_noise_units(b, flags)
    assert set(got) == set(want)
    for k in got:
        if isinstance(got[k], list):
            assert len(got[k]) == len(want[k]) == (C if k == 'channel_covariance_mats' else 3)
            for x, y in zip(got[k], want[k]):
                np.testing.assert_array_equal(x, y)
        else:
            np.testing.assert_array_equal(got[k], want[k])
    assert got['templates'].shape == (3, C * N)
    np.testing.assert_array_equal(got['template_SS'], [1.0, 2.0, 4.0])


def test_pair_lags_rejects_bad_arguments():
    lib = _lib.load()
    tm = np.zeros((2, 8))
    pairs = np.array([[0, 1]], dtype=np.uint32)
    out = np.zeros(3)
    f64 = _lib._f64p
    assert lib.fbp_template_pair_lags(0, f64(tm), 2, 2, 4, _lib._u32p(pairs), 1, -8, 1, f64(out), None) == \
        _lib.FBP_EINVAL
    bad = np.array([[0, 2]], dtype=np.uint32)
    assert lib.fbp_template_pair_lags(0, f64(tm), 2, 2, 4, _lib._u32p(bad), 1, -1, 1, f64(out), None) == \
        _lib.FBP_EINVAL
    assert lib.fbp_template_pair_lags(0, f64(tm), 2, 2, 4, _lib._u32p(pairs), 0, -1, 1, f64(out), None) == _lib.FBP_OK
