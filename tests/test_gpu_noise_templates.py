"""Noise-template pruning on the GPU (fbp_template_pair_lags behind
optimal_align_templates and check_noise_templates; fbp_template_variances
behind check_template_pairs) against np.correlate and the oracle.  The
reference evaluates the cross-correlations and dot products with numpy/BLAS,
whose summation order is not fixed, so values agree to float64 rounding
(relative 1e-12 of the largest magnitude involved) and the decisions built on
them (lags, kept templates) are equal."""
import numpy as np
import pytest

from spikesorting_fullpursuit_amd import _lib, synthetic
from spikesorting_fullpursuit_amd.consolidate import optimal_align_templates
from spikesorting_fullpursuit_amd.neuron_separability import (check_noise_templates, check_template_pair,
                                                              check_template_pairs, compute_separability_metrics,
                                                              find_noisy_templates, set_bp_threshold)

pytestmark = pytest.mark.gpu

RTOL = 1e-12


@pytest.mark.parametrize("C,N,lo,hi", [(4, 61, -14, 14), (1, 40, -39, 39), (3, 16, -47, 47), (16, 61, -60, 5),
                                       (128, 61, -14, 14)])
def test_pair_lags_match_np_correlate(C, N, lo, hi):
    rng = np.random.default_rng(C + N)
    tm = rng.standard_normal((5, C * N)) * rng.uniform(0.5, 20, (5, 1))
    pairs = np.array([[0, 1], [2, 2], [4, 3], [1, 0]], dtype=np.uint32)
    print("case pair lags", C, N, lo, hi, flush=True)
    xc, al = _lib.template_pair_lags(tm, C, pairs, lo, hi)
    L = C * N
    for p, (a, b) in enumerate(pairs):
        full = np.correlate(tm[a], tm[b], mode='full')
        want = full[lo + L - 1:hi + L]
        scale = np.sqrt(np.sum(tm[a] ** 2) * np.sum(tm[b] ** 2))
        np.testing.assert_allclose(xc[p], want, rtol=0, atol=RTOL * scale)
        for k, lag in enumerate(range(lo, hi + 1)):   # per-channel shifted and truncated dot products
            A, B = tm[a].reshape(C, N), tm[b].reshape(C, N)
            if lag >= 0:
                ref = np.sum(A[:, lag:] * B[:, :N - lag]) if lag < N else 0.0
            else:
                ref = np.sum(A[:, :N + lag] * B[:, -lag:]) if -lag < N else 0.0
            assert abs(al[p, k] - ref) <= RTOL * scale
    # one fixed summation order: a second call gives the same bits
    xc2, al2 = _lib.template_pair_lags(tm, C, pairs, lo, hi)
    np.testing.assert_array_equal(xc.view(np.uint64), xc2.view(np.uint64))
    np.testing.assert_array_equal(al.view(np.uint64), al2.view(np.uint64))


@pytest.mark.parametrize("C,N,d,max_shift,align_abs,zero_pad", [
    (4, 61, 3, None, False, False), (4, 61, -5, 14, True, False), (2, 31, 7, None, True, True),
    (16, 61, -9, 14, False, True), (3, 16, 0, None, False, False), (1, 40, 12, 5, False, False)])
def test_optimal_align_templates_matches_oracle(oracle_mod, C, N, d, max_shift, align_abs, zero_pad):
    rng = np.random.default_rng(31 + N + d)
    t1 = rng.standard_normal(C * N)
    A = t1.reshape(C, N)
    t2 = np.zeros_like(A)
    if d >= 0:
        t2[:, :N - d] = A[:, d:]
    else:
        t2[:, -d:] = A[:, :N + d]
    t2 = (-1.0 if align_abs else 1.0) * t2.reshape(-1) + 0.05 * rng.standard_normal(C * N)
    print("case align", C, N, d, max_shift, align_abs, zero_pad, flush=True)
    got = optimal_align_templates(t1, t2, C, max_shift=max_shift, align_abs=align_abs, zero_pad=zero_pad)
    want = oracle_mod.optimal_align_templates(t1, t2, C, max_shift=max_shift, align_abs=align_abs,
                                              zero_pad=zero_pad)
    assert got[2] == want[2] and got[3] == want[3]
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])


def _metrics(seed, C, T, duration_s=0.2, scale=None):
    rec = synthetic.config("16ch-300s" if C == 16 else "4ch-60s", seed=seed, duration_s=duration_s)
    tm = rec['separability_metrics']['templates'].astype(np.float64)
    if scale is not None:
        tm = tm * scale[:, None]
    N = tm.shape[1] // C
    si = {'n_channels': C, 'n_samples_per_chan': N, 'n_samples': 12000,
          'n_threshold_crossings': np.arange(C) * 3, 'sigma_bp_noise': 3.719, 'sigma_bp_CI': 12.0}
    return tm, si, [np.eye(N)] * C


@pytest.mark.parametrize("seed,C,mixed", [(4, 16, True), (6, 16, True), (3, 4, False), (5, 4, False)])
def test_check_noise_templates_matches_oracle(oracle_mod, seed, C, mixed):
    """About half of the templates scaled down (log-uniform 10^-2.7..10^-0.5)
    so that some are noisy; the pruning steps of full_binary_pursuit.py:
    321-328 on the product and on the oracle.  The 16-channel seeds keep some
    noisy templates and delete others."""
    rng = np.random.default_rng(seed)
    T = 20 if C == 16 else 5
    scale = np.where(rng.uniform(size=T) < 0.6, 10 ** rng.uniform(-2.7, -0.5, T), 1.0)
    tm, si, covs = _metrics(seed, C, T, scale=scale)
    sm = compute_separability_metrics(list(tm), covs, si)
    noisy = find_noisy_templates(sm, si)
    want_noisy = oracle_mod.find_noisy_templates(sm)
    np.testing.assert_array_equal(noisy, want_noisy)
    assert 0 < noisy.sum() < noisy.size
    sm = set_bp_threshold(sm)
    print("case noise check", seed, C, int(noisy.sum()), "noisy of", noisy.size, flush=True)
    _, flags = check_noise_templates(sm, si, noisy)
    _, want = oracle_mod.check_noise_templates(sm, si, noisy)
    np.testing.assert_array_equal(flags, want)
    if mixed:
        assert 0 < flags.sum() < noisy.sum()


def test_noise_spike_probabilities_match_oracle_terms(oracle_mod):
    """The per-pair expectations behind check_noise_templates equal the
    oracle's np.dot of the aligned templates to float64 rounding."""
    from scipy.stats import norm
    from spikesorting_fullpursuit_amd.neuron_separability import noise_spike_probabilities
    scale = np.where(np.arange(20) % 3 == 0, 0.1, 1.0)
    tm, si, covs = _metrics(4, 16, 20, scale=scale)
    sm = set_bp_threshold(compute_separability_metrics(list(tm), covs, si))
    good, noisy = np.arange(0, 20, 2), np.arange(1, 20, 2)
    p = noise_spike_probabilities(sm, si, good, noisy)
    C, N = 16, si['n_samples_per_chan']
    for i, n in enumerate(good):
        for j, m in enumerate(noisy):
            a, b, _, _ = oracle_mod.optimal_align_templates(tm[n], tm[m], C, max_shift=N // 4 - 1, align_abs=True)
            e = np.dot(a, b) - 0.5 * sm['template_SS'][n]
            want = norm.sf(sm['neuron_lower_thresholds'][n], e, np.sqrt(sm['neuron_variances'][n]))
            np.testing.assert_allclose(p[i, j], want, rtol=1e-9, atol=1e-300)


def test_check_template_pairs_matches_oracle(oracle_mod):
    rng = np.random.default_rng(12)
    C, N = 4, 61
    covs = []
    for _ in range(C):
        a = rng.standard_normal((N, N)) * 0.1
        covs.append(a @ a.T + np.eye(N))
    t1 = rng.standard_normal((6, C * N)) * 3
    t2 = t1 + rng.standard_normal((6, C * N)) * np.array([0.0, 0.1, 0.5, 1.0, 3.0, 10.0])[:, None]
    si = {'n_channels': C, 'n_samples_per_chan': N}
    print("case template pairs", flush=True)
    got = check_template_pairs(t1, t2, covs, si)
    want = np.array([oracle_mod.check_template_pair(t1[k], t2[k], covs, si) for k in range(6)])
    assert got[0] == 1.0 == want[0]   # identical templates: no variance of the difference
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-300)
    assert check_template_pair(t1[3], t2[3], covs, si) == got[3]
    # float32 templates: the dot products in float32, as the reference evaluates them
    f1, f2 = t1.astype(np.float32), t2.astype(np.float32)
    got32 = check_template_pairs(f1, f2, covs, si)
    want32 = np.array([oracle_mod.check_template_pair(f1[k], f2[k], covs, si) for k in range(6)])
    np.testing.assert_allclose(got32, want32, rtol=1e-9, atol=1e-300)
