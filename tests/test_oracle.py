"""Pins of the oracle (oracle/fbp_oracle.c) -- CPU only.

The reference ships no tests or golden vectors for binary pursuit, and its
OpenCL kernels cannot run here, so the oracle is pinned by:
  * closed forms and exact known answers (isolated templates in zero noise);
  * float64 brute force within a rigorous float32 rounding bound;
  * an independent pure-Python restatement (tests/pyref.py) with exactly
    emulated float32 fused multiply-adds, compared bit for bit;
  * the tie rule of the reference's work-group tree reduction, emulated
    literally and compared with the closed-form rule the GPU uses.
"""
import numpy as np
import pytest

import pyref
from helpers import small_case

U32 = 2.0 ** -24


def _tiny(seed, C=2, T=3, N=9, K=300, peak=(4.0, 9.0), noise=True, n_spikes=8):
    rng = np.random.default_rng(seed)
    x = np.arange(N)
    tm = np.zeros((T, C, N), dtype=np.float32)
    for t in range(T):
        w = -np.exp(-(x - N // 2) ** 2 / (2 * rng.uniform(0.8, 1.6) ** 2)) \
            + 0.4 * np.exp(-(x - N // 2 - 2) ** 2 / (2 * 2.0 ** 2))
        w *= rng.uniform(*peak) / np.abs(w).max()
        for c in range(C):
            tm[t, c] = w * rng.uniform(0.3, 1.0)
    V = rng.standard_normal((C, K)).astype(np.float32) if noise else np.zeros((C, K), np.float32)
    on = np.sort(rng.choice(np.arange(N, K - 2 * N), size=n_spikes, replace=False))
    lab = rng.integers(0, T, size=n_spikes)
    for o, l in zip(on, lab):
        V[:, o:o + N] += tm[l]
    ss = np.sum(tm.astype(np.float64).reshape(T, -1) ** 2, axis=1)
    bias = np.float32(-0.5 * ss)
    thr = np.float32(np.maximum(-0.5 * ss + 1.645 * np.sqrt(ss), 0.0))
    return V, tm.reshape(T, C * N), bias, thr, ss


def test_likelihood_within_float32_bound_of_float64(oracle_mod):
    V, tm, bias, thr, ss = _tiny(0, C=3, T=4, N=11, K=500)
    N = 11
    L = oracle_mod.likelihood(V, tm, bias, N)
    T, C = tm.shape[0], V.shape[0]
    t3 = tm.reshape(T, C, N).astype(np.float64)
    for t in range(T):
        for i in range(V.shape[1]):
            if i + N >= V.shape[1]:
                assert L[t, i] == 0.0
                continue
            seg = V[:, i:i + N].astype(np.float64)
            exact = float(bias[t]) + np.sum(t3[t] * seg)
            absum = abs(float(bias[t])) + np.sum(np.abs(t3[t] * seg))
            n = C * N + 1
            assert abs(L[t, i] - exact) <= n * U32 / (1 - n * U32) * absum


def test_likelihood_isolated_template_is_half_sumsq(oracle_mod):
    # A template alone in zero noise at onset i0: L = -SS/2 + <t,t> = +SS/2.
    V, tm, bias, thr, ss = _tiny(1, C=2, T=3, N=9, K=200, noise=False, n_spikes=0)
    N, T, C = 9, 3, 2
    i0 = 57
    V[:, i0:i0 + N] = tm[1].reshape(C, N)
    L = oracle_mod.likelihood(V, tm, bias, N)
    assert L[1, i0] == pytest.approx(0.5 * ss[1], rel=1e-6)
    t3 = tm.reshape(T, C, N).astype(np.float64)
    for t in range(T):
        cross = np.sum(t3[t] * t3[1])
        assert L[t, i0] == pytest.approx(-0.5 * ss[t] + cross, rel=1e-5, abs=1e-5)
    # far from the spike only the bias remains
    assert L[0, 150] == bias[0]


def test_likelihood_linear_in_voltage(oracle_mod):
    V, tm, bias, thr, ss = _tiny(2, C=2, T=2, N=7, K=150)
    V2 = np.random.default_rng(5).standard_normal(V.shape).astype(np.float32)
    zero = np.zeros(2, np.float32)
    a = oracle_mod.likelihood(V, tm, zero, 7).astype(np.float64)
    b = oracle_mod.likelihood(V2, tm, zero, 7).astype(np.float64)
    ab = oracle_mod.likelihood(V + V2, tm, zero, 7).astype(np.float64)
    assert np.max(np.abs(ab - a - b)) < 1e-4


def test_likelihood_fma_order_bitwise(oracle_mod):
    V, tm, bias, thr, ss = _tiny(3, C=2, T=2, N=7, K=60)
    N, T, C = 7, 2, 2
    L = oracle_mod.likelihood(V, tm, bias, N)
    t3 = tm.reshape(T, C, N)
    for t in range(T):
        for i in range(V.shape[1]):
            ref = pyref.likelihood_at(V, t3, bias, t, i, V.shape[1], N, C)
            assert np.float32(L[t, i]).view(np.uint32) == np.float32(ref).view(np.uint32), (t, i)


def test_fma32_emulation_known_cases():
    # 1 + 2^-24 is a float32 midpoint: fma32 must round the exact value.
    a = np.float32(1.0 + 2.0 ** -12)
    b = np.float32(1.0 + 2.0 ** -12)
    c = np.float32(-1.0)
    exact = (1.0 + 2.0 ** -12) ** 2 - 1.0   # exact in double
    assert pyref.fma32(a, b, c) == np.float32(exact)
    assert pyref.fma32(np.float32(3.0), np.float32(5.0), np.float32(7.0)) == np.float32(22.0)
    # product rounding alone would lose the low bit; the fused result keeps it
    x = np.float32(1.0 + 2.0 ** -23)
    assert pyref.fma32(x, x, np.float32(-1.0)) == np.float32(2.0 ** -22 + 2.0 ** -46)


@pytest.mark.parametrize("seed,M,WG,ov", [(10, 4, 16, True), (11, 3, 8, True), (12, 4, 16, False),
                                          (13, 8, 64, True), (14, 1, 16, True)])
def test_pass_loop_matches_independent_restatement(oracle_mod, seed, M, WG, ov):
    V, tm, bias, thr, ss = _tiny(seed, C=2, T=3, N=9, K=240, n_spikes=10)
    o = oracle_mod.pursuit_chunk(V, tm, bias, thr, 9, M, do_overlap=ov, reduce_wg=WG)
    pi, pl, pres, passes = pyref.pursuit(V, tm, bias, thr, 9, M, do_overlap=ov, WG=WG)
    assert o['n_passes'] == passes
    np.testing.assert_array_equal(o['idx'], pi)
    np.testing.assert_array_equal(o['lab'], pl)
    np.testing.assert_array_equal(o['residual'].view(np.uint32), pres.view(np.uint32))


def test_single_unit_zero_noise_exact_recovery(oracle_mod):
    rec = small_case(3, C=4, T=1, N=61, duration_s=0.5, noise=False, rate=(20.0, 20.0))
    r = oracle_mod.pursuit_chunk(rec['voltage'], rec['templates'], rec['bias'], rec['thr'], 61, 30)
    truth = np.sort(rec['events'][0] - rec['clip_init'])
    K = rec['voltage'].shape[1]
    truth = truth[truth + 2 * 61 < K]   # the last windows are never processed (bp.cl:380-383)
    got = np.sort(r['idx'].astype(np.int64))
    np.testing.assert_array_equal(got[got + 2 * 61 < K], truth)
    assert np.all(r['lab'] == 0)
    keep = np.ones(K, bool)
    keep[K - 3 * 61:] = False
    assert np.all(r['residual'][:, keep] == 0.0)


def test_overlapping_pair_is_resolved(oracle_mod):
    # Two units summed 12 samples apart in zero noise: both are recovered at
    # their exact onsets (the overlap recheck moves the best single match).
    rec = small_case(4, C=4, T=2, N=61, duration_s=0.05, noise=False, rate=(0.0, 0.0))
    K = rec['voltage'].shape[1]
    tm = rec['templates'].reshape(2, 4, 61)
    V = np.zeros_like(rec['voltage'])
    V[:, 600:661] += tm[0]
    V[:, 612:673] += tm[1]
    for ov in (True, False):
        r = oracle_mod.pursuit_chunk(V, rec['templates'], rec['bias'], rec['thr'], 61, 30, do_overlap=ov)
        pairs = sorted(zip(r['idx'].tolist(), r['lab'].tolist()))
        assert pairs == [(600, 0), (612, 1)], (ov, pairs)
        assert np.max(np.abs(r['residual'])) < 1e-4


def test_exact_ties_pick_lowest_template_then_index(oracle_mod):
    # Two identical templates: every likelihood ties; the argmax keeps template 0.
    V, tm, bias, thr, ss = _tiny(6, C=2, T=1, N=9, K=200, noise=False, n_spikes=0)
    tm2 = np.vstack([tm, tm])
    V[:, 90:99] = tm.reshape(2, 9)
    r = oracle_mod.pursuit_chunk(V, tm2, np.concatenate([bias, bias]), np.concatenate([thr, thr]), 9, 4)
    assert r['idx'].tolist() == [90] and r['lab'].tolist() == [0]


def _tree(vals, wg):
    vals = list(vals)
    ids = list(range(wg))
    half = wg // 2
    while half:
        for l in range(half):
            if vals[l + half] > vals[l]:
                vals[l], ids[l] = vals[l + half], ids[l + half]
        half //= 2
    return vals[0], ids[0]


@pytest.mark.parametrize("wg", [2, 8, 64, 256])
def test_tree_reduction_tie_rule_is_min_bitreversed_lane(wg):
    # max_local_reduction (bp.cl:167-186) keeps, among equal maxima, the lane
    # whose bit-reversed id is smallest; the GPU kernels reduce with that key.
    rng = np.random.default_rng(wg)
    bits = wg.bit_length() - 1
    rev = lambda x: int(format(x, "0%db" % bits)[::-1], 2) if bits else 0
    for _ in range(300):
        vals = rng.integers(0, 4, size=wg).astype(np.float32)
        v, lid = _tree(vals, wg)
        ties = [l for l in range(wg) if vals[l] == vals.max()]
        assert v == vals.max()
        assert lid == min(ties, key=rev)


def test_residual_invariant_float64(oracle_mod):
    rec = small_case(8, C=4, T=5, N=61, duration_s=0.3, overlap_frac=0.2)
    r = oracle_mod.pursuit_chunk(rec['voltage'], rec['templates'], rec['bias'], rec['thr'], 61, 30)
    V = rec['voltage'].astype(np.float64)
    tm = rec['templates'].reshape(5, 4, 61).astype(np.float64)
    K = V.shape[1]
    for i, l in zip(r['idx'], r['lab']):
        n = min(61, K - i)
        V[:, i:i + n] -= tm[l][:, :n]
    assert np.max(np.abs(V - r['residual'])) < 1e-3


def test_clips_are_residual_plus_template(oracle_mod):
    rec = small_case(9, C=4, T=5, N=61, duration_s=0.3)
    r = oracle_mod.pursuit_chunk(rec['voltage'], rec['templates'], rec['bias'], rec['thr'], 61, 30)
    clips = oracle_mod.adjusted_clips(r['residual'], rec['templates'], 61, r['idx'], r['lab'])
    tm = rec['templates'].reshape(5, 4, 61)
    for k, (i, l) in enumerate(zip(r['idx'], r['lab'])):
        np.testing.assert_array_equal(clips[k].reshape(4, 61), r['residual'][:, i:i + 61] + tm[l])


def test_pure_noise_finds_nothing_and_keeps_voltage(oracle_mod):
    rec = small_case(12, C=4, T=5, N=61, duration_s=0.2, rate=(0.0, 0.0))
    r = oracle_mod.pursuit_chunk(rec['voltage'], rec['templates'], rec['bias'], rec['thr'], 61, 30)
    assert r['idx'].size == 0 and r['n_passes'] == 1
    np.testing.assert_array_equal(r['residual'], rec['voltage'])
