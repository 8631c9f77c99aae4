"""Known-answer pins of the oracle's decision rules (CPU only).

The scenarios of tests/scenarios.py fix, from their construction and the
reference's stated rules, which spikes every pass adds and which windows the
next pass evaluates; the oracle must reproduce them pass by pass.  A
misreading of the +/-2-window give-up (binary_pursuit.cl:847-864), the
id + 2 < K/N flag (:867-871), the strict threshold of the add (:982) or the
flag pruning (:1024-1044) changes one of these lists.
"""
import numpy as np
import pytest

import scenarios


def _run(oracle_mod, sc, max_passes):
    return oracle_mod.pursuit_chunk(sc['voltage'], sc['templates'], sc['bias'], sc['thr'], sc['N'], sc['M'],
                                    do_overlap=sc['do_overlap'], reduce_wg=256, max_passes=max_passes)


@pytest.mark.parametrize("make", scenarios.SCENARIOS, ids=lambda f: f.__name__)
def test_scenario_pass_by_pass(oracle_mod, make):
    sc = make()
    added = []
    for p, expect in enumerate(sc['passes'], start=1):
        r = _run(oracle_mod, sc, p)
        got = list(zip(r['idx'].tolist(), r['lab'].tolist()))
        assert got[:len(added)] == added, (p, got)
        assert got[len(added):] == expect, (p, got[len(added):])
        added = got
        if p < len(sc['passes']):
            assert r['capped'], p
            assert r['next_windows'].tolist() == sc['lists'][p - 1], (p, r['next_windows'].tolist())
    final = _run(oracle_mod, sc, 0)
    assert final['n_passes'] == len(sc['passes'])
    assert list(zip(final['idx'].tolist(), final['lab'].tolist())) == added
    if sc['residual_zero']:
        assert np.all(final['residual'] == 0.0)
    assert sorted(added) == sorted(sc['truth'])


@pytest.mark.parametrize("make", scenarios.SCENARIOS, ids=lambda f: f.__name__)
def test_scenario_likelihoods_are_exact(oracle_mod, make):
    # The construction's likelihood values are dyadic: the float32 likelihood
    # equals the float64 one at every position, so no rounding can move a tie.
    sc = make()
    L32 = oracle_mod.likelihood(sc['voltage'], sc['templates'], sc['bias'], sc['N']).astype(np.float64)
    T = sc['templates'].shape[0]
    C, K = sc['voltage'].shape
    tm = sc['templates'].reshape(T, C, sc['N']).astype(np.float64)
    V = sc['voltage'].astype(np.float64)
    for t in range(T):
        for i in range(K - sc['N']):
            exact = float(sc['bias'][t]) + np.sum(tm[t] * V[:, i:i + sc['N']])
            assert L32[t, i] == exact


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6])
def test_chained_train_fully_recovered(oracle_mod, seed):
    # Pairs of spikes 5..M+1 samples apart, pairs 22..31 samples apart: the
    # overlap recheck resolves every pair; all true spikes are found and the
    # residual is exactly zero.
    sc = scenarios.chained_train(seed)
    r = _run(oracle_mod, sc, 0)
    got = sorted(zip(r['idx'].tolist(), r['lab'].tolist()))
    assert got == sorted(sc['truth'])
    assert np.all(r['residual'] == 0.0)
    assert r['n_passes'] >= 2
