"""Overlap-template search (find_overlap_templates): the oracle against
known answers on the CPU; the GPU path against the oracle, bit for bit."""
import numpy as np
import pytest

from spikesorting_fullpursuit_amd.overlap_templates import _shifted, find_overlap_templates


def _templates(seed, T, C, N):
    rng = np.random.default_rng(seed)
    x = np.linspace(-3, 3, N)
    t = np.zeros((T, C, N), dtype=np.float32)
    for k in range(T):
        home, width = rng.integers(0, C), rng.uniform(0.5, 1.5)
        gain = np.exp(-0.5 * ((np.arange(C) - home) / 1.5) ** 2)
        shape = -np.exp(-0.5 * (x / width) ** 2) + 0.4 * np.exp(-0.5 * ((x - 1.2) / (2 * width)) ** 2)
        t[k] = (rng.uniform(3, 12) * gain[:, None] * shape[None, :]).astype(np.float32)
    return t.reshape(T, C * N)


def test_oracle_finds_an_exact_sum(oracle_mod):
    C, N = 3, 15
    t = np.round(_templates(0, 5, C, N) * 1024) / np.float32(1024)   # sums of two are exact in float32
    # template 4 = template 0 shifted by +3 plus template 2 shifted by -2
    t[4] = (_shifted(t[0], 3, C, N) + _shifted(t[2], -2, C, N)).astype(np.float32)
    spikes = np.array([1000, 50, 1000, 40, 10], dtype=np.int64)
    to_check, res = oracle_mod.find_overlap_templates(t, N, C, 5, 5, spikes)
    entry = [e for e in to_check if e[0] == 4]
    assert entry and entry[0][2] == [0, 2]
    np.testing.assert_array_equal(entry[0][1], _shifted(t[0], 3, C, N) + _shifted(t[2], -2, C, N))
    assert res[[e[0] for e in to_check].index(4)] == 0.0
    assert oracle_mod.find_overlap_templates(t[:2], N, C, 5, 5, spikes[:2]) == ([], [])


@pytest.mark.gpu
@pytest.mark.parametrize("seed,T,C,N,pre,post", [(1, 6, 3, 15, 6, 6), (2, 8, 4, 21, 9, 10), (3, 7, 2, 31, 14, 15)])
def test_gpu_matches_oracle(oracle_mod, seed, T, C, N, pre, post):
    t = _templates(seed, T, C, N)
    rng = np.random.default_rng(seed)
    t[T - 1] = (_shifted(t[0], 2, C, N) + _shifted(t[1], -3, C, N)).astype(np.float32)
    spikes = rng.integers(10, 5000, T).astype(np.int64)
    spikes[T - 1] = 5
    print("case overlap templates", seed, T, C, N, pre, post, flush=True)
    got, gres = find_overlap_templates(t, N, C, pre, post, spikes, return_residuals=True)
    ref, rres = oracle_mod.find_overlap_templates(t, N, C, pre, post, spikes)
    assert len(got) == len(ref) > 0
    for (gu, gt, gp), (ru, rt, rp), gv, rv in zip(got, ref, gres, rres):
        assert gu == ru and gp == rp
        np.testing.assert_array_equal(gt, rt)
        assert np.float64(gv).view(np.uint64) == np.float64(rv).view(np.uint64)
