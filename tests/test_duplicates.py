"""Duplicate removal after binary pursuit: the oracle against known answers
(CPU) and the GPU path (unit templates and clip distances on the device)
against the oracle."""
import numpy as np
import pytest


def test_oracle_duplicate_known_answers(oracle_mod):
    ev = np.array([0, 5, 100, 104, 300], dtype=np.int64)
    tmpl = np.zeros(4, dtype=np.float32)
    clips = np.array([[2, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0], [0, 3, 0, 0], [9, 9, 9, 9]], dtype=np.float32)
    bp = np.array([True, True, True, True, True])
    # 0/5: the second is closer to the template -> drop the first; 100/104: keep the first
    np.testing.assert_array_equal(oracle_mod.remove_binary_pursuit_duplicates(ev, clips, tmpl, bp, 10),
                                  [False, True, True, False, True])
    # a threshold spike followed by a pursuit spike: both kept
    bp2 = np.array([False, True, True, True, True])
    np.testing.assert_array_equal(oracle_mod.remove_binary_pursuit_duplicates(ev, clips, tmpl, bp2, 10),
                                  [True, True, True, False, True])


def _sort_data(seed, n_chans=3, n_segments=2, n=400, width=2 * 21):
    rng = np.random.default_rng(seed)
    sd = []
    for c in range(n_chans):
        row = []
        for s in range(n_segments):
            if c == 1 and s == 0:
                row.append([[], [], [], [], 7])
                continue
            ev = np.sort(rng.integers(0, 20000, n)).astype(np.int64)
            ev[1::7] = ev[0::7][:ev[1::7].size] + rng.integers(0, 12, ev[1::7].size)   # near duplicates
            ev.sort()
            lab = rng.integers(0, 4, n).astype(np.int64)
            clips = (rng.standard_normal((n, width)) * rng.uniform(1, 30)).astype(np.float32)
            bp = rng.random(n) < 0.8
            row.append([ev, lab, clips, bp, 7])
        sd.append(row)
    return sd


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1])
def test_gpu_duplicate_removal_matches_oracle(oracle_mod, seed):
    from spikesorting_fullpursuit_amd import _lib
    from spikesorting_fullpursuit_amd.postprocessing import remove_segment_binary_pursuit_duplicates
    print("case duplicates", seed, flush=True)
    got = remove_segment_binary_pursuit_duplicates(_sort_data(seed), 3, 2, 10)
    ref = oracle_mod.remove_segment_binary_pursuit_duplicates(_sort_data(seed), 3, 2, 10)
    removed = 0
    for c in range(3):
        for s in range(2):
            for a, b in zip(got[c][s][:4], ref[c][s][:4]):
                np.testing.assert_array_equal(np.asarray(a), np.asarray(b))
            removed += 400 - len(got[c][s][0]) if len(got[c][s][0]) else 0
    assert removed > 0
    # the distances themselves are numpy's, bit for bit
    d = _sort_data(seed)[0][0]
    tm = np.mean(d[2][d[1] == 2], axis=0)
    sse = _lib.clip_template_sse(d[2], np.zeros(400, np.uint32), tm[None, :])
    ref_sse = np.array([np.sum((d[2][k] - tm) ** 2) for k in range(400)], dtype=np.float32)
    np.testing.assert_array_equal(sse.view(np.uint32), ref_sse.view(np.uint32))
