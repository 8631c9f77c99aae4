"""Host-side logic of the drop-in (CPU only): chunk planner, sample windows,
output assembly, the n_spikes side effect, and the C-ABI library surface."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import small_case
from spikesorting_fullpursuit_amd import _lib, binary_pursuit_parallel as bpp
from spikesorting_fullpursuit_amd.neuron_separability import add_n_spikes
from spikesorting_fullpursuit_amd.segment import time_window_to_samples

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_time_window_to_samples_known_answers():
    # 1.5 ms clips at 40 kHz and 2 ms clips at 30 kHz both give 61 samples.
    assert time_window_to_samples([-7.5e-4, 7.5e-4], 40000)[0] == [-30, 31]
    assert time_window_to_samples([1e-3, 1e-3], 30000)[0] == [-30, 31]
    w, tw = time_window_to_samples([2e-4, 1e-5], 10000)
    assert w == [-2, 2] and tw[0] == -2e-4
    assert time_window_to_samples([0.0, 0.0], 1000)[0] == [0, 2]


def test_plan_chunks_memory_model():
    # One chunk when memory is plentiful; the reference's per-second model
    # (bpp.py:306-341) otherwise.
    n, on = bpp.plan_chunks(16, 12_000_000, 40000, 20, 61, 30, 288 * 2 ** 30)
    assert on == [0] and n >= 12_000_000
    items = 61 * 61 * 20
    nl = (items - 1) // 256 + 1
    per_s = 16 * 40000 * 4 + 20 * 40000 * 4 + (40000 / 61) * (20 + 8 * nl / 3)
    budget = 200e6
    n, on = bpp.plan_chunks(16, 12_000_000, 40000, 20, 61, 30, 288 * 2 ** 30, max_gpu_memory=budget)
    assert n == int(np.floor((budget - (20 * 16 * 61 * 4 + 20 * 4)) / per_s * 40000))
    assert on[1] - on[0] == n - 3 * 61
    assert on[-1] < 12_000_000 - 61 and on[-1] + n >= 12_000_000
    with pytest.raises(ValueError):
        bpp.plan_chunks(16, 12_000_000, 40000, 20, 61, 30, 288 * 2 ** 30, max_gpu_memory=100e3)


def test_add_n_spikes():
    sm = {'templates': np.zeros((4, 10))}
    add_n_spikes(sm, np.array([0, 2, 2, 3, 3, 3], dtype=np.int64))
    np.testing.assert_array_equal(sm['n_spikes'], [1, 0, 2, 3])
    add_n_spikes(sm, [])
    np.testing.assert_array_equal(sm['n_spikes'], [0, 0, 0, 0])


def test_synthetic_is_seeded():
    a = small_case(3, duration_s=0.1)
    b = small_case(3, duration_s=0.1)
    np.testing.assert_array_equal(a['voltage'], b['voltage'])
    np.testing.assert_array_equal(a['templates'], b['templates'])
    ss = a['separability_metrics']['template_SS']
    np.testing.assert_allclose(a['separability_metrics']['neuron_lower_thresholds'],
                               np.maximum(-0.5 * ss + 3.719 * np.sqrt(ss), 0))


class _OracleSegment:
    """Stand-in for _lib.Segment that runs the oracle: lets the CPU suite check
    the drop-in's chunking and output assembly against oracle.binary_pursuit."""

    def __init__(self, C, K, T, N, M, do_overlap, reduce_wg, precision, overlap_mode, max_passes, device):
        self.args = (C, K, T, N, M, do_overlap, reduce_wg)

    def set_templates(self, tm, bias, thr):
        self.tm, self.bias, self.thr = tm, bias, thr

    def set_voltage(self, v):
        self.v = np.float32(v)

    def run(self):
        import oracle
        C, K, T, N, M, ov, wg = self.args
        self.r = oracle.pursuit_chunk(self.v, self.tm, self.bias, self.thr, N, M, ov, wg)

    def spikes(self):
        return self.r['idx'], self.r['lab']

    def clips(self):
        import oracle
        return oracle.adjusted_clips(self.r['residual'], self.tm, self.args[3], self.r['idx'], self.r['lab'])

    def residual(self):
        return self.r['residual']

    def stats(self):
        return {}

    def close(self):
        pass


@pytest.mark.parametrize("clips,budget", [(True, None), (False, None), (True, 0.3e6)])
def test_dropin_assembly_matches_oracle_driver(monkeypatch, oracle_mod, clips, budget):
    monkeypatch.setattr(_lib, "Segment", _OracleSegment)
    monkeypatch.setattr(_lib, "device_count", lambda: 1)
    monkeypatch.setattr(_lib, "device_memory", lambda device=0: (288 * 2 ** 30, 288 * 2 ** 30))
    rec = small_case(21, C=4, T=5, N=61, duration_s=0.4, overlap_frac=0.1)
    si = dict(rec['sort_info'], get_adjusted_clips=clips)
    sm = dict(rec['separability_metrics'])
    out = bpp.binary_pursuit(rec['voltage'].astype(np.float64), np.float32, si, sm, None,
                             max_gpu_memory=budget, return_residual=True)
    chunk_len, onsets = bpp.plan_chunks(4, rec['voltage'].shape[1], 40000, 5, 61, 30, 288 * 2 ** 30, budget)
    if budget is not None:
        assert len(onsets) > 1
    ref = oracle_mod.binary_pursuit(rec['voltage'].astype(np.float64), np.float32, si, sm, None, chunk_len)
    for a, b in zip(out[:4], ref[:4]):
        a, b = np.asarray(a), np.asarray(b)
        assert a.dtype == b.dtype and a.shape == b.shape
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(out[4], ref[4])
    assert out[0].dtype == np.int64 and out[1].dtype == np.int64 and out[2].dtype == bool
    np.testing.assert_array_equal(sm['n_spikes'], np.bincount(out[1], minlength=5))


def test_dropin_empty_result(monkeypatch):
    monkeypatch.setattr(_lib, "Segment", _OracleSegment)
    monkeypatch.setattr(_lib, "device_count", lambda: 1)
    monkeypatch.setattr(_lib, "device_memory", lambda device=0: (288 * 2 ** 30, 288 * 2 ** 30))
    rec = small_case(22, C=4, T=5, N=61, duration_s=0.2, rate=(0.0, 0.0))
    sm = dict(rec['separability_metrics'])
    out = bpp.binary_pursuit(rec['voltage'], np.float32, rec['sort_info'], sm)
    assert out == ([], [], [], [])
    assert np.all(sm['n_spikes'] == 0)


def test_dropin_missing_sort_info_key(monkeypatch):
    monkeypatch.setattr(_lib, "device_count", lambda: 1)
    rec = small_case(23, duration_s=0.05)
    si = dict(rec['sort_info'])
    del si['do_overlap_recheck']
    with pytest.raises(KeyError):
        bpp.binary_pursuit(rec['voltage'], np.float32, si, rec['separability_metrics'])


def _header_functions():
    text = open(os.path.join(ROOT, "include", "fbp.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(fbp_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _header_functions()
    assert len(declared) >= 16
    assert set(declared) == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.fbp_version() >= 100


def test_no_gpu_raises_like_the_reference():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="No GPU available"):
        _lib.device_count()
    with pytest.raises(RuntimeError, match="No GPU available"):
        _lib.Segment(4, 1000, 2, 61, 30)


def test_config_validation_messages():
    lib = _lib.load()
    cfg = _lib.FbpConfig(4, 50, 2, 61, 30, 1, 256, 0, 0, 0)
    h = ctypes.c_void_p()
    assert lib.fbp_segment_create(0, ctypes.byref(cfg), ctypes.byref(h)) == _lib.FBP_EINVAL
    assert b"not longer than the template" in lib.fbp_last_error()
    cfg = _lib.FbpConfig(4, 5000, 2, 61, 30, 1, 100, 0, 0, 0)
    assert lib.fbp_segment_create(0, ctypes.byref(cfg), ctypes.byref(h)) == _lib.FBP_EINVAL
    assert b"power of two" in lib.fbp_last_error()


def test_multichunk_residual_is_each_chunks_own_subtraction(monkeypatch, oracle_mod):
    # Several chunks with a shorter tail: the residual the drop-in returns
    # equals, on the samples each chunk supplies, the voltage minus exactly
    # the templates of that chunk's spikes (float64, float32 rounding bound);
    # the supplied spans tile the recording, and every sample a chunk
    # supplies lies at least N samples after its onset, so no spike covering
    # it starts before the chunk (the residual of the first N samples of a
    # chunk lacks the tails of the previous chunk's spikes).
    monkeypatch.setattr(_lib, "Segment", _OracleSegment)
    monkeypatch.setattr(_lib, "device_count", lambda: 1)
    monkeypatch.setattr(_lib, "device_memory", lambda device=0: (288 * 2 ** 30, 288 * 2 ** 30))
    rec = small_case(24, C=4, T=5, N=61, duration_s=0.5, overlap_frac=0.2, rate=(40.0, 80.0))
    V = rec['voltage']
    C, K = V.shape
    N = 61
    budget = 0.3e6
    chunk_len, onsets = bpp.plan_chunks(C, K, 40000, 5, N, 30, 288 * 2 ** 30, budget)
    assert len(onsets) >= 3 and onsets[-1] + chunk_len > K   # a shorter tail chunk
    sm = dict(rec['separability_metrics'])
    out = bpp.binary_pursuit(V, np.float32, rec['sort_info'], sm, None, max_gpu_memory=budget,
                             return_residual=True)
    res = out[4]
    tm = rec['templates'].reshape(5, C, N).astype(np.float64)
    covered = np.zeros(K, dtype=int)
    for k, start in enumerate(onsets):
        stop = min(K, start + chunk_len)
        lo, hi = bpp.residual_span(onsets, k, N, K)
        assert start <= lo < hi <= stop
        if k > 0:
            assert lo >= start + N
        covered[lo:hi] += 1
        r = oracle_mod.pursuit_chunk(V[:, start:stop], rec['templates'], rec['bias'], rec['thr'], N, 30)
        assert r['idx'].size > 0
        expect = V[:, start:stop].astype(np.float64)
        for i, l in zip(r['idx'], r['lab']):
            n = min(N, stop - start - i)
            expect[:, i:i + n] -= tm[l][:, :n]
        err = np.abs(res[:, lo:hi] - expect[:, lo - start:hi - start])
        assert err.max() < 1e-3, (k, err.max())
    assert np.all(covered == 1)
