"""Configurations beyond the small parity cases, HIP path against the oracle:

* the large probe (128 channels, 200 templates of 61 samples each nonzero on
  8 contiguous channels, 30 kHz) on a 1,220-sample slice, both the bf16
  screened route and the float32 direct route;
* the production shift range n_max_shift_inds = N - 1 = 60
  (full_binary_pursuit.py:145) at 16 channels x 20 templates, every route;
* the drop-in binary_pursuit() split into three chunks with a shorter tail
  (binary_pursuit_parallel.py:359-368) through the HIP segment, against the
  oracle's host driver at the same chunk length.

Spike lists (onset, label, record order) and residuals are bitwise equal.
"""
import numpy as np
import pytest

from spikesorting_fullpursuit_amd import synthetic

pytestmark = pytest.mark.gpu


def _prep(rec, N=61):
    sm = rec['separability_metrics']
    rec['templates'] = sm['templates']
    rec['bias'] = np.float32(-0.5 * sm['template_SS'])
    rec['thr'] = np.float32(sm['neuron_lower_thresholds'])
    rec['N'] = N
    return rec


def _gpu(rec, M, precision, mode):
    from spikesorting_fullpursuit_amd import _lib
    V = rec['voltage']
    C, K = V.shape
    T = rec['templates'].shape[0]
    with _lib.Segment(C, K, T, rec['N'], M, True, 256, precision, mode) as seg:
        seg.set_templates(rec['templates'], rec['bias'], rec['thr'])
        seg.set_voltage(V)
        seg.run()
        idx, lab = seg.spikes()
        return idx, lab, seg.residual(), seg.stats()


@pytest.fixture(scope="module")
def large_probe(oracle_mod):
    rec = _prep(synthetic.config("128ch-600s", seed=11, duration_s=1220 / 30000))
    print("large probe slice", rec['voltage'].shape, "oracle ...", flush=True)
    o = oracle_mod.pursuit_chunk(rec['voltage'], rec['templates'], rec['bias'], rec['thr'], 61, 30, True, 256)
    print("oracle: %d spikes in %d passes" % (o['idx'].size, o['n_passes']), flush=True)
    return rec, o


@pytest.mark.parametrize("precision,mode", [("bf16", "screened"), ("fp32", "direct")])
def test_large_probe_slice_bitwise(large_probe, precision, mode):
    rec, o = large_probe
    print("large probe", precision, mode, flush=True)
    idx, lab, res, st = _gpu(rec, 30, precision, mode)
    assert o['idx'].size > 100
    assert st['n_passes'] == o['n_passes']
    np.testing.assert_array_equal(idx, o['idx'])
    np.testing.assert_array_equal(lab, o['lab'])
    np.testing.assert_array_equal(res.view(np.uint32), o['residual'].view(np.uint32))


@pytest.fixture(scope="module")
def production_shift(oracle_mod):
    rec = _prep(synthetic.config("16ch-300s", seed=13, duration_s=1.0))
    print("16 ch x 1 s, M = 60: oracle ...", flush=True)
    o = oracle_mod.pursuit_chunk(rec['voltage'], rec['templates'], rec['bias'], rec['thr'], 61, 60, True, 256)
    return rec, o


@pytest.mark.parametrize("precision,mode", [("fp32", "direct"), ("fp32", "screened"), ("bf16", "direct"),
                                            ("bf16", "screened")])
def test_production_shift_range_16ch(production_shift, precision, mode):
    rec, o = production_shift
    print("16 ch M=60", precision, mode, flush=True)
    idx, lab, res, st = _gpu(rec, 60, precision, mode)
    W = -(-rec['voltage'].shape[1] // 61)
    assert o['idx'].size > W // 2
    assert st['n_passes'] == o['n_passes']
    np.testing.assert_array_equal(idx, o['idx'])
    np.testing.assert_array_equal(lab, o['lab'])
    np.testing.assert_array_equal(res.view(np.uint32), o['residual'].view(np.uint32))


@pytest.mark.parametrize("precision,mode", [("fp32", "direct"), ("bf16", "screened")])
def test_dropin_three_chunks_through_hip(oracle_mod, precision, mode):
    from helpers import small_case
    from spikesorting_fullpursuit_amd import _lib, binary_pursuit_parallel as bpp
    rec = small_case(24, C=4, T=5, N=61, duration_s=0.5, overlap_frac=0.2, rate=(40.0, 80.0))
    V = rec['voltage']
    budget = 0.3e6
    total, _ = _lib.device_memory(0)
    chunk_len, onsets = bpp.plan_chunks(4, V.shape[1], 40000, 5, 61, 30, total, budget)
    assert len(onsets) >= 3 and onsets[-1] + chunk_len > V.shape[1]
    print("drop-in", len(onsets), "chunks of", chunk_len, precision, mode, flush=True)
    si = dict(rec['sort_info'], get_adjusted_clips=True)
    sm = dict(rec['separability_metrics'])
    out = bpp.binary_pursuit(V, np.float32, si, sm, None, max_gpu_memory=budget, return_residual=True,
                             precision=precision, overlap_mode=mode)
    ref = oracle_mod.binary_pursuit(V, np.float32, si, dict(rec['separability_metrics']), None, chunk_len)
    for a, b in zip(out[:4], ref[:4]):
        a, b = np.asarray(a), np.asarray(b)
        assert a.dtype == b.dtype and a.shape == b.shape
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(out[4].view(np.uint32), ref[4].view(np.uint32))
