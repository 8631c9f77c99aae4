"""Serving through the native segment server (fbp_server_*), on the GPU:
results identical alone and in a concurrent mix, every request of a mix of
the served shapes (up to 32 channels and 40 templates) equal to the oracle,
dense requests whose spike lists and clips outgrow their reservation,
cancellation, and the C-ABI's job states.  The scheduler's own logic
(admission order, slots, budget, errors, cancellation, shutdown) is checked
on the CPU with a mock device under thread and address sanitizers
(tests/test_sanitizers.py)."""
import concurrent.futures
import time

import numpy as np
import pytest

from spikesorting_fullpursuit_amd import serving, synthetic

pytestmark = pytest.mark.gpu


def _req(spec, **kw):
    rec = synthetic.recording_for(spec)
    sm = rec['separability_metrics']
    return serving.SegmentRequest(rec['voltage'], sm['templates'], sm['template_SS'], sm['neuron_lower_thresholds'],
                                  spec['n_samples_per_chan'], tag=spec['seed'], **kw)


def _same(a, m):
    np.testing.assert_array_equal(a[0], m.onsets)
    np.testing.assert_array_equal(a[1], m.labels)
    if a[2] is not None:
        np.testing.assert_array_equal(a[2].view(np.uint32), m.clips.view(np.uint32))
    if a[3] is not None:
        np.testing.assert_array_equal(a[3].view(np.uint32), m.residual.view(np.uint32))


def test_results_identical_alone_and_in_the_mix():
    specs = synthetic.serving_mix(seed=3, n_requests=10, rate_hz=200.0, duration_range=(0.5, 2.0),
                                  channel_range=(4, 16), template_range=(3, 12), length_range=(41, 81))
    reqs = [_req(s, return_residual=True, get_clips=True) for s in specs]
    print("serving mix of", len(reqs), "requests", flush=True)
    alone = [serving.run_segment(r) for r in reqs]
    with serving.PursuitServer(max_inflight=6) as srv:
        futs = []
        t0 = time.perf_counter()
        for s, r in zip(specs, reqs):
            delay = s['arrival_s'] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            futs.append(srv.submit(r))
        mixed = [f.result(timeout=120) for f in futs]
        summary = srv.latency_summary()
    print("latency", summary, flush=True)
    assert summary['peak_inflight'] > 1
    assert summary['pass_steps'] == sum(m.stats['n_passes'] for m in mixed)
    for a, m in zip(alone, mixed):
        _same(a, m)


def test_served_shapes_match_the_oracle(oracle_mod):
    # up to 32 channels and 40 templates of 41-81 samples at 30/40 kHz, all
    # submitted at once (the server saturated), each request compared with
    # the oracle on its whole (short) segment
    specs = synthetic.serving_mix(seed=17, n_requests=8, rate_hz=1e6, duration_range=(0.15, 0.3),
                                  channel_range=(16, 32), template_range=(20, 40), length_range=(41, 81))
    specs[0].update(n_chans=32, n_templates=40)
    reqs = [_req(s, return_residual=True) for s in specs]
    print("served shapes:", [(r.C, r.K, r.T, r.N) for r in reqs], flush=True)
    with serving.PursuitServer(max_inflight=4) as srv:
        futs = [srv.submit(r) for r in reqs]
        got = [f.result(timeout=300) for f in futs]
    for r, g in zip(reqs, got):
        o = oracle_mod.pursuit_chunk(r.voltage, r.templates, r.bias, r.thresholds, r.N, r.M, True, 256)
        print("request", r.tag, "spikes", o['idx'].size, "passes", o['n_passes'], flush=True)
        assert o['idx'].size > 0
        np.testing.assert_array_equal(g.onsets, o['idx'])
        np.testing.assert_array_equal(g.labels, o['lab'])
        np.testing.assert_array_equal(g.residual.view(np.uint32), o['residual'].view(np.uint32))


def test_dense_requests_outgrow_their_reservation():
    # more than two spikes per window: the spike list grows past its initial
    # capacity and the clips past the two-per-window reservation; the server
    # counts what they hold, and the results equal the standalone runs
    reqs = []
    for k in range(4):
        rec = synthetic.make_recording(40 + k, 6, 30000, 0.4, 14, 41, peak_range=(8.0, 14.0),
                                       rate_range=(250.0, 400.0), gains="gauss")
        sm = rec['separability_metrics']
        reqs.append(serving.SegmentRequest(rec['voltage'], sm['templates'], sm['template_SS'],
                                           sm['neuron_lower_thresholds'], 41, get_clips=True, tag=k))
    alone = [serving.run_segment(r) for r in reqs]
    W = -(-reqs[0].K // 41)
    assert all(a[0].size > 2 * W for a in alone), [a[0].size for a in alone]
    need = max(serving.footprint(r) for r in reqs)
    with serving.PursuitServer(max_inflight=4, hbm_budget=int(2.5 * need)) as srv:
        futs = [srv.submit(r) for r in reqs]
        got = [f.result(timeout=120) for f in futs]
        st = srv.stats()
    print("dense:", [g.onsets.size for g in got], "reserve", need, "held", [g.stats['device_bytes'] for g in got],
          st, flush=True)
    assert any(g.stats['device_bytes'] > srv.footprint(r) for g, r in zip(got, reqs))
    assert st['peak_running'] <= 4 and st['completed'] == 4
    for a, g in zip(alone, got):
        _same(a, g)


def test_cancel_queued_and_running():
    specs = synthetic.serving_mix(seed=5, n_requests=4, duration_range=(1.0, 1.0), channel_range=(8, 8),
                                  template_range=(8, 8), length_range=(61, 61))
    # the first request takes the single slot for a while (32 ch x 10 s): it
    # is cancelled while it is staged or running, the last one while queued
    specs[0].update(n_chans=32, n_templates=40, duration_s=10.0)
    reqs = [_req(s) for s in specs]
    with serving.PursuitServer(max_inflight=1) as srv:
        futs = [srv.submit(reqs[0])]
        assert srv.cancel(futs[0])
        futs += [srv.submit(r) for r in reqs[1:]]
        assert srv.cancel(futs[3])
        with pytest.raises(concurrent.futures.CancelledError):
            futs[3].result(timeout=60)
        with pytest.raises(concurrent.futures.CancelledError):
            futs[0].result(timeout=60)
        done = [futs[1].result(timeout=120), futs[2].result(timeout=120)]
        st = srv.stats()
    assert st['cancelled'] == 2 and st['completed'] == 2
    alone = serving.run_segment(reqs[1])
    np.testing.assert_array_equal(alone[0], done[0].onsets)


def test_server_c_abi_job_states():
    from spikesorting_fullpursuit_amd import _lib
    rec = synthetic.config("4ch-60s", seed=9, duration_s=0.5)
    sm = rec['separability_metrics']
    bias = np.float32(-0.5 * sm['template_SS'])
    thr = np.float32(sm['neuron_lower_thresholds'])
    with _lib.Server(0, 0, 2) as srv:
        t1 = srv.submit(rec['voltage'], sm['templates'], bias, thr, 61, 30)
        t2 = srv.submit(rec['voltage'], sm['templates'], bias, thr, 61, 30, max_passes=1)
        info = srv.wait(t1, timeout=60)
        assert info['state'] == "done" and info['rc'] == 0 and info['passes'] >= 2
        assert info['t_submit'] <= info['t_admit'] <= info['t_start'] <= info['t_finish']
        seg = srv.segment(t1)
        idx, lab = seg.spikes()
        assert seg.device_bytes() == info['bytes'] > 0
        info2 = srv.wait(t2, timeout=60)
        assert info2['state'] == "done" and info2['rc'] == _lib.FBP_EPASSES and info2['passes'] == 1
        idx2, _ = srv.segment(t2).spikes()
        assert 0 < idx2.size < idx.size
        ended = {srv.wait_any(timeout=10)[0], srv.wait_any(timeout=10)[0]}
        assert ended == {t1, t2}
        assert srv.wait_any(timeout=0.01) is None
        assert not srv.cancel(t1)
        srv.release(t1)
        srv.release(t2)
        with pytest.raises(ValueError):
            srv.poll(t1)
        with pytest.raises(MemoryError):
            srv.submit(rec['voltage'], sm['templates'], bias, thr, 61, 30, extra_bytes=1 << 62)
        st = srv.stats()
        assert st['completed'] == 2 and st['held_bytes'] == 0
    o = __import__("oracle").pursuit_chunk(rec['voltage'], sm['templates'], bias, thr, 61, 30, True, 256)
    np.testing.assert_array_equal(idx, o['idx'])
    np.testing.assert_array_equal(lab, o['lab'])
