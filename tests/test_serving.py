"""Serving: admission by slots and HBM budget (CPU, with a stand-in runner),
and identical results alone and in a concurrent mix (GPU)."""
import threading
import time

import numpy as np
import pytest

from spikesorting_fullpursuit_amd import serving, synthetic


def _req(spec, **kw):
    rec = synthetic.recording_for(spec)
    sm = rec['separability_metrics']
    return serving.SegmentRequest(rec['voltage'], sm['templates'], sm['template_SS'], sm['neuron_lower_thresholds'],
                                  spec['n_samples_per_chan'], tag=spec['seed'], **kw)


def test_admission_respects_slots_and_budget():
    specs = synthetic.serving_mix(seed=1, n_requests=12, duration_range=(0.05, 0.2), channel_range=(2, 6),
                                  template_range=(2, 5), length_range=(21, 41))
    reqs = [_req(s) for s in specs]
    state = {"run": 0, "peak": 0, "bytes": 0, "peak_bytes": 0}
    lock = threading.Lock()
    srv_holder = {}

    def runner(req):
        need = srv_holder['srv'].footprint(req)
        with lock:
            state["run"] += 1
            state["bytes"] += need
            state["peak"] = max(state["peak"], state["run"])
            state["peak_bytes"] = max(state["peak_bytes"], state["bytes"])
        time.sleep(0.02)
        with lock:
            state["run"] -= 1
            state["bytes"] -= need
        return np.zeros(0, np.uint32), np.zeros(0, np.uint32), None, None, {}

    sizes = None
    srv = serving.PursuitServer(hbm_budget=1, max_inflight=3, runner=runner)
    srv_holder['srv'] = srv
    sizes = [srv.footprint(r) for r in reqs]
    srv.close()
    budget = int(sorted(sizes)[-1] + sorted(sizes)[-2] * 0.99)   # at most ~2 of the big ones together
    srv = serving.PursuitServer(hbm_budget=budget, max_inflight=3, runner=runner)
    srv_holder['srv'] = srv
    futs = [srv.submit(r) for r in reqs]
    res = [f.result(timeout=30) for f in futs]
    srv.close()
    assert len(res) == 12 and [r.tag for r in res] == [r.tag for r in reqs]
    assert state["peak"] <= 3 and state["peak_bytes"] <= budget
    assert srv.peak_bytes <= budget and srv.peak_inflight <= 3
    # a request larger than the budget is refused, others still run
    small = serving.PursuitServer(hbm_budget=min(sizes) - 1, max_inflight=2, runner=runner)
    with pytest.raises(MemoryError):
        small.submit(reqs[0]).result(timeout=5)
    small.close()


def test_runner_errors_reach_the_caller():
    spec = synthetic.serving_mix(seed=2, n_requests=1, duration_range=(0.05, 0.05), channel_range=(2, 2),
                                 template_range=(2, 2), length_range=(21, 21))[0]

    def runner(req):
        raise ValueError("boom")

    srv = serving.PursuitServer(hbm_budget=10 ** 12, max_inflight=2, runner=runner)
    with pytest.raises(ValueError, match="boom"):
        srv.submit(_req(spec)).result(timeout=10)
    srv.close()


@pytest.mark.gpu
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
    for a, m in zip(alone, mixed):
        np.testing.assert_array_equal(a[0], m.onsets)
        np.testing.assert_array_equal(a[1], m.labels)
        np.testing.assert_array_equal(a[2].view(np.uint32), m.clips.view(np.uint32))
        np.testing.assert_array_equal(a[3].view(np.uint32), m.residual.view(np.uint32))
