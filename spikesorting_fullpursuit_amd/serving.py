"""Serving many recording segments on one GPU at once.

The reference runs binary pursuit one segment at a time, serially, in the
main process (spikesorting_fullpursuit/parallel/spikesorting_parallel.py:1312-1322).
PursuitServer instead admits segment requests as they arrive and runs them
concurrently: each admitted request owns one fbp_segment (its own HIP stream,
its own device-resident voltage, likelihood buffer and pass state), so the
passes of different requests overlap on the device.  A request is admitted
when a slot is free and its device footprint (fbp_segment_bytes) fits in the
HBM budget next to the requests already running; it releases both when it
finishes, and the next waiting request is admitted at once.  Results do not
depend on what else runs: every segment's kernels touch only its own buffers.
"""
import collections
import concurrent.futures
import threading
import time

import numpy as np

from . import _lib


class SegmentRequest:
    """One recording segment to sort.

    voltage: [C][K] array (any real dtype; cast to float32).
    templates: [T][C*N]; template_ss: [T] squared norms; thresholds: [T].
    n_max_shift_inds: overlap shift range (None -> N // 2).
    """

    def __init__(self, voltage, templates, template_ss, thresholds, n_samples_per_chan, n_max_shift_inds=None,
                 do_overlap=True, get_clips=False, return_residual=False, tag=None):
        self.voltage = voltage
        self.templates = np.ascontiguousarray(np.float32(templates))
        if self.templates.ndim == 1:
            self.templates = self.templates.reshape(1, -1)
        self.bias = np.float32(-0.5 * np.asarray(template_ss))
        self.thresholds = np.float32(thresholds)
        self.N = int(n_samples_per_chan)
        self.M = self.N // 2 if (n_max_shift_inds is None or n_max_shift_inds < 0) else int(n_max_shift_inds)
        self.do_overlap = bool(do_overlap)
        self.get_clips = bool(get_clips)
        self.return_residual = bool(return_residual)
        self.tag = tag
        self.C, self.K = int(voltage.shape[0]), int(voltage.shape[1])
        self.T = self.templates.shape[0]


class SegmentResult:
    def __init__(self, tag, onsets, labels, clips, residual, stats, submitted, admitted, finished):
        self.tag = tag
        self.onsets = onsets        # uint32 template onsets within the segment
        self.labels = labels        # uint32 template rows
        self.clips = clips          # float32 [n][C*N] or None
        self.residual = residual    # float32 [C][K] or None
        self.stats = stats
        self.submitted = submitted
        self.admitted = admitted
        self.finished = finished

    @property
    def latency(self):
        return self.finished - self.submitted

    @property
    def queued(self):
        return self.admitted - self.submitted


def run_segment(req, precision="bf16", overlap_mode="screened", reduce_wg=256, device=0):
    """Run one request to completion on the calling thread."""
    with _lib.Segment(req.C, req.K, req.T, req.N, req.M, req.do_overlap, reduce_wg, precision, overlap_mode, 0,
                      device) as seg:
        seg.set_templates(req.templates, req.bias, req.thresholds)
        seg.set_voltage(req.voltage)
        seg.run()
        onsets, labels = seg.spikes()
        clips = seg.clips() if req.get_clips else None
        residual = seg.residual() if req.return_residual else None
        return onsets, labels, clips, residual, seg.stats()


class PursuitServer:
    """Continuous admission of segment requests on one GPU.

    hbm_budget: device bytes the server may hold for running segments (default:
        85 % of the free HBM at start).
    max_inflight: segments running at once (one host thread and one HIP stream
        each).
    """

    def __init__(self, device=0, hbm_budget=None, max_inflight=8, precision="bf16", overlap_mode="screened",
                 reduce_wg=256, runner=None):
        self.device = device
        self.precision = precision
        self.overlap_mode = overlap_mode
        self.reduce_wg = reduce_wg
        self.max_inflight = int(max_inflight)
        self._runner = runner or (lambda req: run_segment(req, precision, overlap_mode, reduce_wg, device))
        if hbm_budget is None:
            _, free = _lib.device_memory(device)
            hbm_budget = int(0.85 * free)
        self.hbm_budget = int(hbm_budget)
        self._pending = collections.deque()
        self._cv = threading.Condition()
        self._used = 0
        self._inflight = 0
        self._closed = False
        self.peak_inflight = 0
        self.peak_bytes = 0
        self.results = []
        self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=self.max_inflight)
        self._dispatcher = threading.Thread(target=self._dispatch, name="pursuit-admission", daemon=True)
        self._dispatcher.start()

    def footprint(self, req):
        """Device bytes of one request: the segment's buffers plus, when clips
        are asked for, room for clips of up to two spikes per window."""
        b = _lib.segment_bytes(req.C, req.K, req.T, req.N, req.M, req.do_overlap, self.reduce_wg,
                               self.precision, self.overlap_mode)
        if req.get_clips:
            b += 4 * req.C * req.N * 2 * (-(-req.K // req.N))
        return b

    def submit(self, req):
        """Queue a request; returns a Future resolving to a SegmentResult."""
        fut = concurrent.futures.Future()
        need = self.footprint(req)
        if need > self.hbm_budget:
            fut.set_exception(MemoryError("segment needs %d bytes, above the server's HBM budget of %d"
                                          % (need, self.hbm_budget)))
            return fut
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            self._pending.append((req, need, fut, time.perf_counter()))
            self._cv.notify_all()
        return fut

    def _dispatch(self):
        while True:
            with self._cv:
                while not self._closed and not (
                        self._pending and self._inflight < self.max_inflight
                        and self._used + self._pending[0][1] <= self.hbm_budget):
                    self._cv.wait()
                if self._closed and not self._pending:
                    return
                if not (self._pending and self._inflight < self.max_inflight
                        and self._used + self._pending[0][1] <= self.hbm_budget):
                    self._cv.wait(0.05)
                    continue
                req, need, fut, t_sub = self._pending.popleft()
                self._used += need
                self._inflight += 1
                self.peak_inflight = max(self.peak_inflight, self._inflight)
                self.peak_bytes = max(self.peak_bytes, self._used)
            self._pool.submit(self._work, req, need, fut, t_sub, time.perf_counter())

    def _work(self, req, need, fut, t_sub, t_adm):
        try:
            onsets, labels, clips, residual, stats = self._runner(req)
            res = SegmentResult(req.tag, onsets, labels, clips, residual, stats, t_sub, t_adm, time.perf_counter())
        except BaseException as exc:   # noqa: BLE001 -- handed to the caller through the future
            res = exc
        with self._cv:
            self._used -= need
            self._inflight -= 1
            if not isinstance(res, BaseException):
                self.results.append(res)
            self._cv.notify_all()
        if isinstance(res, BaseException):
            fut.set_exception(res)
        else:
            fut.set_result(res)

    def close(self, wait=True):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            self._dispatcher.join()
            self._pool.shutdown(wait=True)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def latency_summary(self):
        lat = np.array([r.latency for r in self.results])
        q = np.array([r.queued for r in self.results])
        if lat.size == 0:
            return {}
        return {"n": int(lat.size), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "mean_s": float(lat.mean()), "max_s": float(lat.max()), "queue_p50_s": float(np.percentile(q, 50)),
                "queue_p99_s": float(np.percentile(q, 99)), "peak_inflight": self.peak_inflight,
                "peak_bytes": int(self.peak_bytes)}
