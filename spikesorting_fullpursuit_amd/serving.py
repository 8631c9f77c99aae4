"""Serving many recording segments on one GPU at once.

The reference runs binary pursuit one segment at a time, serially, in the
main process (spikesorting_fullpursuit/parallel/spikesorting_parallel.py:1312-1322).
PursuitServer admits segment requests as they arrive and runs them
concurrently through the native segment server of libfbp (fbp_server_*):
each admitted request owns one fbp_segment (its own HIP stream,
device-resident voltage, likelihood buffer and pass state) and one native
driver thread advances every running segment pass by pass, so the passes of
different requests overlap on the device and no thread blocks on a pass.  A
request is admitted, in arrival order, when a slot is free and its
reservation fits in the HBM budget next to what the admitted requests hold
(the native side counts each request for its reservation or for what it holds
now, whichever is larger, so spike lists and clips that outgrow the
reservation hold back further admissions).  A collector thread takes finished
requests in completion order, copies their results to the host and releases
their device memory at once, whatever order the caller asks for results in.
Results do not depend on what else runs: every segment's kernels touch only
its own buffers.
"""
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
        self.submitted = submitted  # time.perf_counter() values
        self.admitted = admitted
        self.finished = finished

    @property
    def latency(self):
        return self.finished - self.submitted

    @property
    def queued(self):
        return self.admitted - self.submitted


def _outputs(seg, req):
    onsets, labels = seg.spikes()
    clips = seg.clips() if req.get_clips else None
    residual = seg.residual() if req.return_residual else None
    return onsets, labels, clips, residual, seg.stats()


def run_segment(req, precision="bf16", overlap_mode="screened", reduce_wg=256, device=0):
    """Run one request to completion on the calling thread, alone."""
    with _lib.Segment(req.C, req.K, req.T, req.N, req.M, req.do_overlap, reduce_wg, precision, overlap_mode, 0,
                      device) as seg:
        seg.set_templates(req.templates, req.bias, req.thresholds)
        seg.set_voltage(req.voltage)
        seg.run()
        return _outputs(seg, req)


def clips_bytes(req):
    """Device bytes reserved for a request's clips: two spikes per window (a
    dense segment may need more; what it then holds is counted as it grows)."""
    return 4 * req.C * req.N * 2 * (-(-req.K // req.N)) if req.get_clips else 0


def footprint(req, precision="bf16", overlap_mode="screened", reduce_wg=256):
    """Device bytes a request reserves at admission: its segment's buffers
    (fbp_segment_bytes) and room for its clips."""
    return _lib.segment_bytes(req.C, req.K, req.T, req.N, req.M, req.do_overlap, reduce_wg, precision,
                              overlap_mode) + clips_bytes(req)


class PursuitServer:
    """Continuous admission of segment requests on one GPU.

    hbm_budget: device bytes the server may hold for admitted segments
        (default: 85 % of the free HBM at start).
    max_inflight: segments staged or running at once.
    """

    def __init__(self, device=0, hbm_budget=None, max_inflight=8, precision="bf16", overlap_mode="screened",
                 reduce_wg=256):
        self.device = device
        self.precision = precision
        self.overlap_mode = overlap_mode
        self.reduce_wg = reduce_wg
        self.max_inflight = int(max_inflight)
        self._srv = _lib.Server(device, 0 if hbm_budget is None else int(hbm_budget), self.max_inflight)
        self.hbm_budget = self._srv.stats()["budget_bytes"]
        self._cv = threading.Condition()
        self._pending = {}          # ticket -> (request, future, submit time)
        self._closing = False
        self.results = []
        self._collector = threading.Thread(target=self._collect, name="pursuit-results", daemon=True)
        self._collector.start()

    def clips_bytes(self, req):
        return clips_bytes(req)

    def footprint(self, req):
        """Device bytes a request reserves at admission."""
        return footprint(req, self.precision, self.overlap_mode, self.reduce_wg)

    def submit(self, req):
        """Queue a request; returns a Future resolving to a SegmentResult.
        The future's `ticket` attribute names the native job."""
        fut = concurrent.futures.Future()
        t_sub = time.perf_counter()
        with self._cv:
            if self._closing:
                raise RuntimeError("server is closed")
            try:
                ticket = self._srv.submit(req.voltage, req.templates, req.bias, req.thresholds, req.N, req.M,
                                          req.do_overlap, self.reduce_wg, self.precision, self.overlap_mode,
                                          extra_bytes=self.clips_bytes(req))
            except (MemoryError, ValueError) as exc:
                fut.set_exception(exc)
                return fut
            fut.ticket = ticket
            self._pending[ticket] = (req, fut, t_sub)
        return fut

    def cancel(self, fut):
        """Cancel a submitted request (queued: at once; running: at its next
        pass boundary).  Its future then raises CancelledError.  False if it
        had already ended."""
        return self._srv.cancel(fut.ticket)

    def _collect(self):
        while True:
            got = self._srv.wait_any(timeout=0.05)
            if got is None:
                with self._cv:
                    if self._closing and not self._pending:
                        return
                continue
            ticket, info = got
            with self._cv:
                req, fut, t_sub = self._pending[ticket]
            res = None
            try:
                if info["state"] == "done":
                    seg = self._srv.segment(ticket)
                    onsets, labels, clips, residual, stats = _outputs(seg, req)
                    stats["passes_driven"] = info["passes"]
                    stats["device_bytes"] = seg.device_bytes()   # clips included
                    stats["reserved_bytes"] = info["reserve_bytes"]
                    admitted = t_sub + (info["t_admit"] - info["t_submit"])
                    res = SegmentResult(req.tag, onsets, labels, clips, residual, stats, t_sub, admitted,
                                        time.perf_counter())
                elif info["state"] == "failed":
                    res = _lib.FbpError(info["rc"], self._srv.error(ticket))
                else:
                    res = concurrent.futures.CancelledError()
            except BaseException as exc:   # noqa: BLE001 -- handed to the caller through the future
                res = exc
            self._srv.release(ticket)
            with self._cv:
                del self._pending[ticket]
                if isinstance(res, SegmentResult):
                    self.results.append(res)
                self._cv.notify_all()
            if isinstance(res, BaseException):
                fut.set_exception(res)
            else:
                fut.set_result(res)

    def close(self, wait=True):
        """Stop accepting requests; with wait, finish the submitted ones first
        (otherwise they are cancelled)."""
        with self._cv:
            self._closing = True
            if wait:
                while self._pending:
                    self._cv.wait(0.1)
            else:
                for ticket in list(self._pending):
                    self._srv.cancel(ticket)
        self._collector.join()
        self._final_stats = self._srv.stats()
        self._srv.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def stats(self):
        """Counters of the native server (jobs, bytes, driver thread time)."""
        if self._srv._h is None:
            return dict(self._final_stats)
        return self._srv.stats()

    def latency_summary(self):
        lat = np.array([r.latency for r in self.results])
        q = np.array([r.queued for r in self.results])
        if lat.size == 0:
            return {}
        st = self.stats()
        return {"n": int(lat.size), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "mean_s": float(lat.mean()), "max_s": float(lat.max()), "queue_p50_s": float(np.percentile(q, 50)),
                "queue_p99_s": float(np.percentile(q, 99)), "peak_inflight": st.get("peak_running"),
                "peak_bytes": st.get("peak_held_bytes"), "pass_steps": st.get("pass_steps"),
                "driver_busy_s": st.get("driver_busy_s"), "uptime_s": st.get("uptime_s")}
