#!/usr/bin/env python
"""Binary-pursuit throughput on one MI355X.

One step = one full binary pursuit (every pass to convergence) over the
16-channel, 40 kHz, 300 s recording with 20 templates of 61 samples, the
voltage already resident in HBM (uploaded once, before the timed region; each
step restores it on the device and turns it into the residual again).

Prints one JSON line with the whole-job throughput in channel-samples per
second, the roofline position of the dominant kernel (HIP events recorded by
the library on the stream its kernels run on) and a CPU baseline: the
oracle's C restatement of the reference algorithm, timed on the host's cores
over a bounded leading slice of the same recording.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pursuit ch·samples/s FP32 vs BF16-screened at equal events; p99 segment latency"

# MI355X published peaks.
PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {"fp32": 157.3, "bf16": 2500.0, "fp16": 2500.0, "fp8": 5000.0, "fp64": 78.6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="16ch-300s")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--duration", type=float, default=None, help="override the recording length (s)")
    ap.add_argument("--max-shift", type=int, default=None, help="n_max_shift_inds (default N//2)")
    ap.add_argument("--precision", default="bf16", help="correlation arithmetic of the timed path")
    ap.add_argument("--overlap-mode", default="screened", help="overlap arithmetic of the timed path")
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="steps of the plain float32 path (fp32 correlation, direct overlap) timed for comparison")
    ap.add_argument("--full-shift-steps", type=int, default=2,
                    help="steps of the timed path at n_max_shift_inds = N - 1 (the full pipeline's setting)")
    ap.add_argument("--large-probe-s", type=float, default=120.0,
                    help="seconds of the 128-channel, 200-template recording timed as its own entry (0 = skip)")
    ap.add_argument("--no-overlap", action="store_true")
    ap.add_argument("--cpu-target-s", type=float, default=12.0, help="CPU baseline budget (s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=64, help="requests of the served mix (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=16.0, help="Poisson arrival rate of the mix (1/s)")
    ap.add_argument("--serve-inflight", type=int, default=8)
    ap.add_argument("--no-precision-report", action="store_true",
                    help="skip the likelihood error report against the float64 oracle (4 ch x 60 s)")
    ap.add_argument("--no-post", action="store_true",
                    help="skip the post-pursuit clip gather and unit means (get_clips on the input voltage)")
    ap.add_argument("--segments", type=int, default=0,
                    help="also run the recording as this many equal time segments concurrently (one HIP stream "
                         "each) and report their throughput")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: every box in this setup has a single MI355X, "
              "so --gpus %d is not supported (run with --gpus 1)." % args.gpus, file=sys.stderr)
        return 2

    from spikesorting_fullpursuit_amd import _lib, synthetic

    t0 = time.time()
    rec = synthetic.config(args.workload, seed=args.seed, duration_s=args.duration)
    sm = rec['separability_metrics']
    V = rec['voltage']
    C, K = V.shape
    T = sm['templates'].shape[0]
    N = rec['sort_info']['n_samples_per_chan']
    fs = rec['sort_info']['sampling_rate']
    M = N // 2 if args.max_shift is None else args.max_shift
    do_overlap = not args.no_overlap
    bias = np.float32(-0.5 * sm['template_SS'])
    thr = np.float32(sm['neuron_lower_thresholds'])
    log("workload %s: C=%d K=%d T=%d N=%d M=%d, %d true spikes (generated in %.1f s)"
        % (args.workload, C, K, T, N, M, sum(e.size for e in rec['events']), time.time() - t0))

    seg = _lib.Segment(C, K, T, N, M, do_overlap, 256, args.precision, args.overlap_mode)
    seg.set_templates(sm['templates'], bias, thr)
    seg.set_voltage(V)
    for w in range(args.warmup):
        t = time.perf_counter()
        seg.run()
        log("warmup %d: %.3f s, %d spikes, %d passes" % (w, time.perf_counter() - t, seg.num_spikes(),
                                                          seg.stats()['n_passes']))
    sums = {"ms_correlation": 0.0, "ms_overlap": 0.0, "ms_total": 0.0, "corr_launches": 0,
            "overlap_launches": 0, "corr_positions": 0, "overlap_combos": 0, "screened_combos": 0,
            "corr_pairs": 0,
            "overlap_windows": 0, "rescored_positions": 0}
    t_start = time.perf_counter()
    for k in range(args.steps):
        seg.run()
        st = seg.stats()
        for key in sums:
            sums[key] += st[key]
        log("step %d done: %d spikes, %d passes, device %.1f ms" % (k, st['n_spikes'], st['n_passes'], st['ms_total']))
    elapsed = time.perf_counter() - t_start
    n_spikes = seg.num_spikes()
    last = seg.stats()
    ms_per_step = 1000.0 * elapsed / args.steps
    value = C * K * args.steps / elapsed

    stages, roofline = roofline_of(args, sums, C, T, N, M, K, int(last["n_passes"]),
                                   support_channels(sm['templates'], C, N))

    plain = None
    if args.plain_steps > 0 and (args.precision != "fp32" or args.overlap_mode != "direct"):
        plain, _ = pursuit_entry(args, rec, "fp32", "direct", M, args.plain_steps, warmup=0,
                                 ref_spikes=seg.spikes())
        plain["path"] = "fp32 correlation + direct overlap (reference arithmetic)"
        plain["speedup_of_timed_path"] = value / plain["value"]
        log("plain fp32 path: %.1f ms per step, spike lists identical: %s" % (plain["ms_per_step"],
                                                                           plain["events_equal"]))

    m_full = None
    if args.full_shift_steps > 0 and M != N - 1:
        try:
            m_full, m_spk = pursuit_entry(args, rec, args.precision, args.overlap_mode, N - 1, args.full_shift_steps)
        except ValueError as exc:   # e.g. the screened overlap's per-window tables exceed LDS
            m_full = {"skipped": str(exc)}
            log("n_max_shift_inds = %d: skipped (%s)" % (N - 1, exc))
    if m_full is not None and "skipped" not in m_full:
        if args.plain_steps > 0:
            ref, _ = pursuit_entry(args, rec, "fp32", "direct", N - 1, 1, warmup=0, ref_spikes=m_spk)
            m_full["fp32_plain"] = {k: ref[k] for k in ("value", "ms_per_step", "n_spikes", "n_passes", "stages",
                                                        "roofline")}
            m_full["events_equal_fp32_plain"] = ref["events_equal"]
        log("n_max_shift_inds = %d: %.1f ms per step (%.3g ch*samples/s), %d passes"
            % (N - 1, m_full["ms_per_step"], m_full["value"], m_full["n_passes"]))

    large = None
    if args.large_probe_s > 0:
        lrec = synthetic.config("128ch-600s", seed=args.seed, duration_s=args.large_probe_s)
        LN = lrec['sort_info']['n_samples_per_chan']
        large, _ = pursuit_entry(args, lrec, args.precision, args.overlap_mode, LN // 2, 1, warmup=1,
                                 workload="128ch-600s", duration=args.large_probe_s)
        large["workload"] = "128ch-600s"
        large["duration_s"] = lrec['voltage'].shape[1] / lrec['sort_info']['sampling_rate']
        log("large probe (128 ch, 200 templates, %.0f s): %.2f s per pursuit, %d passes, %.3g ch*samples/s"
            % (large["duration_s"], large["ms_per_step"] / 1000, large["n_passes"], large["value"]))
        del lrec

    prec = None
    if not args.no_precision_report:
        prec = precision_report(args)

    serve = None
    if args.serve_requests > 0:
        serve = serving_mix_run(args)

    concurrent = None
    if args.segments > 1:
        concurrent = concurrent_segments_run(args, V, sm, bias, thr, N, M, do_overlap)

    post = None
    if not args.no_post:
        post = post_pursuit_run(seg, V, rec['sort_info'], T)
        post["overlap_templates"] = overlap_templates_run(sm['templates'], C, N, rec['sort_info'], seg,
                                                          cpu=not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(rec, bias, thr, N, M, do_overlap, args.cpu_target_s, seg_cls=_lib.Segment,
                           precision=args.precision, overlap_mode=args.overlap_mode)
    seg.close()

    out = {
        "metric": METRIC, "value": value, "unit": "ch·samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.precision if args.precision != "fp32" else "f32", "data": "synthetic (seeded N(0,1) noise + DoG templates)",
        "config": {"workload": args.workload, "n_chans": C, "sampling_rate": fs, "duration_s": K / fs,
                   "n_templates": T, "template_len": N, "n_max_shift_inds": M, "do_overlap_recheck": do_overlap,
                   "precision": args.precision, "overlap_mode": args.overlap_mode, "seed": args.seed,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "result": {"n_spikes": int(n_spikes), "n_passes": int(last["n_passes"]),
                   "windows_evaluated": int(last["windows_evaluated"]),
                   "overlap_windows": int(last["overlap_windows"]), "device_ms_per_step": sums["ms_total"] / args.steps,
                   "rescored_fraction": sums["rescored_positions"] / max(1, sums["corr_positions"] * T)},
        "stages": stages,
        "fp32_plain": plain,
        "full_shift_range": m_full,
        "large_probe": large,
        "serving": serve,
        "concurrent_segments": concurrent,
        "post_pursuit": post,
        "precision_errors": prec,
    }
    print(json.dumps(out), flush=True)
    return 0


def pursuit_entry(args, rec, precision, overlap_mode, M, steps, warmup=1, workload=None, ref_spikes=None,
                  duration=None):
    """Time `steps` full pursuits of one recording on one segment (voltage
    resident in HBM) with the given arithmetic and shift range; returns the
    throughput, the stage split and the roofline entry of its dominant stage.
    With ref_spikes = (onsets, labels), also whether the spike list is equal."""
    from spikesorting_fullpursuit_amd import _lib
    sm = rec['separability_metrics']
    V = rec['voltage']
    C, K = V.shape
    T = sm['templates'].shape[0]
    N = rec['sort_info']['n_samples_per_chan']
    bias = np.float32(-0.5 * sm['template_SS'])
    thr = np.float32(sm['neuron_lower_thresholds'])
    sums = {"ms_correlation": 0.0, "ms_overlap": 0.0, "ms_total": 0.0, "corr_launches": 0,
            "overlap_launches": 0, "corr_positions": 0, "overlap_combos": 0, "screened_combos": 0,
            "corr_pairs": 0,
            "overlap_windows": 0, "rescored_positions": 0}
    with _lib.Segment(C, K, T, N, M, True, 256, precision, overlap_mode) as sg:
        sg.set_templates(sm['templates'], bias, thr)
        sg.set_voltage(V)
        for _ in range(warmup):
            sg.run()
        t = time.perf_counter()
        for _ in range(steps):
            sg.run()
            st = sg.stats()
            for key in sums:
                sums[key] += st[key]
        dt = (time.perf_counter() - t) / steps
        idx, lab = sg.spikes()
    ns = argparse.Namespace(precision=precision, overlap_mode=overlap_mode, workload=workload or args.workload,
                            duration=duration if workload else args.duration)
    stages, roof = roofline_of(ns, sums, C, T, N, M, K, int(st["n_passes"]), support_channels(sm['templates'], C, N))
    out = {"value": C * K / dt, "unit": "ch·samples/s", "ms_per_step": 1000.0 * dt, "steps": steps,
           "precision": precision, "overlap_mode": overlap_mode, "n_max_shift_inds": M,
           "n_spikes": int(idx.size), "n_passes": int(st["n_passes"]),
           "device_ms_per_step": sums["ms_total"] / steps, "stages": stages, "roofline": roof}
    if ref_spikes is not None:
        out["events_equal"] = bool(np.array_equal(idx, ref_spikes[0]) and np.array_equal(lab, ref_spikes[1]))
    return out, (idx, lab)


def post_pursuit_run(seg, V, sort_info, T, repeats=3, max_clip_bytes=4 << 30):
    """The caller's clips after binary pursuit (full_binary_pursuit.py:357-368):
    get_clips of every found spike on all channels of the resident input
    voltage (fbp_get_clips, clips copied to the host) and the per-unit mean
    clips (fbp_unit_mean_clips, clips uploaded).  Times are whole calls,
    PCIe copies included; the numpy gather of the same clips is timed beside
    them on the host.  Host memory holds a few copies of the clips, so a
    recording whose clips exceed max_clip_bytes uses its earliest spikes up to
    that size."""
    from spikesorting_fullpursuit_amd import _lib
    from spikesorting_fullpursuit_amd.segment import get_clips, time_window_to_samples
    idx, lab = seg.spikes()
    C, K = V.shape
    window, _ = time_window_to_samples(sort_info['clip_width'], sort_info['sampling_rate'])
    n_found = idx.size
    n_keep = min(n_found, max(1, max_clip_bytes // (4 * C * (window[1] - window[0]))))
    if n_keep < n_found:
        first = np.argsort(idx, kind="stable")[:n_keep]
        idx, lab = idx[first], lab[first]
    ev = idx.astype(np.int64) - window[0]   # event indices = onsets + clip_init_samples
    pd = {'sampling_rate': sort_info['sampling_rate'], 'n_samples': K, 'v_dtype': np.float32}
    chans = np.arange(C)
    t_clip, t_mean = [], []
    for _ in range(repeats):
        t = time.perf_counter()
        clips, valid = get_clips(pd, seg, chans, ev, sort_info['clip_width'])
        t_clip.append(time.perf_counter() - t)
        t = time.perf_counter()
        means, counts = _lib.unit_mean_clips(clips, lab, T)
        t_mean.append(time.perf_counter() - t)
    W = window[1] - window[0]
    rows = np.arange(0, ev.size, max(1, ev.size // 2000))
    sample_ok = all(np.array_equal(clips[r], V[:, ev[r] + window[0]:ev[r] + window[1]].ravel()) for r in rows)
    unit = int(np.argmax(counts))
    mean_ok = bool(np.array_equal(means[unit].view(np.uint32),
                                  np.mean(clips[lab == unit], axis=0).view(np.uint32)))
    # duplicate removal after binary pursuit (postprocessing.py:521-546) on
    # the found spikes as one channel and segment, events ascending
    from spikesorting_fullpursuit_amd.postprocessing import remove_segment_binary_pursuit_duplicates
    order = np.argsort(ev, kind="stable")
    t = time.perf_counter()
    sd = [[[ev[order], lab[order].astype(np.int64), clips[order], np.ones(ev.size, dtype=bool), 0]]]
    remove_segment_binary_pursuit_duplicates(sd, 1, 1, -window[0])
    t_dup = time.perf_counter() - t
    n_dup = ev.size - sd[0][0][0].size
    t = time.perf_counter()
    host = np.empty_like(clips)
    for r in range(ev.size):
        host[r] = V[:, ev[r] + window[0]:ev[r] + window[1]].ravel()
    t_host = time.perf_counter() - t
    del host
    thresholds = threshold_inputs_run(clips, lab, means, counts, C, W, K)
    nbytes = clips.nbytes
    tc, tm = float(np.median(t_clip)), float(np.median(t_mean))
    log("post-pursuit: %d clips x %d floats, get_clips %.1f ms, unit means %.1f ms, host gather %.1f ms"
        % (ev.size, C * W, 1e3 * tc, 1e3 * tm, 1e3 * t_host))
    return {"n_clips": int(ev.size), "n_spikes_found": int(n_found), "clip_floats": int(C * W),
            "clip_bytes": int(nbytes),
            "get_clips_ms": 1e3 * tc, "get_clips_GBps": nbytes / tc / 1e9,
            "unit_means_ms": 1e3 * tm, "n_units": int(np.count_nonzero(counts)),
            "threshold_inputs": thresholds,
            "host_numpy_gather_ms": 1e3 * t_host, "all_valid": bool(valid.all()),
            "duplicate_removal_ms": 1e3 * t_dup, "duplicates_removed": int(n_dup),
            "sampled_clips_equal_voltage": bool(sample_ok), "largest_unit_mean_equals_numpy": mean_ok,
            "note": "whole calls: get_clips includes the device-to-host copy of the clips, unit means the "
                    "host-to-device copy; the gather and mean kernels are in the rocprofv3 kernel table"}


def threshold_inputs_run(clips, lab, means, counts, C, W, n_samples, n_cov_samples=100000):
    """The inputs of the binary-pursuit thresholds from the found spikes:
    residual clips (clip - unit mean), per-channel noise covariance over
    n_cov_samples rows drawn per channel (full_binary_pursuit.py:243-252;
    fbp_channel_covariance) and the separability metrics of the unit means
    (neuron_separability.py:88-152; likelihood variances in
    fbp_template_variances).  numpy's np.cov over the same rows is timed on
    the host beside it."""
    from spikesorting_fullpursuit_amd.neuron_separability import (clip_noise_covariance,
                                                                  compute_separability_metrics)
    units = np.flatnonzero(counts)
    resid = clips - means[lab]
    si = {'n_channels': C, 'n_samples_per_chan': W, 'n_samples': n_samples,
          'n_threshold_crossings': np.zeros(C), 'sigma_bp_noise': 3.719, 'sigma_bp_CI': np.inf}
    t_cov, t_sep = [], []
    for _ in range(3):
        t = time.perf_counter()
        covs = clip_noise_covariance(resid, C, W, n_cov_samples, rng=np.random.RandomState(0))
        t_cov.append(time.perf_counter() - t)
        t = time.perf_counter()
        sm = compute_separability_metrics(list(means[units].astype(np.float64)), covs, si)
        t_sep.append(time.perf_counter() - t)
    draw = np.random.RandomState(0)
    n_rows = min(n_cov_samples, resid.shape[0])
    t = time.perf_counter()
    host = []
    for c in range(C):
        rows = draw.randint(0, resid.shape[0], n_cov_samples) if n_cov_samples < resid.shape[0] else slice(None)
        host.append(np.cov(resid[rows, c * W:(c + 1) * W], rowvar=False, ddof=0))
    t_host = time.perf_counter() - t
    err = max(float(np.max(np.abs(covs[c] - host[c])) / np.max(np.abs(host[c]))) for c in range(C))
    log("threshold inputs: covariance of %d channels x %d rows %.1f ms (numpy %.1f ms), metrics of %d units %.1f ms"
        % (C, n_rows, 1e3 * np.median(t_cov), 1e3 * t_host, units.size, 1e3 * np.median(t_sep)))
    return {"covariance_ms": 1e3 * float(np.median(t_cov)), "rows_per_channel": int(n_rows),
            "numpy_cov_ms": 1e3 * t_host, "max_rel_diff_vs_numpy": err,
            "separability_metrics_ms": 1e3 * float(np.median(t_sep)), "units": int(units.size),
            "note": "whole calls, residual clips uploaded per call; np.cov of the same rows on the host"}


def overlap_templates_run(templates, C, N, sort_info, seg, cpu=True, repeats=3):
    """find_overlap_templates (sort_cython.pyx:333-456) on the workload's
    templates with the shift range full_binary_pursuit uses
    (full_binary_pursuit.py:257-262: |chan_win| - 1 on each side) and the
    found spike counts per unit.  The GPU call is timed whole; the oracle
    (the reference's loops in numpy) is timed on the first test unit's
    candidate pairs and scaled by the number of pairs."""
    from spikesorting_fullpursuit_amd.overlap_templates import _shifted as _shift_f64
    from spikesorting_fullpursuit_amd.overlap_templates import candidate_pairs, find_overlap_templates
    from spikesorting_fullpursuit_amd.segment import time_window_to_samples
    chan_win, _ = time_window_to_samples(sort_info['clip_width'], sort_info['sampling_rate'])
    n_pre, n_post = abs(chan_win[0]) - 1, abs(chan_win[1]) - 1
    tm = np.ascontiguousarray(templates, dtype=np.float32)
    T = tm.shape[0]
    _, lab = seg.spikes()
    spikes = np.bincount(lab, minlength=T).astype(np.int64)
    times = []
    for _ in range(repeats):
        t = time.perf_counter()
        found, res = find_overlap_templates(tm, N, C, n_pre, n_post, spikes, return_residuals=True)
        times.append(time.perf_counter() - t)
    order = np.argsort(np.sum(tm ** 2, axis=1))[-1::-1]
    pairs = [(u, candidate_pairs(tm, u, N, spikes)) for u in order]
    n_triples = sum(len(p) for _, p in pairs)
    S = n_pre + n_post + 1
    out = {"n_templates": int(T), "shift_range": [-int(n_pre), int(n_post)], "candidate_triples": int(n_triples),
           "residuals": int(n_triples * S * S), "gpu_ms": 1e3 * float(np.median(times)), "found": len(found),
           "work": "per residual: C*N float64 squared differences summed in numpy's pairwise order"}
    if cpu and n_triples:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        # the oracle's shift-pair loops (the reference's, in numpy) for the
        # first candidate triple, scaled by the number of triples
        u, (n1, n2) = next((u, p[0]) for u, p in pairs if p)
        sh = [np.stack([_shift_f64(tm[n], s, C, N) for s in range(-n_pre, n_post + 1)]) for n in (n1, n2)]
        t = time.perf_counter()
        oracle._triple_scan(tm[u], sh[0], sh[1], np.zeros(C * N), np.inf, None, n1, n2)
        dt = time.perf_counter() - t
        out["cpu_oracle"] = {"sample": "one candidate triple (%d shift pairs)" % (S * S), "seconds": dt,
                             "estimated_seconds_all": dt * n_triples, "cores": 1}
    log("overlap templates: %d candidate triples x %d shift pairs, GPU %.2f ms, %d found"
        % (n_triples, S * S, out["gpu_ms"], len(found)))
    return out


def precision_report(args):
    """Errors of the first-pass likelihood against float64 on the 4-channel,
    60 s workload (5 templates x 61 samples): the float32 chain, the raw bf16
    and fp16 matrix-core screens (before any float32 re-score) and the screen
    bound; and whether each path's full run returns the float32 spike list."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from spikesorting_fullpursuit_amd import _lib, synthetic
    rec = synthetic.config("4ch-60s", seed=args.seed)
    sm = rec['separability_metrics']
    V = rec['voltage']
    C, K = V.shape
    T = sm['templates'].shape[0]
    N = 61
    M = N // 2
    bias = np.float32(-0.5 * sm['template_SS'])
    thr = np.float32(sm['neuron_lower_thresholds'])
    t = time.perf_counter()
    L64 = oracle.likelihood_f64(V, sm['templates'], sm['template_SS'].astype(np.float32), N)
    log("float64 likelihood (oracle, host) in %.1f s" % (time.perf_counter() - t))
    valid = np.arange(K) + N < K

    def run(precision, debug=0, max_passes=0, overlap="screened"):
        with _lib.Segment(C, K, T, N, M, True, 256, precision, overlap, max_passes, debug_flags=debug) as seg:
            seg.set_templates(sm['templates'], bias, thr)
            seg.set_voltage(V)
            seg.run(allow_cap=max_passes > 0)
            return seg.likelihood(), seg.spikes(), seg.stats()

    L32, ev32, _ = run("fp32", overlap="direct")
    L32p1, _, _ = run("fp32", max_passes=1)
    out = {"workload": "4ch-60s (4 ch, 40 kHz, 60 s, 5 templates x 61)", "first_pass": {}}
    tn = np.sqrt(np.sum(sm['templates'].astype(np.float64) ** 2, axis=1))
    s2 = np.concatenate([np.sum(V.astype(np.float64) ** 2, axis=0), np.zeros(N)])
    cs = np.concatenate([[0.0], np.cumsum(s2)])
    vn = np.sqrt(cs[N:N + K] - cs[:K])
    n32 = C * N + 1
    g32 = n32 * 2.0 ** -24 / (1 - n32 * 2.0 ** -24)
    b32 = g32 * (np.abs(bias.astype(np.float64))[:, None] + tn[:, None] * vn[None, :])
    e32 = np.abs(L32p1.astype(np.float64) - L64)[:, valid]
    out["first_pass"]["fp32"] = {"max_abs_err": float(e32.max()), "median_abs_err": float(np.median(e32)),
                                 "max_err_over_gamma_bound": float((e32 / b32[:, valid]).max())}
    for prec in ("bf16", "fp16", "fp8"):
        Lr, _, _ = run(prec, debug=1, max_passes=1)
        e = np.abs(Lr.astype(np.float64) - L64)[:, valid]
        e_vs32 = np.abs(Lr.astype(np.float64) - L32p1.astype(np.float64))[:, valid]
        _, evp, stp = run(prec)
        same = bool(np.array_equal(evp[0], ev32[0]) and np.array_equal(evp[1], ev32[1]))
        out["first_pass"][prec + "_screen"] = {
            "max_abs_err": float(e.max()), "median_abs_err": float(np.median(e)),
            "max_abs_err_vs_fp32": float(e_vs32.max()),
            "events_equal_fp32_after_rescore": same,
            "rescored_fraction": stp["rescored_positions"] / max(1, stp["corr_positions"] * T)}
        log("%s screen: max |L - L64| = %.4g, events equal after re-score: %s" % (prec, e.max(), same))
    out["spikes_fp32"] = int(ev32[0].size)
    return out


def concurrent_segments_run(args, V, sm, bias, thr, N, M, do_overlap):
    """The recording cut into args.segments equal time segments, each its own
    fbp_segment (own HIP stream) driven from its own host thread, all at once.
    Segment seams are not merged: this measures throughput of concurrent
    segments of the same total size."""
    import concurrent.futures
    from spikesorting_fullpursuit_amd import _lib
    C, K = V.shape
    T = sm['templates'].shape[0]
    nseg = args.segments
    bounds = np.linspace(0, K, nseg + 1).astype(np.int64)
    segs = []
    for a, b in zip(bounds[:-1], bounds[1:]):
        sg = _lib.Segment(C, int(b - a), T, N, M, do_overlap, 256, args.precision, args.overlap_mode)
        sg.set_templates(sm['templates'], bias, thr)
        sg.set_voltage(np.ascontiguousarray(V[:, a:b]))
        segs.append(sg)
    with concurrent.futures.ThreadPoolExecutor(max_workers=nseg) as ex:
        def run_all():
            list(ex.map(lambda sg: sg.run(), segs))
        for _ in range(args.warmup):
            run_all()
        t = time.perf_counter()
        for _ in range(args.steps):
            run_all()
        dt = (time.perf_counter() - t) / args.steps
    spikes = sum(sg.num_spikes() for sg in segs)
    passes = max(sg.stats()['n_passes'] for sg in segs)
    for sg in segs:
        sg.close()
    log("concurrent segments: %d x %d samples in %.3f s per step, %d spikes" % (nseg, K // nseg, dt, spikes))
    return {"segments": nseg, "samples_per_segment": int(bounds[1] - bounds[0]), "ms_per_step": 1000.0 * dt,
            "value": C * K / dt, "unit": "ch·samples/s", "spikes": int(spikes), "max_passes": int(passes),
            "note": "equal time segments run concurrently, one HIP stream each; seams not merged"}


def _serve(reqs, specs, args, paced, inflight):
    """Run the requests through a PursuitServer: paced = submitted at their
    Poisson arrival times, otherwise all at once (arrivals faster than the
    server can serve them).  Returns (makespan, latency summary, results)."""
    from spikesorting_fullpursuit_amd import serving
    with serving.PursuitServer(max_inflight=inflight, precision=args.precision,
                               overlap_mode=args.overlap_mode) as srv:
        futs = []
        t_start = time.perf_counter()
        for sp, r in zip(specs, reqs):
            if paced:
                delay = sp['arrival_s'] - (time.perf_counter() - t_start)
                if delay > 0:
                    time.sleep(delay)
            futs.append(srv.submit(r))
        done = [f.result() for f in futs]
        makespan = time.perf_counter() - t_start
        summ = srv.latency_summary()
    return makespan, summ, done


def serving_mix_run(args):
    """The served mix: Poisson arrivals of segments of 5-120 s at 30/40 kHz,
    4-32 channels, 3-40 templates of 41-81 samples; admitted by HBM budget and
    driven concurrently by the native segment server.  Latency = submission
    to result on the host.  Three legs: every request alone, one after the
    other (the reference's serial order); the mix at its arrival rate; and
    the mix submitted at once, so the server is saturated and its throughput
    is its capacity."""
    from spikesorting_fullpursuit_amd import serving, synthetic
    specs = synthetic.serving_mix(seed=args.seed + 100, n_requests=args.serve_requests, rate_hz=args.serve_rate)
    t0 = time.time()
    reqs = []
    for sp in specs:
        rec = synthetic.recording_for(sp)
        sm = rec['separability_metrics']
        reqs.append(serving.SegmentRequest(rec['voltage'], sm['templates'], sm['template_SS'],
                                           sm['neuron_lower_thresholds'], sp['n_samples_per_chan'], tag=len(reqs)))
    chs = sum(r.C * r.K for r in reqs)
    log("served mix: %d requests, %.2f G ch*samples, generated in %.1f s" % (len(reqs), chs / 1e9, time.time() - t0))
    # warm the kernels of every shape class once, outside the timed legs
    serving.run_segment(reqs[0], args.precision, args.overlap_mode)
    t = time.perf_counter()
    alone = [serving.run_segment(r, args.precision, args.overlap_mode)[0].size for r in reqs]
    standalone_s = time.perf_counter() - t
    log("served mix alone, one after the other: %.2f s" % standalone_s)
    makespan, summ, done = _serve(reqs, specs, args, True, args.serve_inflight)
    spikes = sum(int(d.onsets.size) for d in done)
    assert [d.onsets.size for d in done] == alone, "served spike counts differ from the standalone runs"
    log("served mix at %.0f/s: makespan %.2f s, p50 %.3f s, p99 %.3f s, peak %d in flight"
        % (args.serve_rate, makespan, summ['p50_s'], summ['p99_s'], summ['peak_inflight']))
    saturated = []
    for inflight in sorted({4, args.serve_inflight, 16}):
        ms, su, dn = _serve(reqs, specs, args, False, inflight)
        assert [d.onsets.size for d in dn] == alone, "saturated spike counts differ from the standalone runs"
        saturated.append({"max_inflight": inflight, "makespan_s": ms, "throughput": chs / ms,
                          "vs_standalone": standalone_s / ms, "latency_p50_s": su['p50_s'],
                          "latency_p99_s": su['p99_s'], "peak_inflight": su['peak_inflight'],
                          "peak_hbm_bytes": su['peak_bytes'], "pass_steps": su['pass_steps'],
                          "driver_busy_fraction": su['driver_busy_s'] / max(ms, 1e-9)})
        log("served mix saturated, %d in flight: makespan %.2f s (%.2fx the serial runs), p50 %.3f s, p99 %.3f s, "
            "driver thread busy %.1f %%" % (inflight, ms, standalone_s / ms, su['p50_s'], su['p99_s'],
                                            100 * su['driver_busy_s'] / ms))
    return {"requests": len(reqs), "arrival_rate_hz": args.serve_rate, "max_inflight": args.serve_inflight,
            "ch_samples": chs, "spikes": spikes, "makespan_s": makespan, "throughput": chs / makespan,
            "unit": "ch·samples/s", "latency_p50_s": summ['p50_s'], "latency_p99_s": summ['p99_s'],
            "latency_mean_s": summ['mean_s'], "queue_p99_s": summ['queue_p99_s'],
            "peak_inflight": summ['peak_inflight'], "peak_hbm_bytes": summ['peak_bytes'],
            "arrivals_span_s": specs[-1]['arrival_s'], "standalone_serial_s": standalone_s,
            "standalone_throughput": chs / standalone_s, "saturated": saturated,
            "events_equal_standalone": True}


def profile_name(workload, duration_s, precision, overlap_mode, M):
    """Base name of the committed rocprofv3 summary of one pursuit
    configuration under profiles/ (kernel table, calibrated HBM traffic,
    matrix-core counters)."""
    if duration_s is not None:
        workload = "%s-%ds" % (workload.split("-")[0], int(round(duration_s)))
    return "kernels_%s_%s-%s_M%d" % (workload, precision, overlap_mode, M)


def load_profile(workload, duration_s, precision, overlap_mode, M):
    try:
        return json.load(open(os.path.join(ROOT, "profiles",
                                           profile_name(workload, duration_s, precision, overlap_mode, M) + ".json")))
    except (OSError, ValueError):
        return None


# kernels of each stage, by name prefix (template arguments vary with the shapes)
STAGE_KERNELS = {"correlation": ("k_corr_mfma", "k_rescore", "k_corr_fp32"),
                 "overlap": ("k_overlap_q", "k_overlap_screened", "k_overlap_exact", "k_overlap_direct",
                             "k_sort_ov_by_label", "k_check_overlap")}


def measured_traffic(prof, stage, passes):
    """HBM bytes per pass of a stage (FETCH_SIZE + WRITE_SIZE of its kernels,
    calibrated against a copy of known size) from a committed rocprofv3
    summary, or None.  The summary covers one warm-up and one timed run."""
    if not prof:
        return None
    total, seen = 0.0, False
    for k, t in prof.get("traffic", {}).items():
        if not k.startswith(STAGE_KERNELS[stage]):
            continue
        if t["fetch_bytes_per_dispatch"] is None or t["write_bytes_per_dispatch"] is None:
            return None
        total += (t["fetch_bytes_per_dispatch"] + t["write_bytes_per_dispatch"]) * t["dispatches"]
        seen = True
    runs = prof.get("runs", 2)
    return total / runs / max(1, passes) if seen else None


def measured_mfma(prof, stage):
    """Matrix-core utilisation (percent of SIMD cycles) of a stage's matrix-core
    kernels from a committed summary: {kernel: percent}."""
    if not prof:
        return None
    out = {k: v["mfma_util_pct"] for k, v in prof.get("mfma", {}).items()
           if k.startswith(STAGE_KERNELS[stage]) and v.get("mfma_util_pct") is not None}
    return out or None


def support_channels(templates, C, N):
    """Channels per template between its first and last nonzero one, summed
    over the templates (the channels the kernels compute on)."""
    nz = np.abs(np.asarray(templates, dtype=np.float32).reshape(-1, C, N)).max(axis=2) > 0
    total = 0
    for row in nz:
        idx = np.flatnonzero(row)
        total += int(idx[-1] - idx[0] + 1) if idx.size else 0
    return total


def roofline_of(args, sums, C, T, N, M, K, passes_per_step, sup_total=None):
    """Algorithmic work of the two stages (HIP-event time summed over the timed
    steps) and the roofline entry of the slower one.  The correlation's work
    counts the (position, template) likelihoods it computed (listed blocks
    whose voltage changed since their last pass), over each template's
    channel support; the Q tables' over the supports too."""
    S = 2 * M + 1
    R = N + 2 * M
    sup_total = C * T if sup_total is None else sup_total
    sup_mean = sup_total / max(1, T)
    pairs = sums["corr_pairs"]
    stages = {}
    t_corr = sums["ms_correlation"] * 1e-3
    if args.precision == "fp32":
        flop = 2.0 * C * N * pairs
        stages["correlation"] = {"kernel": "k_corr_fp32", "ms": sums["ms_correlation"],
                                 "launches": sums["corr_launches"], "bound": "mfma", "unit": "TFLOP/s",
                                 "achieved": flop / t_corr / 1e12 if t_corr else 0.0, "peak": PEAK_TFLOPS["fp32"],
                                 "work": "2*C*N flop per computed (position, template) likelihood (float32 "
                                         "FMA chain over every channel)"}
    else:
        nbytes = (4.0 * C / max(1, T) + 4.0) * pairs   # the position's voltage once over its T templates
        flop = 2.0 * sup_mean * N * pairs
        # float32 re-score with the templates resident in LDS when they fit (150 KB)
        tp = -(-T // 8) * 8
        rescore = "k_rescore_res" if (C * N * tp + 16 * 512) * 4 <= 150 * 1024 else "k_rescore_sup"
        stages["correlation"] = {"kernel": "k_corr_mfma + " + rescore, "ms": sums["ms_correlation"],
                                 "launches": sums["corr_launches"], "bound": "hbm", "unit": "GB/s",
                                 "achieved": nbytes / t_corr / 1e9 if t_corr else 0.0, "peak": PEAK_HBM_GBS,
                                 "mfma_tflops": flop / t_corr / 1e12 if t_corr else 0.0,
                                 "mfma_peak": PEAK_TFLOPS[args.precision],
                                 "work": "per computed (position, template) likelihood: 4 bytes written "
                                         "and 4*C/T bytes of voltage read (a position's voltage once over "
                                         "its T templates); 2*N matrix-core flop per channel of the "
                                         "template's support"}
    t_ov = sums["ms_overlap"] * 1e-3
    if args.overlap_mode == "direct":
        flop = (5.0 * C * N + C) * sums["overlap_combos"]
        stages["overlap"] = {"kernel": "k_overlap_direct", "ms": sums["ms_overlap"],
                             "launches": sums["overlap_launches"], "bound": "mfma", "unit": "TFLOP/s",
                             "achieved": flop / t_ov / 1e12 if t_ov else 0.0, "peak": PEAK_TFLOPS["fp32"],
                             "work": "5*C*N + C flop per scored (t,f,s) combination"}
    else:
        # Per overlap window: the window voltage (read by the Q-table kernel and
        # the screen), the main template, the likelihood signs at every (t, s),
        # the Q table written and read back, and one 4-byte AB entry per
        # unskipped combination.  The Q table itself is 3 bf16 products per
        # multiply-add of the truncated correlations.
        win = sums.get("overlap_windows", 0)
        nbytes = 4.0 * sums["screened_combos"] + win * (2.0 * C * N * 4 + C * R * 4 + 3.0 * T * S * 4)
        valid = sum(min(N, N + s - M) - max(0, s - M) for s in range(S))
        qflop = 3.0 * 2.0 * sup_total * valid * win
        stages["overlap"] = {"kernel": "k_overlap_q + k_overlap_screened + k_overlap_exact", "ms": sums["ms_overlap"],
                             "launches": sums["overlap_launches"], "bound": "hbm", "unit": "GB/s",
                             "achieved": nbytes / t_ov / 1e9 if t_ov else 0.0, "peak": PEAK_HBM_GBS,
                             "mfma_tflops": qflop / t_ov / 1e12 if t_ov else 0.0, "mfma_peak": PEAK_TFLOPS["bf16"],
                             "work": "4 B of the AB table per unskipped combination + per window voltage, "
                                     "main template, likelihood signs and the Q table (bytes); split-bf16 "
                                     "Q table (matrix-core flop)"}
    prof = load_profile(args.workload, getattr(args, "duration", None), args.precision, args.overlap_mode, M)
    for name, stg in stages.items():
        util = measured_mfma(prof, name)
        if util:
            stg["mfma_util_pct"] = util
        stg["traffic_per_pass"] = measured_traffic(prof, name, passes_per_step)
    dom = max(stages, key=lambda k: stages[k]["ms"])
    st = stages[dom]
    traffic = st["traffic_per_pass"]
    per_launch = st["achieved"] * (st["ms"] * 1e-3 / max(1, st["launches"])) * (1e9 if st["unit"] == "GB/s" else 1e12)
    roofline = {"bound": st["bound"], "kernel": st["kernel"], "achieved": st["achieved"], "peak": st["peak"],
                "unit": st["unit"], "frac": st["achieved"] / st["peak"],
                "algorithmic_per_launch": per_launch, "ms_per_launch": st["ms"] / max(1, st["launches"]),
                "traffic": traffic, "traffic_unit": "bytes per pass (rocprofv3 FETCH_SIZE + WRITE_SIZE)",
                "profile": ("profiles/" + profile_name(args.workload, getattr(args, "duration", None), args.precision,
                                                       args.overlap_mode, M) + ".md") if prof else None,
                "work": st["work"]}
    return stages, roofline


def cpu_baseline(rec, bias, thr, N, M, do_overlap, target_s, seg_cls, precision, overlap_mode):
    """Oracle (C restatement of the reference) on a leading slice of the
    recording, sized to about target_s seconds of host time; the GPU runs the
    same slice and the spike lists are compared."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    V = rec['voltage']
    tm = rec['separability_metrics']['templates']
    C, K = V.shape
    fs = rec['sort_info']['sampling_rate']
    # first slice: 0.5 s for the 16-channel, 20-template workload, shorter in
    # proportion for larger C*T (the oracle's overlap recheck is dense)
    n = min(K, max(20 * N, int(fs * 0.5 * min(1.0, 320.0 / (C * tm.shape[0])))))
    while True:
        t = time.perf_counter()
        o = oracle.pursuit_chunk(V[:, :n], tm, bias, thr, N, M, do_overlap=do_overlap)
        dt = time.perf_counter() - t
        log("cpu oracle: %d samples in %.2f s" % (n, dt))
        if dt >= target_s * 0.4 or n >= K:
            break
        n = min(K, int(n * min(8.0, max(1.5, target_s / max(dt, 1e-3)))))
    with seg_cls(C, n, tm.shape[0], N, M, do_overlap, 256, precision, overlap_mode) as seg:
        seg.set_templates(tm, bias, thr)
        seg.set_voltage(V[:, :n])
        seg.run()
        gi, gl = seg.spikes()
    parity = bool(np.array_equal(gi, o['idx']) and np.array_equal(gl, o['lab']))
    return {"value": C * n / dt, "unit": "ch·samples/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": "first %.2f s (%d samples x %d ch) of the same recording; %d spikes; GPU spike list "
                      "identical on this slice: %s" % (n / fs, n, C, o['idx'].size, parity),
            "seconds": dt, "gpu_parity_on_sample": parity}


if __name__ == "__main__":
    sys.exit(main())
