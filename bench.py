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
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--overlap-mode", default="direct")
    ap.add_argument("--no-overlap", action="store_true")
    ap.add_argument("--cpu-target-s", type=float, default=12.0, help="CPU baseline budget (s)")
    ap.add_argument("--no-cpu", action="store_true")
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
            "overlap_launches": 0, "corr_positions": 0, "overlap_combos": 0}
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

    # Roofline of the dominant kernel (algorithmic work / HIP-event time).
    corr_flop = 2.0 * T * C * N * sums["corr_positions"]
    ovl_flop = (5.0 * C * N + C) * sums["overlap_combos"]
    stages = {
        "correlation": {
            "ms": sums["ms_correlation"], "launches": sums["corr_launches"], "flop": corr_flop,
            "tflops": corr_flop / (sums["ms_correlation"] * 1e-3) / 1e12 if sums["ms_correlation"] > 0 else 0.0,
            "flop_per_unit": "2*T*C*N per likelihood position",
        },
        "overlap": {
            "ms": sums["ms_overlap"], "launches": sums["overlap_launches"], "flop": ovl_flop,
            "tflops": ovl_flop / (sums["ms_overlap"] * 1e-3) / 1e12 if sums["ms_overlap"] > 0 else 0.0,
            "flop_per_unit": "5*C*N + C per scored (t,f,s) combination",
        },
    }
    dom = "overlap" if sums["ms_overlap"] > sums["ms_correlation"] else "correlation"
    ds = stages[dom]
    peak = PEAK_TFLOPS["fp32"] if args.precision == "fp32" else PEAK_TFLOPS[args.precision]
    roofline = {
        "bound": "mfma", "kernel": "k_overlap_direct" if dom == "overlap" else "k_corr_fp32",
        "achieved": ds["tflops"], "peak": peak, "unit": "TFLOP/s", "frac": ds["tflops"] / peak,
        "traffic": None,
        "note": "float32 VALU fused multiply-adds (v_fmac_f32); MI355X float32 peak 157.3 TFLOP/s",
    }

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(rec, bias, thr, N, M, do_overlap, args.cpu_target_s, seg_cls=_lib.Segment,
                           precision=args.precision, overlap_mode=args.overlap_mode)
    seg.close()

    out = {
        "metric": METRIC, "value": value, "unit": "ch·samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded N(0,1) noise + DoG templates)",
        "config": {"workload": args.workload, "n_chans": C, "sampling_rate": fs, "duration_s": K / fs,
                   "n_templates": T, "template_len": N, "n_max_shift_inds": M, "do_overlap_recheck": do_overlap,
                   "precision": args.precision, "overlap_mode": args.overlap_mode, "seed": args.seed,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "result": {"n_spikes": int(n_spikes), "n_passes": int(last["n_passes"]),
                   "windows_evaluated": int(last["windows_evaluated"]),
                   "overlap_windows": int(last["overlap_windows"]), "device_ms_per_step": sums["ms_total"] / args.steps},
        "stages": stages,
    }
    print(json.dumps(out), flush=True)
    return 0


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
    n = min(K, int(fs * 0.5))
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
