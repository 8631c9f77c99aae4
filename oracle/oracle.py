"""CPU oracle for the Full Binary Pursuit hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker.  The product package never imports it.

`pursuit_chunk` drives the C restatement in fbp_oracle.c (one chunk, the pass
loop of spikesorting_fullpursuit/parallel/binary_pursuit_parallel.py:524-736
and the kernels of spikesorting_fullpursuit/kernels/binary_pursuit.cl).
`binary_pursuit` restates the host driver around it (chunk onsets
binary_pursuit_parallel.py:359-368, per-chunk slicing :371-377, output
assembly :814-845) for a caller-fixed chunk length.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfbp_oracle.so")
_lib = None


class Params(ctypes.Structure):
    _fields_ = [
        ("n_chans", ctypes.c_uint32),
        ("n_samples", ctypes.c_uint32),
        ("n_templates", ctypes.c_uint32),
        ("template_len", ctypes.c_uint32),
        ("max_shift", ctypes.c_uint32),
        ("do_overlap", ctypes.c_uint32),
        ("reduce_wg", ctypes.c_uint32),
        ("max_passes", ctypes.c_uint32),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        f32p, u32p, u64p, f64p = P(ctypes.c_float), P(ctypes.c_uint32), P(ctypes.c_uint64), P(ctypes.c_double)
        lib.fbo_pursuit_chunk.argtypes = [P(Params), f32p, f32p, f32p, f32p, u32p, u32p, ctypes.c_uint64,
                                          u64p, u32p, u32p, ctypes.c_uint32, f32p, u32p, u64p]
        lib.fbo_pursuit_chunk.restype = ctypes.c_int
        lib.fbo_likelihood.argtypes = [P(Params), f32p, f32p, f32p, u32p, ctypes.c_uint64, f32p]
        lib.fbo_likelihood.restype = ctypes.c_int
        lib.fbo_likelihood_f64.argtypes = [P(Params), f32p, f32p, f32p, u32p, ctypes.c_uint64, f64p]
        lib.fbo_likelihood_f64.restype = ctypes.c_int
        lib.fbo_residual.argtypes = [P(Params), f32p, f32p, u32p, u32p, ctypes.c_uint64]
        lib.fbo_residual.restype = ctypes.c_int
        lib.fbo_adjusted_clips.argtypes = [P(Params), f32p, f32p, u32p, u32p, ctypes.c_uint64, f32p]
        lib.fbo_adjusted_clips.restype = ctypes.c_int
        lib.fbo_num_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def num_threads():
    return int(load().fbo_num_threads())


def make_params(C, K, T, N, M=0, do_overlap=True, reduce_wg=256, max_passes=0):
    return Params(C, K, T, N, M, 1 if do_overlap else 0, reduce_wg, max_passes)


def likelihood(V, templates, bias, N, windows=None):
    """Likelihood buffer [T][K] for the listed windows (all windows by default)."""
    V = np.ascontiguousarray(V, dtype=np.float32)
    C, K = V.shape
    tm = np.ascontiguousarray(templates, dtype=np.float32).reshape(-1)
    T = tm.size // (C * N)
    bias = np.ascontiguousarray(bias, dtype=np.float32)
    if windows is None:
        windows = np.arange((K + N - 1) // N, dtype=np.uint32)
    windows = np.ascontiguousarray(windows, dtype=np.uint32)
    L = np.zeros((T, K), dtype=np.float32)
    p = make_params(C, K, T, N)
    rc = load().fbo_likelihood(ctypes.byref(p), _ptr(V, ctypes.c_float), _ptr(tm, ctypes.c_float),
                               _ptr(bias, ctypes.c_float), _ptr(windows, ctypes.c_uint32),
                               windows.size, _ptr(L, ctypes.c_float))
    assert rc == 0, rc
    return L


def likelihood_f64(V, templates, template_ss, N, windows=None):
    """Same sums as `likelihood`, accumulated in float64 from -0.5*template_SS."""
    V = np.ascontiguousarray(V, dtype=np.float32)
    C, K = V.shape
    tm = np.ascontiguousarray(templates, dtype=np.float32).reshape(-1)
    T = tm.size // (C * N)
    ss = np.ascontiguousarray(template_ss, dtype=np.float32)
    if windows is None:
        windows = np.arange((K + N - 1) // N, dtype=np.uint32)
    windows = np.ascontiguousarray(windows, dtype=np.uint32)
    L = np.zeros((T, K), dtype=np.float64)
    p = make_params(C, K, T, N)
    rc = load().fbo_likelihood_f64(ctypes.byref(p), _ptr(V, ctypes.c_float), _ptr(tm, ctypes.c_float),
                                   _ptr(ss, ctypes.c_float), _ptr(windows, ctypes.c_uint32),
                                   windows.size, _ptr(L, ctypes.c_double))
    assert rc == 0, rc
    return L


def adjusted_clips(residual, templates, N, idx, lab):
    residual = np.ascontiguousarray(residual, dtype=np.float32)
    C, K = residual.shape
    tm = np.ascontiguousarray(templates, dtype=np.float32).reshape(-1)
    T = tm.size // (C * N)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    lab = np.ascontiguousarray(lab, dtype=np.uint32)
    clips = np.zeros((idx.size, C * N), dtype=np.float32)
    p = make_params(C, K, T, N)
    rc = load().fbo_adjusted_clips(ctypes.byref(p), _ptr(residual, ctypes.c_float), _ptr(tm, ctypes.c_float),
                                   _ptr(idx, ctypes.c_uint32), _ptr(lab, ctypes.c_uint32), idx.size,
                                   _ptr(clips, ctypes.c_float))
    assert rc == 0, rc
    return clips


def pursuit_chunk(V, templates, bias, thr, N, M, do_overlap=True, reduce_wg=256, max_passes=0,
                  want_L=False):
    """Run the pass loop on one chunk.

    Returns dict(idx, lab (uint32, record order), residual [C][K], n_passes,
    pass_counts, L (final likelihood buffer, if want_L), next_windows (the
    window list of the next pass when max_passes stopped the loop, else
    empty)).
    """
    V = np.array(V, dtype=np.float32, order="C", copy=True)
    C, K = V.shape
    tm = np.ascontiguousarray(templates, dtype=np.float32).reshape(-1)
    T = tm.size // (C * N)
    bias = np.ascontiguousarray(bias, dtype=np.float32)
    thr = np.ascontiguousarray(thr, dtype=np.float32)
    W = (K + N - 1) // N
    cap = max(16, 4 * W)
    pc_cap = 4096
    while True:
        out_idx = np.zeros(cap, dtype=np.uint32)
        out_lab = np.zeros(cap, dtype=np.uint32)
        n_out = ctypes.c_uint64(0)
        n_passes = ctypes.c_uint32(0)
        pass_counts = np.zeros(pc_cap, dtype=np.uint32)
        L = np.zeros((T, K), dtype=np.float32) if want_L else None
        nxt = np.zeros(W, dtype=np.uint32)
        n_next = ctypes.c_uint64(0)
        Vw = V.copy()
        p = make_params(C, K, T, N, M, do_overlap, reduce_wg, max_passes)
        rc = load().fbo_pursuit_chunk(
            ctypes.byref(p), _ptr(Vw, ctypes.c_float), _ptr(tm, ctypes.c_float), _ptr(bias, ctypes.c_float),
            _ptr(thr, ctypes.c_float), _ptr(out_idx, ctypes.c_uint32), _ptr(out_lab, ctypes.c_uint32), cap,
            ctypes.byref(n_out), ctypes.byref(n_passes), _ptr(pass_counts, ctypes.c_uint32), pc_cap,
            _ptr(L, ctypes.c_float) if want_L else None, _ptr(nxt, ctypes.c_uint32), ctypes.byref(n_next))
        if rc == -3:
            cap *= 4
            continue
        if rc not in (0, -4):
            raise RuntimeError("oracle failed with code %d" % rc)
        break
    n = int(n_out.value)
    npass = int(n_passes.value)
    return dict(idx=out_idx[:n].copy(), lab=out_lab[:n].copy(), residual=Vw, n_passes=npass,
                pass_counts=pass_counts[:min(npass, pc_cap)].copy(), L=L, capped=(rc == -4),
                next_windows=nxt[:int(n_next.value)].copy())


def time_window_to_samples(time_window, sampling_rate):
    """parallel/segment_parallel.py:82-98 (restated)."""
    tw = list(time_window)
    if tw[0] > 0:
        tw[0] *= -1
    first = min(int(round(tw[0] * sampling_rate)), 0)
    second = max(int(round(tw[1] * sampling_rate)), 1) + 1
    return [first, second], tw


def binary_pursuit(voltage, v_dtype, sort_info, separability_metrics, n_max_shift_inds,
                   chunk_len, reduce_wg=256):
    """Reference host-driver semantics around `pursuit_chunk` for a fixed chunk
    length (the reference derives it from device memory, bpp.py:276-344).
    Returns (event_indices, neuron_labels, bp_bool, adjusted_clips, residual)."""
    n_samples = voltage.shape[1]
    chan_win, _ = time_window_to_samples(sort_info['clip_width'], sort_info['sampling_rate'])
    N = int(sort_info['n_samples_per_chan'])
    templates = separability_metrics['templates']
    if templates.ndim == 1:
        templates = templates.reshape(1, -1)
    templates = np.float32(templates)
    clip_init = int(abs(chan_win[0]))
    bias = np.float32(-0.5 * separability_metrics['template_SS'])
    thr = np.float32(separability_metrics['neuron_lower_thresholds'])
    M = N // 2 if (n_max_shift_inds is None or n_max_shift_inds < 0) else int(n_max_shift_inds)
    onsets = []
    cur = 0
    while cur < n_samples - N:
        onsets.append(cur)
        cur += chunk_len - 3 * N
    idx_all, lab_all, clips_all, bool_all = [], [], [], []
    residual = np.array(voltage, dtype=np.float32, copy=True)
    for k, start in enumerate(onsets):
        stop = min(n_samples, start + chunk_len)
        V = np.float32(voltage[:, start:stop])
        r = pursuit_chunk(V, templates, bias, thr, N, M, sort_info['do_overlap_recheck'], reduce_wg)
        # chunk k supplies [onset_k + N, onset_{k+1} + N): the samples whose
        # covering spikes all start inside chunk k
        lo = 0 if k == 0 else min(start + N, n_samples)
        hi = min(onsets[k + 1] + N, n_samples) if k + 1 < len(onsets) else n_samples
        residual[:, lo:hi] = r['residual'][:, lo - start:hi - start]
        if r['idx'].size > 0:
            if sort_info['get_adjusted_clips']:
                clips = adjusted_clips(r['residual'], templates, N, r['idx'], r['lab'])
            else:
                clips = []
            idx_all.append(r['idx'].astype(np.int64) + start)
            lab_all.append(r['lab'])
            clips_all.append(clips)
            bool_all.append(np.ones(r['idx'].size, dtype=bool))
    if idx_all:
        ev = np.int64(np.hstack(idx_all)) + clip_init
        lab = np.int64(np.hstack(lab_all))
        clips = np.vstack(clips_all).astype(v_dtype)
        bpb = np.hstack(bool_all)
    else:
        ev, lab, bpb, clips = [], [], [], []
    return ev, lab, bpb, clips, residual


def get_clips(probe_dict, voltage, neighbors, event_indices, clip_width, check_valid=True):
    """parallel/segment_parallel.py:451-509 (restated, without memmap): the
    front/back trimming scans and the per-channel slices, on the host."""
    window, _ = time_window_to_samples(clip_width, probe_dict['sampling_rate'])
    W = window[1] - window[0]
    if len(event_indices) == 0:
        return np.zeros((0, W * len(neighbors)), dtype=probe_dict['v_dtype']), np.ones(0, dtype=bool)
    valid = np.ones(event_indices.shape[0], dtype=bool)
    start = 0
    stop = event_indices.shape[0] - 1
    if check_valid:
        while event_indices[start] + window[0] < 0:
            valid[start] = False
            start += 1
            if start == event_indices.size:
                valid[:] = False
                return None, valid
        while event_indices[stop] + window[1] >= probe_dict['n_samples']:
            valid[stop] = False
            stop -= 1
            if stop < 0:
                valid[:] = False
                return None, valid
    out = np.empty((np.count_nonzero(valid), W * len(neighbors)), dtype=probe_dict['v_dtype'])
    for row, k in enumerate(range(start, stop + 1)):
        for j, chan in enumerate(neighbors):
            out[row, j * W:(j + 1) * W] = voltage[chan, event_indices[k] + window[0]:event_indices[k] + window[1]]
    return out, valid


def unit_mean_clips(clips, labels, n_units):
    """full_binary_pursuit.py:368: np.mean(clips[labels == unit], axis=0) of
    every unit (NaN rows for units without clips)."""
    out = np.full((n_units, clips.shape[1]), np.nan, dtype=np.float32)
    for u in range(n_units):
        sel = labels == u
        if np.any(sel):
            out[u] = np.mean(clips[sel, :], axis=0)
    return out


def unit_channels(clips, labels, n_channels, n_samples_per_chan, thresholds, sigma):
    """full_binary_pursuit.py:365-383 (restated): {channel: [units]}, each
    unit on the first channel with the best range / (3 * noise) of its mean
    clip."""
    out = {c: [] for c in range(n_channels)}
    for unit in np.unique(labels):
        tmpl = np.mean(clips[labels == unit, :], axis=0)
        best, best_c = -1.0, None
        for c in range(n_channels):
            seg = tmpl[c * n_samples_per_chan:(c + 1) * n_samples_per_chan]
            snr = (np.amax(seg) - np.amin(seg)) / (3 * (thresholds[c] / sigma))
            if snr > best:
                best, best_c = snr, c
        out[best_c].append(unit)
    return out


def _triple_scan(test, shifts1, shifts2, acc, best_val, best, n1, n2):
    """The shift-pair loops of sort_cython.pyx:428-450 for one pair."""
    for i1 in range(shifts1.shape[0]):
        for i2 in range(shifts2.shape[0]):
            acc[:] = shifts1[i1, :]
            acc += shifts2[i2, :]
            r = np.sum((test - acc) ** 2)
            if r < best_val:
                best_val, best = r, ([n1, n2], np.copy(acc))
    return best_val, best


def find_overlap_templates(templates, n_samples_per_chan, n_chans, n_pre_inds, n_post_inds, n_template_spikes):
    """c_cython/sort_cython.pyx:333-456 (restated with numpy, loops as in the
    reference).  Returns (templates_to_check, smallest residual of each)."""
    N = n_samples_per_chan

    def shifted(t, shift):
        out = np.zeros_like(t)
        for c in range(n_chans):
            ct = t[c * N:(c + 1) * N]
            if shift > 0:
                out[c * N + shift:(c + 1) * N] = ct[:-shift]
            else:
                out[c * N:(c + 1) * N + shift] = ct[-shift:]
        return out

    to_check, residuals = [], []
    T = templates.shape[0]
    if T < 3:
        return to_check, residuals
    ss = np.sum(templates ** 2, axis=1)
    acc = np.zeros(templates.shape[1])
    all_shifts = []
    for t in range(T):
        a = np.zeros((n_post_inds + 1 + n_pre_inds, templates.shape[1]))
        for i, s in enumerate(range(-n_pre_inds, n_post_inds + 1)):
            a[i, :] = shifted(templates[t, :], s)
        all_shifts.append(a)
    for test_unit in np.argsort(ss)[-1::-1]:
        test = templates[test_unit, :]
        min_c, max_c = np.argmin(test) // N, np.argmax(test) // N
        best_val, best = np.inf, None
        for n1 in range(T):
            if n_template_spikes[n1] < 4 * n_template_spikes[test_unit] or n1 == test_unit:
                continue
            for n2 in range(n1 + 1, T):
                if n_template_spikes[n2] < 4 * n_template_spikes[test_unit] or n2 == test_unit:
                    continue
                lo, hi = slice(min_c * N, (min_c + 1) * N), slice(max_c * N, (max_c + 1) * N)
                if np.amin(templates[n1, lo]) + np.amin(templates[n2, lo]) > 0.75 * np.amin(test[lo]):
                    continue
                if np.amax(templates[n1, hi]) + np.amax(templates[n2, hi]) < 0.75 * np.amax(test[hi]):
                    continue
                best_val, best = _triple_scan(test, all_shifts[n1], all_shifts[n2], acc, best_val, best, n1, n2)
        if best is not None:
            to_check.append([test_unit, best[1], best[0]])
            residuals.append(best_val)
    return to_check, residuals


def remove_binary_pursuit_duplicates(event_indices, clips, unit_template, binary_pursuit_bool, tol_inds):
    """analyze_spike_timing.py:116-156 (restated)."""
    keep = np.ones(event_indices.size, dtype=bool)
    sse = np.zeros(2)
    cur, nxt = 0, 1
    while nxt < event_indices.size:
        if event_indices[nxt] - event_indices[cur] <= tol_inds:
            if binary_pursuit_bool[cur] and binary_pursuit_bool[nxt]:
                sse[0] = np.sum((clips[cur, :] - unit_template) ** 2)
                sse[1] = np.sum((clips[nxt, :] - unit_template) ** 2)
                if sse[0] <= sse[1]:
                    keep[nxt] = False
                    nxt += 1
                else:
                    keep[cur] = False
                    cur = nxt
                    nxt += 1
            elif binary_pursuit_bool[cur] and ~binary_pursuit_bool[nxt]:
                nxt += 1
            else:
                cur = nxt
                nxt += 1
        else:
            cur = nxt
            nxt += 1
    return keep


def remove_segment_binary_pursuit_duplicates(sort_data, n_chans, n_segments, half_clip_inds):
    """postprocessing.py:521-546 (restated) on sort_data[chan][seg]."""
    for chan in range(n_chans):
        for seg in range(n_segments):
            d = sort_data[chan][seg]
            if len(d[0]) == 0:
                continue
            keep = np.ones(d[0].shape[0], dtype=bool)
            for lab in np.unique(d[1]):
                sel = d[1] == lab
                keep[sel] = remove_binary_pursuit_duplicates(d[0][sel], d[2][sel, :], np.mean(d[2][sel, :], axis=0),
                                                             d[3][sel], 2 * half_clip_inds)
            d[0], d[1], d[2], d[3] = d[0][keep], d[1][keep], d[2][keep, :], d[3][keep]
    return sort_data


def channel_covariance(clips, n_chans, n_samples_per_chan, rows=None):
    """Per-channel noise covariance of clip rows, np.cov(..., rowvar=False,
    ddof=0) in float64 (full_binary_pursuit.py:245-252); rows [n_chans][n_rows]
    or None for every clip."""
    N = n_samples_per_chan
    out = []
    for c in range(n_chans):
        sel = clips if rows is None else clips[np.asarray(rows[c], dtype=np.int64)]
        out.append(np.cov(sel[:, c * N:(c + 1) * N], rowvar=False, ddof=0))
    return np.stack(out)


def separability_metrics(templates, channel_covariance_mats, sort_info):
    """neuron_separability.py:88-152 restated: per template the likelihood
    variance sum_c (t_c[None, :] @ cov_c) @ t_c[:, None] (:126-135), the lower
    threshold and confidence bound (:137-142), peak channel (:144-146) and
    contamination (:147-150); channel_p_noise (:114-121)."""
    from scipy.stats import norm
    C, N = sort_info['n_channels'], sort_info['n_samples_per_chan']
    tm = np.vstack(templates)
    T = tm.shape[0]
    ss = np.sum(tm ** 2, axis=1)
    var = np.zeros(T)
    thr, ci, peak, cont = np.zeros(T), np.zeros(T), np.zeros(T, dtype=np.int64), np.zeros(T)
    for n in range(T):
        for c in range(C):
            t = tm[n, c * N:(c + 1) * N].astype(np.float64)
            var[n] += ((t[None, :] @ channel_covariance_mats[c]) @ t[:, None])[0, 0]
        e = 0.5 * ss[n]
        ci[n] = e - sort_info['sigma_bp_CI'] * np.sqrt(var[n])
        thr[n] = -1 * e + sort_info['sigma_bp_noise'] * np.sqrt(var[n])
        peak[n] = np.argmax(np.abs(tm[n])) // N
        cont[n] = norm.sf(thr[n], -e, np.sqrt(var[n]))
    p_noise = np.array([(sort_info['n_samples'] - sort_info['n_threshold_crossings'][c]) / sort_info['n_samples']
                        for c in range(C)], dtype=np.float64)
    return {'template_SS': ss, 'neuron_variances': var, 'neuron_lower_thresholds': thr, 'neuron_lower_CI': ci,
            'peak_channel': peak, 'contamination': cont, 'channel_p_noise': p_noise}


def optimal_align_templates(temp_1, temp_2, n_chans, max_shift=None, align_abs=False, zero_pad=False):
    """consolidate.py:9-71 restated: the lag of the largest (absolute, with
    align_abs) value of np.correlate(temp_1, temp_2, 'full'), within
    +-max_shift of zero lag when given (first of equal maxima), then every
    channel of both templates shifted by that lag and truncated (:34-54), or
    zero padded back to full length (:56-69).  Returns (shifted 1, shifted 2,
    lag, samples per channel)."""
    L = temp_1.shape[0]
    N = L // n_chans
    if temp_1.shape[0] != temp_2.shape[0] or temp_1.ndim > 1 or N == 0:
        raise ValueError("Input templates must be 1D vectors of the same size")
    if L % n_chans != 0:
        raise ValueError("Template shape[0] must be evenly divisible by n_chans")
    xc = np.correlate(temp_1, temp_2, mode='full')
    if align_abs:
        xc = np.abs(xc)
    centre = xc.shape[0] // 2
    if max_shift is None:
        lag = int(np.argmax(xc)) - centre
    else:
        lag = int(np.argmax(xc[centre - max_shift:centre + max_shift + 1])) - max_shift
    if lag == 0:
        return temp_1, temp_2, 0, N
    a = [temp_1[c * N:(c + 1) * N] for c in range(n_chans)]
    b = [temp_2[c * N:(c + 1) * N] for c in range(n_chans)]
    if lag > 0:
        s1 = np.hstack([x[lag:] for x in a])
        s2 = np.hstack([x[:-lag] for x in b])
    else:
        s1 = np.hstack([x[:lag] for x in a])
        s2 = np.hstack([x[-lag:] for x in b])
    n_shift = s1.shape[0] // n_chans
    if not zero_pad:
        return s1, s2, lag, n_shift
    p1, p2 = np.zeros(L), np.zeros(L)
    for c in range(n_chans):
        dst = slice(c * N + lag, (c + 1) * N) if lag >= 0 else slice(c * N, (c + 1) * N + lag)
        p1[dst] = s1[c * n_shift:(c + 1) * n_shift]
        p2[dst] = s2[c * n_shift:(c + 1) * n_shift]
    return p1, p2, lag, N


def find_noisy_templates(separability_metrics):
    """neuron_separability.py:155-169: noisy where 2 * lower threshold
    exceeds half the template's sum of squares."""
    thr = separability_metrics['neuron_lower_thresholds']
    ss = separability_metrics['template_SS']
    return np.array([2 * thr[n] > 0.5 * ss[n] for n in range(ss.shape[0])], dtype=bool)


def check_noise_templates(separability_metrics, sort_info, noisy_templates):
    """neuron_separability.py:172-224 restated loop for loop: a noisy
    template stays (is no longer marked) as soon as one template that is not
    noisy, with nonzero likelihood variance, would take its spikes above
    threshold with probability above norm.sf(sigma_bp_noise)."""
    from scipy.stats import norm
    C, N = sort_info['n_channels'], sort_info['n_samples_per_chan']
    max_shift = N // 4 - 1
    tm = separability_metrics['templates']
    p_limit = norm.sf(sort_info['sigma_bp_noise'], 0, 1)
    out = np.copy(noisy_templates)
    for noise_n in range(tm.shape[0]):
        if not noisy_templates[noise_n]:
            continue
        for n in range(tm.shape[0]):
            if noisy_templates[n]:
                continue
            a, b, _, _ = optimal_align_templates(tm[n, :], tm[noise_n, :], C, max_shift=max_shift, align_abs=True)
            expect = np.dot(a, b) - 0.5 * separability_metrics['template_SS'][n]
            var = separability_metrics['neuron_variances'][n]
            if var == 0.0:
                continue
            if norm.sf(separability_metrics['neuron_lower_thresholds'][n], expect, np.sqrt(var)) > p_limit:
                out[noise_n] = False
                break
    return separability_metrics, out


def delete_noise_units(separability_metrics, noisy_templates):
    """neuron_separability.py:227-248: rows of the noisy templates removed
    from every per-template entry, the per-channel entries kept."""
    keep = ~noisy_templates
    for key in list(separability_metrics.keys()):
        if key in ('channel_covariance_mats', 'channel_p_noise', 'bp_n_samples_per_chan'):
            continue
        if key in ('templates', 'template_SS_by_chan'):
            separability_metrics[key] = separability_metrics[key][keep, :]
        elif key in ('template_SS', 'neuron_variances', 'neuron_lower_thresholds', 'neuron_lower_CI',
                     'contamination', 'peak_channel'):
            separability_metrics[key] = separability_metrics[key][keep]
        elif key == 'template_covariance_mats':
            for x in reversed(range(len(separability_metrics[key]))):
                if noisy_templates[x]:
                    del separability_metrics[key][x]
    return separability_metrics


def check_template_pair(template_1, template_2, chan_covariance_mats, sort_info):
    """neuron_separability.py:56-85: probability that the likelihood of
    template_2 exceeds that of template_1 when the voltage is template_1,
    norm.cdf(0, E1 - E2, sqrt(sum_c d_c^T cov_c d_c)), d = template_1 -
    template_2; 1 when that variance is not positive."""
    from scipy.stats import norm
    C, N = sort_info['n_channels'], sort_info['n_samples_per_chan']
    e1 = 0.5 * np.dot(template_1, template_1)
    e2 = np.dot(template_1, template_2) - 0.5 * np.dot(template_2, template_2)
    var = 0
    for c in range(C):
        d = template_1[c * N:(c + 1) * N] - template_2[c * N:(c + 1) * N]
        var += (d[None, :] @ chan_covariance_mats[c] @ d[:, None])[0, 0]
    if var > 0:
        return float(norm.cdf(0, e1 - e2, np.sqrt(var)))
    return 1.


def binary_pursuit_clip_width(events, voltage, sort_info):
    """full_binary_pursuit.py:14-115 restated for the crossings `events` of a
    segment: the mean clip over the widened clip width, trimmed at each end
    while every channel's end sample is below 5 % of the median absolute
    voltage (np.random.choice of 10 s of samples), never inside the original
    width; returns ([-w, w], original column starts, stops)."""
    C, N, fs = sort_info['n_channels'], sort_info['n_samples_per_chan'], sort_info['sampling_rate']
    factor = sort_info['max_binary_pursuit_clip_width_factor']
    plain = (np.arange(0, N * C, N, dtype=np.int64), np.arange(N, (N + 1) * C, N, dtype=np.int64))
    w0 = max(np.abs(sort_info['clip_width']))
    if factor <= 1.0 or len(events) == 0:
        return [-1 * w0, w0], plain[0], plain[1]
    events = np.sort(np.hstack(events))
    bp_cw = [factor * v for v in sort_info['clip_width']]
    pd = {'sampling_rate': fs, 'n_samples': voltage.shape[1], 'v_dtype': voltage.dtype}
    clips, valid = get_clips(pd, voltage, np.arange(C), events, bp_cw)
    if np.count_nonzero(valid) == 0:
        return [-1 * w0, w0], plain[0], plain[1]
    mean_clip = np.mean(clips, axis=0)
    bp_n = clips.shape[1] // C
    first = np.arange(0, bp_n * (C - 1) + 1, bp_n, dtype=np.int64)
    last = np.arange(bp_n - 1, bp_n * C + 1, bp_n, dtype=np.int64)
    inds = np.random.choice(voltage.shape[1], 10 * fs)
    tol = 0.05 * np.median(np.median(np.abs(voltage[:, inds]), axis=1))
    bp_win, _ = time_window_to_samples(bp_cw, fs)
    win, _ = time_window_to_samples(sort_info['clip_width'], fs)
    max_pre = -1 * bp_win[0] + win[0]
    while np.all(np.abs(mean_clip[first]) < tol):
        if first[0] >= max_pre:
            break
        first += 1
    bp_cw[0] = -1 * (-1 * bp_win[0] - first[0]) / fs
    min_post = (bp_n - bp_win[1]) + win[1] - 1
    while np.all(np.abs(mean_clip[last]) < tol):
        if last[0] <= min_post:
            break
        last -= 1
    bp_cw[1] = (bp_win[1] - (bp_n - last[0])) / fs
    start = (-1 * bp_win[0] + win[0]) - first[0]
    stop = start + (win[1] - win[0])
    n_new = last[0] - first[0]
    m = max(np.abs(bp_cw))
    return ([-1 * m, m], np.arange(start, n_new * C, n_new, dtype=np.int64),
            np.arange(stop, (n_new + 1) * C, n_new, dtype=np.int64))


def full_binary_pursuit_units(voltage, v_dtype, sort_info, unit_spikes, seg_w_items, chunk_len):
    """full_binary_pursuit.py:117-438 restated for units given as spike index
    arrays (one unit per array, the clustering output after consolidation),
    each unit's template the mean of its clips: binary-pursuit clip width,
    residual-clip noise covariance (np.random.randint rows per channel),
    sum-of-two-templates removal (check_template_pair > norm.sf(sigma)),
    separability metrics, noise-template removal, binary pursuit (one chunk
    length), clips of the input voltage and the per-channel output lists.
    Returns (output list, separability metrics, clip width used)."""
    from scipy.stats import norm
    C = sort_info['n_channels']
    fs = sort_info['sampling_rate']
    si = dict(sort_info)
    original_n = si['n_samples_per_chan']
    chan_win, _ = time_window_to_samples(si['clip_width'], fs)
    width, starts, stops = binary_pursuit_clip_width([u for u in unit_spikes if len(u)], voltage, si)
    si['clip_width'] = width
    bp_win, _ = time_window_to_samples(width, fs)
    N = bp_win[1] - bp_win[0]
    si['n_samples_per_chan'] = N
    pd = {'sampling_rate': fs, 'n_samples': voltage.shape[1], 'v_dtype': v_dtype}
    templates, resid, counts = [], [], []
    for u in unit_spikes:
        clips, _ = get_clips(pd, voltage, np.arange(C), u, width)
        t = np.mean(clips, axis=0)
        templates.append(t)
        resid.append(clips - t)
        counts.append(u.shape[0])
    templates = np.vstack(templates)
    resid = np.vstack(resid)
    n = resid.shape[0]
    rows = None if si['n_cov_samples'] >= n else [np.random.randint(0, n, si['n_cov_samples']) for _ in range(C)]
    covs = list(channel_covariance(resid, C, N, rows))
    delete = np.zeros(templates.shape[0], dtype=bool)
    sums, _ = find_overlap_templates(templates, N, C, np.abs(chan_win[0]) - 1, np.abs(chan_win[1]) - 1,
                                     np.array(counts, dtype=np.int64))
    for t_ind, tmpl, _ in sums:
        if check_template_pair(templates[t_ind], tmpl, covs, si) > norm.sf(si['sigma']):
            delete[t_ind] = True
    templates = templates[~delete]
    sm = separability_metrics(list(templates), covs, si)
    sm['templates'] = templates
    noisy = find_noisy_templates(sm)
    sm['neuron_lower_thresholds'] = np.maximum(np.maximum(sm['neuron_lower_thresholds'], 0), sm['neuron_lower_CI'])
    sm, noisy = check_noise_templates(sm, si, noisy)
    keep = ~noisy
    for key in ('templates', 'template_SS', 'neuron_variances', 'neuron_lower_thresholds', 'neuron_lower_CI',
                'contamination', 'peak_channel'):
        sm[key] = sm[key][keep]
    ev, lab, bpb, _, _ = binary_pursuit(voltage, v_dtype, si, sm, original_n - 1, chunk_len)
    ev, lab = np.asarray(ev, dtype=np.int64), np.asarray(lab, dtype=np.int64)
    clips, _ = get_clips(pd, voltage, np.arange(C), ev, width)
    out = []
    assign = unit_channels(clips, lab, C, N, seg_w_items[0]['thresholds'], si['sigma']) if ev.size else \
        {c: [] for c in range(C)}
    for c in range(C):
        item = [w for w in seg_w_items if w['channel'] == c][0]
        units = assign[c]
        if not units:
            out.append([[], [], [], [], item['ID']])
            continue
        sel = np.concatenate([np.flatnonzero(lab == u) for u in units])
        out.append([ev[sel] + item['index_window'][0], lab[sel], None, np.ones(sel.size, dtype=bool), item['ID']])
    return out, sm, width
