"""Drop-in replacement for the reference's binary_pursuit() entry point.

spikesorting_fullpursuit/parallel/binary_pursuit_parallel.py:128-845 drives
the OpenCL kernels of kernels/binary_pursuit.cl through pyopencl.  This module
keeps that function's signature, argument meaning, chunk planning, output
assembly and errors, and runs each chunk on an MI355X through libfbp.so (hand-
written gfx950 HIP kernels behind a C-ABI, bound with ctypes).  There is no CPU
path: without the library or a GPU the call raises.
"""
import numpy as np

from . import _lib
from .neuron_separability import add_n_spikes
from .segment import time_window_to_samples


def plan_chunks(n_chans, n_samples, sampling_rate, n_templates, template_samples_per_chan,
                n_max_shift_inds, gpu_total_bytes, max_gpu_memory=None, overlap_work_group=256):
    """Chunk length and onsets, with the reference's memory model
    (binary_pursuit_parallel.py:276-368).

    The device budget is max(0.5*total, total - 1 GiB), capped by
    max_gpu_memory.  One second of data is charged for the voltage, the
    likelihood buffer and per-window state including the overlap group buffers
    (a third of the per-window group count, twice).  Onsets advance by chunk - 3N while
    onset < n_samples - N.  Raises ValueError when a chunk would hold fewer
    than 4N samples (bpp.py:343-344).
    """
    budget = max(gpu_total_bytes * 0.5, gpu_total_bytes - (1024 * 1024 * 1024))
    if max_gpu_memory is not None:
        budget = min(budget, max_gpu_memory)
    assert budget > 0
    N = int(template_samples_per_chan)
    combos_per_window = np.int64((2 * n_max_shift_inds + 1) * (2 * n_max_shift_inds + 1) * n_templates)
    groups_per_window = np.int64(((combos_per_window - 1) / overlap_work_group) + 1)
    fixed_bytes = n_templates * n_chans * N * 4 + n_templates * 4
    bytes_per_second = n_chans * sampling_rate * 4
    bytes_per_second += n_templates * sampling_rate * 4
    bytes_per_second += (sampling_rate / N) * (4 + 4 + 4 + 4 + 4 + (groups_per_window / 3) * 4 + (groups_per_window / 3) * 4)
    seconds_per_chunk = (budget - fixed_bytes) / bytes_per_second
    device_address_bits = 64
    addressable_seconds = (((1 << device_address_bits) - 1) / (4 * sampling_rate * n_chans))
    seconds_per_chunk = min(seconds_per_chunk, np.floor(addressable_seconds))
    samples_per_chunk = int(np.floor(seconds_per_chunk * sampling_rate))
    if samples_per_chunk < 4 * N:
        raise ValueError("Cannot fit enough data on GPU to run binary pursuit. "
                         "Decrease template samples or get a new GPU.")
    onsets = []
    cur = 0
    while cur < n_samples - N:
        onsets.append(cur)
        cur += samples_per_chunk - 3 * N
    return samples_per_chunk, onsets


def residual_span(chunk_onsets, k, template_samples_per_chan, n_samples):
    """Samples [lo, hi) of the returned residual taken from chunk k.

    Chunks overlap by 3N samples (binary_pursuit_parallel.py:363-367).  A
    chunk sees only spikes whose onsets lie inside it, and a spike changes
    the N samples from its onset on, so the first N samples of chunk k + 1
    can still carry tails of chunk k's spikes that chunk k + 1 never
    subtracts.  The switch from chunk k to chunk k + 1 is therefore N samples
    after chunk k + 1's onset: every spike touching the samples a chunk
    supplies starts inside that chunk.
    """
    N = int(template_samples_per_chan)
    lo = 0 if k == 0 else min(chunk_onsets[k] + N, n_samples)
    hi = min(chunk_onsets[k + 1] + N, n_samples) if k + 1 < len(chunk_onsets) else n_samples
    return lo, hi


AUTO_ROUTES = (("bf16", "screened"), ("fp32", "screened"), ("bf16", "direct"), ("fp32", "direct"))


def open_segment(n_chans, K, n_templates, N, M, do_overlap, reduce_work_group, precision, overlap_mode,
                 max_passes, device, templates, bias, thresholds):
    """A segment for one chunk with its templates set.  "auto" precision or
    overlap mode takes the first route of AUTO_ROUTES that the library
    accepts for these shapes (it refuses a shape with ValueError, before any
    work); explicit choices are used as given."""
    routes = [(p, o) for p, o in AUTO_ROUTES
              if precision in ("auto", p) and overlap_mode in ("auto", o)] or [(precision, overlap_mode)]
    for k, (prec, mode) in enumerate(routes):
        seg = None
        try:
            seg = _lib.Segment(n_chans, K, n_templates, N, M, do_overlap, reduce_work_group, prec, mode,
                               max_passes, device)
            seg.set_templates(templates, bias, thresholds)
            return seg
        except ValueError:
            if seg is not None:
                seg.close()
            if k + 1 == len(routes):
                raise


def binary_pursuit(voltage, v_dtype, sort_info, separability_metrics, n_max_shift_inds=None,
                   kernels_path=None, max_gpu_memory=None, *, return_residual=False, precision="auto",
                   overlap_mode="auto", reduce_work_group=256, device=0, max_passes=0, stats=None,
                   verbose=False):
    """Full binary pursuit over `voltage` with the templates in
    `separability_metrics`; same arguments and returns as the reference's
    binary_pursuit (binary_pursuit_parallel.py:128-130, :829-845).

    voltage: [n_chans][n_samples] array of any real dtype (cast to float32 per
        chunk).  v_dtype: dtype of the returned adjusted clips.
    sort_info keys read: 'sampling_rate', 'clip_width', 'n_samples_per_chan',
        'do_overlap_recheck', 'get_adjusted_clips' (KeyError if absent, as in
        the reference).
    separability_metrics keys read: 'templates' [T][C*N], 'template_SS' [T],
        'neuron_lower_thresholds' [T]; 'n_spikes' is written.
    n_max_shift_inds: overlap shift range M (None or < 0 -> N // 2).
    kernels_path: accepted for compatibility and ignored (the kernels are in
        libfbp.so).
    max_gpu_memory: cap, in bytes, of the chunk planner's device budget.

    Keyword-only extensions (no reference counterpart):
    return_residual: also return the residual voltage float32 [C][n_samples].
    precision: correlation arithmetic: "bf16", "fp16" or "fp8" matrix-core
        screen whose candidates are re-scored in float32, or "fp32" (the
        reference's float32 chain for every position).  Every choice gives
        the same spikes, labels, clips and residual.
    overlap_mode: overlap recheck arithmetic: "screened" (float64 bounds from
        matrix-core tables, exact float32 score of every combination that can
        win) or "direct" (the reference's float32 score of every
        combination).  Both give the same decisions.
        With "auto" (the default for both) a chunk runs bf16 + screened when
        its shapes fit those kernels (templates of at most 96 samples, the
        screened overlap's LDS tables), else the first of fp32 + screened,
        bf16 + direct, fp32 + direct that fits; all are GPU routes with the
        same results.
    reduce_work_group: width of the overlap reduction whose tie-break the
        results follow (the reference uses its device's work-group size).
    stats: optional list; per-chunk engine statistics dicts are appended.

    Returns (event_indices int64, neuron_labels int64, binary_pursuit_spike_bool,
    adjusted_clips) -- four empty lists when nothing is found -- plus the
    residual when return_residual is set.
    """
    del kernels_path
    n_chans = int(voltage.shape[0])
    n_samples = int(voltage.shape[1])
    sampling_rate = sort_info['sampling_rate']
    chan_win, _ = time_window_to_samples(sort_info['clip_width'], sampling_rate)
    template_samples_per_chan = int(sort_info['n_samples_per_chan'])
    N = template_samples_per_chan

    templates = separability_metrics['templates']
    if templates.ndim == 1:
        templates = np.reshape(templates, (1, -1))
    templates = np.ascontiguousarray(np.float32(templates))
    n_templates = templates.shape[0]
    if templates.shape[1] != n_chans * N:
        raise ValueError("templates have %d samples per row; expected n_chans * n_samples_per_chan = %d"
                         % (templates.shape[1], n_chans * N))
    clip_init_samples = int(np.abs(chan_win[0]))
    template_sum_squared = np.float32(-0.5 * separability_metrics['template_SS'])
    neuron_lower_thresholds = np.float32(separability_metrics['neuron_lower_thresholds'])
    do_overlap = bool(sort_info['do_overlap_recheck'])
    get_clips = bool(sort_info['get_adjusted_clips'])

    _lib.device_count()   # RuntimeError("No GPU available") when there is none
    total_mem, _ = _lib.device_memory(device)
    if n_max_shift_inds is None or n_max_shift_inds < 0:
        n_max_shift_inds = template_samples_per_chan // 2
    n_max_shift_inds = int(np.uint32(n_max_shift_inds))
    chunk_len, chunk_onsets = plan_chunks(n_chans, n_samples, sampling_rate, n_templates, N,
                                          n_max_shift_inds, total_mem, max_gpu_memory, reduce_work_group)
    if verbose:
        print("Using", len(chunk_onsets), "chunks for binary pursuit.", flush=True)

    chunk_onset_lists, chunk_label_lists, chunk_bool_lists, chunk_clip_lists = [], [], [], []
    residual = np.zeros((n_chans, n_samples), dtype=np.float32) if return_residual else None
    if return_residual and chunk_onsets:
        residual[:, :] = np.float32(voltage)
    seg = None
    seg_len = None
    try:
        for k, start_index in enumerate(chunk_onsets):
            stop_index = min(n_samples, start_index + chunk_len)
            K = stop_index - start_index
            if seg is None or seg_len != K:
                if seg is not None:
                    seg.close()
                    seg = None
                seg = open_segment(n_chans, K, n_templates, N, n_max_shift_inds, do_overlap, reduce_work_group,
                                   precision, overlap_mode, max_passes, device, templates, template_sum_squared,
                                   neuron_lower_thresholds)
                seg_len = K
            chunk = voltage[:, start_index:stop_index]
            seg.set_voltage(chunk)
            seg.run()
            if stats is not None:
                stats.append(seg.stats())
            idx, lab = seg.spikes()
            if return_residual:
                lo, hi = residual_span(chunk_onsets, k, N, n_samples)
                residual[:, lo:hi] = seg.residual()[:, lo - start_index:hi - start_index]
            if verbose:
                print("Found", idx.size, "secret spikes in chunk", k + 1, flush=True)
            if idx.size > 0:
                clips = seg.clips() if get_clips else []
                chunk_onset_lists.append(idx.astype(np.int64) + start_index)
                chunk_label_lists.append(lab)
                chunk_clip_lists.append(clips)
                chunk_bool_lists.append(np.ones(idx.size, dtype=bool))
    finally:
        if seg is not None:
            seg.close()

    if len(chunk_onset_lists) > 0:
        event_indices = np.int64(np.hstack(chunk_onset_lists))
        neuron_labels = np.int64(np.hstack(chunk_label_lists))
        adjusted_clips = (np.vstack(chunk_clip_lists)).astype(v_dtype)
        binary_pursuit_spike_bool = np.hstack(chunk_bool_lists)
        event_indices += clip_init_samples
    else:
        event_indices, neuron_labels, binary_pursuit_spike_bool, adjusted_clips = [], [], [], []

    add_n_spikes(separability_metrics, neuron_labels)
    if return_residual:
        return event_indices, neuron_labels, binary_pursuit_spike_bool, adjusted_clips, residual
    return event_indices, neuron_labels, binary_pursuit_spike_bool, adjusted_clips
