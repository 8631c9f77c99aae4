"""The segment driver around binary pursuit (spikesorting_fullpursuit/
full_binary_pursuit.py).

full_binary_pursuit() takes one segment's clustering results to the sorter's
per-channel output, in the reference's order (full_binary_pursuit.py:117-438):
the binary-pursuit clip width from the mean clip of every crossing, one
template per consolidated unit, the per-channel noise covariance of the
residual clips, removal of templates that are sums of two others, the
separability metrics and the removal of noise templates, then binary pursuit
and the per-channel output.  The clip gathers, unit means, covariance,
variances, overlap-template search, template-pair cross-correlations and
binary pursuit itself run on the GPU through libfbp.so; the decisions on top
of them are the reference's, on the host.

Consolidation of the clustering output (the reference's SegSummary, with its
merging and sharpening) and its robust templates belong to the sorter's
clustering stage, not to binary pursuit: they are plug-ins here
(`segment_summary`, `robust_template`).  Without them every clustered unit
is its own template, the mean of its clips.

binary_pursuit_segment() is the tail alone (full_binary_pursuit.py:351-438),
for callers whose templates are already final.
"""
from copy import deepcopy

import numpy as np

from . import _lib
from .binary_pursuit_parallel import binary_pursuit
from .segment import get_clips, time_window_to_samples


def assign_units_to_channels(neuron_labels, unit_means, n_channels, n_samples_per_chan, thresholds, sigma):
    """Channel of every unit (the rule of full_binary_pursuit.py:367-383).

    Each unit present in `neuron_labels` (ascending) is placed on the channel
    where its mean clip (unit_means[unit], channel blocks of
    n_samples_per_chan) has the largest peak-to-peak range relative to three
    noise standard deviations, noise = thresholds[channel] / sigma; ties go to
    the lowest channel, and ratios that are not above -1 (NaN) never win.
    Returns {channel: [units in ascending order]} with every channel present.
    """
    noise = np.asarray([thresholds[c] / sigma for c in range(n_channels)], dtype=np.float64)
    units_of = {c: [] for c in range(n_channels)}
    for unit in np.unique(neuron_labels):
        blocks = np.asarray(unit_means[unit])[:n_channels * n_samples_per_chan].reshape(n_channels,
                                                                                      n_samples_per_chan)
        spread = blocks.max(axis=1) - blocks.min(axis=1)   # in the clips' dtype, per channel
        ratio = np.asarray([spread[c] / (3 * noise[c]) for c in range(n_channels)])
        eligible = np.flatnonzero(ratio > -1.0)
        if eligible.size == 0:
            raise KeyError("unit %d has no channel with a finite range-to-noise ratio" % unit)
        best = eligible[np.argmax(ratio[eligible])]   # first of the largest
        units_of[int(best)].append(unit)
    return units_of


def _work_item_of(seg_w_items, channel):
    for item in seg_w_items:
        if item['channel'] == channel:
            return item
    raise RuntimeError("no work item of this segment covers channel %d" % channel)


def _neighbourhood_clips(clips, rows, neighbors, starts, stops, width, v_dtype):
    """Clips of the selected rows mapped onto a work item's neighbourhood:
    block k of each row is columns starts[neighbors[k]]:stops[neighbors[k]]."""
    out = np.zeros((rows.size, len(neighbors) * width), dtype=v_dtype)
    for k, chan in enumerate(neighbors):
        out[:, k * width:(k + 1) * width] = clips[rows, starts[chan]:stops[chan]]
    return out


def segment_output(crossings, neuron_labels, bp_bool, clips, chans_to_template_labels, seg_w_items, n_channels,
                   n_samples_per_chan, original_clip_starts, original_clip_stops, v_dtype, output_clips=False):
    """Sorter output of one segment, one entry per channel (the layout of
    full_binary_pursuit.py:385-438): [event indices shifted by the work
    item's window onset, labels, clips (mapped to the work item's
    neighbourhood with output_clips, else an empty column per unit), binary
    pursuit flags, work item ID], the channel's units in assignment order, or
    [[], [], [], [], ID] for a channel without units."""
    result = []
    for chan in range(n_channels):
        item = _work_item_of(seg_w_items, chan)
        units = chans_to_template_labels[chan]
        if not units:
            result.append([[], [], [], [], item['ID']])
            continue
        rows = [np.flatnonzero(neuron_labels == unit) for unit in units]
        events = np.concatenate([crossings[r] for r in rows]) + item['index_window'][0]
        labels = np.concatenate([neuron_labels[r] for r in rows])
        flags = np.concatenate([bp_bool[r] for r in rows])
        if output_clips:
            clip_rows = np.vstack([_neighbourhood_clips(clips, r, item['neighbors'], original_clip_starts,
                                                        original_clip_stops, n_samples_per_chan, v_dtype)
                                   for r in rows])
        else:
            clip_rows = np.zeros((len(units), 0))
        result.append([events, labels, clip_rows, flags, item['ID']])
    return result


def binary_pursuit_segment(voltage, v_dtype, sort_info, separability_metrics, seg_w_items,
                           original_clip_starts, original_clip_stops, original_clip_width,
                           bp_reduction_samples_per_chan, n_max_shift_inds=None, max_gpu_memory=None,
                           output_clips=False, device=0, **bp_kwargs):
    """full_binary_pursuit.py:351-438 for a segment whose templates are final.

    sort_info holds the binary pursuit clip width and samples per channel on
    entry (full_binary_pursuit.py:207-216) and the original ones on return,
    as in the reference.  Returns the per-channel output list.
    """
    crossings, neuron_labels, bp_bool, clips = binary_pursuit(
        voltage, v_dtype, sort_info, separability_metrics, n_max_shift_inds=n_max_shift_inds,
        kernels_path=None, max_gpu_memory=max_gpu_memory, device=device, **bp_kwargs)
    n_chans = sort_info['n_channels']
    clips_dict = {'sampling_rate': sort_info['sampling_rate'], 'n_samples': seg_w_items[0]['n_samples'],
                  'v_dtype': v_dtype}
    if not sort_info['get_adjusted_clips']:
        clips, _ = get_clips(clips_dict, voltage, np.arange(0, n_chans, dtype=np.int64), crossings,
                             clip_width=sort_info['clip_width'], device=device)
    if sort_info['output_separability_metrics']:
        sort_info['separability_metrics'][seg_w_items[0]['seg_number']] = separability_metrics
    if len(crossings) > 0:
        n_units = int(np.amax(neuron_labels)) + 1
        unit_means, _ = _lib.unit_mean_clips(clips, neuron_labels, n_units, device=device)
    else:
        unit_means = None
    chans_to_template_labels = assign_units_to_channels(
        neuron_labels, unit_means, n_chans, sort_info['n_samples_per_chan'], seg_w_items[0]['thresholds'],
        sort_info['sigma'])
    sort_info['clip_width'] = original_clip_width
    sort_info['n_samples_per_chan'] = bp_reduction_samples_per_chan
    return segment_output(crossings, neuron_labels, bp_bool, clips, chans_to_template_labels, seg_w_items, n_chans,
                          sort_info['n_samples_per_chan'], original_clip_starts, original_clip_stops, v_dtype,
                          output_clips=output_clips)


def _output_without_units(seg_w_items, n_channels):
    """[[], [], [], [], ID] for every channel (full_binary_pursuit.py:304-316)."""
    return [[[], [], [], [], _work_item_of(seg_w_items, chan)['ID']] for chan in range(n_channels)]


def get_binary_pursuit_clip_width(seg_w_items, clips_dict, voltage, data_dict, sort_info, device=0):
    """Clip width for binary pursuit (full_binary_pursuit.py:14-115).

    Clips of every crossing of the segment are taken over the clip width
    scaled by sort_info['max_binary_pursuit_clip_width_factor'], and each end
    of their mean clip (the first, then the last sample of every channel) is
    trimmed while it stays within 5 % of the median absolute voltage of a
    random 10 s of the segment on every channel, but never inside the
    original clip width.  Returns the symmetric width [-w, w] (seconds) and
    the column ranges of the original clip width inside clips of the new
    one (start and stop per channel).  A factor <= 1, no crossings or no
    valid clip keep the original width.  The clip gather and the mean clip
    run on the GPU; the median uses the global numpy generator, as the
    reference."""
    C = sort_info['n_channels']
    N = sort_info['n_samples_per_chan']
    fs = sort_info['sampling_rate']

    def unchanged():
        w = max(np.abs(sort_info['clip_width']))
        starts = np.arange(0, N * C, N, dtype=np.int64)
        return [-1 * w, w], starts, starts + N

    if sort_info['max_binary_pursuit_clip_width_factor'] <= 1.0:
        return unchanged()
    events = [data_dict['results_dict'][w['ID']][0] for w in seg_w_items
              if w['ID'] in data_dict['results_dict'] and len(data_dict['results_dict'][w['ID']][0]) > 0]
    if not events:
        return unchanged()
    events = np.sort(np.hstack(events))
    factor = sort_info['max_binary_pursuit_clip_width_factor']
    wide = [factor * v for v in sort_info['clip_width']]
    clips, valid = get_clips(clips_dict, voltage, np.arange(0, voltage.shape[0]), events, clip_width=wide,
                             device=device)
    if np.count_nonzero(valid) == 0:
        return unchanged()
    mean_clip = _lib.unit_mean_clips(clips, np.zeros(clips.shape[0], dtype=np.uint32), 1, device=device)[0][0]
    n_wide = clips.shape[1] // C
    first = np.arange(C, dtype=np.int64) * n_wide
    last = first + n_wide - 1
    noise_inds = np.random.choice(voltage.shape[1], 10 * fs)
    tolerance = 0.05 * np.median(np.median(np.abs(voltage[:, noise_inds]), axis=1))
    wide_win, _ = time_window_to_samples(wide, fs)
    win, _ = time_window_to_samples(sort_info['clip_width'], fs)
    # trim the front while every channel's first sample is flat, not past the original window
    front_limit = win[0] - wide_win[0]
    while np.all(np.abs(mean_clip[first]) < tolerance) and first[0] < front_limit:
        first += 1
    back_limit = (n_wide - wide_win[1]) + win[1] - 1
    while np.all(np.abs(mean_clip[last]) < tolerance) and last[0] > back_limit:
        last -= 1
    new_width = [-1 * (-1 * wide_win[0] - first[0]) / fs, (wide_win[1] - (n_wide - last[0])) / fs]
    start = front_limit - first[0]
    n_new = last[0] - first[0]
    starts = np.arange(start, n_new * C, n_new, dtype=np.int64)
    stops = np.arange(start + (win[1] - win[0]), (n_new + 1) * C, n_new, dtype=np.int64)
    w = max(np.abs(new_width))
    return [-1 * w, w], starts, stops


class UnitSummary:
    """Default consolidation: every (work item, label) cluster of the segment
    is one unit, with its crossings as spike indices, and no sharpening.
    The reference's SegSummary (consolidate.py) has the same interface
    (`summaries`, `set_bp_templates`, `sharpen_across_chans`) and merges
    units across channels; pass it as full_binary_pursuit's
    segment_summary to reproduce the reference's template set."""

    def __init__(self, seg_data, seg_w_items, sort_info, v_dtype, absolute_refractory_period=None, verbose=False):
        self.summaries = []
        for crossings, labels, _clips, _bp, item_id in seg_data:
            if len(crossings) == 0:
                continue
            crossings = np.asarray(crossings)
            labels = np.asarray(labels)
            for label in np.unique(labels):
                self.summaries.append({'spike_indices': np.sort(crossings[labels == label]), 'ID': item_id,
                                       'label': label})

    def set_bp_templates(self, templates):
        for unit, tmpl in zip(self.summaries, templates):
            unit['bp_template'] = tmpl

    def sharpen_across_chans(self, chan_covariance_mats):
        return None


def full_binary_pursuit(work_items, data_dict, seg_number, sort_info, v_dtype, overlap_ratio_threshold,
                        absolute_refractory_period, kernels_path=None, max_gpu_memory=None, use_memmap=False,
                        output_clips=False, *, segment_summary=None, robust_template=None, device=0, **bp_kwargs):
    """One segment from clustering results to the sorter's per-channel output
    (full_binary_pursuit.py:117-438), with the reference's arguments.

    data_dict['segment_voltages'][seg_number] = (buffer, shape) of the
    segment's voltage in v_dtype; data_dict['results_dict'][ID] = (crossings,
    labels, ...) of each work item.  segment_summary(seg_data, seg_w_items,
    sort_info, v_dtype, absolute_refractory_period=..., verbose=False) builds
    the units (default UnitSummary); robust_template(clips) gives a unit's
    template (default: the mean clip, on the GPU).  overlap_ratio_threshold
    is accepted for the reference's signature (its consolidation reads it
    from sort_info).  Returns the per-channel list [events, labels, clips,
    bp_bool, work item ID], or [[[], [], [], [], None]] when the segment has
    no unit.  sort_info['clip_width'] and ['n_samples_per_chan'] hold the
    original values again on return; sort_info['binary_pursuit_clip_width']
    is widened to the width used."""
    from scipy.stats import norm
    from .neuron_separability import (check_noise_templates, check_template_pairs, clip_noise_covariance,
                                      compute_separability_metrics, delete_noise_units, find_noisy_templates,
                                      set_bp_threshold)
    from .overlap_templates import find_overlap_templates
    del kernels_path, overlap_ratio_threshold
    if use_memmap:
        raise NotImplementedError("use_memmap: the segment voltage is read from data_dict['segment_voltages']")
    buf, shape = data_dict['segment_voltages'][seg_number][0], data_dict['segment_voltages'][seg_number][1]
    voltage = np.frombuffer(buf, dtype=v_dtype).reshape(shape)
    C = voltage.shape[0]
    all_chans = np.arange(0, C, dtype=np.int64)
    original_clip_width = [x for x in sort_info['clip_width']]
    original_n = int(sort_info['n_samples_per_chan'])
    n_max_shift_inds = original_n - 1
    seg_w_items = [w for w in work_items if w['seg_number'] == seg_number]
    clips_dict = {'sampling_rate': sort_info['sampling_rate'], 'n_samples': seg_w_items[0]['n_samples'],
                  'v_dtype': v_dtype}

    # clustering results of the segment, every work item widened to all channels
    seg_data, saved_neighbors = [], []
    for item in seg_w_items:
        result = data_dict['results_dict'].get(item['ID'])
        if result is None:
            seg_data.append([[], [], [], [], item['ID']])
            continue
        saved_neighbors.append(item['neighbors'])
        item['neighbors'] = np.copy(all_chans)
        if len(result[0]) == 0:
            seg_data.append([[], [], [], [], item['ID']])
            continue
        clips, _ = get_clips(clips_dict, voltage, item['neighbors'], result[0], clip_width=sort_info['clip_width'],
                             device=device)
        entry = [result[0], result[1], clips, np.zeros(len(result[0]), dtype="bool"), item['ID']]
        if isinstance(entry[0][0], np.ndarray) and entry[0][0].size > 0:
            entry[0][0] += item['index_window'][0]   # nested crossing lists (full_binary_pursuit.py:175-178)
        seg_data.append(entry)

    make_summary = segment_summary or UnitSummary
    summary = make_summary(seg_data, seg_w_items, deepcopy(sort_info), v_dtype,
                           absolute_refractory_period=absolute_refractory_period, verbose=False)
    if len(summary.summaries) == 0:
        print("Found no neuron templates for binary pursuit")
        return [[[], [], [], [], None]]
    chan_win, _ = time_window_to_samples(sort_info['clip_width'], sort_info['sampling_rate'])

    # binary-pursuit clip width (the sort_info values are restored on return)
    sort_info['clip_width'], clip_starts, clip_stops = get_binary_pursuit_clip_width(
        seg_w_items, clips_dict, voltage, data_dict, sort_info, device=device)
    bp_width = sort_info.setdefault('binary_pursuit_clip_width', [0, 0])
    bp_width[0] = min(sort_info['clip_width'][0], bp_width[0])
    bp_width[1] = max(sort_info['clip_width'][1], bp_width[1])
    bp_win, _ = time_window_to_samples(sort_info['clip_width'], sort_info['sampling_rate'])
    sort_info['n_samples_per_chan'] = bp_win[1] - bp_win[0]
    reduced_n = int(clip_stops[0] - clip_starts[0])
    if reduced_n != original_n:
        raise RuntimeError("Template reduction from binary pursuit does not have the same number of samples as "
                           "original!")
    N = sort_info['n_samples_per_chan']

    def restore():
        sort_info['clip_width'] = original_clip_width
        sort_info['n_samples_per_chan'] = reduced_n

    # one template per unit; residual clips for the noise covariance
    templates, resid, n_spikes = [], [], []
    for unit in summary.summaries:
        clips, _ = get_clips(clips_dict, voltage, all_chans, unit['spike_indices'],
                             clip_width=sort_info['clip_width'], device=device)
        if robust_template is not None:
            tmpl = robust_template(clips)
        else:
            tmpl = _lib.unit_mean_clips(clips, np.zeros(clips.shape[0], dtype=np.uint32), 1, device=device)[0][0]
        templates.append(tmpl)
        resid.append(clips - tmpl)
        n_spikes.append(unit['spike_indices'].shape[0])
    templates = np.vstack(templates)
    resid = np.vstack(resid)
    covs = clip_noise_covariance(np.ascontiguousarray(resid, dtype=np.float32) if resid.dtype != np.float32
                                 else resid, C, N, sort_info['n_cov_samples'], device=device)

    # templates that are the sum of two others (full_binary_pursuit.py:254-285)
    sums = find_overlap_templates(templates, N, C, np.int64(np.abs(chan_win[0]) - 1),
                                  np.int64(np.abs(chan_win[1]) - 1), np.array(n_spikes, dtype=np.int64),
                                  device=device)
    delete = np.zeros(templates.shape[0], dtype="bool")
    if sums:
        p = check_template_pairs(np.vstack([templates[s[0]] for s in sums]), np.vstack([s[1] for s in sums]),
                                 covs, sort_info, device=device)
        for s, pc in zip(sums, p):
            if pc > norm.sf(sort_info['sigma']):
                delete[s[0]] = True
    templates = templates[~delete, :]
    summary.summaries = [u for u, d in zip(summary.summaries, delete) if not d]
    for item, nb in zip([w for w in seg_w_items if w['ID'] in data_dict['results_dict']], saved_neighbors):
        item['neighbors'] = nb
    if len(summary.summaries) == 0:
        restore()
        return _output_without_units(seg_w_items, sort_info['n_channels'])

    summary.set_bp_templates(templates)
    summary.sharpen_across_chans(covs)
    templates = [u['bp_template'] for u in summary.summaries]
    del summary

    # separability and noise templates (full_binary_pursuit.py:321-330)
    metrics = compute_separability_metrics(templates, covs, sort_info, device=device)
    noisy = find_noisy_templates(metrics, sort_info)
    metrics = set_bp_threshold(metrics)
    metrics, noisy = check_noise_templates(metrics, sort_info, noisy, device=device)
    metrics = delete_noise_units(metrics, noisy)
    if metrics['templates'].shape[0] == 0:
        restore()
        return _output_without_units(seg_w_items, sort_info['n_channels'])

    return binary_pursuit_segment(voltage, v_dtype, sort_info, metrics, seg_w_items, clip_starts, clip_stops,
                                  original_clip_width, reduced_n, n_max_shift_inds=n_max_shift_inds,
                                  max_gpu_memory=max_gpu_memory, output_clips=output_clips, device=device,
                                  **bp_kwargs)
