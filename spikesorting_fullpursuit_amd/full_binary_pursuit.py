"""The caller's side of binary pursuit, after the templates are final.

Mirrors the part of full_binary_pursuit() (spikesorting_fullpursuit/
full_binary_pursuit.py:351-438) that runs binary pursuit and turns its output
into per-channel sorter output: binary pursuit on the GPU, clips of the input
voltage gathered on the GPU (get_clips, full_binary_pursuit.py:357-359), the
mean clip of every unit reduced on the GPU (full_binary_pursuit.py:366-368),
each unit assigned to the channel where its mean clip has the best
range-to-noise ratio (full_binary_pursuit.py:369-383), and the output list
[events, labels, clips, bp_bool, work item ID] per channel
(full_binary_pursuit.py:385-438).  The template set itself (consolidation,
robust templates, noise covariance, overlap-template removal, separability
metrics: full_binary_pursuit.py:117-341) is the input.
"""
import numpy as np

from . import _lib
from .binary_pursuit_parallel import binary_pursuit
from .segment import get_clips


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
