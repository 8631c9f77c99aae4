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
    """Channel of every unit (full_binary_pursuit.py:367-383).

    unit_means[u] is the mean clip of unit u over all channels; the unit goes
    to the first channel with the largest (max - min) / (3 * thresholds[chan]
    / sigma) of its mean clip on that channel.  Returns {channel: [units]}.
    """
    chans_to_template_labels = {chan: [] for chan in range(n_channels)}
    for unit in np.unique(neuron_labels):
        curr_template = unit_means[unit]
        unit_best_snr, unit_best_chan = -1.0, None
        for chan in range(n_channels):
            background_noise_std = thresholds[chan] / sigma
            chan_template = curr_template[n_samples_per_chan * chan:n_samples_per_chan * (chan + 1)]
            chan_snr = (np.amax(chan_template) - np.amin(chan_template)) / (3 * background_noise_std)
            if chan_snr > unit_best_snr:
                unit_best_snr, unit_best_chan = chan_snr, chan
        chans_to_template_labels[unit_best_chan].append(unit)
    return chans_to_template_labels


def segment_output(crossings, neuron_labels, bp_bool, clips, chans_to_template_labels, seg_w_items, n_channels,
                   n_samples_per_chan, original_clip_starts, original_clip_stops, v_dtype, output_clips=False):
    """Per-channel output (full_binary_pursuit.py:385-438): for each channel,
    [events + work item onset, labels, clips mapped to the work item's
    neighbourhood (or []), bp_bool, work item ID], units in the order they were
    assigned; empty lists for a channel without units."""
    seg_data = []
    for chan in range(n_channels):
        curr_item = next((w for w in seg_w_items if w['channel'] == chan), None)
        if curr_item is None:
            raise RuntimeError("Could not find a matching work item for unit")
        if len(chans_to_template_labels[chan]) == 0:
            seg_data.append([[], [], [], [], curr_item['ID']])
            continue
        chan_events, chan_labels, chan_bp_bool, chan_clips = [], [], [], []
        for unit in chans_to_template_labels[chan]:
            select = neuron_labels == unit
            chan_events.append(crossings[select])
            chan_labels.append(neuron_labels[select])
            chan_bp_bool.append(bp_bool[select])
            if output_clips:
                neighbors = curr_item['neighbors']
                unit_clips = np.zeros((np.count_nonzero(select), neighbors.shape[0] * n_samples_per_chan),
                                      dtype=v_dtype)
                for neigh in range(neighbors.shape[0]):
                    chan_ind = neighbors[neigh]
                    unit_clips[:, neigh * n_samples_per_chan:(neigh + 1) * n_samples_per_chan] = \
                        clips[select, original_clip_starts[chan_ind]:original_clip_stops[chan_ind]]
                chan_clips.append(unit_clips)
            else:
                chan_clips.append([])
        chan_events = np.hstack(chan_events)
        chan_events += curr_item['index_window'][0]
        seg_data.append([chan_events, np.hstack(chan_labels), np.vstack(chan_clips), np.hstack(chan_bp_bool),
                         curr_item['ID']])
    return seg_data


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
        unit_means, _ = _lib.unit_mean_clips(np.asarray(clips, dtype=np.float32), neuron_labels, n_units,
                                             device=device)
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
