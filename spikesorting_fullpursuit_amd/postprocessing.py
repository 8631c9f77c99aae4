"""Duplicate removal after binary pursuit.

Drop-ins for remove_binary_pursuit_duplicates
(spikesorting_fullpursuit/analyze_spike_timing.py:116-156) and the per-channel,
per-segment driver WorkItemSummary.remove_segment_binary_pursuit_duplicates
(spikesorting_fullpursuit/postprocessing.py:521-546).  The unit templates
(np.mean of the unit's clips) and every clip's squared distance to its unit's
template are computed on the GPU, bit-exact to numpy's float32 arithmetic
(fbp_unit_mean_clips, fbp_clip_template_sse); the pairwise scan over
consecutive events is the reference's, on the host.
"""
import numpy as np

from . import _lib


def _scan(event_indices, sse, binary_pursuit_bool, tol_inds):
    """analyze_spike_timing.py:124-154 with the squared distances given."""
    keep_bool = np.ones(event_indices.size, dtype="bool")
    curr_index, next_index = 0, 1
    while next_index < event_indices.size:
        if event_indices[next_index] - event_indices[curr_index] <= tol_inds:
            if binary_pursuit_bool[curr_index] and binary_pursuit_bool[next_index]:
                if np.float64(sse[curr_index]) <= np.float64(sse[next_index]):
                    keep_bool[next_index] = False
                    next_index += 1
                else:
                    keep_bool[curr_index] = False
                    curr_index = next_index
                    next_index += 1
            elif binary_pursuit_bool[curr_index] and ~binary_pursuit_bool[next_index]:
                next_index += 1
            else:
                curr_index = next_index
                next_index += 1
        else:
            curr_index = next_index
            next_index += 1
    return keep_bool


def remove_binary_pursuit_duplicates(event_indices, clips, unit_template, binary_pursuit_bool, tol_inds, device=0):
    """Which spikes of one unit to keep: of two spikes within tol_inds that
    were both found by binary pursuit, the one whose clip is farther from the
    unit template (float32 sum of squares) is dropped; same scan and rules as
    analyze_spike_timing.py:116-156.  clips must be float32."""
    clips = np.asarray(clips)
    if clips.dtype != np.float32:
        raise TypeError("clips must be float32 (the device reproduces float32 arithmetic)")
    n = np.asarray(event_indices).size
    sse = _lib.clip_template_sse(clips, np.zeros(n, dtype=np.uint32), np.asarray(unit_template, np.float32)[None, :],
                                 device=device) if n else np.zeros(0, np.float32)
    return _scan(np.asarray(event_indices), sse, np.asarray(binary_pursuit_bool), tol_inds)


def remove_segment_binary_pursuit_duplicates(sort_data, n_chans, n_segments, half_clip_inds, device=0):
    """postprocessing.py:521-546 on sort_data[chan][seg] = [events, labels,
    clips, bp_bool, ...] (modified in place, as the reference's method does):
    for every unit of every channel and segment, the unit template is the
    mean of its clips and duplicates within 2 * half_clip_inds are removed."""
    for chan in range(n_chans):
        for seg in range(n_segments):
            data = sort_data[chan][seg]
            if len(data[0]) == 0:
                continue
            labels = np.asarray(data[1])
            clips = np.asarray(data[2])
            if clips.dtype != np.float32:
                raise TypeError("clips must be float32 (the device reproduces float32 arithmetic)")
            units, index = np.unique(labels, return_inverse=True)
            index = index.astype(np.uint32).ravel()
            means, _ = _lib.unit_mean_clips(clips, index, units.size, device=device)
            sse = _lib.clip_template_sse(clips, index, means, device=device)
            keep_bool = np.ones(data[0].shape[0], dtype="bool")
            for u in range(units.size):
                unit_select = index == u
                keep_bool[unit_select] = _scan(data[0][unit_select], sse[unit_select], data[3][unit_select],
                                               2 * half_clip_inds)
            data[0] = data[0][keep_bool]
            data[1] = data[1][keep_bool]
            data[2] = data[2][keep_bool, :]
            data[3] = data[3][keep_bool]
    return sort_data
