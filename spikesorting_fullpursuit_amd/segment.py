"""Sample-window arithmetic and clip gathers shared with the reference's segment utilities."""
import copy

import numpy as np


def time_window_to_samples(time_window, sampling_rate):
    """Clip window in seconds -> [first, last+1] in samples.

    Same contract as spikesorting_fullpursuit/parallel/segment_parallel.py:82-98:
    the first element is made negative, rounded and capped at 0; the second is
    rounded, floored at 1 and has 1 added so it slices inclusively.  Returns
    the sample window and the (possibly sign-flipped) copy of the time window.
    """
    window = copy.copy(time_window)
    if window[0] > 0:
        window[0] *= -1
    first = min(int(round(window[0] * sampling_rate)), 0)
    last = max(int(round(window[1] * sampling_rate)), 1) + 1
    return [first, last], window


# Voltage dtypes whose values float32 holds exactly (the device gathers float32).
_EXACT_IN_F32 = (np.float32, np.int8, np.uint8, np.int16, np.uint16)


def get_clips(probe_dict, voltage, neighbors, event_indices, clip_width, use_memmap=False, check_valid=True,
              device=0):
    """Clips of `voltage` around `event_indices` on the GPU.

    Same contract as spikesorting_fullpursuit/parallel/segment_parallel.py:440-509:
    clips of every channel in `neighbors` (in that order) over the window
    time_window_to_samples(clip_width) gives, one row per valid event, in event
    order, dtype probe_dict['v_dtype']; events are trimmed from the front and
    the back only (the reference requires ascending indices); returns
    (clips, valid mask), (None, all-False mask) when no event is valid, and a
    (0, width) array for no events.

    `voltage` is a [C][K] numpy array (the selected channels are uploaded)
    or a Segment, whose resident input voltage is used without a copy.  The gather runs on the
    device (fbp_get_clips); float32 and 8/16-bit integer voltages are exact,
    other dtypes are refused.  use_memmap (clips written to a file) is not
    supported.
    """
    from . import _lib
    event_indices = np.asarray(event_indices)
    if event_indices.ndim > 1:
        raise ValueError("Event_indices must be one dimensional array of indices")
    if use_memmap:
        raise NotImplementedError("use_memmap: clips are returned in memory")
    window, _ = time_window_to_samples(clip_width, probe_dict['sampling_rate'])
    neighbors = np.asarray(neighbors, dtype=np.int64).ravel()
    width = (window[1] - window[0]) * neighbors.size
    if len(event_indices) == 0:
        return np.zeros((0, width), dtype=probe_dict['v_dtype']), np.ones(0, dtype="bool")
    if isinstance(voltage, _lib.Segment):
        if np.any(neighbors < 0) or np.any(neighbors >= voltage.C):
            raise IndexError("neighbors index a channel outside the voltage")
        ptr, stride = voltage.voltage_device()
        clips, valid = _lib.get_clips_device(ptr, stride, voltage.C, int(probe_dict['n_samples']),
                                             event_indices.astype(np.int64), neighbors.astype(np.uint32),
                                             window[0], window[1], check_valid=check_valid, device=voltage.device)
    else:
        voltage = np.asarray(voltage)
        if voltage.ndim != 2:
            raise ValueError("voltage must be [channels][samples]")
        if voltage.dtype.type not in _EXACT_IN_F32:
            raise TypeError("get_clips on the device takes float32 or 8/16-bit integer voltage, not %s"
                            % voltage.dtype)
        if np.any(neighbors < 0) or np.any(neighbors >= voltage.shape[0]):
            raise IndexError("neighbors index a channel outside the voltage")
        clips, valid = _lib.get_clips_host(voltage, int(probe_dict['n_samples']), event_indices.astype(np.int64),
                                           neighbors.astype(np.uint32), window[0], window[1],
                                           check_valid=check_valid, device=device)
    if check_valid:
        # the reference returns None when a trimming scan runs off the end:
        # every event starts before the data, or every event ends after it
        ev = event_indices.astype(np.int64)
        if np.all(ev + window[0] < 0) or np.all(ev + window[1] >= probe_dict['n_samples']):
            return None, valid
    return clips.astype(probe_dict['v_dtype'], copy=False), valid
