"""Sample-window arithmetic shared with the reference's segment utilities."""
import copy


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
