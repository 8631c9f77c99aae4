"""Template alignment used when deciding which templates reach binary pursuit.

optimal_align_templates has the reference's signature and returns
(spikesorting_fullpursuit/consolidate.py:9-71); the cross-correlation over the
lags it searches runs on the GPU (fbp_template_pair_lags), the per-channel
shift and truncation of the two templates on the host.
"""
import numpy as np

from . import _lib


def alignment_lags(n_samples, max_shift):
    """Lag range [lo, hi] searched for templates of n_samples flat samples:
    every lag of the full cross-correlation without max_shift, else
    -max_shift..max_shift (clipped to the full range; a negative max_shift
    leaves no lag, where the reference's argmax of an empty slice raises)."""
    full = n_samples - 1
    if max_shift is None:
        return -full, full
    if max_shift < 0:
        raise ValueError("attempt to get argmax of an empty sequence (max_shift = %d)" % max_shift)
    m = min(int(max_shift), full)
    return -m, m


def best_lag(xcorr, lag_lo, align_abs):
    """Lag of the first largest entry of each row of xcorr (absolute values
    with align_abs), xcorr[:, k] holding lag lag_lo + k."""
    score = np.abs(xcorr) if align_abs else xcorr
    return np.argmax(score, axis=1).astype(np.int64) + lag_lo


def shift_truncate(temp_1, temp_2, n_chans, lag, zero_pad=False):
    """Both templates with every channel shifted by `lag` samples against each
    other and cut to the overlapping part (consolidate.py:34-54), or placed
    back at full length with zeros where nothing overlaps (:56-69)."""
    N = temp_1.shape[0] // n_chans
    a = temp_1.reshape(n_chans, N)
    b = temp_2.reshape(n_chans, N)
    if lag > 0:
        s1, s2 = a[:, lag:], b[:, :N - lag]
    else:
        s1, s2 = a[:, :N + lag], b[:, -lag:]
    width = s1.shape[1]
    if not zero_pad:
        return s1.reshape(-1).copy(), s2.reshape(-1).copy(), width
    p1 = np.zeros((n_chans, N))
    p2 = np.zeros((n_chans, N))
    cols = slice(lag, N) if lag >= 0 else slice(0, N + lag)
    p1[:, cols] = s1
    p2[:, cols] = s2
    return p1.reshape(-1), p2.reshape(-1), N


def optimal_align_templates(temp_1, temp_2, n_chans, max_shift=None, align_abs=False, zero_pad=False, device=0):
    """Align temp_2 to temp_1 (consolidate.py:9-71): the lag of the largest
    (absolute, with align_abs) flat cross-correlation sum_i temp_1[i + lag]
    temp_2[i], first of equal maxima, searched over every lag or within
    +-max_shift; then both templates shifted by that lag on every channel and
    truncated (or zero padded with zero_pad).  Returns (shifted temp_1,
    shifted temp_2, lag, samples per channel of the result); the inputs
    themselves when the lag is 0.  The cross-correlation is float64 in one
    fixed order on the GPU; np.correlate's BLAS order can differ in the last
    bits, so lags with equal scores to that precision may be chosen
    differently."""
    temp_1 = np.asarray(temp_1)
    temp_2 = np.asarray(temp_2)
    n_samples_per_chan = int(temp_1.shape[0] / n_chans) if temp_1.ndim == 1 else 0
    if temp_1.shape[0] != temp_2.shape[0] or temp_1.ndim > 1 or n_samples_per_chan == 0:
        raise ValueError("Input templates must be 1D vectors of the same size")
    if temp_1.shape[0] % n_chans != 0:
        raise ValueError("Template shape[0] must be evenly divisible by n_chans (i.e. there are the same number "
                         "of samples per channel).")
    lo, hi = alignment_lags(temp_1.shape[0], max_shift)
    xc, _ = _lib.template_pair_lags(np.vstack([temp_1, temp_2]).astype(np.float64), n_chans, [[0, 1]], lo, hi,
                                    aligned=False, device=device)
    lag = int(best_lag(xc, lo, align_abs)[0])
    if lag == 0:
        return temp_1, temp_2, 0, n_samples_per_chan
    s1, s2, width = shift_truncate(temp_1, temp_2, n_chans, lag, zero_pad)
    return s1, s2, lag, width
