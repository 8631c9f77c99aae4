"""Which templates look like the sum of two other shifted templates.

Drop-in for find_overlap_templates (spikesorting_fullpursuit/c_cython/
sort_cython.pyx:333-456), which full_binary_pursuit() calls before binary
pursuit (full_binary_pursuit.py:257-262).  The candidate pairs, their order
and the result list follow the reference; the residual of every shift pair of
every candidate (the O(T^3 S^2 C N) part) runs on the GPU
(fbp_overlap_template_residuals), bit-exact to numpy's float64 sum.
"""
import numpy as np

from . import _lib


def _shifted(template, shift, n_chans, n_samples_per_chan):
    """shift(t, s)[c][j] = t[c][j - s] inside the channel's window, else 0
    (sort_cython.pyx:340-350), as float64."""
    N = n_samples_per_chan
    t = np.asarray(template, dtype=np.float64).reshape(n_chans, N)
    out = np.zeros_like(t)
    if shift > 0:
        out[:, shift:] = t[:, :N - shift]
    elif shift < 0:
        out[:, :N + shift] = t[:, -shift:]
    else:
        out[:] = t
    return out.ravel()


def candidate_pairs(templates, test_unit, n_samples_per_chan, n_template_spikes):
    """Pairs (n1 < n2) the reference scores for `test_unit`
    (sort_cython.pyx:390-426): neither is the test unit, both have at least
    four times its spikes, and their extremes on the test template's min and
    max channels can reach 75 % of the test template's."""
    N = n_samples_per_chan
    test = templates[test_unit, :]
    min_chan = np.argmin(test) // N
    max_chan = np.argmax(test) // N
    lo, hi = slice(min_chan * N, (min_chan + 1) * N), slice(max_chan * N, (max_chan + 1) * N)
    testmin, testmax = np.amin(test[lo]), np.amax(test[hi])
    pairs = []
    T = templates.shape[0]
    for n1 in range(T):
        if n_template_spikes[n1] < 4 * n_template_spikes[test_unit] or n1 == test_unit:
            continue
        for n2 in range(n1 + 1, T):
            if n_template_spikes[n2] < 4 * n_template_spikes[test_unit] or n2 == test_unit:
                continue
            if np.amin(templates[n1, lo]) + np.amin(templates[n2, lo]) > 0.75 * testmin:
                continue
            if np.amax(templates[n1, hi]) + np.amax(templates[n2, hi]) < 0.75 * testmax:
                continue
            pairs.append((n1, n2))
    return pairs


def find_overlap_templates(templates, n_samples_per_chan, n_chans, n_pre_inds, n_post_inds, n_template_spikes,
                           device=0, return_residuals=False):
    """[test_unit, best shifted-pair sum (float64 [C*N]), [n1, n2]] for every
    template, in descending order of its sum of squares, that some candidate
    pair can be shifted to approximate; the pair and shifts minimise the
    float64 sum of squared differences (first minimum in the reference's
    (n1, n2, shift1, shift2) loop order).  templates: float32 [T][C*N] (other
    dtypes are refused, as the reference's typed buffer does).  With
    return_residuals, also the list of those smallest residuals."""
    if not isinstance(templates, np.ndarray) or templates.dtype != np.float32 or templates.ndim != 2:
        raise ValueError("templates must be a two-dimensional float32 array")
    n_template_spikes = np.asarray(n_template_spikes, dtype=np.int64)
    out, residuals = [], []
    T = templates.shape[0]
    if T < 3:
        return (out, residuals) if return_residuals else out
    n_pre, n_post = int(n_pre_inds), int(n_post_inds)
    S = n_pre + n_post + 1
    templates_SS = np.sum(templates ** 2, axis=1)
    order = np.argsort(templates_SS)[-1::-1]
    per_unit = [(u, candidate_pairs(templates, u, n_samples_per_chan, n_template_spikes)) for u in order]
    triples = np.array([(u, n1, n2) for u, pairs in per_unit for n1, n2 in pairs], dtype=np.uint32).reshape(-1, 3)
    res, combo = _lib.overlap_template_residuals(templates, n_chans, n_samples_per_chan, n_pre, n_post, triples,
                                                 device=device)
    k = 0
    for u, pairs in per_unit:
        best, best_k = np.inf, None
        for _ in pairs:
            if res[k] < best:   # strict <: the first pair holding the minimum
                best, best_k = res[k], k
            k += 1
        if best_k is None:
            continue
        n1, n2 = int(triples[best_k, 1]), int(triples[best_k, 2])
        s1, s2 = int(combo[best_k]) // S - n_pre, int(combo[best_k]) % S - n_pre
        tmpl = (_shifted(templates[n1], s1, n_chans, n_samples_per_chan)
                + _shifted(templates[n2], s2, n_chans, n_samples_per_chan))
        out.append([u, tmpl, [n1, n2]])
        residuals.append(best)
    return (out, residuals) if return_residuals else out
