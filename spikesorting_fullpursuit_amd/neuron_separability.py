"""Separability metrics around binary pursuit: the threshold inputs and the
n_spikes side effect."""
import numpy as np


def add_n_spikes(separability_metrics, neuron_labels):
    """Store the number of spikes assigned to each template row in
    separability_metrics['n_spikes'] (float array, one entry per template), as
    spikesorting_fullpursuit/neuron_separability.py:264-273 does after binary
    pursuit."""
    n_templates = separability_metrics['templates'].shape[0]
    labels = np.asarray(neuron_labels)
    counts = np.zeros(n_templates)
    if labels.size:
        valid = labels[(labels >= 0) & (labels < n_templates)].astype(np.int64)
        counts[:] = np.bincount(valid, minlength=n_templates)[:n_templates]
    separability_metrics['n_spikes'] = counts
    return None


def compute_separability_metrics(templates, channel_covariance_mats, sort_info, device=0):
    """The separability metrics binary pursuit reads its thresholds from, as
    spikesorting_fullpursuit/neuron_separability.py:88-152 builds them: the
    same dictionary keys, with the likelihood variances
    sum_c t_c^T cov_c t_c computed on the GPU (fbp_template_variances, float64)
    and the scalar statistics (thresholds, confidence bounds, contamination)
    on the host.  Reads sort_info['n_channels', 'n_samples_per_chan',
    'n_samples', 'n_threshold_crossings', 'sigma_bp_noise', 'sigma_bp_CI']
    (KeyError if missing, as the reference)."""
    from scipy.stats import norm
    from . import _lib
    n_chans = sort_info['n_channels']
    N = sort_info['n_samples_per_chan']
    sm = {'bp_n_samples_per_chan': N}
    tm = np.vstack(templates)
    n_templates = tm.shape[0]
    sm['templates'] = tm
    sm['template_SS'] = np.sum(tm ** 2, axis=1)
    sm['template_SS_by_chan'] = np.zeros((n_templates, n_chans))
    sm['channel_covariance_mats'] = channel_covariance_mats
    n_samples = sort_info['n_samples']
    crossings = np.asarray(sort_info['n_threshold_crossings'], dtype=np.float64)
    sm['channel_p_noise'] = np.array([(n_samples - crossings[c]) / n_samples for c in range(n_chans)],
                                     dtype=np.float64)
    var = _lib.template_variances(tm, channel_covariance_mats, device=device) if n_templates else np.zeros(0)
    expectation = 0.5 * sm['template_SS']
    sd = np.sqrt(var)
    sm['neuron_variances'] = var
    sm['neuron_lower_CI'] = expectation - sort_info['sigma_bp_CI'] * sd
    sm['neuron_lower_thresholds'] = -1 * expectation + sort_info['sigma_bp_noise'] * sd
    sm['peak_channel'] = (np.argmax(np.abs(tm), axis=1) // N).astype(np.int64) if n_templates else \
        np.zeros(0, dtype=np.int64)
    sm['contamination'] = np.array([norm.sf(sm['neuron_lower_thresholds'][n], -expectation[n], sd[n])
                                    for n in range(n_templates)], dtype=np.float64)
    return sm


def set_bp_threshold(separability_metrics):
    """Binary-pursuit threshold = max(lower threshold, 0, lower confidence
    bound) per template, in place (neuron_separability.py:251-261)."""
    thr = separability_metrics['neuron_lower_thresholds']
    thr[:] = np.maximum(np.maximum(thr, 0), separability_metrics['neuron_lower_CI'])
    return separability_metrics


def clip_noise_covariance(all_resid_clips, n_chans, n_samples_per_chan, n_cov_samples, rng=None, device=0):
    """Per-channel noise covariance of the residual clips
    (full_binary_pursuit.py:243-252): every clip when n_cov_samples covers
    them, else n_cov_samples rows drawn per channel with rng.randint (the
    global numpy generator by default, as the reference), then
    np.cov(rows, rowvar=False, ddof=0) of the channel's samples, on the GPU
    (fbp_channel_covariance).  Returns a list of float64 [N][N] matrices."""
    from . import _lib
    n = all_resid_clips.shape[0]
    if n_cov_samples >= n:
        cov = _lib.channel_covariance(all_resid_clips, n_chans, n_samples_per_chan, None, device=device)
    else:
        draw = (rng or np.random).randint
        rows = np.stack([draw(0, n, n_cov_samples) for _ in range(n_chans)])
        cov = _lib.channel_covariance(all_resid_clips, n_chans, n_samples_per_chan, rows, device=device)
    return [cov[c] for c in range(n_chans)]


def find_noisy_templates(separability_metrics, sort_info=None):
    """Templates too close to the noise to sort (neuron_separability.py:155-169):
    boolean [T], True where twice the lower threshold exceeds half the
    template's sum of squares.  sort_info is accepted for the reference's
    signature and not read."""
    thr = np.asarray(separability_metrics['neuron_lower_thresholds'])
    ss = np.asarray(separability_metrics['template_SS'])
    return np.asarray(2 * thr > 0.5 * ss, dtype=bool).reshape(ss.shape[0])


def noise_spike_probabilities(separability_metrics, sort_info, good, noisy, device=0):
    """For every (good, noisy) template pair: the probability that a spike of
    the noisy template lifts the good template's likelihood above its
    threshold (neuron_separability.py:192-210).  The noisy template is
    aligned to the good one within +-(N // 4 - 1) samples by the largest
    absolute flat cross-correlation, and the expected likelihood is the
    aligned dot product minus half the good template's sum of squares; both
    come from one GPU call over all pairs (fbp_template_pair_lags).  Returns
    float64 [len(good)][len(noisy)]; NaN where the good template's likelihood
    variance is 0 (the reference skips those pairs)."""
    from scipy.stats import norm
    from .consolidate import alignment_lags, best_lag
    good = np.asarray(good, dtype=np.int64)
    noisy = np.asarray(noisy, dtype=np.int64)
    out = np.full((good.size, noisy.size), np.nan)
    if good.size == 0 or noisy.size == 0:
        return out
    tm = np.asarray(separability_metrics['templates'])
    C = sort_info['n_channels']
    lo, hi = alignment_lags(tm.shape[1], sort_info['n_samples_per_chan'] // 4 - 1)
    pairs = np.stack(np.meshgrid(good, noisy, indexing='ij'), axis=-1).reshape(-1, 2)
    xc, al = _lib_pair_lags(tm, C, pairs, lo, hi, device)
    k = best_lag(xc, lo, True) - lo
    dots = al[np.arange(al.shape[0]), k].reshape(good.size, noisy.size)
    ss = np.asarray(separability_metrics['template_SS'], dtype=np.float64)[good]
    var = np.asarray(separability_metrics['neuron_variances'], dtype=np.float64)[good]
    thr = np.asarray(separability_metrics['neuron_lower_thresholds'], dtype=np.float64)[good]
    expect = dots - 0.5 * ss[:, None]
    ok = var != 0.0
    out[ok, :] = norm.sf(thr[ok, None], expect[ok, :], np.sqrt(var[ok, None]))
    return out


def _lib_pair_lags(templates, n_chans, pairs, lo, hi, device):
    from . import _lib
    return _lib.template_pair_lags(templates, n_chans, pairs, lo, hi, aligned=True, device=device)


def check_noise_templates(separability_metrics, sort_info, noisy_templates, device=0):
    """Keep the noisy templates that good templates would otherwise absorb
    (neuron_separability.py:172-224): a noisy template stays (its flag is
    cleared) when, for some template not marked noisy whose likelihood
    variance is nonzero, a spike of the noisy template exceeds that
    template's threshold with probability above norm.sf(sigma_bp_noise).
    Returns (separability_metrics, new boolean flags); the input flags are
    not modified."""
    from scipy.stats import norm
    noisy_templates = np.asarray(noisy_templates, dtype=bool)
    new_flags = np.copy(noisy_templates)
    noisy = np.flatnonzero(noisy_templates)
    good = np.flatnonzero(~noisy_templates)
    if noisy.size == 0 or good.size == 0:
        return separability_metrics, new_flags
    p = noise_spike_probabilities(separability_metrics, sort_info, good, noisy, device=device)
    limit = norm.sf(sort_info['sigma_bp_noise'], 0, 1)
    useful = np.any(np.nan_to_num(p, nan=-1.0) > limit, axis=0)
    new_flags[noisy[useful]] = False
    return separability_metrics, new_flags


_PER_TEMPLATE_ROWS = ('templates', 'template_SS_by_chan')
_PER_TEMPLATE = ('template_SS', 'neuron_variances', 'neuron_lower_thresholds', 'neuron_lower_CI', 'contamination',
                 'peak_channel')
_PER_CHANNEL = ('channel_covariance_mats', 'channel_p_noise', 'bp_n_samples_per_chan')


def delete_noise_units(separability_metrics, noisy_templates, verbose=True):
    """Drop the flagged templates from every per-template entry of
    separability_metrics, in place (neuron_separability.py:227-248): rows of
    'templates' and 'template_SS_by_chan', entries of the per-template
    vectors, items of the 'template_covariance_mats' list; per-channel entries
    are kept.  An unknown key is left as it is, with the reference's notice.
    Returns separability_metrics."""
    flags = np.asarray(noisy_templates, dtype=bool)
    keep = ~flags
    for key in list(separability_metrics.keys()):
        if key in _PER_CHANNEL:
            continue
        if key in _PER_TEMPLATE_ROWS:
            separability_metrics[key] = separability_metrics[key][keep, :]
        elif key in _PER_TEMPLATE:
            separability_metrics[key] = separability_metrics[key][keep]
        elif key == 'template_covariance_mats':
            separability_metrics[key][:] = [m for m, drop in zip(separability_metrics[key], flags) if not drop]
        elif verbose:
            print("!!! Could not find a condition for a key in separability_metrics!!!", key)
    if verbose:
        print("Removed", np.count_nonzero(flags), "templates as noise")
    return separability_metrics


def check_template_pairs(templates_1, templates_2, chan_covariance_mats, sort_info, device=0):
    """Probability of confusing each template_1 with its template_2
    (neuron_separability.py:56-85, one value per row): with d = t1 - t2,
    norm.cdf(0, E1 - E2, sqrt(sum_c d_c^T cov_c d_c)), E1 = t1.t1 / 2,
    E2 = t1.t2 - t2.t2 / 2, and 1 where that variance is not positive.  The
    dot products are numpy's on the host, as in the reference; the variances
    come from the GPU (fbp_template_variances, float64)."""
    from scipy.stats import norm
    from . import _lib
    t1 = np.asarray(templates_1)
    t2 = np.asarray(templates_2)
    if t1.ndim == 1:
        t1, t2 = t1[None, :], t2[None, :]
    n = t1.shape[0]
    if n == 0:
        return np.zeros(0)
    # E1 - E2 in the templates' own arithmetic, as the reference evaluates it
    ediff = np.array([0.5 * np.dot(t1[k], t1[k]) - (np.dot(t1[k], t2[k]) - 0.5 * np.dot(t2[k], t2[k]))
                      for k in range(n)], dtype=np.float64)
    var = _lib.template_variances(np.asarray(t1 - t2, dtype=np.float64), chan_covariance_mats, device=device)
    p = np.ones(n)
    pos = var > 0
    p[pos] = norm.cdf(0, ediff[pos], np.sqrt(var[pos]))
    return p


def check_template_pair(template_1, template_2, chan_covariance_mats, sort_info, device=0):
    """One pair of check_template_pairs (the reference's scalar form)."""
    return float(check_template_pairs(template_1, template_2, chan_covariance_mats, sort_info, device=device)[0])
