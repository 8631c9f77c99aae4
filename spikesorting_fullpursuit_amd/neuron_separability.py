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
