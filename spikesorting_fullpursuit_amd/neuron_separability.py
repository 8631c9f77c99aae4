"""The one separability-metrics side effect of binary pursuit."""
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
