"""Seeded synthetic recordings for the binary-pursuit workloads.

Voltage is i.i.d. N(0,1) float32 noise, channel-major [C][K].  Templates are
biphasic difference-of-Gaussian (DoG) waveforms, one per unit, scaled per
channel by a spatial gain and flattened to the reference's [T][C*N] layout
(spikesorting_fullpursuit/kernels/binary_pursuit.cl:11-29).  Spike trains are
Poisson with an absolute refractory period; every spike adds its template at
onset = event - clip_init_samples, overlaps included.

Thresholds use the closed form of the reference's threshold for white noise
(neuron_separability.py:126-142 with channel covariance = I, so the
likelihood variance equals template_SS, then set_bp_threshold :251-261 with
sigma_bp_CI = inf): thr = max(-SS/2 + sigma_bp_noise*sqrt(SS), 0).
"""
import numpy as np

from .segment import time_window_to_samples


def dog_waveform(rng, n, peak):
    """Biphasic DoG of length n: a sharp trough at the centre followed by a
    broader positive lobe, scaled so that max |w| == peak."""
    x = np.arange(n, dtype=np.float64)
    c = n // 2
    s1 = rng.uniform(1.5, 3.0) * n / 61.0
    s2 = rng.uniform(3.0, 6.0) * n / 61.0
    d = rng.uniform(4.0, 10.0) * n / 61.0
    w = -np.exp(-((x - c) ** 2) / (2 * s1 ** 2)) + 0.45 * np.exp(-((x - c - d) ** 2) / (2 * s2 ** 2))
    return w * (peak / np.max(np.abs(w)))


def poisson_train(rng, rate_hz, n_samples, fs, refractory_samples):
    """Onset-free event times (sample indices) of a Poisson process with an
    absolute refractory period."""
    if rate_hz <= 0:
        return np.zeros(0, dtype=np.int64)
    mean_isi = fs / rate_hz
    expected = int(n_samples / mean_isi * 1.3) + 16
    isi = rng.exponential(mean_isi, size=expected) + refractory_samples
    t = np.cumsum(isi)
    while t[-1] < n_samples:
        more = np.cumsum(rng.exponential(mean_isi, size=expected) + refractory_samples) + t[-1]
        t = np.concatenate([t, more])
    return np.floor(t[t < n_samples]).astype(np.int64)


def iid_thresholds(template_ss, sigma_bp_noise=3.719):
    ss = np.asarray(template_ss, dtype=np.float64)
    return np.maximum(-0.5 * ss + sigma_bp_noise * np.sqrt(ss), 0.0)


def make_recording(seed, n_chans, fs, duration_s, n_templates, n_samples_per_chan=61,
                   peak_range=(6.0, 15.0), rate_range=(5.0, 40.0), gains="ladder",
                   spatial_sigma=2.0, support=None, overlap_frac=0.0, refractory_ms=1.5,
                   sigma_bp_noise=3.719, noise=True, dtype=np.float32):
    """Build one seeded recording.

    gains: "ladder" -> 1.0/0.6/0.3/0.1 (then 0) by channel distance from a random
    home channel; "gauss" -> exp(-d^2/(2 spatial_sigma^2)).  support: if set,
    each template is nonzero only on `support` contiguous channels around home.
    overlap_frac: fraction of spikes re-timed to within +-1 ms of a spike of
    another unit.
    Returns dict(voltage, templates [T][C*N], events (list per unit, event
    sample), sort_info, separability_metrics, clip_init).
    """
    rng = np.random.default_rng(seed)
    N = int(n_samples_per_chan)
    K = int(round(fs * duration_s))
    half = (N - 1) // 2
    clip_width = [-half / fs, (N - half - 1) / fs]
    chan_win, _ = time_window_to_samples(clip_width, fs)
    if chan_win[1] - chan_win[0] != N:
        # rounding guard: pick the widths straight from the sample counts
        clip_width = [-(half + 0.25) / fs, (N - half - 1 + 0.25) / fs]
        chan_win, _ = time_window_to_samples(clip_width, fs)
    assert chan_win[1] - chan_win[0] == N, (chan_win, N)
    clip_init = abs(chan_win[0])
    C, T = int(n_chans), int(n_templates)

    templates = np.zeros((T, C, N), dtype=np.float64)
    for t in range(T):
        home = rng.integers(0, C)
        w = dog_waveform(rng, N, rng.uniform(*peak_range))
        if support is not None:
            lo = int(np.clip(home - support // 2, 0, max(C - support, 0)))
            chans = range(lo, min(lo + support, C))
        else:
            chans = range(C)
        for c in chans:
            dist = abs(c - home)
            if gains == "ladder":
                g = [1.0, 0.6, 0.3, 0.1][dist] if dist < 4 else 0.0
            else:
                g = float(np.exp(-dist ** 2 / (2 * spatial_sigma ** 2)))
            templates[t, c] = g * w
    templates = templates.reshape(T, C * N).astype(np.float32)

    refr = int(round(refractory_ms * 1e-3 * fs))
    events = []
    for t in range(T):
        events.append(poisson_train(rng, rng.uniform(*rate_range), K, fs, refr))
    if overlap_frac > 0 and T > 1:
        win = int(round(1e-3 * fs))
        for t in range(T):
            ev = events[t]
            n_move = int(round(overlap_frac * ev.size))
            if n_move == 0:
                continue
            sel = rng.choice(ev.size, size=n_move, replace=False)
            for k in sel:
                other = rng.integers(0, T - 1)
                other = other + 1 if other >= t else other
                if events[other].size == 0:
                    continue
                anchor = events[other][rng.integers(0, events[other].size)]
                ev[k] = np.clip(anchor + rng.integers(-win, win + 1), 0, K - 1)
            events[t] = np.sort(ev)

    if noise:
        voltage = rng.standard_normal((C, K), dtype=np.float32)
    else:
        voltage = np.zeros((C, K), dtype=np.float32)
    tm3 = templates.reshape(T, C, N)
    onsets, labels = [], []
    for t in range(T):
        on = events[t] - clip_init
        on = on[(on >= 0) & (on + N <= K)]
        events[t] = on + clip_init
        onsets.append(on)
        labels.append(np.full(on.size, t, dtype=np.int64))
    if onsets:
        on_all = np.concatenate(onsets)
        lab_all = np.concatenate(labels)
        # Sum every spike's contribution per channel in float64 (bincount is
        # deterministic), then add it to the noise once.  Only the spikes of
        # templates that are nonzero on the channel are gathered.
        nz = np.any(tm3 != 0.0, axis=2)   # [T][C]
        for c in range(C):
            keep = nz[lab_all, c]
            if not np.any(keep):
                continue
            vals = tm3[lab_all[keep], c, :].astype(np.float64)
            idx = (on_all[keep, None] + np.arange(N)[None, :]).ravel()
            add = np.bincount(idx, weights=vals.ravel(), minlength=K)
            voltage[c] += add.astype(np.float32)
    voltage = voltage.astype(dtype, copy=False)

    ss = np.sum(templates.astype(np.float64) ** 2, axis=1)
    sm = {
        'templates': templates,
        'template_SS': ss,
        'template_SS_by_chan': np.sum(templates.astype(np.float64).reshape(T, C, N) ** 2, axis=2),
        'neuron_lower_thresholds': iid_thresholds(ss, sigma_bp_noise),
    }
    sort_info = {
        'sampling_rate': fs,
        'clip_width': clip_width,
        'n_samples_per_chan': N,
        'n_channels': C,
        'do_overlap_recheck': True,
        'get_adjusted_clips': False,
    }
    return dict(voltage=voltage, templates=templates, events=events, sort_info=sort_info,
                separability_metrics=sm, clip_init=clip_init)


def config(name, seed=0, duration_s=None):
    """Recording for a named workload.

    "4ch-60s": 4 ch, 40 kHz, 60 s, 5 DoG templates x 61 samples, 6-15 sigma
        peaks, channel gains 1.0/0.6/0.3/0.1, rates 5-40 Hz, 1.5 ms refractory.
    "16ch-300s": 16 ch, 40 kHz, 300 s, 20 templates x 61 x 16, 5-20 sigma,
        Gaussian falloff sigma = 2 channels, rates 0.5-80 Hz, 10% of spikes
        moved to within +-1 ms of another unit.
    "128ch-600s": 128 ch, 30 kHz, 600 s, 200 templates x 61 on 8 channels,
        6-25 sigma, rates 0.5-100 Hz.
    """
    if name == "4ch-60s":
        return make_recording(seed, 4, 40000, duration_s or 60.0, 5, 61, (6.0, 15.0), (5.0, 40.0),
                              gains="ladder")
    if name == "16ch-300s":
        return make_recording(seed, 16, 40000, duration_s or 300.0, 20, 61, (5.0, 20.0), (0.5, 80.0),
                              gains="gauss", spatial_sigma=2.0, overlap_frac=0.10)
    if name == "128ch-600s":
        return make_recording(seed, 128, 30000, duration_s or 600.0, 200, 61, (6.0, 25.0), (0.5, 100.0),
                              gains="gauss", spatial_sigma=2.0, support=8)
    raise ValueError("unknown workload %r" % name)


def serving_mix(seed=0, n_requests=64, rate_hz=16.0, duration_range=(5.0, 120.0), channel_range=(4, 32),
                template_range=(3, 40), length_range=(41, 81), peak_range=(4.0, 20.0), rates=(1.0, 60.0)):
    """Request specs of the served mix: Poisson arrivals at rate_hz; per
    request a length uniform in duration_range at 30 or 40 kHz, channels and
    templates uniform in their ranges, an odd template length in
    length_range, DoG peaks in peak_range (sigma) and Poisson trains at rates
    (Hz).  Returns a list of dicts (arrival_s, seed and make_recording
    arguments); build the data with recording_for(spec)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    specs = []
    for k in range(n_requests):
        t += rng.exponential(1.0 / rate_hz)
        n = int(rng.integers(length_range[0] // 2, length_range[1] // 2 + 1)) * 2 + 1
        n = min(max(n, length_range[0] | 1), length_range[1])
        specs.append(dict(arrival_s=t, seed=int(rng.integers(0, 2 ** 31)),
                          n_chans=int(rng.integers(channel_range[0], channel_range[1] + 1)),
                          fs=int(rng.choice([30000, 40000])),
                          duration_s=float(rng.uniform(*duration_range)),
                          n_templates=int(rng.integers(template_range[0], template_range[1] + 1)),
                          n_samples_per_chan=n, peak_range=peak_range, rate_range=rates))
    return specs


def recording_for(spec):
    return make_recording(spec['seed'], spec['n_chans'], spec['fs'], spec['duration_s'], spec['n_templates'],
                          spec['n_samples_per_chan'], spec['peak_range'], spec['rate_range'], gains="gauss",
                          spatial_sigma=2.0)
