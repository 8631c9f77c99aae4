"""ctypes binding of libfbp.so (include/fbp.h).

The library is loaded lazily, on first use, so that importing the package
never initialises HIP (the reference imports pyopencl inside binary_pursuit()
for the same reason: a parent process that forks workers must not hold a GPU
context, spikesorting_fullpursuit/parallel/binary_pursuit_parallel.py:166-169).
There is no fallback: if the library is missing or the GPU is absent the
calls raise.
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfbp.so")

FBP_OK = 0
FBP_EINVAL = -1
FBP_ENOMEM = -2
FBP_ENODEV = -3
FBP_EHIP = -4
FBP_EPASSES = -5
FBP_ESTATE = -6
FBP_ETIMEDOUT = -7

JOB_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "cancelled"}

PRECISIONS = {"fp32": 0, "bf16": 1, "fp16": 2, "fp8": 3}
OVERLAP_MODES = {"direct": 0, "screened": 1}

# Names every build of libfbp.so must export (checked by the CPU test suite).
EXPORTS = (
    "fbp_version", "fbp_last_error", "fbp_device_count", "fbp_device_memory",
    "fbp_segment_create", "fbp_segment_bytes", "fbp_segment_destroy", "fbp_segment_set_templates",
    "fbp_segment_set_voltage", "fbp_segment_set_voltage_device", "fbp_segment_run",
    "fbp_segment_num_spikes", "fbp_segment_get_spikes", "fbp_segment_get_clips",
    "fbp_segment_get_residual", "fbp_segment_get_likelihood", "fbp_segment_get_stats",
    "fbp_segment_voltage_device", "fbp_get_clips", "fbp_unit_mean_clips", "fbp_overlap_template_residuals",
    "fbp_clip_template_sse", "fbp_get_clips_host", "fbp_channel_covariance", "fbp_template_variances",
    "fbp_segment_get_window_list", "fbp_channel_covariance_noalloc", "fbp_unit_mean_clips_f64",
    "fbp_segment_device_bytes", "fbp_server_create", "fbp_server_destroy", "fbp_server_submit", "fbp_server_poll",
    "fbp_server_wait", "fbp_server_wait_any", "fbp_server_cancel", "fbp_server_segment", "fbp_server_release",
    "fbp_server_get_stats", "fbp_template_pair_lags",
)


class FbpConfig(ctypes.Structure):
    _fields_ = [
        ("n_chans", ctypes.c_uint32),
        ("n_samples", ctypes.c_uint32),
        ("n_templates", ctypes.c_uint32),
        ("template_len", ctypes.c_uint32),
        ("max_shift", ctypes.c_uint32),
        ("do_overlap", ctypes.c_uint32),
        ("reduce_wg", ctypes.c_uint32),
        ("precision", ctypes.c_uint32),
        ("overlap_mode", ctypes.c_uint32),
        ("max_passes", ctypes.c_uint32),
        ("debug_flags", ctypes.c_uint32),
    ]


class FbpStats(ctypes.Structure):
    _fields_ = [
        ("n_passes", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("n_spikes", ctypes.c_uint64),
        ("windows_evaluated", ctypes.c_uint64),
        ("overlap_windows", ctypes.c_uint64),
        ("rescored_positions", ctypes.c_uint64),
        ("screened_combos", ctypes.c_uint64),
        ("ms_total", ctypes.c_double),
        ("ms_correlation", ctypes.c_double),
        ("ms_overlap", ctypes.c_double),
        ("corr_launches", ctypes.c_uint64),
        ("overlap_launches", ctypes.c_uint64),
        ("corr_positions", ctypes.c_uint64),
        ("overlap_combos", ctypes.c_uint64),
        ("corr_pairs", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class FbpServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("failed", ctypes.c_uint64),
        ("cancelled", ctypes.c_uint64),
        ("queued", ctypes.c_uint32),
        ("running", ctypes.c_uint32),
        ("peak_running", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("held_bytes", ctypes.c_uint64),
        ("peak_held_bytes", ctypes.c_uint64),
        ("budget_bytes", ctypes.c_uint64),
        ("pass_steps", ctypes.c_uint64),
        ("driver_busy_s", ctypes.c_double),
        ("uptime_s", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class FbpJobInfo(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_int32),
        ("rc", ctypes.c_int32),
        ("reserve_bytes", ctypes.c_uint64),
        ("bytes", ctypes.c_uint64),
        ("passes", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("t_submit", ctypes.c_double),
        ("t_admit", ctypes.c_double),
        ("t_start", ctypes.c_double),
        ("t_finish", ctypes.c_double),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}
        d["state"] = JOB_STATES.get(self.state, str(self.state))
        return d


class FbpError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


_lib = None
_lock = threading.Lock()


def _declare(lib):
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    f32p = P(ctypes.c_float)
    u32p = P(ctypes.c_uint32)
    u64p = P(ctypes.c_uint64)
    lib.fbp_version.restype = ctypes.c_int
    lib.fbp_last_error.restype = ctypes.c_char_p
    lib.fbp_device_count.argtypes = [P(ctypes.c_int)]
    lib.fbp_device_memory.argtypes = [ctypes.c_int, u64p, u64p]
    lib.fbp_segment_create.argtypes = [ctypes.c_int, P(FbpConfig), P(vp)]
    lib.fbp_segment_bytes.argtypes = [P(FbpConfig), u64p]
    lib.fbp_segment_destroy.argtypes = [vp]
    lib.fbp_segment_destroy.restype = None
    lib.fbp_segment_set_templates.argtypes = [vp, f32p, f32p, f32p]
    lib.fbp_segment_set_voltage.argtypes = [vp, f32p, ctypes.c_uint64]
    lib.fbp_segment_set_voltage_device.argtypes = [vp, vp, ctypes.c_uint64]
    lib.fbp_segment_run.argtypes = [vp, vp]
    lib.fbp_segment_num_spikes.argtypes = [vp, u64p]
    lib.fbp_segment_get_spikes.argtypes = [vp, u32p, u32p]
    lib.fbp_segment_get_clips.argtypes = [vp, f32p]
    lib.fbp_segment_get_residual.argtypes = [vp, f32p]
    lib.fbp_segment_get_likelihood.argtypes = [vp, f32p]
    lib.fbp_segment_get_stats.argtypes = [vp, P(FbpStats)]
    lib.fbp_segment_get_window_list.argtypes = [vp, u32p, ctypes.c_uint64, u64p]
    lib.fbp_segment_voltage_device.argtypes = [vp, P(vp), u64p]
    i64p = P(ctypes.c_int64)
    u8p = P(ctypes.c_uint8)
    lib.fbp_get_clips.argtypes = [ctypes.c_int, vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, i64p,
                                  ctypes.c_uint64, u32p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32,
                                  ctypes.c_int, f32p, u8p, u64p]
    lib.fbp_get_clips_host.argtypes = [ctypes.c_int, f32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64,
                                       i64p, ctypes.c_uint64, u32p, ctypes.c_uint32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_int, f32p, u8p, u64p]
    lib.fbp_unit_mean_clips.argtypes = [ctypes.c_int, f32p, ctypes.c_uint64, ctypes.c_uint32, u32p,
                                        ctypes.c_uint32, f32p, u64p]
    lib.fbp_unit_mean_clips_f64.argtypes = [ctypes.c_int, P(ctypes.c_double), ctypes.c_uint64, ctypes.c_uint32,
                                            u32p, ctypes.c_uint32, P(ctypes.c_double), u64p]
    lib.fbp_clip_template_sse.argtypes = [ctypes.c_int, f32p, ctypes.c_uint64, ctypes.c_uint32, u32p, f32p,
                                          ctypes.c_uint32, f32p]
    lib.fbp_overlap_template_residuals.argtypes = [ctypes.c_int, f32p, ctypes.c_uint32, ctypes.c_uint32,
                                                   ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32, u32p,
                                                   ctypes.c_uint64, P(ctypes.c_double), u32p]
    f64p = P(ctypes.c_double)
    lib.fbp_channel_covariance.argtypes = [ctypes.c_int, f32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                           u64p, ctypes.c_uint64, f64p]
    lib.fbp_channel_covariance_noalloc.argtypes = [ctypes.c_int, vp, ctypes.c_uint64, ctypes.c_uint32,
                                                   ctypes.c_uint32, vp, ctypes.c_uint64, vp, vp]
    lib.fbp_template_variances.argtypes = [ctypes.c_int, f64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                           f64p, f64p]
    lib.fbp_template_pair_lags.argtypes = [ctypes.c_int, f64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                           u32p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, f64p, f64p]
    lib.fbp_segment_device_bytes.argtypes = [vp, u64p]
    lib.fbp_server_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, P(vp)]
    lib.fbp_server_destroy.argtypes = [vp]
    lib.fbp_server_destroy.restype = None
    lib.fbp_server_submit.argtypes = [vp, P(FbpConfig), f32p, f32p, f32p, f32p, ctypes.c_uint64, ctypes.c_uint64,
                                      u64p]
    lib.fbp_server_poll.argtypes = [vp, ctypes.c_uint64, P(FbpJobInfo)]
    lib.fbp_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_int64, P(FbpJobInfo)]
    lib.fbp_server_wait_any.argtypes = [vp, ctypes.c_int64, u64p, P(FbpJobInfo)]
    lib.fbp_server_cancel.argtypes = [vp, ctypes.c_uint64]
    lib.fbp_server_segment.argtypes = [vp, ctypes.c_uint64, P(vp)]
    lib.fbp_server_release.argtypes = [vp, ctypes.c_uint64]
    lib.fbp_server_get_stats.argtypes = [vp, P(FbpServerStats)]
    for name in EXPORTS:
        if name not in ("fbp_last_error", "fbp_segment_destroy", "fbp_server_destroy"):
            getattr(lib, name).restype = ctypes.c_int


def load():
    """Load libfbp.so (raises if it is not built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError("libfbp.so is not built (%s); run "
                                   "`python -m spikesorting_fullpursuit_amd.build`" % LIB_PATH)
            lib = ctypes.CDLL(LIB_PATH)
            _declare(lib)
            _lib = lib
    return _lib


def check(rc, what=""):
    if rc == FBP_OK:
        return
    msg = load().fbp_last_error().decode(errors="replace")
    if rc == FBP_ENODEV:
        raise RuntimeError("No GPU available" + (": " + msg if msg else ""))
    if rc == FBP_EINVAL:
        raise ValueError(msg or what)
    if rc == FBP_ENOMEM:
        raise MemoryError(msg or what)
    if rc == FBP_ETIMEDOUT:
        raise TimeoutError(msg or what)
    raise FbpError(rc, "%s failed (%d): %s" % (what, rc, msg))


def _f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _u32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _f64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def device_count():
    lib = load()
    n = ctypes.c_int(0)
    rc = lib.fbp_device_count(ctypes.byref(n))
    check(rc, "fbp_device_count")
    return n.value


def device_memory(device=0):
    lib = load()
    tot, fr = ctypes.c_uint64(0), ctypes.c_uint64(0)
    check(lib.fbp_device_memory(device, ctypes.byref(tot), ctypes.byref(fr)), "fbp_device_memory")
    return tot.value, fr.value


def segment_bytes(n_chans, n_samples, n_templates, template_len, max_shift, do_overlap=True, reduce_wg=256,
                  precision="fp32", overlap_mode="direct"):
    """Device bytes a Segment of this shape allocates (raises ValueError if
    the shape is not supported).  Needs no GPU."""
    lib = load()
    cfg = FbpConfig(n_chans, n_samples, n_templates, template_len, max_shift, 1 if do_overlap else 0, reduce_wg,
                    PRECISIONS[precision], OVERLAP_MODES[overlap_mode], 0, 0)
    b = ctypes.c_uint64(0)
    check(lib.fbp_segment_bytes(ctypes.byref(cfg), ctypes.byref(b)), "fbp_segment_bytes")
    return b.value


class Segment:
    """Device-resident state of one chunk (owner of an fbp_segment*)."""

    def __init__(self, n_chans, n_samples, n_templates, template_len, max_shift, do_overlap=True,
                 reduce_wg=256, precision="fp32", overlap_mode="direct", max_passes=0, device=0,
                 debug_flags=0):
        self._lib = load()
        self.cfg = FbpConfig(n_chans, n_samples, n_templates, template_len, max_shift,
                             1 if do_overlap else 0, reduce_wg, PRECISIONS[precision],
                             OVERLAP_MODES[overlap_mode], max_passes, debug_flags)
        self._h = ctypes.c_void_p(None)
        check(self._lib.fbp_segment_create(device, ctypes.byref(self.cfg), ctypes.byref(self._h)),
              "fbp_segment_create")
        self.C, self.K, self.T, self.N = n_chans, n_samples, n_templates, template_len
        self.device = device

    @classmethod
    def _view(cls, handle, cfg, device):
        """A non-owning Segment over a handle owned by someone else (a
        server's finished job): close() does not destroy it."""
        self = cls.__new__(cls)
        self._lib = load()
        self.cfg = cfg
        self._h = handle
        self._owned = False
        self.C, self.K, self.T, self.N = cfg.n_chans, cfg.n_samples, cfg.n_templates, cfg.template_len
        self.device = device
        return self

    def close(self):
        if self._h is not None and self._h.value and getattr(self, "_owned", True):
            self._lib.fbp_segment_destroy(self._h)
        self._h = None

    def device_bytes(self):
        """Device bytes the segment holds now (buffers, spike list, clips)."""
        b = ctypes.c_uint64(0)
        check(self._lib.fbp_segment_device_bytes(self._h, ctypes.byref(b)), "fbp_segment_device_bytes")
        return b.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_templates(self, templates, bias, thresholds):
        tm = np.ascontiguousarray(templates, dtype=np.float32)
        b = np.ascontiguousarray(bias, dtype=np.float32)
        th = np.ascontiguousarray(thresholds, dtype=np.float32)
        assert tm.size == self.T * self.C * self.N and b.size == self.T and th.size == self.T
        check(self._lib.fbp_segment_set_templates(self._h, _f32p(tm), _f32p(b), _f32p(th)),
              "fbp_segment_set_templates")

    def set_voltage(self, voltage):
        """voltage: float32 [C][>=K] host array whose rows are contiguous."""
        v = voltage
        if v.dtype != np.float32 or v.strides[1] != 4 or v.shape[0] != self.C or v.shape[1] < self.K:
            v = np.ascontiguousarray(v[:, :self.K], dtype=np.float32)
        stride = v.strides[0] // 4
        check(self._lib.fbp_segment_set_voltage(self._h, _f32p(v), stride), "fbp_segment_set_voltage")

    def set_voltage_device(self, ptr, row_stride):
        check(self._lib.fbp_segment_set_voltage_device(self._h, ctypes.c_void_p(ptr), row_stride),
              "fbp_segment_set_voltage_device")

    def run(self, stream=None, allow_cap=False):
        """Run the pass loop.  With allow_cap, reaching max_passes returns
        False (results of the passes run stay readable) instead of raising."""
        rc = self._lib.fbp_segment_run(self._h, ctypes.c_void_p(stream) if stream else None)
        if allow_cap and rc == FBP_EPASSES:
            return False
        check(rc, "fbp_segment_run")
        return True

    def num_spikes(self):
        n = ctypes.c_uint64(0)
        check(self._lib.fbp_segment_num_spikes(self._h, ctypes.byref(n)), "fbp_segment_num_spikes")
        return n.value

    def spikes(self):
        n = self.num_spikes()
        idx = np.zeros(n, dtype=np.uint32)
        lab = np.zeros(n, dtype=np.uint32)
        if n:
            check(self._lib.fbp_segment_get_spikes(self._h, _u32p(idx), _u32p(lab)), "fbp_segment_get_spikes")
        return idx, lab

    def clips(self):
        n = self.num_spikes()
        out = np.zeros((n, self.C * self.N), dtype=np.float32)
        if n:
            check(self._lib.fbp_segment_get_clips(self._h, _f32p(out)), "fbp_segment_get_clips")
        return out

    def residual(self):
        out = np.zeros((self.C, self.K), dtype=np.float32)
        check(self._lib.fbp_segment_get_residual(self._h, _f32p(out)), "fbp_segment_get_residual")
        return out

    def likelihood(self):
        out = np.zeros((self.T, self.K), dtype=np.float32)
        check(self._lib.fbp_segment_get_likelihood(self._h, _f32p(out)), "fbp_segment_get_likelihood")
        return out

    def window_list(self):
        """Windows the next pass would evaluate after a run stopped by
        max_passes (empty after a run that converged)."""
        n = ctypes.c_uint64(0)
        W = -(-self.K // self.N)
        out = np.zeros(W, dtype=np.uint32)
        check(self._lib.fbp_segment_get_window_list(self._h, _u32p(out), W, ctypes.byref(n)),
              "fbp_segment_get_window_list")
        return out[:n.value].copy()

    def voltage_device(self):
        """(device address, row stride) of the segment's input voltage."""
        ptr, stride = ctypes.c_void_p(None), ctypes.c_uint64(0)
        check(self._lib.fbp_segment_voltage_device(self._h, ctypes.byref(ptr), ctypes.byref(stride)),
              "fbp_segment_voltage_device")
        return ptr.value, stride.value

    def stats(self):
        st = FbpStats()
        check(self._lib.fbp_segment_get_stats(self._h, ctypes.byref(st)), "fbp_segment_get_stats")
        return st.as_dict()


def get_clips_device(d_voltage, row_stride, n_chans, n_samples, event_indices, channels, w0, w1,
                     check_valid=True, device=0):
    """fbp_get_clips on device-resident float32 voltage: (clips [n_valid][len(channels)*(w1-w0)],
    valid mask [n_events])."""
    lib = load()
    idx = np.ascontiguousarray(event_indices, dtype=np.int64)
    ch = np.ascontiguousarray(channels, dtype=np.uint32)
    width = ch.size * (w1 - w0)
    valid = np.zeros(idx.size, dtype=np.uint8)
    n_valid = ctypes.c_uint64(0)
    # the valid events are a contiguous run of at most n_events
    out = np.empty((idx.size, width), dtype=np.float32)
    check(lib.fbp_get_clips(device, ctypes.c_void_p(d_voltage), row_stride, n_chans, n_samples,
                            idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), idx.size, _u32p(ch), ch.size,
                            w0, w1, 1 if check_valid else 0, _f32p(out),
                            valid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(n_valid)),
          "fbp_get_clips")
    return out[:n_valid.value], valid.astype(bool)


def get_clips_host(voltage, n_samples, event_indices, channels, w0, w1, check_valid=True, device=0):
    """fbp_get_clips_host on a host float32 voltage [C][row length]."""
    lib = load()
    v = np.ascontiguousarray(voltage, dtype=np.float32)
    idx = np.ascontiguousarray(event_indices, dtype=np.int64)
    ch = np.ascontiguousarray(channels, dtype=np.uint32)
    valid = np.zeros(idx.size, dtype=np.uint8)
    n_valid = ctypes.c_uint64(0)
    out = np.empty((idx.size, ch.size * (w1 - w0)), dtype=np.float32)
    check(lib.fbp_get_clips_host(device, _f32p(v), v.shape[1], v.shape[0], n_samples,
                                 idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), idx.size, _u32p(ch), ch.size,
                                 w0, w1, 1 if check_valid else 0, _f32p(out),
                                 valid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(n_valid)),
          "fbp_get_clips_host")
    return out[:n_valid.value], valid.astype(bool)


def unit_mean_clips(clips, labels, n_units, device=0):
    """Mean clip of every unit, as np.mean(clips[labels == u], axis=0) gives it:
    float32 clips are averaged in float32 (fbp_unit_mean_clips), clips of any
    other real dtype in float64 (fbp_unit_mean_clips_f64), numpy's result type
    for integer and float64 input.  Returns (means [n_units][width], counts)."""
    lib = load()
    c = np.asarray(clips)
    if c.ndim != 2:
        raise ValueError("clips must be two-dimensional")
    lab = np.ascontiguousarray(labels, dtype=np.uint32)
    if lab.size != c.shape[0]:
        raise ValueError("one label per clip is required")
    counts = np.zeros(n_units, dtype=np.uint64)
    cp = counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    if c.dtype == np.float32:
        c = np.ascontiguousarray(c)
        means = np.empty((n_units, c.shape[1]), dtype=np.float32)
        check(lib.fbp_unit_mean_clips(device, _f32p(c), c.shape[0], c.shape[1], _u32p(lab), n_units, _f32p(means),
                                      cp), "fbp_unit_mean_clips")
    else:
        c = np.ascontiguousarray(c, dtype=np.float64)
        means = np.empty((n_units, c.shape[1]), dtype=np.float64)
        check(lib.fbp_unit_mean_clips_f64(device, _f64p(c), c.shape[0], c.shape[1], _u32p(lab), n_units,
                                          _f64p(means), cp), "fbp_unit_mean_clips_f64")
    return means, counts


def overlap_template_residuals(templates, n_chans, n_samples_per_chan, n_pre, n_post, triples, device=0):
    """fbp_overlap_template_residuals: (best residual float64 [n], best combination uint32 [n])."""
    lib = load()
    tm = np.ascontiguousarray(templates, dtype=np.float32)
    tri = np.ascontiguousarray(triples, dtype=np.uint32).reshape(-1, 3)
    res = np.zeros(tri.shape[0], dtype=np.float64)
    combo = np.zeros(tri.shape[0], dtype=np.uint32)
    check(lib.fbp_overlap_template_residuals(device, _f32p(tm), tm.shape[0], n_chans, n_samples_per_chan, n_pre,
                                             n_post, _u32p(tri), tri.shape[0],
                                             res.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _u32p(combo)),
          "fbp_overlap_template_residuals")
    return res, combo


def clip_template_sse(clips, template_index, templates, device=0):
    """fbp_clip_template_sse: float32 [n], np.sum((clips[k] - templates[template_index[k]]) ** 2)."""
    lib = load()
    c = np.ascontiguousarray(clips, dtype=np.float32)
    tm = np.ascontiguousarray(templates, dtype=np.float32).reshape(-1, c.shape[1] if c.ndim == 2 else 0)
    idx = np.ascontiguousarray(template_index, dtype=np.uint32)
    if c.ndim != 2 or idx.size != c.shape[0]:
        raise ValueError("clips must be [n][width] with one template index per clip")
    out = np.zeros(c.shape[0], dtype=np.float32)
    check(lib.fbp_clip_template_sse(device, _f32p(c), c.shape[0], c.shape[1], _u32p(idx), _f32p(tm), tm.shape[0],
                                    _f32p(out)), "fbp_clip_template_sse")
    return out


def channel_covariance(clips, n_chans, n_samples_per_chan, rows=None, device=0):
    """fbp_channel_covariance: float64 [n_chans][N][N], per channel
    np.cov(clips[rows[c], c*N:(c+1)*N], rowvar=False, ddof=0); rows is
    [n_chans][n_rows] (row indices drawn per channel) or None for every clip."""
    lib = load()
    c = np.ascontiguousarray(clips, dtype=np.float32)
    N = int(n_samples_per_chan)
    if c.ndim != 2 or c.shape[1] != n_chans * N:
        raise ValueError("clips must be [n_clips][n_chans * n_samples_per_chan]")
    out = np.zeros((n_chans, N, N), dtype=np.float64)
    if rows is None:
        r, n_rows, rp = None, c.shape[0], None
    else:
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        if r.ndim != 2 or r.shape[0] != n_chans:
            raise ValueError("rows must be [n_chans][n_rows]")
        n_rows = r.shape[1]
        rp = r.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    check(lib.fbp_channel_covariance(device, _f32p(c), c.shape[0], n_chans, N, rp, n_rows, _f64p(out)),
          "fbp_channel_covariance")
    return out


def channel_covariance_noalloc(d_clips, n_clips, n_chans, n_samples_per_chan, d_rows, n_rows, d_cov, stream=None,
                               device=0):
    """fbp_channel_covariance_noalloc on device addresses (ints): enqueues the
    covariance on `stream` and returns without waiting; allocates nothing."""
    lib = load()
    check(lib.fbp_channel_covariance_noalloc(device, ctypes.c_void_p(d_clips), n_clips, n_chans, n_samples_per_chan,
                                             ctypes.c_void_p(d_rows) if d_rows else None, n_rows,
                                             ctypes.c_void_p(d_cov), ctypes.c_void_p(stream) if stream else None),
          "fbp_channel_covariance_noalloc")


def template_variances(templates, channel_covariance_mats, device=0):
    """fbp_template_variances: float64 [T], sum over channels of t_c^T cov_c t_c."""
    lib = load()
    cov = np.ascontiguousarray(np.stack([np.asarray(m, dtype=np.float64) for m in channel_covariance_mats]))
    n_chans, N = cov.shape[0], cov.shape[1]
    if cov.shape != (n_chans, N, N):
        raise ValueError("covariance matrices must be square, one per channel")
    tm = np.ascontiguousarray(templates, dtype=np.float64).reshape(-1, n_chans * N)
    out = np.zeros(tm.shape[0], dtype=np.float64)
    check(lib.fbp_template_variances(device, _f64p(tm), tm.shape[0], n_chans, N, _f64p(cov), _f64p(out)),
          "fbp_template_variances")
    return out


def template_pair_lags(templates, n_chans, pairs, lag_lo, lag_hi, aligned=True, device=0):
    """fbp_template_pair_lags: for each row (a, b) of `pairs`, the flat
    cross-correlation sum_i a[i + lag] b[i] and (with aligned) the
    channel-aligned dot product at every lag of [lag_lo, lag_hi], float64
    [n_pairs][n_lags] each (aligned is None without it)."""
    lib = load()
    tm = np.ascontiguousarray(templates, dtype=np.float64)
    tm = tm.reshape(tm.shape[0] if tm.ndim > 1 else 1, -1)
    if tm.shape[1] % n_chans:
        raise ValueError("template length %d is not a multiple of n_chans = %d" % (tm.shape[1], n_chans))
    N = tm.shape[1] // n_chans
    pr = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    n_lags = int(lag_hi) - int(lag_lo) + 1
    xc = np.zeros((pr.shape[0], max(n_lags, 0)), dtype=np.float64)
    al = np.zeros_like(xc) if aligned else None
    check(lib.fbp_template_pair_lags(device, _f64p(tm), tm.shape[0], n_chans, N, _u32p(pr), pr.shape[0],
                                     int(lag_lo), int(lag_hi), _f64p(xc), _f64p(al) if aligned else None),
          "fbp_template_pair_lags")
    return xc, al


class Server:
    """The segment server of libfbp (fbp_server_*): jobs admitted by slot and
    HBM budget, every running job advanced pass by pass by one native driver
    thread.  submit() keeps the host arrays alive until the job is released."""

    def __init__(self, device=0, hbm_budget=0, max_running=8):
        self._lib = load()
        self.device = device
        self._h = ctypes.c_void_p(None)
        check(self._lib.fbp_server_create(device, int(hbm_budget), int(max_running), ctypes.byref(self._h)),
              "fbp_server_create")
        self._inputs = {}
        self._cfgs = {}
        self._lock = threading.Lock()

    def close(self):
        if self._h is not None and self._h.value:
            self._lib.fbp_server_destroy(self._h)
        self._h = None
        with self._lock:
            self._inputs.clear()
            self._cfgs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, voltage, templates, bias, thresholds, template_len, max_shift, do_overlap=True, reduce_wg=256,
               precision="bf16", overlap_mode="screened", max_passes=0, extra_bytes=0):
        """Queue one segment; returns its ticket."""
        v = voltage
        if v.dtype != np.float32 or v.ndim != 2 or v.strides[1] != 4:
            v = np.ascontiguousarray(v, dtype=np.float32)
        C, K = v.shape
        tm = np.ascontiguousarray(templates, dtype=np.float32)
        b = np.ascontiguousarray(bias, dtype=np.float32)
        th = np.ascontiguousarray(thresholds, dtype=np.float32)
        T = b.size
        if tm.size != T * C * template_len or th.size != T:
            raise ValueError("templates, bias and thresholds do not match %d templates of %d x %d"
                             % (T, C, template_len))
        cfg = FbpConfig(C, K, T, template_len, max_shift, 1 if do_overlap else 0, reduce_wg, PRECISIONS[precision],
                        OVERLAP_MODES[overlap_mode], max_passes, 0)
        t = ctypes.c_uint64(0)
        with self._lock:
            check(self._lib.fbp_server_submit(self._h, ctypes.byref(cfg), _f32p(tm), _f32p(b), _f32p(th), _f32p(v),
                                              v.strides[0] // 4, int(extra_bytes), ctypes.byref(t)),
                  "fbp_server_submit")
            self._inputs[t.value] = (v, tm, b, th)
            self._cfgs[t.value] = cfg
        return t.value

    def _info(self, rc, info, what):
        if rc == FBP_ETIMEDOUT:
            return None
        check(rc, what)
        return info.as_dict()

    def poll(self, ticket):
        info = FbpJobInfo()
        check(self._lib.fbp_server_poll(self._h, ticket, ctypes.byref(info)), "fbp_server_poll")
        return info.as_dict()

    def wait(self, ticket, timeout=None):
        """Job info once it has ended, or None after timeout seconds."""
        info = FbpJobInfo()
        us = -1 if timeout is None else int(timeout * 1e6)
        return self._info(self._lib.fbp_server_wait(self._h, ticket, us, ctypes.byref(info)), info,
                          "fbp_server_wait")

    def wait_any(self, timeout=None):
        """(ticket, info) of the next job to end, or None after timeout seconds."""
        info = FbpJobInfo()
        t = ctypes.c_uint64(0)
        us = -1 if timeout is None else int(timeout * 1e6)
        d = self._info(self._lib.fbp_server_wait_any(self._h, us, ctypes.byref(t), ctypes.byref(info)), info,
                       "fbp_server_wait_any")
        return None if d is None else (t.value, d)

    def cancel(self, ticket):
        """True if the job was queued or running (it ends cancelled)."""
        rc = self._lib.fbp_server_cancel(self._h, ticket)
        if rc == FBP_ESTATE:
            return False
        check(rc, "fbp_server_cancel")
        return True

    def segment(self, ticket):
        """Non-owning Segment of a finished job (valid until release)."""
        h = ctypes.c_void_p(None)
        check(self._lib.fbp_server_segment(self._h, ticket, ctypes.byref(h)), "fbp_server_segment")
        with self._lock:
            cfg = self._cfgs[ticket]
        return Segment._view(h, cfg, self.device)

    def error(self, ticket):
        """Message of a failed job."""
        info = FbpJobInfo()
        check(self._lib.fbp_server_poll(self._h, ticket, ctypes.byref(info)), "fbp_server_poll")
        return self._lib.fbp_last_error().decode(errors="replace") if info.state == 3 else ""

    def release(self, ticket):
        check(self._lib.fbp_server_release(self._h, ticket), "fbp_server_release")
        with self._lock:
            self._inputs.pop(ticket, None)
            self._cfgs.pop(ticket, None)

    def stats(self):
        st = FbpServerStats()
        check(self._lib.fbp_server_get_stats(self._h, ctypes.byref(st)), "fbp_server_get_stats")
        return st.as_dict()
