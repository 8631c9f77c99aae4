"""Shared helpers of the test suite: seeded cases and runners for the HIP
path (through the C-ABI) and the oracle."""
import numpy as np

from spikesorting_fullpursuit_amd import synthetic


def small_case(seed, C=4, T=5, N=61, duration_s=0.5, fs=40000, peak=(6.0, 15.0), rate=(5.0, 40.0),
               gains="ladder", overlap_frac=0.0, noise=True):
    rec = synthetic.make_recording(seed, C, fs, duration_s, T, N, peak, rate, gains=gains,
                                   overlap_frac=overlap_frac, noise=noise)
    sm = rec['separability_metrics']
    rec['bias'] = np.float32(-0.5 * sm['template_SS'])
    rec['thr'] = np.float32(sm['neuron_lower_thresholds'])
    rec['N'] = N
    return rec


def run_gpu(rec, M, do_overlap=True, reduce_wg=256, max_passes=0, precision="fp32", overlap_mode="direct",
            debug_flags=0):
    from spikesorting_fullpursuit_amd import _lib
    V = rec['voltage']
    C, K = V.shape
    T = rec['templates'].shape[0]
    with _lib.Segment(C, K, T, rec['N'], M, do_overlap, reduce_wg, precision, overlap_mode, max_passes,
                      debug_flags=debug_flags) as seg:
        seg.set_templates(rec['templates'], rec['bias'], rec['thr'])
        seg.set_voltage(V)
        seg.run(allow_cap=max_passes > 0)
        idx, lab = seg.spikes()
        return dict(idx=idx, lab=lab, residual=seg.residual(), L=seg.likelihood(), stats=seg.stats(),
                    clips=seg.clips())


def run_oracle(oracle, rec, M, do_overlap=True, reduce_wg=256, max_passes=0):
    return oracle.pursuit_chunk(rec['voltage'], rec['templates'], rec['bias'], rec['thr'], rec['N'], M,
                                do_overlap=do_overlap, reduce_wg=reduce_wg, max_passes=max_passes, want_L=True)


def screen_bound(rec, precision):
    """Float64 restatement of the correlation screen's bound e[t, i] (the
    kernel's k1..k4 coefficients, exact template norms and per-position window
    norms)."""
    V = rec['voltage'].astype(np.float64)
    C, K = V.shape
    N = rec['N']
    T = rec['templates'].shape[0]
    u, delta = {"bf16": (2.0 ** -8, 2.0 ** -126), "fp16": (2.0 ** -11, 2.0 ** -14),
                "fp8": (2.0 ** -4, 2.0 ** -6)}[precision]
    kstep = 32
    n = C * kstep * ((N + kstep - 1) // kstep) + 2.0
    gp = n * 2.0 ** -23 / (1 - n * 2.0 ** -23)
    cn = C * N
    k1 = (2 * u + u * u + 2 * gp * (1 + u) ** 2) * 1.01
    k2 = delta * (1 + u) * np.sqrt(cn) * (1 + gp) * 1.01
    k3 = cn * delta * delta * (1 + gp) * 1.01
    k4 = 2 * gp * 1.01
    tn = np.sqrt(np.sum(rec['templates'].astype(np.float64) ** 2, axis=1))
    s2 = np.concatenate([np.sum(V ** 2, axis=0), np.zeros(N)])
    cs = np.concatenate([[0.0], np.cumsum(s2)])
    vn = np.sqrt(cs[N:N + K] - cs[:K])
    bias = np.abs(rec['bias'].astype(np.float64))
    inv = np.ones(T)
    if precision == "fp8":
        mx = np.max(np.abs(rec['templates']), axis=1).astype(np.float64)
        inv = np.where(mx > 0, 1.0 / 2.0 ** np.floor(np.log2(224.0 / np.where(mx > 0, mx, 1.0))), 1.0)
    return (k1 * tn[:, None] * vn[None, :] + k2 * (tn[:, None] + vn[None, :] * inv[:, None]) + k3 * inv[:, None]
            + k4 * bias[:, None])


class HipMemory:
    """Device buffers for tests that call device-pointer entry points: the
    HIP runtime libfbp.so itself links (libamdhip64.so.7), bound with ctypes,
    so the buffers live in the same runtime instance as the library's."""

    H2D, D2H = 1, 2

    def __init__(self):
        import ctypes
        from spikesorting_fullpursuit_amd import _lib
        _lib.load()
        self.ct = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.live = []

    def _ok(self, rc, what):
        assert rc == 0, "%s failed with HIP error %d" % (what, rc)

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.ct.c_void_p()
        self._ok(self.hip.hipMalloc(self.ct.byref(p), max(a.nbytes, 1)), "hipMalloc")
        self.live.append(p.value)
        self._ok(self.hip.hipMemcpy(p.value, a.ctypes.data, a.nbytes, self.H2D), "hipMemcpy")
        return p.value

    def empty(self, nbytes):
        p = self.ct.c_void_p()
        self._ok(self.hip.hipMalloc(self.ct.byref(p), max(nbytes, 1)), "hipMalloc")
        self.live.append(p.value)
        return p.value

    def download(self, ptr, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        self._ok(self.hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self._ok(self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, self.D2H), "hipMemcpy")
        return out

    def close(self):
        for p in self.live:
            self.hip.hipFree(p)
        self.live = []
