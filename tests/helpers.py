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


def run_gpu(rec, M, do_overlap=True, reduce_wg=256, max_passes=0, precision="fp32", overlap_mode="direct"):
    from spikesorting_fullpursuit_amd import _lib
    V = rec['voltage']
    C, K = V.shape
    T = rec['templates'].shape[0]
    with _lib.Segment(C, K, T, rec['N'], M, do_overlap, reduce_wg, precision, overlap_mode, max_passes) as seg:
        seg.set_templates(rec['templates'], rec['bias'], rec['thr'])
        seg.set_voltage(V)
        seg.run(allow_cap=max_passes > 0)
        idx, lab = seg.spikes()
        return dict(idx=idx, lab=lab, residual=seg.residual(), L=seg.likelihood(), stats=seg.stats(),
                    clips=seg.clips())


def run_oracle(oracle, rec, M, do_overlap=True, reduce_wg=256, max_passes=0):
    return oracle.pursuit_chunk(rec['voltage'], rec['templates'], rec['bias'], rec['thr'], rec['N'], M,
                                do_overlap=do_overlap, reduce_wg=reduce_wg, max_passes=max_passes, want_L=True)
