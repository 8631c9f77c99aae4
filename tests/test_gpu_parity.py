"""HIP path (libfbp.so through its C-ABI) against the oracle, bit for bit.

Float32 arithmetic of both sides follows the reference's fused multiply-add
order, so the likelihood buffer, the spike list (onset, label, record order),
the residual and the adjusted clips must be identical, not merely close.
"""
import numpy as np
import pytest

from helpers import run_gpu, run_oracle, screen_bound, small_case

pytestmark = pytest.mark.gpu

CASES = [
    # (seed, C, T, N, duration_s, M, do_overlap, reduce_wg)
    (1, 4, 5, 61, 0.5, 30, True, 256),
    (2, 4, 5, 61, 0.5, 60, True, 256),
    (3, 4, 5, 61, 0.5, 30, False, 256),
    (4, 8, 7, 41, 0.4, 20, True, 64),
    (5, 3, 3, 81, 0.4, 40, True, 1024),
    (6, 2, 1, 9, 0.05, 4, True, 16),
]


def _case(seed, C, T, N, dur):
    gains = "ladder" if C <= 4 else "gauss"
    return small_case(seed, C=C, T=T, N=N, duration_s=dur, gains=gains, overlap_frac=0.1 if T > 1 else 0.0)


@pytest.mark.parametrize("seed,C,T,N,dur,M,ov,wg", CASES)
def test_first_pass_likelihood_bitwise(oracle_mod, seed, C, T, N, dur, M, ov, wg):
    rec = _case(seed, C, T, N, dur)
    print("case", seed, C, T, N, dur, M, ov, wg, flush=True)
    g = run_gpu(rec, M, ov, wg, max_passes=1)
    o = run_oracle(oracle_mod, rec, M, ov, wg, max_passes=1)
    np.testing.assert_array_equal(g['L'].view(np.uint32), o['L'].view(np.uint32))
    np.testing.assert_array_equal(g['idx'], o['idx'])
    np.testing.assert_array_equal(g['lab'], o['lab'])


@pytest.mark.parametrize("mode", ["direct", "screened"])
@pytest.mark.parametrize("seed,C,T,N,dur,M,ov,wg", CASES)
def test_full_pursuit_bitwise(oracle_mod, seed, C, T, N, dur, M, ov, wg, mode):
    rec = _case(seed, C, T, N, dur)
    print("case", seed, C, T, N, dur, M, ov, wg, mode, flush=True)
    g = run_gpu(rec, M, ov, wg, overlap_mode=mode)
    o = run_oracle(oracle_mod, rec, M, ov, wg)
    assert g['stats']['n_passes'] == o['n_passes']
    np.testing.assert_array_equal(g['idx'], o['idx'])
    np.testing.assert_array_equal(g['lab'], o['lab'])
    np.testing.assert_array_equal(g['residual'].view(np.uint32), o['residual'].view(np.uint32))
    np.testing.assert_array_equal(g['L'].view(np.uint32), o['L'].view(np.uint32))
    clips = oracle_mod.adjusted_clips(o['residual'], rec['templates'], rec['N'], o['idx'], o['lab'])
    np.testing.assert_array_equal(g['clips'].view(np.uint32), clips.view(np.uint32))


@pytest.mark.parametrize("mode", ["direct", "screened"])
def test_dense_16ch_spike_list_grows(oracle_mod, mode):
    # 16 channels, 20 templates, 1 s at 40 kHz: more spikes than windows, so
    # the device spike list grows during the run.
    from spikesorting_fullpursuit_amd import synthetic
    rec = synthetic.config("16ch-300s", seed=3, duration_s=1.0)
    sm = rec['separability_metrics']
    rec['bias'] = np.float32(-0.5 * sm['template_SS'])
    rec['thr'] = np.float32(sm['neuron_lower_thresholds'])
    rec['N'] = 61
    print("case 16ch 1s", mode, flush=True)
    g = run_gpu(rec, 30, True, 256, overlap_mode=mode)
    o = run_oracle(oracle_mod, rec, 30, True, 256)
    W = -(-rec['voltage'].shape[1] // 61)
    assert o['idx'].size > W
    np.testing.assert_array_equal(g['idx'], o['idx'])
    np.testing.assert_array_equal(g['lab'], o['lab'])
    np.testing.assert_array_equal(g['residual'].view(np.uint32), o['residual'].view(np.uint32))
    if mode == "screened":
        assert g['stats']['screened_combos'] > 20 * g['stats']['overlap_combos']


@pytest.mark.parametrize("debug", [0, 2, 4, 8, 12, 16, 18, 32, 50, 64, 66, 68, 72, 76, 80, 128, 130, 136, 160,
                                   256, 258, 264, 272, 288, 384, 386, 512])
def test_screened_overlap_candidate_paths(oracle_mod, debug):
    # The screened overlap finds its exact-score candidates from per-thread
    # lists and settles single candidates from their score interval, unless
    # forced onto the second traversal (2) or exact scores for single
    # candidates (8); rows of the traversal are skipped by their upper bound
    # unless 16 is set; 32 splits each pass's overlap windows into batches of
    # 37 (the Q tables and the screen run per batch).  The candidates are
    # scored by the deferred exact-score kernel (LDS templates, one wave per
    # window), or in the screen's block with 64 (partner templates staged in
    # LDS, or read from global memory with 4); 128 leaves 64 deferred slots
    # per batch, so windows of both kinds meet in one batch; 256 takes the
    # route of templates too large for LDS (eight-wave screen blocks without
    # template rows, exact scores read from the zero-padded templates in
    # global memory, channels no lane's templates touch skipped); 512
    # recomputes every listed window's likelihood, also the template chunks
    # whose channels saw no voltage change since the last pass.  Every route
    # gives the oracle's events.
    from spikesorting_fullpursuit_amd import synthetic
    rec = synthetic.config("16ch-300s", seed=7, duration_s=0.5)
    sm = rec['separability_metrics']
    rec['bias'] = np.float32(-0.5 * sm['template_SS'])
    rec['thr'] = np.float32(sm['neuron_lower_thresholds'])
    rec['N'] = 61
    print("case 16ch 0.5s debug", debug, flush=True)
    g = run_gpu(rec, 30, True, 256, precision="bf16", overlap_mode="screened", debug_flags=debug)
    o = run_oracle(oracle_mod, rec, 30, True, 256)
    np.testing.assert_array_equal(g['idx'], o['idx'])
    np.testing.assert_array_equal(g['lab'], o['lab'])
    np.testing.assert_array_equal(g['residual'].view(np.uint32), o['residual'].view(np.uint32))
    if debug & 8:
        assert g['stats']['overlap_combos'] > 0
    if debug & 32:
        assert g['stats']['overlap_windows'] > 2 * 37   # several batches
    if debug == 0:
        full = run_gpu(rec, 30, True, 256, precision="bf16", overlap_mode="screened", debug_flags=16)
        assert g['stats']['screened_combos'] < full['stats']['screened_combos']


@pytest.mark.parametrize("precision", ["bf16", "fp16", "fp8"])
@pytest.mark.parametrize("seed,C,T,N,dur", [(31, 4, 5, 61, 0.5), (32, 8, 20, 41, 0.3), (33, 3, 3, 81, 0.3)])
def test_screen_error_bound_holds(precision, seed, C, T, N, dur):
    # Without the float32 re-score, the matrix-core likelihood must lie within
    # the stated bound of the float32 likelihood at every position.
    rec = _case(seed, C, T, N, dur)
    print("case", precision, seed, C, T, N, dur, flush=True)
    lo = run_gpu(rec, N // 2, True, 256, max_passes=1, precision=precision, debug_flags=1)
    hi = run_gpu(rec, N // 2, True, 256, max_passes=1)
    K = rec['voltage'].shape[1]
    err = np.abs(lo['L'].astype(np.float64) - hi['L'].astype(np.float64))
    e = screen_bound(rec, precision)
    valid = np.arange(K) + N < K
    assert np.all(err[:, valid] <= e[:, valid]), float(np.max(err[:, valid] / e[:, valid]))
    assert np.all(lo['L'][:, ~valid] == 0.0)
    # the bound is not vacuous: it is within a small factor of the observed error
    assert np.max(err[:, valid] / e[:, valid]) > 1e-3


@pytest.mark.parametrize("mode", ["direct", "screened"])
@pytest.mark.parametrize("precision", ["bf16", "fp16", "fp8"])
@pytest.mark.parametrize("seed,C,T,N,dur,M,ov,wg", CASES[:5])
def test_low_precision_decisions_equal_fp32(oracle_mod, seed, C, T, N, dur, M, ov, wg, precision, mode):
    rec = _case(seed, C, T, N, dur)
    print("case", seed, C, T, N, dur, M, ov, wg, precision, mode, flush=True)
    g = run_gpu(rec, M, ov, wg, precision=precision, overlap_mode=mode)
    o = run_oracle(oracle_mod, rec, M, ov, wg)
    assert g['stats']['n_passes'] == o['n_passes']
    np.testing.assert_array_equal(g['idx'], o['idx'])
    np.testing.assert_array_equal(g['lab'], o['lab'])
    np.testing.assert_array_equal(g['residual'].view(np.uint32), o['residual'].view(np.uint32))
    # stored likelihood: positive values exact, every other value non-positive
    pos = o['L'] > 0
    np.testing.assert_array_equal(g['L'][pos].view(np.uint32), o['L'][pos].view(np.uint32))
    assert np.all(g['L'][~pos] <= 0.0)
    assert g['stats']['rescored_positions'] > 0


@pytest.mark.parametrize("precision", ["bf16", "fp16", "fp8"])
def test_low_precision_dense_16ch(oracle_mod, precision):
    from spikesorting_fullpursuit_amd import synthetic
    rec = synthetic.config("16ch-300s", seed=5, duration_s=1.0)
    sm = rec['separability_metrics']
    rec['bias'] = np.float32(-0.5 * sm['template_SS'])
    rec['thr'] = np.float32(sm['neuron_lower_thresholds'])
    rec['N'] = 61
    print("case 16ch 1s", precision, flush=True)
    g = run_gpu(rec, 30, True, 256, precision=precision, overlap_mode="screened")
    o = run_oracle(oracle_mod, rec, 30, True, 256)
    np.testing.assert_array_equal(g['idx'], o['idx'])
    np.testing.assert_array_equal(g['lab'], o['lab'])
    np.testing.assert_array_equal(g['residual'].view(np.uint32), o['residual'].view(np.uint32))
    frac = g['stats']['rescored_positions'] / g['stats']['corr_positions'] / 20
    print("re-scored fraction of (t, i):", frac, flush=True)


@pytest.mark.parametrize("precision,mode,M", [("fp32", "screened", 20), ("bf16", "screened", 20), ("fp32", "direct", 20),
                                              ("bf16", "screened", 40)])
def test_sparse_channel_support(oracle_mod, precision, mode, M):
    # Templates nonzero on 6 of 32 channels (a dense probe): the screened
    # overlap stores and scores only the channels a template is nonzero on.
    # At n_max_shift_inds = N - 1 the screen's bound of a template pair is
    # taken over the channels of their two supports.
    from spikesorting_fullpursuit_amd import synthetic
    rec = synthetic.make_recording(41, 32, 30000, 0.4, 12, 41, (8.0, 20.0), (5.0, 60.0), gains="gauss",
                                   spatial_sigma=2.0, support=6, overlap_frac=0.1)
    sm = rec['separability_metrics']
    rec['bias'] = np.float32(-0.5 * sm['template_SS'])
    rec['thr'] = np.float32(sm['neuron_lower_thresholds'])
    rec['N'] = 41
    print("case sparse support", precision, mode, M, flush=True)
    g = run_gpu(rec, M, True, 256, precision=precision, overlap_mode=mode)
    o = run_oracle(oracle_mod, rec, M, True, 256)
    assert o['idx'].size > 0
    np.testing.assert_array_equal(g['idx'], o['idx'])
    np.testing.assert_array_equal(g['lab'], o['lab'])
    np.testing.assert_array_equal(g['residual'].view(np.uint32), o['residual'].view(np.uint32))


def test_large_template_set_rescore(oracle_mod):
    # 64 channels x 30 templates of 61 samples (0.47 MB) do not fit in LDS: the
    # float32 re-score reads template rows and voltage from global memory over
    # each template's 8-channel support.
    from spikesorting_fullpursuit_amd import synthetic
    rec = synthetic.make_recording(43, 64, 30000, 0.25, 30, 61, (8.0, 20.0), (5.0, 60.0), gains="gauss",
                                   spatial_sigma=2.0, support=8, overlap_frac=0.1)
    sm = rec['separability_metrics']
    rec['bias'] = np.float32(-0.5 * sm['template_SS'])
    rec['thr'] = np.float32(sm['neuron_lower_thresholds'])
    rec['N'] = 61
    print("case 64ch x 30 templates", flush=True)
    g = run_gpu(rec, 30, True, 256, precision="bf16", overlap_mode="screened")
    o = run_oracle(oracle_mod, rec, 30, True, 256)
    assert o['idx'].size > 0
    assert g['stats']['rescored_positions'] > 0
    np.testing.assert_array_equal(g['idx'], o['idx'])
    np.testing.assert_array_equal(g['lab'], o['lab'])
    np.testing.assert_array_equal(g['residual'].view(np.uint32), o['residual'].view(np.uint32))
    pos = o['L'] > 0
    np.testing.assert_array_equal(g['L'][pos].view(np.uint32), o['L'][pos].view(np.uint32))


@pytest.mark.parametrize("precision,mode,C,T,support", [("bf16", "screened", 96, 30, 8), ("fp16", "screened", 40, 16, 5),
                                                        ("fp32", "direct", 32, 12, 6)])
def test_incremental_correlation_equals_full(precision, mode, C, T, support):
    # The correlation skips the (window, template chunk) blocks whose channels
    # saw no voltage change since their last pass (with 96 channels a mask bit
    # covers two).  What it leaves in place must be exactly what recomputing
    # gives: the whole likelihood buffer (negative screened values included),
    # the spike list and the residual equal a run that recomputes every
    # listed block (FBP_DEBUG_FULL_CORRELATION), with fewer likelihoods
    # computed.
    from spikesorting_fullpursuit_amd import synthetic
    rec = synthetic.make_recording(47 + C, C, 30000, 0.4, T, 41, (8.0, 20.0), (5.0, 60.0), gains="gauss",
                                   spatial_sigma=2.0, support=support, overlap_frac=0.1)
    sm = rec['separability_metrics']
    rec['bias'] = np.float32(-0.5 * sm['template_SS'])
    rec['thr'] = np.float32(sm['neuron_lower_thresholds'])
    rec['N'] = 41
    print("case incremental correlation", precision, mode, C, T, support, flush=True)
    g = run_gpu(rec, 20, True, 256, precision=precision, overlap_mode=mode)
    full = run_gpu(rec, 20, True, 256, precision=precision, overlap_mode=mode, debug_flags=512)
    assert g['idx'].size > 0 and g['stats']['n_passes'] > 2
    assert g['stats']['corr_pairs'] < full['stats']['corr_pairs']
    np.testing.assert_array_equal(g['idx'], full['idx'])
    np.testing.assert_array_equal(g['lab'], full['lab'])
    np.testing.assert_array_equal(g['residual'].view(np.uint32), full['residual'].view(np.uint32))
    np.testing.assert_array_equal(g['L'].view(np.uint32), full['L'].view(np.uint32))
