"""fbp_channel_covariance_noalloc (device inputs and outputs, partial sums in
the library's fixed scratch array) against fbp_channel_covariance on the same
rows: bit-identical, on two calls in a row, the second smaller than the
first (the scratch keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(clips, C, N, rows):
    from helpers import HipMemory
    from spikesorting_fullpursuit_amd import _lib
    mem = HipMemory()
    try:
        dc = mem.upload(clips.astype(np.float32))
        dr = mem.upload(rows.astype(np.uint64)) if rows is not None else 0
        n_rows = rows.shape[1] if rows is not None else clips.shape[0]
        dcov = mem.upload(np.full((C, N, N), -1.0))
        _lib.channel_covariance_noalloc(dc, clips.shape[0], C, N, dr, n_rows, dcov)
        return mem.download(dcov, (C, N, N), np.float64)
    finally:
        mem.close()


def test_noalloc_covariance_equals_allocating_entry_point():
    from spikesorting_fullpursuit_amd import _lib
    rng = np.random.default_rng(5)
    # first call: 16 channels x 61 samples, 100,000 drawn rows (64 slices:
    # 31 MB of partial sums); second: 4 channels x 41 samples, 3,000 rows
    for C, N, n_clips, n_rows in [(16, 61, 30000, 100000), (4, 41, 2000, 3000)]:
        print("no-alloc covariance", C, N, n_clips, n_rows, flush=True)
        clips = rng.standard_normal((n_clips, C * N)).astype(np.float32)
        rows = rng.integers(0, n_clips, size=(C, n_rows)).astype(np.uint64)
        ref = _lib.channel_covariance(clips, C, N, rows)
        got = _noalloc(clips, C, N, rows)
        np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))
    # every clip (rows = NULL)
    clips = rng.standard_normal((700, 3 * 20)).astype(np.float32)
    np.testing.assert_array_equal(_noalloc(clips, 3, 20, None).view(np.uint64),
                                  _lib.channel_covariance(clips, 3, 20).view(np.uint64))


def test_noalloc_covariance_rejects_requests_beyond_scratch():
    from spikesorting_fullpursuit_amd import _lib
    lib = _lib.load()
    # 128 channels x 61 samples over 64 slices needs 244 MB of partial sums
    rc = lib.fbp_channel_covariance_noalloc(0, ctypes.c_void_p(16), 100000, 128, 61, None, 100000,
                                            ctypes.c_void_p(16), None)
    assert rc == _lib.FBP_EINVAL
    assert b"scratch" in lib.fbp_last_error()
