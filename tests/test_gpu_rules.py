"""The hand-built scenarios of tests/scenarios.py through the HIP path: every
correlation precision and both overlap routes must add the same spikes in the
same pass and leave the same window list for the next pass as the
construction fixes (binary_pursuit.cl:822-875, :918-1063)."""
import numpy as np
import pytest

import scenarios

pytestmark = pytest.mark.gpu

ROUTES = [("fp32", "direct"), ("fp32", "screened"), ("bf16", "screened"), ("fp16", "screened"),
          ("fp8", "screened"), ("bf16", "direct")]


def _run(sc, precision, mode, max_passes):
    from spikesorting_fullpursuit_amd import _lib
    C, K = sc['voltage'].shape
    T = sc['templates'].shape[0]
    with _lib.Segment(C, K, T, sc['N'], sc['M'], sc['do_overlap'], 256, precision, mode, max_passes) as seg:
        seg.set_templates(sc['templates'], sc['bias'], sc['thr'])
        seg.set_voltage(sc['voltage'])
        capped = not seg.run(allow_cap=max_passes > 0)
        idx, lab = seg.spikes()
        return dict(spikes=list(zip(idx.tolist(), lab.tolist())), capped=capped,
                    windows=seg.window_list().tolist(), residual=seg.residual(), stats=seg.stats())


@pytest.mark.parametrize("precision,mode", ROUTES)
@pytest.mark.parametrize("make", scenarios.SCENARIOS, ids=lambda f: f.__name__)
def test_scenario_pass_by_pass(make, precision, mode):
    sc = make()
    print("scenario", sc['name'], precision, mode, flush=True)
    added = []
    for p, expect in enumerate(sc['passes'], start=1):
        r = _run(sc, precision, mode, p)
        assert r['spikes'][:len(added)] == added, (p, r['spikes'])
        assert r['spikes'][len(added):] == expect, (p, r['spikes'][len(added):])
        added = r['spikes']
        if p < len(sc['passes']):
            assert r['capped'], p
            assert r['windows'] == sc['lists'][p - 1], (p, r['windows'])
    r = _run(sc, precision, mode, 0)
    assert r['stats']['n_passes'] == len(sc['passes'])
    assert r['spikes'] == added and r['windows'] == []
    if sc['residual_zero']:
        assert np.all(r['residual'] == 0.0)


@pytest.mark.parametrize("precision,mode", ROUTES)
@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6])
def test_chained_train_fully_recovered(seed, precision, mode):
    sc = scenarios.chained_train(seed)
    print("chained train", seed, precision, mode, flush=True)
    r = _run(sc, precision, mode, 0)
    assert sorted(r['spikes']) == sorted(sc['truth'])
    assert np.all(r['residual'] == 0.0)
