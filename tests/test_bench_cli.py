"""bench.py's command line on the CPU: the one-GPU rule, the --model table
(each model's workload has the fittype it names) and the per-evaluation flop
count the roofline divides by."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr


@pytest.mark.parametrize("model", sorted(bench.FIT_MODELS))
def test_fit_models_generate_their_fittype(model):
    from bsfit import gen
    cfg, ft, label, grid = bench.FIT_MODELS[model]
    w = gen.config(cfg, N=4, S=13 if ft == 8 else None)
    w = w[0] if isinstance(w, list) else w
    assert w.N == 4
    if ft == 104:                       # 4Pi: four linked channels, the colour-4D machinery per channel
        assert w.imstack.ndim == 4 and w.nch == 4
        assert bench.f_alg(7, w.S, nch=4, NV=6) == 4 * bench.f_alg(7, w.S)
        return
    assert w.fittype == ft
    assert (w.coeff3 is not None) == (ft == 8)
    assert bench.f_alg(ft, w.S) > 0


def test_default_headline_is_the_mixed_size_vortex_fit():
    """With no flags the headline is fittype 8 over sizes 11-17 in one launch;
    the metric is the workload's own."""
    import argparse
    assert bench.DEFAULT_N["vortex5d"] == 100000
    src = open(os.path.join(ROOT, "bench.py")).read()
    assert 'default="vortex5d"' in src
    assert bench.METRIC.startswith("ROI fits/sec per fittype")


def test_f_alg_values():
    # colour-4D at S = 15: P = 18, T = 16, K = 2, NV = 6:
    # 2*18^2*16*3 + 32*225*5 + 225*(42 + 12 + 8) = 31104 + 36000 + 13950
    assert bench.f_alg(7, 15) == 81054
    # 3D at S = 13: P = 16, T = 4, K = 1, NV = 5: 2*256*4*2 + 32*169*4 + 169*(30 + 10 + 8)
    assert bench.f_alg(6, 13) == 4096 + 21632 + 8112
    # fittype 8 adds the 3D model and the blend to the 5D terms (D = 5, NV = 8)
    S, P = 13, 16
    base = 2 * P * P * 64 * 4 + 32 * S * S * 6 + S * S * (72 + 16 + 8)
    assert bench.f_alg(8, S) == base + 2 * P * P * 8 + 2 * S * S * 64 + 20 * S * S
