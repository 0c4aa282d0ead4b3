"""Multi-start fits (bsf_fit_multistart_*) against the reference callers'
loop: K single fits (bsf_fit_host, whose parity with the oracle the other
GPU tests establish) combined on the host with the callers' rule
(demo_orientation/store/fitb5n3.m:38-45: replace where LLh - LL >= 1e-4;
MultiD_fitter/simplefitter_spline.m:488: > 1e-4), compared in FP32.  Every
ROI's fit from every start is the same computation either way, so the
results must be identical bit for bit."""
import numpy as np
import pytest

import bsfit
from bsfit import gen

pytestmark = pytest.mark.gpu


def _host_loop(fittype, w, starts, coeff3, strict, precision):
    P, CR, LL = bsfit.fit(fittype, w.imstack, w.iterations, w.coeff, starts[0], coeff3, precision=precision)
    for s in starts[1:]:
        Ph, CRh, LLh = bsfit.fit(fittype, w.imstack, w.iterations, w.coeff, s, coeff3, precision=precision)
        d = (LLh - LL).astype(np.float32)
        better = d > np.float32(1e-4) if strict else d >= np.float32(1e-4)
        P[better] = Ph[better]
        CR[better] = CRh[better]
        LL[better] = LLh[better]
    return P, CR, LL, better


@pytest.mark.parametrize("precision", [bsfit.PREC_MIX, bsfit.PREC_F64])
def test_colour_4d_gui_starts(precision):
    """Fittype 7 with the GUI's start pattern (z0, w) pairs, strict rule."""
    w = gen.config(1, N=256)
    nz, nw = [d - 3 for d in w.coeff.shape[2:4]]
    starts = np.array([[nz / 2, 1.5], [0.8 * nz / 2, 2.5], [1.2 * nz / 2, 1.0]], np.float32)
    got = bsfit.fit_multistart(7, w.imstack, w.iterations, w.coeff, starts, strict=True, precision=precision)
    want = _host_loop(7, w, starts, None, True, precision)
    for a, b in zip(got, want[:3]):
        np.testing.assert_array_equal(a, b)
    single = bsfit.fit(7, w.imstack, w.iterations, w.coeff, starts[0], precision=precision)
    assert np.all(got[2] >= single[2] - 1e-3)          # the kept log-likelihood never drops
    changed = int(np.sum(got[2] != single[2]))
    print("precision %d: %d of %d ROIs took a later start" % (precision, changed, w.N))
    assert changed > 0          # the later starts are exercised


def test_vortex_5d_demo_starts():
    """Fittype 8 with the 5D demo's two starts (fitb5n3.m:14), >= rule."""
    ws = gen.config(3, N=120, S=13)
    w = ws[0]
    c5, c3 = gen.model_coeffs("dipole5d")
    nd = [d - 3 for d in c5.shape]
    starts = np.array([[nd[2] / 2, nd[3] / 3, nd[4] / 2, 0.75], [nd[2] / 2, 2 * nd[3] / 3, nd[4] / 2, 0.75]],
                      np.float32)
    got = bsfit.fit_multistart(8, w.imstack, w.iterations, c5, starts, coeff3=c3)
    want = _host_loop(8, w, starts, c3, False, bsfit.PREC_MIX)
    for a, b in zip(got, want[:3]):
        np.testing.assert_array_equal(a, b)


def test_multistart_argument_errors():
    w = gen.config(0, N=4)
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.fit_multistart(6, w.imstack, 10, w.coeff, np.zeros((2, 2)))
    assert e.value.code == 2
    w7 = gen.config(1, N=4)
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.fit_multistart(7, w7.imstack, 10, w7.coeff, np.zeros((2, 1)))
    assert e.value.code == 1
