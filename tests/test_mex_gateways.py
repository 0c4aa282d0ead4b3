"""The MATLAB MEX gateways (gpu-based-b-spline-fitter_amd/matlab/*.c),
compiled with -Wall -Wextra -Werror against the test MX API (tests/mex_api)
and called as MATLAB calls them; and the GUI result table
(bsf_gui_results, host code).  Without a GPU the fitter gateway reports the
library's device error through mexErrMsgIdAndTxt, never a crash.

The table's column layout and single-precision arithmetic follow
MultiD_fitter/simplefitter_spline.m:550-630; its meaning (x/y exchange, z
sign and scale) is pinned by planted emitters in tests/test_gpu_mex.py."""
import numpy as np
import pytest

import bsfit
from bsfit import gen
from parity_util import Mex, MexError


@pytest.fixture(scope="module")
def mexdir(tmp_path_factory):
    return tmp_path_factory.mktemp("mex")


def test_fitter_gateway_argument_errors(mexdir):
    import torch
    m = Mex("GPUmleFit_LM_multiD", mexdir)
    w = gen.config(0, N=2)
    for args, ident in (((w.imstack.astype(np.float64), 6.0, 50.0, w.coeff), "GPUmleFit_LM:WrongDataType"),
                        ((w.imstack, 5.0, 50.0, w.coeff), "GPUmleFit_LM:FitType"),
                        ((w.imstack,), "GPUmleFit_LM:WrongNoArgs"),
                        ((w.imstack, 6.0, 50.0, w.coeff.astype(np.float64)), "GPUmleFit_LM:WrongDataType")):
        with pytest.raises(MexError) as e:
            m(3, *args)
        assert e.value.ident == ident
    with pytest.raises(MexError) as e:                                   # P, CRLB, LogL and status at most
        m(5, w.imstack, 6.0, 50.0, w.coeff)
    assert e.value.ident == "GPUmleFit_LM:WrongNoArgs"
    if not torch.cuda.is_available():
        with pytest.raises(MexError) as e:
            m(4, w.imstack, 6.0, 50.0, w.coeff)
        assert e.value.ident == "GPUmleFit_LM:Fit" and "device" in str(e.value)
        with pytest.raises(MexError) as e:
            m(3, w.imstack, 6.0, 50.0, w.coeff)
        assert e.value.ident == "GPUmleFit_LM:Fit" and "device" in str(e.value)


def test_roi_gateway_argument_errors(mexdir):
    import torch
    m = Mex("bsf_find_rois", mexdir)
    fr = np.zeros((32, 32, 1), np.uint16)
    h = bsfit.peak_filter("dog", 1.0)
    with pytest.raises(MexError) as e:
        m(2, fr.astype(np.int32).astype(np.float64), 7.0, h, 1.0, 100.0, 0.5)
    assert e.value.ident == "bsf_find_rois:WrongDataType"
    with pytest.raises(MexError) as e:
        m(2, fr, 8.0, h, 1.0, 100.0, 0.5)
    assert e.value.ident == "bsf_find_rois:Run" and "odd" in str(e.value)
    if not torch.cuda.is_available():
        with pytest.raises(MexError) as e:
            m(2, fr, 7.0, h, 1.0, 100.0, 0.5)
        assert e.value.ident == "bsf_find_rois:Run"


def _block(ft, N, rng):
    NV = bsfit.num_params(ft)
    P = rng.uniform(0.5, 12.0, (N, NV + 1)).astype(np.float32)
    CR = rng.uniform(-0.01, 0.5, (N, NV)).astype(np.float32)        # a few negative variances
    LL = rng.uniform(-300, -100, N).astype(np.float32)
    pk = np.stack([rng.integers(8, 500, N), rng.integers(8, 500, N), rng.integers(1, 40, N)], 1).astype(np.float64)
    return P, CR, LL, pk


@pytest.mark.parametrize("ft", [6, 7, 8])
@pytest.mark.parametrize("mirror", [False, True])
def test_gui_results_gateway_equals_library(mexdir, ft, mirror):
    m = Mex("bsf_gui_results", mexdir)
    rng = np.random.default_rng(ft + 10 * mirror)
    P, CR, LL, pk = _block(ft, 50, rng)
    args = (6.0, 30.7, 20.3, 0.41, 0.17)
    got = m(1, P, CR, LL, pk, float(ft), float(mirror), *args)[0]
    want = bsfit.gui_results(ft, P, CR, LL, pk, *args, mirror=mirror)
    assert got.shape == (50, {6: 12, 7: 17, 8: 19}[ft])
    np.testing.assert_array_equal(got, want)


def test_gui_results_layout_and_single_precision():
    rng = np.random.default_rng(3)
    f = np.float32
    dx, z0, dz, du, dv = 6.0, 30.7, 20.3, 0.41, 0.17
    for ft in (6, 7, 8):
        P, CR, LL, pk = _block(ft, 40, rng)
        R = bsfit.gui_results(ft, P, CR, LL, pk, dx, z0, dz, du, dv)
        px, py = pk[:, 0].astype(f), pk[:, 1].astype(f)
        np.testing.assert_array_equal(R[:, 0], pk[:, 2])
        xcol = 0 if ft in (7, 8) else 1                                   # the 4D/5D x/y exchange
        np.testing.assert_array_equal(R[:, 1], (P[:, xcol] - f(dx)) + px)
        np.testing.assert_array_equal(R[:, 2], (P[:, 1 - xcol] - f(dx)) + py)
        zsign = f(1) if ft == 8 else f(-1)
        zz = zsign * ((P[:, 4] - f(z0)) * f(dz))
        np.testing.assert_array_equal(R[:, 3], zz)
        # single-precision evaluation, not double: the double result differs somewhere
        zd = float(zsign) * ((P[:, 4].astype(np.float64) - z0) * dz)
        assert np.any(R[:, 3] != zd)
        ll_col = {6: 11, 7: 13, 8: 17}[ft]
        np.testing.assert_array_equal(R[:, ll_col], LL)
        err_x = {6: 6, 7: 7, 8: 9}[ft]                                    # real(sqrt(CRLB(:,2))): 0 if negative
        np.testing.assert_array_equal(R[:, err_x], np.where(CR[:, 1] < 0, f(0), np.sqrt(np.maximum(CR[:, 1], 0))))
        if ft != 6:
            np.testing.assert_array_equal(R[:, -1 if ft == 8 else 14], P[:, -1])   # iterations
    Pm, CRm, LLm, pkm = _block(7, 10, rng)
    Rm = bsfit.gui_results(7, Pm, CRm, LLm, pkm, dx, z0, dz, mirror=True)
    np.testing.assert_array_equal(Rm[:, 1], (f(dx) - Pm[:, 1]) + pkm[:, 0].astype(f))


def test_color_class():
    w = np.array([np.nan, 1.0, 2.2, 2.5, 2.75, 3.7], np.float32)
    f = np.float32
    want = np.array([-1, -1, (f(2.2) - f(2.5)) * f(2), 0, 0.5, 1], np.float32)   # NaN -> max(2, NaN) = 2
    np.testing.assert_array_equal(bsfit.color_class(w), want)


def test_gui_results_errors():
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.gui_results(5, np.zeros((2, 6)), np.zeros((2, 5)), np.zeros(2), np.zeros((2, 3)), 6, 30, 20)
    assert e.value.code == 2
