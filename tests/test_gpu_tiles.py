"""Two ROIs per wavefront (tile 2) against one ROI per wavefront (tile 1),
and the autotuner that picks between them (bsf_tune).

In a tile-2 launch each half-wave fits its own ROI; a lane there holds the
pixel sums of two lanes of the one-ROI kernel and adds them in the same
order, so the two tiles must agree bit for bit: parameters, CRLB,
log-likelihood, status flags and evaluation counts, for uniform launches
(including an odd ROI count, whose last wave has an empty half) and for a
launch of mixed sizes, in both precision modes."""
import numpy as np
import pytest

import bsfit
from bsfit import gen
from parity_util import Hip

pytestmark = pytest.mark.gpu


def _model(ft):
    if ft == 6:
        return gen.model_coeffs("astig3d"), None
    if ft == 7:
        return gen.model_coeffs("color4d"), None
    return gen.model_coeffs("dipole5d")


def _run(hip, ft, data, S, N, c, c3, prec, tile, initP=None, rois=None):
    NV = bsfit.num_params(ft)
    d = hip.put(data)
    dc = hip.put(np.asfortranarray(c, np.float32))
    dc3 = hip.put(np.asfortranarray(c3, np.float32)) if c3 is not None else None
    dr = hip.put(rois) if rois is not None else None
    P, CR, LL = hip.alloc((NV + 1) * N * 4), hip.alloc(NV * N * 4), hip.alloc(N * 4)
    st, ev = hip.alloc(N), hip.alloc(N * 4)
    try:
        bsfit.fit_device(ft, d, S, N, 50, dc, c.shape, P, CR, LL, initP=initP, coeff3_ptr=dc3, precision=prec,
                         stream=None, status_ptr=st, rois_ptr=dr, evals_ptr=ev, tile=tile)
        assert hip.h.hipDeviceSynchronize() == 0
        return (hip.get(P, (NV + 1) * N, np.float32), hip.get(CR, NV * N, np.float32), hip.get(LL, N, np.float32),
                hip.get(st, N, np.uint8), hip.get(ev, N, np.int32))
    finally:
        hip.free(d, dc, dc3, dr, P, CR, LL, st, ev)


def _same(a, b):
    for x, y, name in zip(a, b, ("P", "CRLB", "LL", "status", "evals")):
        np.testing.assert_array_equal(x, y, err_msg=name)


@pytest.mark.parametrize("ft", [6, 7])
@pytest.mark.parametrize("prec", [bsfit.PREC_MIX, bsfit.PREC_F64])
@pytest.mark.parametrize("S,N", [(7, 64), (11, 101), (13, 96), (17, 63)])
def test_tile2_equals_tile1_uniform(ft, prec, S, N):
    c, c3 = _model(ft)
    w = gen.make_rois(ft, N, S, 900 + 7 * S + ft, c, c3)
    print("case fittype=%d precision=%d S=%d N=%d" % (ft, prec, S, N), flush=True)
    data = np.asfortranarray(w.imstack, np.float32).ravel(order="F")
    hip = Hip()
    one = _run(hip, ft, data, S, N, c, c3, prec, 1, initP=w.initP)
    two = _run(hip, ft, data, S, N, c, c3, prec, 2, initP=w.initP)
    _same(one, two)
    # and the plain host entry point (tile from the table) gives the same
    ref = bsfit.fit_status(ft, w.imstack, 50, c, w.initP, c3, precision=prec)
    NV = bsfit.num_params(ft)
    np.testing.assert_array_equal(one[0].reshape(NV + 1, N).T, ref[0])


@pytest.mark.parametrize("ft", [6, 7])
@pytest.mark.parametrize("prec", [bsfit.PREC_MIX, bsfit.PREC_F64])
def test_tile2_equals_tile1_mixed_sizes(ft, prec):
    c, c3 = _model(ft)
    sizes = (17, 15, 13, 11)
    groups = [gen.make_rois(ft, 33, S, 950 + S + ft, c, c3) for S in sizes]   # odd counts: halves of mixed size
    print("case fittype=%d precision=%d mixed sizes %s" % (ft, prec, sizes), flush=True)
    data = np.concatenate([np.asfortranarray(w.imstack).ravel(order="F") for w in groups]).astype(np.float32)
    S_list = np.concatenate([[w.S] * w.N for w in groups])
    table = bsfit.roi_table(S_list)
    hip = Hip()
    N = len(S_list)
    one = _run(hip, ft, data, max(sizes), N, c, c3, prec, 1, initP=groups[0].initP, rois=table)
    two = _run(hip, ft, data, max(sizes), N, c, c3, prec, 2, initP=groups[0].initP, rois=table)
    _same(one, two)


def test_tile2_request_for_fittype8_runs_one_roi_per_wave():
    c, c3 = _model(8)
    w = gen.make_rois(8, 24, 11, 977, c, c3)
    data = np.asfortranarray(w.imstack, np.float32).ravel(order="F")
    hip = Hip()
    one = _run(hip, 8, data, 11, w.N, c, c3, bsfit.PREC_MIX, 1, initP=w.initP)
    two = _run(hip, 8, data, 11, w.N, c, c3, bsfit.PREC_MIX, 2, initP=w.initP)
    _same(one, two)


@pytest.mark.parametrize("ft,S", [(6, 11), (7, 13)])
def test_tune_records_a_tile_and_keeps_results(ft, S):
    c, c3 = _model(ft)
    w = gen.make_rois(ft, 2048, S, 990 + S, c, c3)
    N, NV = w.N, bsfit.num_params(ft)
    hip = Hip()
    data = np.asfortranarray(w.imstack, np.float32).ravel(order="F")
    one = _run(hip, ft, data, S, N, c, c3, bsfit.PREC_MIX, 1, initP=w.initP)
    d = hip.put(data)
    dc = hip.put(np.asfortranarray(c, np.float32))
    P, CR, LL = hip.alloc((NV + 1) * N * 4), hip.alloc(NV * N * 4), hip.alloc(N * 4)
    try:
        tile, ms = bsfit.tune(ft, d, S, N, 50, dc, c.shape, P, CR, LL, initP=w.initP, reps=3)
        Pm = hip.get(P, (NV + 1) * N, np.float32)
    finally:
        hip.free(d, dc, P, CR, LL)
    print("case fittype=%d S=%d: tile 1 %.3f ms, tile 2 %.3f ms -> tile %d" % (ft, S, ms[0], ms[1], tile), flush=True)
    try:
        assert tile in (1, 2) and ms[0] > 0 and ms[1] > 0
        assert tile == (1 if ms[0] <= ms[1] else 2)
        assert bsfit.fit_tile(ft, bsfit.PREC_MIX, S) == tile
        np.testing.assert_array_equal(Pm, one[0])
    finally:
        bsfit.set_fit_tile(ft, bsfit.PREC_MIX, S, 0)
