"""Two ROIs per wavefront (tile 2) against one ROI per wavefront (tile 1),
and the autotuner that picks between them (bsf_tune).

Fittypes 6 and 7: in a tile-2 launch each half-wave fits its own ROI; a
lane there holds the pixel sums of two lanes of the one-ROI kernel and adds
them in the same order, so the two tiles must agree bit for bit:
parameters, CRLB, log-likelihood, status flags and evaluation counts, for
uniform launches (including an odd ROI count, whose last wave has an empty
half) and for a launch of mixed sizes, in both precision modes.

Fittype 8 (FP32): a tile-2 lane keeps one set of sums over its half's
pixels (a second set does not fit the registers), so its pixel sums are
added in another order and the bits differ; its fits must then meet the
same bars against the FP64 oracle as the one-ROI kernel's
(test_gpu_parity.py: agreement against the reference's own float-vs-FP64
spread, coordinate, N / bg and CRLB errors at the same optimum, final
log-likelihood).  FP64 fittype 8 has one tile."""
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


def test_tile2_request_for_fp64_fittype8_runs_one_roi_per_wave():
    c, c3 = _model(8)
    w = gen.make_rois(8, 24, 11, 977, c, c3)
    data = np.asfortranarray(w.imstack, np.float32).ravel(order="F")
    hip = Hip()
    one = _run(hip, 8, data, 11, w.N, c, c3, bsfit.PREC_F64, 1, initP=w.initP)
    two = _run(hip, 8, data, 11, w.N, c, c3, bsfit.PREC_F64, 2, initP=w.initP)
    _same(one, two)


def _fit8_bars(P, CR, LL, w, orc):
    """The one-ROI kernel's MIX bars against the FP64 oracle
    (test_gpu_parity.py), on fittype-8 fits."""
    from parity_util import compare, errors, fmt_errors
    of = orc.fit(8, w.imstack, 50, w.coeff, w.initP, w.coeff3, nthreads=16)
    od = orc.fit_f64(8, w.imstack, 50, w.coeff, w.initP, w.coeff3, nthreads=16)
    sref = compare(8, *of, *od)["agree"]
    s64 = compare(8, P, CR, LL, *od)["agree"]
    e = errors(8, P, CR, od[0], od[1])
    e_ref = errors(8, of[0], of[1], od[0], od[1])
    print("tile 2: ~orcF64 %.3f (reference float vs FP64 %.3f) %s" % (s64, sref, fmt_errors(e)), flush=True)
    assert np.all(np.isfinite(P)), "non-finite parameters"
    assert not np.any(np.isnan(CR)), "NaN CRLB"
    assert s64 >= sref - 0.05
    assert e["coord"][0] <= 1e-4 and e["coord"][0] <= 1.5 * e_ref["coord"][0] + 1e-6, (e["coord"], e_ref["coord"])
    assert e["coord"][1] <= 1.5 * e_ref["coord"][1] + 1e-5, (e["coord"], e_ref["coord"])
    assert e["same"] >= e_ref["same"] - 0.05, (e["same"], e_ref["same"])
    for k in ("rel", "crlb"):
        assert e[k][0] <= 1e-4 and e[k][1] <= 1e-3, (k, e[k])
    LLd = np.asarray(LL, np.float64)
    LLo = np.asarray(of[2], np.float64)
    assert np.mean(np.isfinite(LLd) & (LLd >= LLo - 1e-3 * np.abs(LLo) - 1e-3)) >= 0.75


@pytest.mark.parametrize("S,N", [(11, 257), (17, 192)])
def test_fittype8_tile2_meets_the_parity_bars(orc, S, N):
    w = gen.config(3, N=N, S=S)
    w = w[0] if isinstance(w, list) else w
    print("case fittype=8 tile 2 S=%d N=%d" % (S, N), flush=True)
    data = np.asfortranarray(w.imstack, np.float32).ravel(order="F")
    hip = Hip()
    P, CR, LL, st, ev = _run(hip, 8, data, S, N, w.coeff, w.coeff3, bsfit.PREC_MIX, 2, initP=w.initP)
    P = P.reshape(9, N).T
    CR = CR.reshape(8, N).T
    _fit8_bars(P, CR, LL, w, orc)
    # the status flags and evaluation counts describe these fits
    assert np.all(ev >= 2) and np.all((ev == 2) == (P[:, 8] == 0))
    assert np.all(((st & 1) != 0) == (P[:, 8] >= 50))


def test_fittype8_tile2_mixed_sizes(orc):
    """A launch of mixed sizes (largest first, an odd count) in two-ROI waves
    fits each ROI as a uniform tile-2 launch of its size does."""
    c, c3 = _model(8)
    sizes = (17, 15, 13, 11)
    groups = [gen.make_rois(8, 33, S, 960 + S, c, c3) for S in sizes]
    data = np.concatenate([np.asfortranarray(w.imstack).ravel(order="F") for w in groups]).astype(np.float32)
    S_list = np.concatenate([[w.S] * w.N for w in groups])
    hip = Hip()
    N = len(S_list)
    mixed = _run(hip, 8, data, max(sizes), N, c, c3, bsfit.PREC_MIX, 2, initP=groups[0].initP,
                 rois=bsfit.roi_table(S_list))
    off = 0
    for w in groups:
        one = _run(hip, 8, np.asfortranarray(w.imstack, np.float32).ravel(order="F"), w.S, w.N, c, c3,
                   bsfit.PREC_MIX, 2, initP=w.initP)
        sl = slice(off, off + w.N)
        np.testing.assert_array_equal(mixed[0].reshape(9, N)[:, sl], one[0].reshape(9, w.N), err_msg="S=%d" % w.S)
        np.testing.assert_array_equal(mixed[2][sl], one[2], err_msg="S=%d" % w.S)
        off += w.N


@pytest.mark.parametrize("ft,S", [(6, 11), (7, 13), (8, 11)])
def test_tune_records_a_tile_and_keeps_results(ft, S):
    c, c3 = _model(ft)
    w = gen.make_rois(ft, 2048, S, 990 + S, c, c3)
    N, NV = w.N, bsfit.num_params(ft)
    hip = Hip()
    data = np.asfortranarray(w.imstack, np.float32).ravel(order="F")
    d = hip.put(data)
    dc = hip.put(np.asfortranarray(c, np.float32))
    dc3 = hip.put(np.asfortranarray(c3, np.float32)) if c3 is not None else None
    P, CR, LL = hip.alloc((NV + 1) * N * 4), hip.alloc(NV * N * 4), hip.alloc(N * 4)
    try:
        tile, ms = bsfit.tune(ft, d, S, N, 50, dc, c.shape, P, CR, LL, initP=w.initP, reps=3, coeff3_ptr=dc3)
        Pm = hip.get(P, (NV + 1) * N, np.float32)
    finally:
        hip.free(d, dc, dc3, P, CR, LL)
    # res holds the fits of the tile it chose
    one = _run(hip, ft, data, S, N, c, c3, bsfit.PREC_MIX, tile, initP=w.initP)
    print("case fittype=%d S=%d: tile 1 %.3f ms, tile 2 %.3f ms -> tile %d" % (ft, S, ms[0], ms[1], tile), flush=True)
    try:
        assert tile in (1, 2) and ms[0] > 0 and ms[1] > 0
        assert tile == (1 if ms[0] <= ms[1] else 2)
        assert bsfit.fit_tile(ft, bsfit.PREC_MIX, S) == tile
        np.testing.assert_array_equal(Pm, one[0])
    finally:
        bsfit.set_fit_tile(ft, bsfit.PREC_MIX, S, 0)
