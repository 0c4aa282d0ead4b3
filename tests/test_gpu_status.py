"""Per-ROI status flags (bsf_fit_host_status / bsf_fit_device_status).

The status entry points run the same fits as bsf_fit_host / bsf_fit_device
(outputs equal bit for bit) and add one byte per ROI; each flag is checked
against what the outputs themselves show:
  BSF_ROI_MAX_ITER        <=> P[:, NV] == iterations
  BSF_ROI_FISHER_SINGULAR <=> some CRLB entry is not a finite positive number
  BSF_ROI_NONFINITE       <=> some parameter or the log-likelihood is not finite
BSF_ROI_SOLVE_FAILED (the last iteration's damped solve had no solution) and
BSF_ROI_FIRST_SOLVE_FAILED (the first iteration's) are not visible in the
outputs: an all-zero ROI, whose centre-of-mass start is 0/0, must raise both
(with the other three), a normal workload must raise the first rarely, and
with one iteration the two are the same solve."""
import numpy as np
import pytest

import bsfit
from bsfit import gen

pytestmark = pytest.mark.gpu


def _check_flags(P, CR, LL, st, NV, iterations):
    maxit = P[:, NV] == iterations
    sing = ~np.all(np.isfinite(CR) & (CR > 0), axis=1)
    nonf = ~(np.all(np.isfinite(P[:, :NV]), axis=1) & np.isfinite(LL))
    np.testing.assert_array_equal((st & bsfit.ROI_MAX_ITER) != 0, maxit)
    np.testing.assert_array_equal((st & bsfit.ROI_FISHER_SINGULAR) != 0, sing)
    np.testing.assert_array_equal((st & bsfit.ROI_NONFINITE) != 0, nonf)
    assert np.all(st < 32)
    if iterations == 1:
        np.testing.assert_array_equal((st & bsfit.ROI_FIRST_SOLVE_FAILED) != 0, (st & bsfit.ROI_SOLVE_FAILED) != 0)


@pytest.mark.parametrize("ft,cfg,S,N", [(6, 0, None, 256), (7, 1, None, 128), (8, 3, 13, 64)])
@pytest.mark.parametrize("prec", [bsfit.PREC_MIX, bsfit.PREC_F64])
@pytest.mark.parametrize("iterations", [50, 3, 1])
def test_status_flags_match_outputs(ft, cfg, S, N, prec, iterations):
    w = gen.config(cfg, N=N, S=S)
    w = w[0] if isinstance(w, list) else w
    print("case fittype=%d S=%d N=%d precision=%d iterations=%d" % (ft, w.S, w.N, prec, iterations), flush=True)
    NV = bsfit.num_params(ft)
    ref = bsfit.fit(ft, w.imstack, iterations, w.coeff, w.initP, w.coeff3, precision=prec)
    P, CR, LL, st = bsfit.fit_status(ft, w.imstack, iterations, w.coeff, w.initP, w.coeff3, precision=prec)
    for a, b in zip((P, CR, LL), ref):
        np.testing.assert_array_equal(a, b)
    _check_flags(P, CR, LL, st, NV, iterations)
    counts = {name: int(np.sum((st & bit) != 0)) for name, bit in
              (("max_iter", bsfit.ROI_MAX_ITER), ("solve_failed", bsfit.ROI_SOLVE_FAILED),
               ("fisher_singular", bsfit.ROI_FISHER_SINGULAR), ("nonfinite", bsfit.ROI_NONFINITE))}
    print("flags", counts, flush=True)
    if iterations == 50:
        assert counts["solve_failed"] <= 0.25 * N
    else:
        assert counts["max_iter"] >= 0.5 * N


def test_all_zero_roi_raises_every_flag():
    w = gen.config(0, N=8)
    im = np.array(w.imstack, np.float32)
    im[:, :, 3] = 0.0
    P, CR, LL, st = bsfit.fit_status(6, im, 50, w.coeff)
    print("status", st.tolist(), flush=True)
    _check_flags(P, CR, LL, st, 5, 50)
    assert st[3] == (bsfit.ROI_MAX_ITER | bsfit.ROI_SOLVE_FAILED | bsfit.ROI_FISHER_SINGULAR | bsfit.ROI_NONFINITE
                     | bsfit.ROI_FIRST_SOLVE_FAILED)
    others = np.delete(np.arange(8), 3)
    ref = bsfit.fit(6, im, 50, w.coeff)
    for a, b in zip((P, CR, LL), ref):
        np.testing.assert_array_equal(a, b)
    assert np.all((st[others] & bsfit.ROI_NONFINITE) == 0)


def test_device_status_equals_host_status():
    from parity_util import Hip
    hip = Hip()
    w = gen.config(1, N=96)
    N = w.N
    d = hip.put(np.asfortranarray(w.imstack, np.float32))
    c = hip.put(np.asfortranarray(w.coeff, np.float32))
    P, CR, LL, st = hip.alloc(7 * N * 4), hip.alloc(6 * N * 4), hip.alloc(N * 4), hip.put(np.full(N, 255, np.uint8))
    try:
        bsfit.fit_device(7, d, w.S, N, 50, c, w.coeff.shape, P, CR, LL, initP=w.initP, stream=None,
                         status_ptr=st)
        assert hip.h.hipDeviceSynchronize() == 0
        Pd = hip.get(P, 7 * N, np.float32).reshape(7, N).T
        std = hip.get(st, N, np.uint8)
    finally:
        hip.free(d, c, P, CR, LL, st)
    Ph, CRh, LLh, sth = bsfit.fit_status(7, w.imstack, 50, w.coeff, w.initP)
    np.testing.assert_array_equal(Pd, Ph)
    np.testing.assert_array_equal(std, sth)


@pytest.mark.parametrize("ft,strict", [(7, True), (8, False)])
def test_multistart_status_follows_the_kept_result(ft, strict):
    """Multi-start with status: outputs equal bsf_fit_multistart_host's, and
    the flags describe the kept result (the kernel writes them only where it
    writes that ROI's outputs)."""
    if ft == 7:
        w = gen.config(1, N=96)
        starts = np.array([[30.5, 1.5], [24.4, 2.5], [36.6, 1.0]], np.float32)
    else:
        w = gen.config(3, N=48, S=13)[0]
        starts = np.array([[10.5, 6.0, 4.5, 0.75], [10.5, 12.0, 4.5, 0.75]], np.float32)
    ref = bsfit.fit_multistart(ft, w.imstack, 50, w.coeff, starts, coeff3=w.coeff3, strict=strict)
    P, CR, LL, st = bsfit.fit_multistart(ft, w.imstack, 50, w.coeff, starts, coeff3=w.coeff3, strict=strict,
                                         status=True)
    for a, b in zip((P, CR, LL), ref):
        np.testing.assert_array_equal(a, b)
    _check_flags(P, CR, LL, st, bsfit.num_params(ft), 50)


@pytest.mark.parametrize("N", [200, 1500])   # the wave-per-channel kernel below 1,024 ROIs, one wave per ROI above
@pytest.mark.parametrize("prec", [bsfit.PREC_MIX, bsfit.PREC_F64])
def test_4pi_status_flags_match_outputs(N, prec):
    w = gen.make_rois_4pi(N, 13, 99, gen.model_coeffs("4pi"), shared=(1, 1, 1, 0, 1, 1))
    NV = bsfit.num_params_4pi(w.shared, w.nch)
    args = (w.imstack, w.shared, 50, w.coeff, w.phaseShift)
    kw = dict(initZ=w.initZ, initPhi=w.initPhi, precision=prec)
    ref = bsfit.fit_4pi(*args, **kw)
    P, CR, LL, st = bsfit.fit_4pi(*args, status=True, **kw)
    for a, b in zip((P, CR, LL), ref):
        np.testing.assert_array_equal(a, b)
    _check_flags(P, CR, LL, st, NV, 50)
    print("4Pi N=%d precision=%d flags: max_iter %d solve_failed %d" % (
        N, prec, int(np.sum(st & bsfit.ROI_MAX_ITER != 0)), int(np.sum(st & bsfit.ROI_SOLVE_FAILED != 0))), flush=True)
