"""Whole fits on the GPU (libbsfit.so through the C-ABI) against the CPU
oracle on the same seeded inputs.

The reference's LM control (accept / reject on |d err / err| < 1e-6 and
err > 1.5 old err, multiplicative damping, sign-flip clamp fading, bounds
and wraps) makes iteration counts and late trajectory steps sensitive to
rounding: the oracle's own float build and its FP64 build agree (P within
1e-4 px / 1e-4 relative, CRLB within 1e-4 relative, same iteration count) on
only 87 % (fittype 6), 28 % (7) and 20-27 % (8) of ROIs.  So per-ROI
agreement is asserted as rates, each against the reference's own spread:

* FP64 kernel vs FP64 oracle (same arithmetic class, different summation
  order): floors just under the measured 0.99 / 0.53-0.55 / 0.86-0.90;
* FP32 kernel (MIX: FP32 evaluation, FP64 solve) vs the oracle's float
  build (the reference's arithmetic): at least the float build's own
  agreement with the FP64 build, less 0.10;
* MIX vs the FP64 oracle: at least that same self-agreement, less 0.05;
  coordinate errors (median, 90th percentile) no larger than the float
  build's own against the FP64 oracle (x 1.5), the median within 1e-4; at
  the same optimum (coordinates within 1e-3, reached as often as by the
  float build) N, bg and CRLB within 1e-4 relative at the median, 1e-3 at
  the 90th percentile and no worse than the float build's (x 1.5);
* the objective: the GPU's final log-likelihood not worse than the
  oracle's on the large majority of ROIs; accuracy against the simulation
  truth no worse than the oracle's;
* CRLB: no NaN; an infinite entry only for a parameter on a mirror-padded
  bound (no information there).

The model, derivatives and pixel sums themselves are pinned strictly by
test_gpu_eval.py; the first LM steps by the short trajectories below.
"""
import numpy as np
import pytest

from parity_util import compare, errors, fmt_errors, crlb_nan_free

pytestmark = pytest.mark.gpu

# (fittype, cfg, S, N, min FP64-vs-FP64 agreement)
CASES = [
    (6, 0, None, 512, 0.97),
    (7, 1, None, 384, 0.45),
    (8, 3, 11, 256, 0.78),
    (8, 3, 17, 192, 0.78),
]

# parameter index and bound value at which a parameter has no information
# (the spline's derivative vanishes on its mirror-padded first sample)
NO_INFO_BOUNDS = {6: [(4, 0.0)], 7: [(4, 0.0), (5, 0.0)], 8: [(4, 0.0), (6, 1.0)]}


def _ll_not_worse(LLg, LLo):
    LLg = np.asarray(LLg, np.float64)
    LLo = np.asarray(LLo, np.float64)
    ok = np.isfinite(LLg) & (LLg >= LLo - 1e-3 * np.abs(LLo) - 1e-3)
    return float(ok.mean())


def _check_crlb(ft, P, CR):
    assert crlb_nan_free(CR), "NaN CRLB"
    inf = ~np.isfinite(CR)
    on_bound = np.zeros_like(inf)
    for a, v in NO_INFO_BOUNDS[ft]:
        on_bound[:, a] = P[:, a] == v
    # fittype 8 at u = 1: the polar angle is 0, the azimuth has no information either
    if ft == 8:
        on_bound[:, 5] |= P[:, 6] == 1.0
    assert np.all(on_bound[inf]), "infinite CRLB off the bounds"
    return float(np.mean(inf.any(1)))


@pytest.mark.parametrize("ft,cfg,S,N,min_f64", CASES)
def test_fit_matches_oracle(orc, ft, cfg, S, N, min_f64):
    import bsfit
    from bsfit import gen
    w = gen.config(cfg, N=N, S=S)
    w = w[0] if isinstance(w, list) else w
    print("case fittype=%d cfg=%d S=%d N=%d" % (ft, cfg, w.S, w.N), flush=True)
    gm = bsfit.fit(ft, w.imstack, 50, w.coeff, w.initP, w.coeff3)
    gd = bsfit.fit(ft, w.imstack, 50, w.coeff, w.initP, w.coeff3, precision=bsfit.PREC_F64)
    of = orc.fit(ft, w.imstack, 50, w.coeff, w.initP, w.coeff3, nthreads=16)
    od = orc.fit_f64(ft, w.imstack, 50, w.coeff, w.initP, w.coeff3, nthreads=16)
    s64 = compare(ft, *gd, *od)
    smx = compare(ft, *gm, *of)
    smx64 = compare(ft, *gm, *od)
    sref = compare(ft, *of, *od)
    llm = _ll_not_worse(gm[2], of[2])
    e_mix = errors(ft, gm[0], gm[1], od[0], od[1])
    e_ref = errors(ft, of[0], of[1], od[0], od[1])
    print("agreement: gpuF64~orcF64 %.3f  gpuMIX~orcF32 %.3f  gpuMIX~orcF64 %.3f  orcF32~orcF64 %.3f"
          % (s64["agree"], smx["agree"], smx64["agree"], sref["agree"]))
    print("gpuMIX - orcF64:", fmt_errors(e_mix))
    print("orcF32 - orcF64:", fmt_errors(e_ref))
    print("LL(gpuMIX) >= LL(orcF32) on %.3f" % llm, flush=True)
    assert np.all(np.isfinite(gm[0])) and np.all(np.isfinite(gd[0])), "non-finite parameters"
    inf_mix = _check_crlb(ft, gm[0], gm[1])
    inf_f64 = _check_crlb(ft, gd[0], gd[1])
    print("ROIs with a parameter without information (CRLB inf): MIX %.3f, F64 %.3f" % (inf_mix, inf_f64))
    assert s64["agree"] >= min_f64
    assert smx["agree"] >= sref["agree"] - 0.10
    assert smx64["agree"] >= sref["agree"] - 0.05
    # coordinates over all ROIs: median within the tolerance, median and 90th
    # percentile no larger than the reference arithmetic's own (x 1.5)
    assert e_mix["coord"][0] <= 1e-4, e_mix["coord"]
    assert e_mix["coord"][0] <= 1.5 * e_ref["coord"][0] + 1e-6, (e_mix["coord"], e_ref["coord"])
    assert e_mix["coord"][1] <= 1.5 * e_ref["coord"][1] + 1e-5, (e_mix["coord"], e_ref["coord"])
    # at the same optimum (as often as the reference's arithmetic reaches it):
    # N, bg and CRLB within 1e-4 relative at the median and 1e-3 at the 90th
    # percentile, and no worse than the reference's own (x 1.5)
    assert e_mix["same"] >= e_ref["same"] - 0.05, (e_mix["same"], e_ref["same"])
    for k in ("rel", "crlb"):
        assert e_mix[k][0] <= 1e-4, (k, e_mix[k])
        assert e_mix[k][1] <= 1e-3, (k, e_mix[k])
        assert e_mix[k][1] <= 1.5 * e_ref[k][1] + 1e-5, (k, e_mix[k], e_ref[k])
    assert llm >= 0.75
    eg = np.median(np.abs(gm[0][:, :2] - w.truth[:, :2]), 0)
    eo = np.median(np.abs(of[0][:, :2] - w.truth[:, :2]), 0)
    print("median |x,y - truth|: gpu MIX", eg, "oracle F32", eo, flush=True)
    assert np.all(eg <= 1.2 * eo + 1e-3)


def test_c_abi_errors():
    import bsfit
    from bsfit import gen
    w = gen.config(0, N=4)
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.fit(5, w.imstack, 50, w.coeff)
    assert e.value.code == 2
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.fit(7, w.imstack, 50, w.coeff)        # 3-D coefficients for fittype 7
    assert e.value.code == 4
    c5, c3 = gen.model_coeffs("dipole5d")
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.fit(8, w.imstack, 50, c5)             # fittype 8 without coeff3
    assert e.value.code == 1
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.fit(6, np.zeros((53, 53, 2), np.float32), 50, w.coeff)   # sz > 51
    assert e.value.code == 3
    P, CR, LL = bsfit.fit(6, w.imstack[:, :, :0], 50, w.coeff)   # empty stack
    assert P.shape == (0, 6)


def test_single_roi_and_iterations_zero(orc):
    """A 2-D imstack is one ROI (mexFunction.cpp:257); iterations = 0 returns
    the initial estimates with the iteration count 0, as the oracle."""
    import bsfit
    from bsfit import gen
    w = gen.config(0, N=3)
    P1, _, _ = bsfit.fit(6, w.imstack[:, :, 0], 50, w.coeff)
    assert P1.shape == (1, 6)
    P0, CR0, LL0 = bsfit.fit(6, w.imstack, 0, w.coeff, precision=bsfit.PREC_F64)
    Po, CRo, LLo = orc.fit_f64(6, w.imstack, 0, w.coeff)
    assert np.all(P0[:, 5] == 0)
    assert np.allclose(P0[:, :5], Po[:, :5], rtol=1e-6, atol=1e-5)
    assert np.allclose(CR0, CRo, rtol=1e-4)
    assert np.allclose(LL0, LLo, rtol=1e-5, atol=1e-3)


# fittype 8 FP64 at S = 11 and 17 runs the kernel capped at two waves per
# SIMD, at S = 13 the uncapped one
@pytest.mark.parametrize("ft,cfg,S", [(6, 0, None), (7, 1, None), (8, 3, 11), (8, 3, 13), (8, 3, 17)])
@pytest.mark.parametrize("iters", [1, 2, 4])
def test_short_trajectories_match_f64_oracle(orc, ft, cfg, S, iters):
    """The first LM steps (initial estimates, accept / reject, damping,
    solve, limiter, clamp fading, bounds, wraps and the fittype-8 extra
    convergence test) before rounding differences grow: the FP64 kernel
    against the FP64 oracle to 1e-5.  Fittype 7's first pass copies dudt[4]
    into dudt[5] (CPUmleFit_LM.cpp:2896), so the first Hessian has two equal
    rows and whether its Cholesky passes is decided by rounding
    (test_oracle_pins.py::test_ft7_first_hessian_singular): both sides report
    that outcome (status bit 16, a rounding coin flip on each side).  Where
    both failed it, both continue from the same state with the damping
    raised tenfold, where the system is no longer singular: the trajectories
    are compared there."""
    import bsfit
    from bsfit import gen
    w = gen.config(cfg, N=96 if ft == 7 else 48, S=S)
    w = w[0] if isinstance(w, list) else w
    P, CR, LL, st = bsfit.fit_status(ft, w.imstack, iters, w.coeff, w.initP, w.coeff3, precision=bsfit.PREC_F64)
    Po, CRo, LLo, sto = orc.fit_f64(ft, w.imstack, iters, w.coeff, w.initP, w.coeff3, nthreads=16, status=True)
    NV = CR.shape[1]
    rel = np.abs(P[:, :NV] - Po[:, :NV]) / np.maximum(np.abs(Po[:, :NV]), 1.0)
    ok = (rel.max(1) < 1e-5) & (P[:, NV] == Po[:, NV])
    first = (st & bsfit.ROI_FIRST_SOLVE_FAILED) != 0
    firsto = (sto & bsfit.ROI_FIRST_SOLVE_FAILED) != 0
    both = first & firsto
    print("fittype %d S %d iterations %d: first solve failed GPU %.3f oracle %.3f, both %.3f; "
          "agree %.3f (where both failed it %.3f)" % (ft, w.S, iters, first.mean(), firsto.mean(), both.mean(),
                                                     ok.mean(), ok[both].mean() if both.any() else -1), flush=True)
    if ft == 7:
        assert both.mean() >= 0.4
        assert ok[both].mean() >= 0.95
    else:
        assert np.all(first == firsto)
        assert ok.mean() >= 0.95


@pytest.mark.parametrize("ft,cfg,N,shift,min_f64", [(6, 0, 128, 12.0, 0.9), (7, 1, 96, 12.0, 0.25)])
def test_negative_pixels_match_oracle(orc, ft, cfg, N, shift, min_f64):
    """Data shifted below zero, as offset-subtracted camera frames are: the
    initial values use the raw values (centre of mass, minimum), the
    likelihood treats d <= 0 as 0 (CPUmleFit_LM.cpp:2530-2549); the GPU
    against the oracle at the bars of test_fit_matches_oracle."""
    import bsfit
    from bsfit import gen
    w = gen.config(cfg, N=N)
    im = np.asfortranarray(w.imstack - np.float32(shift), np.float32)
    frac_neg = float(np.mean(im <= 0))
    print("case fittype=%d N=%d shift=%g: %.1f%% of pixels <= 0" % (ft, N, shift, 100 * frac_neg), flush=True)
    assert frac_neg > 0.05
    gm = bsfit.fit(ft, im, 50, w.coeff, w.initP)
    gd = bsfit.fit(ft, im, 50, w.coeff, w.initP, precision=bsfit.PREC_F64)
    of = orc.fit(ft, im, 50, w.coeff, w.initP, nthreads=16)
    od = orc.fit_f64(ft, im, 50, w.coeff, w.initP, nthreads=16)
    s64 = compare(ft, *gd, *od)
    llm = _ll_not_worse(gm[2], of[2])
    print("gpuF64 vs orcF64", s64, "LL(gpuMIX) >= LL(orcF32) on %.3f" % llm, flush=True)
    assert np.all(np.isfinite(gm[0])) and np.all(np.isfinite(gd[0]))
    assert s64["agree"] >= min_f64
    assert llm >= 0.75
