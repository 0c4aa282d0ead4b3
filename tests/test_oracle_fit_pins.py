"""Whole-fit pins of the oracle's LM control (fittypes 6, 7, 8) by what the
mathematics fixes.

The reference's fitter translation unit (CPUmleFit_LM.cpp, LM control at
:2555-2712 and its twins :2763-3158, :3163-3620) cannot be compiled here
(it includes MATLAB's mex.h), so no reference run pins the oracle's
fit_one.  These checks do:

* truth recovery: ROIs rendered without noise from the spline model by the
  generator's own numpy evaluation (independent of the oracle) at interior
  truths; the deviance is zero only at the truth, so a working Poisson-MLE
  LM must end there;
* stationarity: on noisy ROIs the fit must end where the Newton decrement
  J^T H^-1 J of the oracle's own evaluation pass is a vanishing fraction of
  its value at the start;
* efficiency: over seeded Poisson repeats of one ROI the spread of the
  estimates must match sqrt(CRLB), the reference's own validation
  (demo_orientation/bspline_based_5D_fitting_v3.m:86-98,
  demo_orientation/store/fitb5n3.m);
* planted mistakes (oracle.planted, bsf_oracle.c orc_plant): a flipped
  limiter sign, a dropped clamp fade, damping never raised, no step limiter,
  no lambda decay, a wrong Hessian weight, a flipped derivative -- each must
  fail at least one of the pins above.

What these do not fix: the trajectory itself (iteration counts, the exact
point where the 1e-6 relative-deviance test stops a fit).  The FP64 kernel's
short trajectories pin that against the oracle (test_gpu_parity.py).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from bsfit import gen

ROI_MAX_ITER = 1


def _coeffs(ft):
    if ft == 6:
        return gen.model_coeffs("astig3d"), None
    if ft == 7:
        return gen.model_coeffs("color4d"), None
    return gen.model_coeffs("dipole5d")


def _render(ft, T, S, c, c3):
    """Noise-free expected counts (S, S, N) Fortran order, float32."""
    xy = T[:, :2]
    if ft == 6:
        f = gen.eval_spline_rois(c, S, xy, T[:, 4:5])
    elif ft == 7:
        f = gen.eval_spline_rois(c, S, xy, T[:, 4:6])
    else:
        g2 = T[:, 7][:, None, None]
        f = (1 - g2) * gen.eval_spline_rois(c3, S, xy, T[:, 4:5]) + g2 * gen.eval_spline_rois(c, S, xy, T[:, 4:7])
    mu = T[:, 3][:, None, None] + T[:, 2][:, None, None] * np.maximum(f, 0.0)
    return np.asfortranarray(np.transpose(mu, (1, 2, 0)).astype(np.float32))


def _interior_truths(ft, N, S, nd, rng, wide=False):
    """Truths off every bound and wrap.  wide: the 3D stress set (positions
    up to 2.5 px off centre, 500-20,000 photons, z across the grid), whose
    far starts make the step limiter, the clamp fade and lambda matter."""
    c0 = (S - 1) / 2
    r = 2.5 if wide else 1.0
    x, y = c0 + rng.uniform(-r, r, N), c0 + rng.uniform(-r, r, N)
    if wide:
        ph, bg = np.exp(rng.uniform(np.log(5e2), np.log(2e4), N)), rng.uniform(1, 60, N)
    else:
        ph, bg = np.exp(rng.uniform(np.log(2e3), np.log(1e4), N)), rng.uniform(5, 30, N)
    if ft == 6:
        return np.stack([x, y, ph, bg, rng.uniform(3, 54, N) if wide else rng.uniform(12, 48, N)], 1)
    if ft == 7:
        return np.stack([x, y, ph, bg, rng.uniform(12, 48, N), rng.uniform(1.2, 2.3, N)], 1)
    return np.stack([x, y, ph, bg, rng.uniform(5, nd[2] - 5, N), rng.uniform(3, nd[3] - 3, N),
                     rng.uniform(2.5, nd[4] - 2.5, N), np.full(N, 0.75)], 1)


def _initP(ft, nd):
    if ft == 7:
        return np.array([nd[2] / 2.0, nd[3] / 2.0], np.float32)
    return None


def _errors(ft, P, T):
    NV = T.shape[1]
    coord = [0, 1] + list(range(4, NV))
    d = np.abs(np.asarray(P[:, :NV], np.float64) - T)
    return d[:, coord].max(1), (d[:, [2, 3]] / T[:, [2, 3]]).max(1)


def _ft8_fits(orc, im, T, c, c3, f64=True, iterations=200):
    """Fittype 8 from a start near each truth (one call per ROI: initP is one
    start per call, as the MEX's).  The start stays in the truth's basin: the
    dipole model is symmetric under azimuth + pi (7.5 samples of the w axis),
    which puts a second optimum there."""
    fit = orc.fit_f64 if f64 else orc.fit

    def one(n):
        ip = np.array([T[n, 4] + 0.7, T[n, 5] - 0.6, T[n, 6] + 0.5, 0.5], np.float32)
        P, _, _, st = fit(8, im[:, :, n:n + 1], iterations, c, ip, c3, status=True)
        return P[0], st[0]

    with ThreadPoolExecutor(8) as ex:
        out = list(ex.map(one, range(T.shape[0])))
    return np.array([o[0] for o in out]), np.array([o[1] for o in out])


def _ft6_stress(orc, f64=True):
    c, _ = _coeffs(6)
    nd = [s - 3 for s in c.shape]
    T = _interior_truths(6, 256, 13, nd, np.random.default_rng(11), wide=True)
    im = _render(6, T, 13, c, None)
    P, _, _ = (orc.fit_f64 if f64 else orc.fit)(6, im, 50, c, None, None, nthreads=8)
    return _errors(6, P, T)


def _ft7_interior(orc, f64=True):
    c, _ = _coeffs(7)
    nd = [s - 3 for s in c.shape]
    T = _interior_truths(7, 64, 15, nd, np.random.default_rng(5))
    im = _render(7, T, 15, c, None)
    P, _, _ = (orc.fit_f64 if f64 else orc.fit)(7, im, 200, c, _initP(7, nd), None, nthreads=8)
    return _errors(7, P, T)


def _ft8_interior(orc, N=32, f64=True):
    c, c3 = _coeffs(8)
    nd = [s - 3 for s in c.shape]
    T = _interior_truths(8, N, 13, nd, np.random.default_rng(5))
    im = _render(8, T, 13, c, c3)
    P, st = _ft8_fits(orc, im, T, c, c3, f64=f64)
    dc, dr = _errors(8, P, T)
    return dc, dr, (st & ROI_MAX_ITER) == 0


# ------------------------------------------------------------ truth recovery
def test_ft6_recovers_noise_free_truths(orc):
    """3D stress set with the reference's 50-iteration budget and its own
    starts (centre of mass, z = nz/2): every fit lands on the truth."""
    dc, dr = _ft6_stress(orc)
    print("ft6 FP64: max |d coord| %.2e, max rel d(N, bg) %.2e" % (dc.max(), dr.max()))
    assert dc.max() < 1e-5 and dr.max() < 1e-5
    dc, dr = _ft6_stress(orc, f64=False)
    ok = (dc < 1e-3) & (dr < 1e-3)
    print("ft6 float: %.3f within 1e-3, max |d coord| %.2e" % (ok.mean(), dc.max()))
    assert ok.mean() >= 0.97 and dc.max() < 1e-2


def test_ft7_recovers_noise_free_truths(orc):
    """Colour 4D at interior truths from the reference's start (initP = nz/2,
    nw/2).  The colour clamp starts at 1e-5 and opens only once the z step is
    small (CPUmleFit_LM.cpp:2989-2993), so the fit needs more than 50
    iterations: 200."""
    dc, dr = _ft7_interior(orc)
    print("ft7 FP64: max |d coord| %.2e, max rel d(N, bg) %.2e" % (dc.max(), dr.max()))
    assert dc.max() < 1e-4 and dr.max() < 1e-4
    dc, dr = _ft7_interior(orc, f64=False)
    print("ft7 float: median |d coord| %.2e, max %.2e" % (np.median(dc), dc.max()))
    assert np.median(dc) < 2e-3 and dc.max() < 0.05


def test_ft8_recovers_noise_free_truths(orc):
    """Vortex 5D (+3D blend) at interior truths from nearby starts: most fits
    land on the truth.  The rest stall short of it: every sign flip of a step
    halves that parameter's clamp for good (fade 0.5,
    CPUmleFit_LM.cpp:3191-3193, 3426-3447), so steps can shrink until the
    iteration limit, or until fittype 8's extra test (summed step < 1e-5,
    :3355) stops the fit where the deviance is not yet zero."""
    dc, dr, conv = _ft8_interior(orc)
    ok = (dc < 1e-3) & (dr < 1e-3)
    print("ft8 FP64: %d of %d ended by the convergence tests; %.3f of all within 1e-3 (max |d coord| %.2e)"
          % (conv.sum(), len(conv), ok.mean(), dc[ok].max()))
    assert ok.mean() >= 0.75
    assert dc[ok].max() < 1e-4


# ------------------------------------------------------------- stationarity
@pytest.mark.parametrize("ft,cfg,S,p90", [(6, 0, None, 1e-7), (7, 1, None, 1e-4), (8, 3, 13, 5e-3)])
def test_fits_end_stationary(orc, ft, cfg, S, p90):
    """Noisy ROIs (the seeded workloads, 200 iterations): at the interior
    fits that ended by their convergence tests, the Newton decrement
    J^T H^-1 J of the oracle's own evaluation pass is a vanishing fraction of
    its value at the start (median below 1e-6 for every fittype; measured
    2e-11, 3e-7, 6e-8).  The 90th percentile bar is per fittype: fittype 7's
    colour and fittype 8's stalls (steps shrunk by the clamp fade until the
    summed-step test stops the fit) leave some fits short of the optimum
    (measured p90 1e-9, 2e-5, 9e-4)."""
    w = gen.config(cfg, N=96, S=S, seed=gen.SEED0 + 1000 * cfg + 501)
    w = w[0] if isinstance(w, list) else w
    P, CR, LL, st = orc.fit_f64(ft, w.imstack, 200, w.coeff, w.initP, w.coeff3, nthreads=8, status=True)
    P0, _, _ = orc.fit_f64(ft, w.imstack, 0, w.coeff, w.initP, w.coeff3, nthreads=8)
    NV = CR.shape[1]
    nd = [s - 3 for s in w.coeff.shape]
    # off the bounds and wraps (CPUmleFit_LM.cpp:2631-2634, 3021-3026, 3449-3472)
    box = {4: (0.5, nd[2] - (4.5 if ft == 6 else 1.0))}
    if ft == 7:
        box[5] = (0.5, nd[3] - 1.0)
    if ft == 8:
        box.update({5: (1.5, nd[3] - 1.5), 6: (1.5, nd[4] - 1.0), 7: (0.02, 0.98)})
    inner = np.ones(w.N, bool)
    for a, (lo, hi) in box.items():
        inner &= (P[:, a] > lo) & (P[:, a] < hi)
    if ft == 8:
        inner &= np.abs(P[:, :2] - (w.S - 1) / 2).max(1) < w.S / 4 - 0.05
    sel = np.nonzero(inner & ((st & ROI_MAX_ITER) == 0))[0]
    ratios = []
    for n in sel:
        _, J1, H1, _ = orc.eval_pass(ft, w.imstack[:, :, n], P[n, :NV], w.coeff, w.coeff3, f64=True)
        _, J0, H0, _ = orc.eval_pass(ft, w.imstack[:, :, n], P0[n, :NV], w.coeff, w.coeff3, f64=True)
        ratios.append(float(J1 @ np.linalg.solve(H1, J1)) / float(J0 @ np.linalg.solve(H0, J0)))
    ratios = np.array(ratios)
    print("fittype %d: %d interior converged fits; decrement ratio median %.2e, p90 %.2e, max %.2e"
          % (ft, len(sel), np.median(ratios), np.percentile(ratios, 90), ratios.max()))
    assert len(sel) >= 0.2 * w.N
    assert np.median(ratios) < 1e-6
    assert np.percentile(ratios, 90) < p90


# --------------------------------------------------------- STD versus CRLB
STD_CASES = {
    # fittype: (truth, S, parameters whose spread is asserted)
    6: (np.array([6.2, 5.7, 3000.0, 20.0, 25.0]), 13, [0, 1, 2, 3, 4]),
    7: (np.array([7.3, 6.8, 3000.0, 20.0, 25.0, 1.8]), 15, [0, 1, 2, 3, 4, 5]),
    8: (np.array([6.3, 5.8, 8000.0, 20.0, 10.0, 9.0, 5.0, 0.75]), 13, [2, 3]),
}


@pytest.mark.parametrize("ft", [6, 7, 8])
def test_spread_matches_crlb(orc, ft):
    """600 Poisson repeats of one interior ROI: the STD of the estimates
    against sqrt(CRLB) at the truth (the inverse of the Fisher matrix that
    the oracle's evaluation pass forms when the data equal the model) within
    [0.88, 1.12] (a 600-sample STD has ~3 % sampling error), bias below a
    quarter of sqrt(CRLB).  Fittype 8: photons and background only; its
    positions, orientation and g2 are strongly coupled in this 5D + 3D
    blend, and the reference's estimator does not reach the bound for them
    here (STD/sqrt(CRLB) 0.2-2)."""
    T, S, params = STD_CASES[ft]
    c, c3 = _coeffs(ft)
    nd = [s - 3 for s in c.shape]
    R = 600
    mu = _render(ft, T[None, :], S, c, c3)[:, :, 0]
    _, _, H, _ = orc.eval_pass(ft, mu, T, c, c3, f64=True)
    crlb_t = np.diag(np.linalg.inv(H))
    rng = np.random.default_rng(2026 + ft)
    im = np.asfortranarray(rng.poisson(np.repeat(mu[:, :, None], R, 2)).astype(np.float32))
    ip = _initP(ft, nd) if ft != 8 else np.array([T[4] + 0.5, T[5] - 0.5, T[6] + 0.5, 0.5], np.float32)
    P, CR, LL = orc.fit_f64(ft, im, 100, c, ip, c3, nthreads=8)
    NV = len(T)
    assert np.all(np.isfinite(P[:, :NV]))
    ratio = P[:, :NV].std(0) / np.sqrt(crlb_t)
    bias = (P[:, :NV].mean(0) - T) / np.sqrt(crlb_t)
    med = np.median(CR, 0) / crlb_t
    print("fittype %d: STD/sqrt(CRLB) %s; bias/sqrt(CRLB) %s; median reported CRLB / CRLB %s"
          % (ft, np.round(ratio, 3), np.round(bias, 2), np.round(med, 3)))
    for a in params:
        assert 0.88 <= ratio[a] <= 1.12, (a, ratio[a])
        assert abs(bias[a]) < 0.25, (a, bias[a])
        assert 0.9 <= med[a] <= 1.1, (a, med[a])


# -------------------------------------------------------- planted mistakes
PLANTS = {1: "limiter sign flipped", 2: "clamp fade dropped", 3: "damping never raised on a rejected step",
          4: "step limiter dropped", 5: "lambda decay dropped", 6: "Hessian weight d/mu",
          7: "x derivative sign flipped"}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_fails_a_pin(orc, plant):
    """Each planted LM-control or evaluation mistake makes the fittype-6
    stress set or the fittype-7 interior set miss its truths (bars as in
    the recovery tests above); the three that break the estimator itself
    (1, 6, 7) make every fittype miss, fittype 8 included."""
    with orc.planted(plant):
        dc6, dr6 = _ft6_stress(orc)
        dc7, dr7 = _ft7_interior(orc)
        caught6 = not (dc6.max() < 1e-5 and dr6.max() < 1e-5)
        caught7 = not (dc7.max() < 1e-4 and dr7.max() < 1e-4)
        caught8 = None
        if plant in (1, 6, 7):
            dc8, dr8, _ = _ft8_interior(orc, N=8)
            caught8 = not (np.mean((dc8 < 1e-3) & (dr8 < 1e-3)) >= 0.75)
    print("plant %d (%s): fittype 6 %s, fittype 7 %s, fittype 8 %s" % (plant, PLANTS[plant], caught6, caught7, caught8))
    assert caught6 or caught7
    if plant in (1, 6, 7):
        assert caught6 and caught7 and caught8
