"""Pins of the coefficient-builder oracle (oracle/bsarray.py), the restatement
of the reference's bsarray (MultiD_bspline/bsarray/*.m, degree 3): MATLAB is
not in this image, so the reference cannot run; what the mathematics fixes
pins it instead.

* The coefficients c of every line solve, with whole-sample mirror ends,
    (c[i-1] + 4 c[i] + c[i+1]) / 6 + lambda (c[i-2] - 4 c[i-1] + 6 c[i] - 4 c[i+1] + c[i+2]) = x[i]
  (the transfer function 1 / (B3(z) + lambda (2 - z - 1/z)^2) that the
  reference's pole polynomial, getBSplineFiltCoeffs.m, factors).  A wrong
  pole, gain, sign, boundary start or axis order breaks it; the last test
  plants a wrong gain in the oracle to show that it does.
* The interpolating pole and gain (sqrt(3) - 2, 6), the smoothing poles as
  the roots of the reference's polynomial inside the unit circle.
* Mirror padding (directFilter.m:286-297): indices -1..M+1 of the
  whole-sample reflection.
* Interpolation at the knots (lambda = 0) and agreement with the model
  generator the fit tests use.
"""
import math

import numpy as np
import pytest

from oracle import bsarray as B


SHAPES = [(7,), (12, 9), (8, 6, 5), (6, 5, 4, 7), (5, 4, 4, 3, 3)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("lam", [0.0, 0.1, 0.2, 1.0, "demo"])
def test_operator_identity(shape, lam):
    rng = np.random.default_rng(len(shape) * 7 + 1)
    x = rng.standard_normal(shape)
    if lam == "demo":   # the colour demo's per-axis lambda, bspline_based_color_classification_v3.m:28
        lam = ([0.0, 0.0, 0.1, 0.2, 0.0])[:len(shape)]
    c = B.prefilter(x, lam)
    r = B.operator(c, lam)
    assert np.abs(r - x).max() < 1e-11


def test_poles_and_gain():
    z, g = B.poles(0.0)
    assert len(z) == 1 and abs(z[0] - (math.sqrt(3.0) - 2.0)) < 1e-16 and g == 6.0
    for lam in (0.01, 0.1, 0.2, 1.0, 10.0):
        z, g = B.poles(lam)
        assert len(z) == 2 and g == pytest.approx(1.0 / lam)
        p = [1.0, -4.0 + 1.0 / (6.0 * lam), 6.0 + 2.0 / (3.0 * lam), -4.0 + 1.0 / (6.0 * lam), 1.0]
        for zi in z:
            assert abs(zi) < 1.0
            assert abs(np.polyval(p, zi)) < 1e-12 * max(1.0, 1.0 / lam)
        # a real polynomial: the pair is conjugate (or both real)
        assert abs(z[0] - np.conj(z[1])) < 1e-12 or (abs(z[0].imag) < 1e-14 and abs(z[1].imag) < 1e-14)


def test_mirror_pad_indices():
    for M in (2, 3, 5, 9):
        a = np.arange(M, dtype=np.float64)
        p = B.mirror_pad(a)
        # sample 1, samples 0..M-1, samples M-2 and M-3 (M = 2: 0 and 1)
        want = [1] + list(range(M)) + ([M - 2, M - 3] if M > 2 else [0, 1])
        assert p.tolist() == [float(v) for v in want]


def test_interpolation_at_knots():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(11)
    c = B.bsarray(x)                       # padded: sample i at padded index i + 1
    r = (c[:-3] + 4.0 * c[1:-2] + c[2:-1]) / 6.0
    assert np.abs(r - x).max() < 1e-13


def test_matches_the_model_generator():
    from bsfit import gen
    rng = np.random.default_rng(5)
    x = rng.standard_normal((9, 8, 7, 5))
    a = gen.bspline_coeffs(x)
    b = B.bsarray(x)
    assert a.shape == b.shape == (12, 11, 10, 8)
    assert np.abs(a - b).max() <= 2e-6 * np.abs(b).max()


def test_identity_catches_a_wrong_gain(monkeypatch):
    """The pin has teeth: the oracle with gain 5 instead of 6 fails it."""
    orig = B.poles
    monkeypatch.setattr(B, "poles", lambda lam: (orig(lam)[0], 5.0) if lam == 0 else orig(lam))
    x = np.random.default_rng(2).standard_normal((8, 6))
    c = B.prefilter(x, 0.0)
    assert np.abs(B.operator(c, 0.0) - x).max() > 1e-2
