"""Oracle (test infrastructure, not shipped): numpy restatement of the
reference's cubic B-spline coefficient builder, `bsarray` with degree 3 on
every axis and a per-axis smoothing weight lambda >= 0.

Follows, for degree 3 with centred basis functions:
  MultiD_bspline/bsarray/bsarray.m:58-173          argument checks, per-axis lambda
  MultiD_bspline/bsarray/getBSplineFiltCoeffs.m    direct filter poles and gain:
        lambda == 0: pole sqrt(3)-2, gain 6
        lambda  > 0: the roots of z^4 + (-4 + 1/(6 lambda)) z^3
                     + (6 + 2/(3 lambda)) z^2 + (-4 + 1/(6 lambda)) z + 1
                     inside the unit circle (a complex pair when the
                     smoothing is strong), gain 1/lambda
  MultiD_bspline/bsarray/symExpFilt.m:26-45        one pole along one line:
        causal start = K0-term sum over the whole-sample mirror extension,
        K0 = ceil(log(eps) / log|z|); causal pass; anticausal start
        (2 y[N-1] - x[N-1]) * (-z / (1 - z^2)); anticausal pass
  MultiD_bspline/bsarray/directFilter.m:1-297      axes in order 0, 1, ...;
        all poles of an axis, then real part times the gain; finally every
        axis mirror-padded from M to M+3 samples (indices -1 .. M+1 of the
        whole-sample reflection)
in FP64 (MATLAB's `double(A)`), complex where a pole is.

The coefficients c of one line solve, with whole-sample mirror ends,
  (c[i-1] + 4 c[i] + c[i+1]) / 6 + lambda (c[i-2] - 4 c[i-1] + 6 c[i] - 4 c[i+1] + c[i+2]) = x[i],
which `operator` evaluates: the tests pin poles, gain, ordering and signs by it.
"""
import math

import numpy as np

EPS = np.finfo(float).eps


def poles(lam):
    """Filter poles (inside the unit circle) and gain for one axis."""
    if lam == 0:
        return [complex(math.sqrt(3.0) - 2.0, 0.0)], 6.0
    p = [1.0, -4.0 + 1.0 / (6.0 * lam), 6.0 + 2.0 / (3.0 * lam), -4.0 + 1.0 / (6.0 * lam), 1.0]
    r = np.roots(p)
    r = [complex(z) for z in r if abs(z) <= 1.0]
    return r, 1.0 / lam


def _reflect(k, M):
    period = 2 * M - 2
    k = k % period
    return k if k < M else period - k


def sym_exp_filt(x, z):
    """One pole along axis 0 of x (complex), the reference's symExpFilt."""
    M = x.shape[0]
    K0 = int(math.ceil(math.log(EPS) / math.log(abs(z))))
    idx = np.array([_reflect(k, M) for k in range(K0)])
    zk = z ** np.arange(K0)
    y = np.empty_like(x)
    y[0] = np.tensordot(zk, x[idx], axes=(0, 0))
    for k in range(1, M):
        y[k] = x[k] + z * y[k - 1]
    c0 = -z / (1.0 - z * z)
    y[M - 1] = (2.0 * y[M - 1] - x[M - 1]) * c0
    for k in range(M - 2, -1, -1):
        y[k] = (y[k + 1] - y[k]) * z
    return y


def prefilter(a, lam):
    """Unpadded coefficients (FP64) of array a, lam one value per axis."""
    c = np.asarray(a, np.float64).copy()
    lam = np.broadcast_to(np.asarray(lam, np.float64), (c.ndim,))
    for d in range(c.ndim):
        if c.shape[d] < 2:
            raise ValueError("every axis needs at least 2 samples")
        zs, gain = poles(float(lam[d]))
        y = np.moveaxis(c, d, 0).astype(np.complex128)
        for z in zs:
            y = sym_exp_filt(y, z)
        c = np.moveaxis(np.real(y) * gain, 0, d).copy()
    return c


def mirror_pad(c):
    """Every axis from M to M+3 samples: whole-sample reflection of -1..M+1."""
    for d in range(c.ndim):
        M = c.shape[d]
        idx = [_reflect(k, M) for k in range(-1, M + 2)]
        c = np.take(c, idx, axis=d)
    return c


def bsarray(a, lam=0.0):
    """Padded cubic B-spline coefficients, as bsarray(a, 'lambda', lam).coeffs."""
    return mirror_pad(prefilter(a, lam))


def operator(c, lam):
    """(B3 + lambda D4) applied along every axis in turn to unpadded
    coefficients with whole-sample mirror ends: returns the data they fit."""
    x = np.asarray(c, np.float64)
    lam = np.broadcast_to(np.asarray(lam, np.float64), (x.ndim,))
    for d in range(x.ndim):
        y = np.moveaxis(x, d, 0)
        M = y.shape[0]
        tap = lambda o: y[[_reflect(i + o, M) for i in range(M)]]
        r = (tap(-1) + 4.0 * tap(0) + tap(1)) / 6.0
        if lam[d] != 0:
            r = r + lam[d] * (tap(-2) - 4.0 * tap(-1) + 6.0 * tap(0) - 4.0 * tap(1) + tap(2))
        x = np.moveaxis(r, 0, d)
    return x
