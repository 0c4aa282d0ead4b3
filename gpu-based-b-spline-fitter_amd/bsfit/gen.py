"""Seeded synthetic workloads: PSF models, cubic B-spline coefficients, ROIs.

The reference builds its coefficient arrays in MATLAB with ``bsarray``
(``MultiD_bspline/bsarray/bsarray.m:58-173``): Unser's recursive cubic
prefilter (pole sqrt(3)-2, gain 6 per axis, ``getBSplineFiltCoeffs.m:219-230``,
``symExpFilt.m:26-45``) with whole-sample mirror boundaries, then each axis
mirror-padded from M to M+3 samples (``directFilter.m:286-297``).
``bspline_coeffs`` restates that transform (smoothing lambda = 0).

Spline coordinates are 1-based sample indices: evaluating at coordinate xi
reproduces sample xi-1 (0-based).  An ROI of size S on a grid with nData
samples maps pixel (row ii, column jj) to column coordinate
``(S+1)//2 - theta0 + jj + off`` and row coordinate
``(S+1)//2 - theta1 + ii + off``, ``off = (nData - S)/2``
(``CPUmleFit_LM.cpp:2512-2516``).
"""
import math

import numpy as np

SEED0 = 20261203

# ----------------------------------------------------------------- prefilter
_POLE = math.sqrt(3.0) - 2.0


def _prefilter_axis(a, axis):
    """Cubic B-spline interpolation prefilter along one axis (mirror ends)."""
    a = np.moveaxis(np.asarray(a, np.float64), axis, 0).copy()
    M = a.shape[0]
    z = _POLE
    if M == 1:
        return np.moveaxis(a * 1.0, 0, axis)
    # causal init: sum_k z^k x[reflect(k)], K0 terms (bsarray's K0Tol = eps)
    K0 = int(math.ceil(math.log(np.finfo(float).eps) / math.log(abs(z))))
    period = 2 * M - 2
    idx = np.arange(K0) % period
    idx = np.where(idx < M, idx, period - idx)
    zk = z ** np.arange(K0)
    y = np.empty_like(a)
    y[0] = np.tensordot(zk, a[idx], axes=(0, 0))
    for k in range(1, M):
        y[k] = a[k] + z * y[k - 1]
    c0 = -z / (1.0 - z * z)
    y[M - 1] = (2.0 * y[M - 1] - a[M - 1]) * c0
    for k in range(M - 2, -1, -1):
        y[k] = (y[k + 1] - y[k]) * z
    return np.moveaxis(y * 6.0, 0, axis)


def _mirror_pad(a):
    """Pad every axis from M to M+3 by whole-sample reflection (indices -1..M+1)."""
    for ax in range(a.ndim):
        M = a.shape[ax]
        k = np.arange(-1, M + 2)
        if M == 1:
            src = np.zeros_like(k)
        else:
            period = 2 * M - 2
            k = np.mod(k, period)
            src = np.where(k < M, k, period - k)
        a = np.take(a, src, axis=ax)
    return a


def bspline_coeffs(samples):
    """Cubic B-spline coefficients of a sampled model, padded to dims+3.

    Returns a Fortran-order float32 array (the MATLAB ``b.coeffs`` layout)."""
    c = np.asarray(samples, np.float64)
    for ax in range(c.ndim):
        c = _prefilter_axis(c, ax)
    return np.asfortranarray(_mirror_pad(c).astype(np.float32))


# ---------------------------------------------------------- spline evaluation
def _beta3(x):
    x = np.abs(x)
    return np.where(x < 1, 2.0 / 3.0 - x * x + 0.5 * x ** 3,
                    np.where(x < 2, (2.0 - x) ** 3 / 6.0, 0.0))


def _taps(coord, n):
    """Clamped base index k (tap window k-2..k+1, 0-based) and the 4 weights."""
    coord = np.asarray(coord, np.float64)
    fl = np.floor(coord)
    kc = np.clip(fl, 1, n - 1)
    kc = np.where(coord == n, n, kc).astype(np.int64)
    d = coord - fl
    w = np.stack([_beta3(d + 1), _beta3(d), _beta3(d - 1), _beta3(d - 2)], -1)
    return kc + 1, w


def eval_spline_rois(coeff, S, theta_xy, nonxy, batch=None):
    """Noise-free spline model f (without N, bg) on S x S ROIs, float64.

    coeff: (nx+3, ny+3, ...) coefficients; theta_xy: (N, 2) ROI-centre
    positions (theta0 = column, theta1 = row); nonxy: (N, D-2) spline
    coordinates of the remaining axes.  Returns (N, S, S) indexed
    [n, row, column]."""
    coeff = np.asarray(coeff, np.float32)
    D = coeff.ndim
    nd = [s - 3 for s in coeff.shape]
    N = theta_xy.shape[0]
    P2 = coeff.shape[0] * coeff.shape[1]
    c2 = coeff.reshape(P2, -1, order="F").astype(np.float64)     # (plane, rest)
    K = D - 2
    if batch is None:
        batch = max(64, int(4e6 // (P2 * 4 ** K)))
    out = np.empty((N, S, S))
    off = (nd[0] - S) / 2.0
    half = (S + 1) // 2
    ar = np.arange(S)
    rest_strides = np.cumprod([1] + list(coeff.shape[2:-1])) if K else np.array([1])
    for b0 in range(0, N, batch):
        sl = slice(b0, min(N, b0 + batch))
        n = sl.stop - sl.start
        flat = np.zeros((n, 1), np.int64)
        W = np.ones((n, 1))
        for a in range(K):
            k, w = _taps(nonxy[sl, a], nd[2 + a])
            taps = k[:, None] - 2 + np.arange(4)[None, :]                 # (n, 4)
            flat = (flat[:, :, None] + taps[:, None, :] * rest_strides[a]).reshape(n, -1)
            W = (W[:, :, None] * w[:, None, :]).reshape(n, -1)
        G = np.einsum("pnt,nt->np", c2[:, flat], W)                      # (n, plane)
        G = G.reshape(n, coeff.shape[1], coeff.shape[0]).transpose(0, 2, 1)  # (n, dim0, dim1)
        colc = (half - theta_xy[sl, 0])[:, None] + ar[None, :] + off
        rowc = (half - theta_xy[sl, 1])[:, None] + ar[None, :] + off
        kx, wx = _taps(colc, nd[0])
        ky, wy = _taps(rowc, nd[1])
        ry = ky[:, :, None] - 2 + np.arange(4)
        rx = kx[:, :, None] - 2 + np.arange(4)
        nn = np.arange(n)[:, None, None, None, None]
        vals = G[nn, ry[:, :, :, None, None], rx[:, None, None, :, :]]  # (n, S, 4, S, 4)
        out[sl] = np.einsum("nitju,nit,nju->nij", vals, wy, wx)
    return out


# ---------------------------------------------------------------- PSF models
def gauss2d(nx, sr, sc, cr=None, cc=None, rho=None):
    """Pixel-centre samples of a normalised 2D Gaussian on an nx x nx grid."""
    c0 = (nx - 1) / 2.0
    r = np.arange(nx) - c0 if cr is None else np.arange(nx) - cr
    q = np.arange(nx) - c0 if cc is None else np.arange(nx) - cc
    R, Q = np.meshgrid(r, q, indexing="ij")
    if rho is None:
        g = np.exp(-0.5 * (R / sr) ** 2 - 0.5 * (Q / sc) ** 2)
    else:
        g = np.exp(-0.5 / (1 - rho ** 2) * ((R / sr) ** 2 - 2 * rho * R * Q / (sr * sc) + (Q / sc) ** 2))
    return g / g.sum()


def psf_astig3d(nx=27, nz=61, dz=20.0, s0=1.3, zr=400.0, za=300.0, scale=1.0):
    """Astigmatic Gaussian: sigma_row, sigma_col = s0*sqrt(1+((z -/+ za)/zr)^2).
    z_k = (k - (nz-1)/2) * dz; each slice sums to 1.  Shape (nx, nx, nz)."""
    out = np.empty((nx, nx, nz))
    for k in range(nz):
        z = (k - (nz - 1) / 2.0) * dz
        sr = scale * s0 * math.sqrt(1 + ((z - za) / zr) ** 2)
        sc = scale * s0 * math.sqrt(1 + ((z + za) / zr) ** 2)
        out[:, :, k] = gauss2d(nx, sr, sc)
    return out


def psf_color4d(nx=27, nz=61, scales=(1.0, 1.1, 1.2)):
    return np.stack([psf_astig3d(nx, nz, scale=s) for s in scales], axis=3)


def _dipole_slice(nx, z, phi, pol, s0=1.3, zr=400.0, za=300.0, elong=0.7, tilt=0.4):
    """Astigmatic focus (as psf_astig3d) convolved with an orientation blur:
    covariance diag(sr^2, sc^2) + sin(pol) * elong * s0^2 * u u^T, u the unit
    vector at azimuth phi in the (row, column) plane, and the centre shifted
    by tilt * sin(pol) * s0 along u (so phi and phi + pi differ)."""
    sr2 = (s0 ** 2) * (1 + ((z - za) / zr) ** 2)
    sc2 = (s0 ** 2) * (1 + ((z + za) / zr) ** 2)
    e = math.sin(pol) * elong * s0 ** 2
    c, s = math.cos(phi), math.sin(phi)
    cov_rr = sr2 + e * c * c
    cov_cc = sc2 + e * s * s
    cov_rc = e * c * s
    sr, sc = math.sqrt(cov_rr), math.sqrt(cov_cc)
    c0 = (nx - 1) / 2.0
    sh = tilt * math.sin(pol) * s0
    return gauss2d(nx, sr, sc, cr=c0 + sh * c, cc=c0 + sh * s, rho=cov_rc / (sr * sc))


def dipole_axes(nz=21, nw=17, nu=9, dz=60.0):
    """Sample positions of the 5D grid: z (nm), azimuth phi (period nw-2), polar."""
    z = (np.arange(nz) - (nz - 1) / 2.0) * dz
    phi = 2 * math.pi * (np.arange(nw) - 1) / (nw - 2)
    pol = (math.pi / 2) * np.arange(nu) / (nu - 1)
    return z, phi, pol


def psf_dipole5d(nx=21, nz=21, nw=17, nu=9, dz=60.0):
    """Dipole-like anisotropic Gaussian: ellipticity grows with sin(polar),
    major axis along the azimuth.  Axis order (x, y, z, azimuth, polar) as
    the fittype-8 coefficient array.  The azimuth samples are
    phi_k = 2*pi*(k-1)/(nw-2), so samples 0 and nw-1 repeat nw-2 and 1 and the
    fitter's wrap of period nw-2 is exact."""
    z, phi, pol = dipole_axes(nz, nw, nu, dz)
    out = np.empty((nx, nx, nz, nw, nu))
    for a in range(nz):
        for b in range(nw):
            for c in range(nu):
                out[:, :, a, b, c] = _dipole_slice(nx, z[a], phi[b], pol[c])
    return out


def psf_dipole3d(nx=21, nz=21, dz=60.0, nphi=24, npol=24):
    """Orientation-averaged 3D model of psf_dipole5d (the fittype-8 coeff3)."""
    z = (np.arange(nz) - (nz - 1) / 2.0) * dz
    out = np.zeros((nx, nx, nz))
    phis = 2 * math.pi * np.arange(nphi) / nphi
    pols = (math.pi / 2) * (np.arange(npol) + 0.5) / npol
    wts = np.sin(pols)
    for a in range(nz):
        acc = np.zeros((nx, nx))
        for p, wp in zip(pols, wts):
            for f in phis:
                acc += wp * _dipole_slice(nx, z[a], f, p)
        out[:, :, a] = acc / acc.sum()
    return out


def phase_axis(nphi=16):
    """Phase samples of the 4Pi model: phi_k = 2*pi*(k-1)/nphi, k = 0..nphi+1,
    so the samples repeat with period nphi (w units) and the fitter's wrap
    (period nw-2) is exact.  Spline coordinate w = 2 + phi*nphi/(2*pi)."""
    return 2 * math.pi * (np.arange(nphi + 2) - 1) / nphi


def psf_4pi(nx=27, nz=61, nphi=16, visibility=0.8):
    """4Pi interference model sampled on (x, y, z, phi): I(x,y,z) * (1 + v cos phi)
    with I the astigmatic 3D PSF of psf_astig3d.  Channel k of a 4Pi ROI sees
    phi + offset_k.  Shape (nx, nx, nz, nphi+2)."""
    I = psf_astig3d(nx, nz)
    ph = phase_axis(nphi)
    return I[:, :, :, None] * (1 + visibility * np.cos(ph))[None, None, None, :]


PHASE_OFFSETS = (0.0, math.pi / 2, math.pi, 3 * math.pi / 2)


def make_rois_4pi(N, S, seed, coeff, shared=(1, 1, 1, 1, 1, 1), noise=True, iterations=50,
                  offsets=PHASE_OFFSETS):
    """Linked 4Pi ROIs: shared x, y, N, z, phi; per-channel background when
    shared[3] == 0.  imstack (S, S, N, nch) Fortran-order."""
    rng = np.random.default_rng(seed)
    nd = [s - 3 for s in coeff.shape]
    nphi = nd[3] - 2
    nch = len(offsets)
    c0 = (S - 1) / 2.0
    x = c0 + rng.uniform(-1.5, 1.5, N)
    y = c0 + rng.uniform(-1.5, 1.5, N)
    photons = np.exp(rng.uniform(math.log(1e3), math.log(1e4), N))
    bg = rng.uniform(5, 50, (N, nch))
    if shared[3]:
        bg[:] = bg[:, :1]
    zc = 1 + (rng.uniform(-600, 600, N) + 600) / 20.0
    phi = rng.uniform(0, 2 * math.pi, N)
    wc = 2 + phi * nphi / (2 * math.pi)
    shift = np.array([o * nphi / (2 * math.pi) for o in offsets], np.float32)
    data = np.empty((S, S, N, nch), np.float32, order="F")
    mu_all = []
    for k in range(nch):
        wk = wc + shift[k]
        wk = np.where(wk > nd[3] - 1, wk - (nd[3] - 2), wk)
        f = eval_spline_rois(coeff, S, np.stack([x, y], 1), np.stack([zc, wk], 1))
        mu = bg[:, k, None, None] + photons[:, None, None] * np.maximum(f, 0.0)
        mu_all.append(mu)
        d = rng.poisson(mu).astype(np.float32) if noise else mu.astype(np.float32)
        data[:, :, :, k] = np.transpose(d, (1, 2, 0))
    truth = dict(x=x, y=y, N=photons, bg=bg, z=zc, w=wc)
    return Workload(fittype="4pi", imstack=data, coeff=coeff, S=S, truth=truth, shared=np.array(shared, np.int32),
                    phaseShift=shift, initZ=np.full(N, nd[2] / 2.0, np.float32),
                    initPhi=np.full(N, 2 + nphi / 2.0, np.float32), iterations=iterations, nch=nch,
                    mu=np.stack(mu_all, -1))


# ------------------------------------------------------------------ configs
class Workload:
    """One fit call: ROI stack plus everything the fitter needs."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    @property
    def N(self):
        return self.imstack.shape[2]


_MODEL_CACHE = {}


def model_coeffs(kind):
    """Coefficient arrays (cached per process) for 'astig3d', 'color4d',
    'dipole5d' (returns (coeff5, coeff3))."""
    if kind not in _MODEL_CACHE:
        if kind == "astig3d":
            _MODEL_CACHE[kind] = bspline_coeffs(psf_astig3d())
        elif kind == "color4d":
            _MODEL_CACHE[kind] = bspline_coeffs(psf_color4d())
        elif kind == "4pi":
            _MODEL_CACHE[kind] = bspline_coeffs(psf_4pi())
        elif kind == "dipole5d":
            _MODEL_CACHE[kind] = (bspline_coeffs(psf_dipole5d()), bspline_coeffs(psf_dipole3d()))
        else:
            raise KeyError(kind)
    return _MODEL_CACHE[kind]


def _poisson(rng, mu):
    return rng.poisson(mu).astype(np.float32)


def make_rois(fittype, N, S, seed, coeff, coeff3=None, noise=True, iterations=50):
    """Seeded ROIs rendered from the spline model itself, then Poisson noise.

    Returns a Workload with imstack (S, S, N) Fortran-order float32, the
    truth theta in fitter coordinates (N, NV), initP and iterations."""
    rng = np.random.default_rng(seed)
    nd = [s - 3 for s in coeff.shape]
    c0 = (S - 1) / 2.0
    x = c0 + rng.uniform(-1.5, 1.5, N)
    y = c0 + rng.uniform(-1.5, 1.5, N)
    if fittype == 8:
        photons = np.exp(rng.uniform(math.log(2e3), math.log(2e4), N))
    else:
        photons = np.exp(rng.uniform(math.log(1e3), math.log(1e4), N))
    bg = rng.uniform(5, 50, N)
    xy = np.stack([x, y], 1)
    if fittype == 6:
        zc = 1 + (rng.uniform(-600, 600, N) + 600) / 20.0
        nonxy = zc[:, None]
        f = eval_spline_rois(coeff, S, xy, nonxy)
        truth = np.stack([x, y, photons, bg, zc], 1)
        initP = None
    elif fittype == 7:
        zc = 1 + (rng.uniform(-600, 600, N) + 600) / 20.0
        wc = 1 + rng.uniform(0, 2, N)
        nonxy = np.stack([zc, wc], 1)
        f = eval_spline_rois(coeff, S, xy, nonxy)
        truth = np.stack([x, y, photons, bg, zc, wc], 1)
        initP = np.array([nd[2] / 2.0, nd[3] / 2.0], np.float32)
    elif fittype == 8:
        dz = 60.0
        zc = 1 + (rng.uniform(-600, 600, N) + 600) / dz
        phi = rng.uniform(0, 2 * math.pi, N)
        pol = rng.uniform(0, math.pi / 2, N)
        wc = 2 + phi * (nd[3] - 2) / (2 * math.pi)
        uc = 1 + pol * (nd[4] - 1) / (math.pi / 2)
        g2 = np.full(N, 0.75)
        f5 = eval_spline_rois(coeff, S, xy, np.stack([zc, wc, uc], 1))
        f3 = eval_spline_rois(coeff3, S, xy, zc[:, None])
        f = (1 - g2)[:, None, None] * f3 + g2[:, None, None] * f5
        truth = np.stack([x, y, photons, bg, zc, wc, uc, g2], 1)
        initP = np.array([nd[2] / 2.0, nd[3] / 2.0, nd[4] / 2.0, 0.5], np.float32)
    else:
        raise ValueError(fittype)
    mu = bg[:, None, None] + photons[:, None, None] * np.maximum(f, 0.0)
    data = _poisson(rng, mu) if noise else mu.astype(np.float32)
    # [n, row, col] -> MATLAB (row, col, n) column-major
    imstack = np.asfortranarray(np.transpose(data, (1, 2, 0)).astype(np.float32))
    return Workload(fittype=fittype, imstack=imstack, coeff=coeff, coeff3=coeff3, S=S,
                    truth=truth, initP=initP, iterations=iterations, mu=mu)


def config(idx, N=None, S=None, noise=True, seed=None, bg_per_channel=False):
    """Workloads for the documented configurations.

    0: 3D astigmatic, 27x27x61 grid, S=13, N=10,000, fittype 6
    1: colour 4D, 27x27x61x3 grid, S=15, N=100,000, fittype 7
    2: 4Pi, four 13x13 channels (phase offsets 0, pi/2, pi, 3pi/2), 27x27x61x18
       grid (16 phase samples + 2 wrap samples), N=200,000, all shared (NV=6);
       bg_per_channel: a background of its own per channel, drawn and fitted
       per channel (x, y, photons, z, phase shared: NV=9)
    3: vortex 5D, 21x21x21x17x9 grid (+21^3 averaged model), S in
       {11,13,15,17}, N=100,000, fittype 8 (one Workload per S group)
    """
    seed = SEED0 + 1000 * idx if seed is None else seed
    if idx == 0:
        return make_rois(6, N or 10000, S or 13, seed, model_coeffs("astig3d"), noise=noise)
    if idx == 1:
        return make_rois(7, N or 100000, S or 15, seed, model_coeffs("color4d"), noise=noise)
    if idx == 2:
        shared = (1, 1, 1, 0, 1, 1) if bg_per_channel else (1, 1, 1, 1, 1, 1)
        return make_rois_4pi(N or 200000, S or 13, seed, model_coeffs("4pi"), shared=shared, noise=noise)
    if idx == 3:
        c5, c3 = model_coeffs("dipole5d")
        Ntot = N or 100000
        sizes = [11, 13, 15, 17] if S is None else [S]
        rng = np.random.default_rng(seed)
        pick = rng.integers(0, len(sizes), Ntot)
        out = []
        for i, s in enumerate(sizes):
            n = int((pick == i).sum())
            if n:
                out.append(make_rois(8, n, s, seed + 17 * s, c5, c3, noise=noise))
        return out
    raise ValueError("configuration %r has no fittype 6/7/8 workload" % idx)


# ------------------------------------------------------------ camera frames
def make_frames(nframes, nrows, ncols, per_frame, seed, coeff=None, offset=100.0, conversion=0.5, patch=21,
                margin=None, dtype=np.uint16):
    """Seeded camera frames (nrows, ncols, nframes), Fortran order, for the
    frame -> ROI front-end: per frame `per_frame` emitters rendered from the
    3D astigmatic spline model (27x27x61 grid, as configuration 0) on a
    patch x patch footprint at random sub-pixel positions and z, photons
    ~ logU(1e3, 1e4), a background ~ U(5, 30) per frame, Poisson noise, then
    ADU = photons / conversion + offset, rounded (uint16) or not (float32).

    Returns (frames, truth) with truth rows (frame index from 0, x, y, z
    spline coordinate, photons, bg[, colour coordinate]), x = column and y =
    row in pixels from 0.  A 4D (colour) coefficient array renders each
    emitter at a colour coordinate ~ U(1, 3) as well."""
    if coeff is None:
        coeff = model_coeffs("astig3d")
    rng = np.random.default_rng(seed)
    h = (patch - 1) // 2
    margin = h + 1 if margin is None else margin
    n = nframes * per_frame
    fi = np.repeat(np.arange(nframes), per_frame)
    x = rng.uniform(margin, ncols - 1 - margin, n)
    y = rng.uniform(margin, nrows - 1 - margin, n)
    zc = 1 + (rng.uniform(-500, 500, n) + 600) / 20.0
    ph = np.exp(rng.uniform(math.log(1e3), math.log(1e4), n))
    bgf = rng.uniform(5, 30, nframes)
    cj, ci = np.rint(x).astype(int), np.rint(y).astype(int)
    c0 = float(h)
    nonxy = zc[:, None]
    if coeff.ndim == 4:                        # colour model: a colour coordinate per emitter
        wc = 1 + rng.uniform(0, 2, n)
        nonxy = np.stack([zc, wc], 1)
    f = eval_spline_rois(coeff, patch, np.stack([c0 + x - cj, c0 + y - ci], 1), nonxy)
    mu = np.repeat(bgf[None, None, :], nrows, 0).repeat(ncols, 1).astype(np.float64)
    for k in range(n):
        mu[ci[k] - h:ci[k] + h + 1, cj[k] - h:cj[k] + h + 1, fi[k]] += ph[k] * np.maximum(f[k], 0.0)
    counts = rng.poisson(mu).astype(np.float64)
    adu = counts / conversion + offset
    if dtype == np.uint16:
        frames = np.clip(np.rint(adu), 0, 65535).astype(np.uint16)
    else:
        frames = adu.astype(np.float32)
    truth = np.stack([fi, x, y, zc, ph, bgf[fi]] + ([nonxy[:, 1]] if coeff.ndim == 4 else []), 1)
    return np.asfortranarray(frames), truth
