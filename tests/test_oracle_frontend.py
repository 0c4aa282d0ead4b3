"""Pins of the front-end oracle (oracle/bsf_frontend.c), the restatement of
the GUI fitter's frame -> ROI steps (MultiD_fitter/simplefitter_spline.m:
197-307, MultiD_fitter/shared/maximumfind2.c:18-65).  The reference's steps
are MATLAB (not in this image) and a MEX that includes mex.h (absent), so
what the definitions fix pins the oracle instead:

* photon conversion: numpy's IEEE single-precision arithmetic, exactly;
* filter2 'same': an FP64 evaluation of its documented definition,
  filter2(h, X) = conv2(X, rot90(h, 2), 'same') with 'same' the central part
  of the full convolution starting at index floor(m/2) + 1 (1-based), built
  here from scipy's full convolution; plus the known answer
  filter2([1 1], [1 2 3]) = [3 5 3].  The plausible mistake for even
  kernels, the anchor on the other side of the centre, is shown to fail;
* maxima: a vectorised numpy brute force of maximumfind2's strict
  8-neighbour test in column-major scan order (the reference's skip of the
  next row after a maximum changes nothing: that pixel is below the maximum);
  a plateau, which a >= test would accept, yields nothing;
* the caller's filters and the ROI cut: numpy slicing of the photon frame,
  mirrored or not;
* end to end: planted emitters are found.
The peak-finding kernels (bsf_peak_filter*, host code of the library) are
checked against numpy restatements of fspecial('gaussian') and of the GUI's
PSF-derived kernels."""
import math

import numpy as np
import pytest
import scipy.signal

import oracle
from bsfit import gen


def _filter2_def(h, x):
    """FP64 filter2(h, x, 'same') from the documented definition."""
    h = np.asarray(h, np.float64)
    full = scipy.signal.convolve2d(np.asarray(x, np.float64), h[::-1, ::-1], mode="full")
    r0, c0 = h.shape[0] // 2, h.shape[1] // 2
    return full[r0:r0 + x.shape[0], c0:c0 + x.shape[1]]


def test_photons_are_single_precision_arithmetic():
    rng = np.random.default_rng(1)
    raw = rng.integers(0, 65535, (37, 29), dtype=np.uint16)
    got = oracle.photons(raw, 100.3, 0.47)
    want = (raw.astype(np.float32) - np.float32(100.3)) * np.float32(0.47)
    np.testing.assert_array_equal(got, want)
    rawf = rng.uniform(0, 3000, (11, 13)).astype(np.float32)
    np.testing.assert_array_equal(oracle.photons(rawf, 99.5, 2.0),
                                  (rawf - np.float32(99.5)) * np.float32(2.0))


def test_filter2_known_answer():
    got = oracle.filter2(np.array([[1.0, 1.0]]), np.array([[1.0, 2.0, 3.0]]))
    np.testing.assert_array_equal(got, [[3.0, 5.0, 3.0]])
    got = oracle.filter2(np.array([[1.0], [1.0]]), np.array([[1.0], [2.0], [3.0]]))
    np.testing.assert_array_equal(got, [[3.0], [5.0], [3.0]])


@pytest.mark.parametrize("hs", [(3, 3), (4, 4), (7, 7), (5, 8), (2, 1), (13, 13), (16, 16), (1, 1)])
def test_filter2_matches_definition(hs):
    rng = np.random.default_rng(sum(hs))
    x = rng.uniform(0, 500, (41, 35)).astype(np.float32)
    h = rng.standard_normal(hs).astype(np.float32)
    got = oracle.filter2(h, x).astype(np.float64)
    want = _filter2_def(h, x)
    scale = np.abs(h).sum() * np.abs(x).max()
    assert np.abs(got - want).max() <= 2e-6 * scale


def test_filter2_even_kernels_anchor():
    """An even kernel has two candidate centres; taking the central part of
    the full convolution from (m-1)//2 instead of m//2 (numpy.convolve's
    'same') moves the anchor by one pixel, and the pin catches it."""
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 100, (20, 20)).astype(np.float32)
    h = rng.uniform(0, 1, (4, 4)).astype(np.float32)
    full = scipy.signal.convolve2d(x.astype(np.float64), h[::-1, ::-1].astype(np.float64), mode="full")
    other = full[1:21, 1:21]              # start (m-1)//2 = 1 instead of m//2 = 2
    got = oracle.filter2(h, x)
    assert np.abs(got - other).max() > 1.0
    assert np.abs(got - _filter2_def(h, x)).max() < 1e-3


def _maxima_brute(F):
    F = np.asarray(F, np.float32)
    c = F[1:-1, 1:-1]
    m = np.ones(c.shape, bool)
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            if di or dj:
                m &= c > F[1 + di:F.shape[0] - 1 + di, 1 + dj:F.shape[1] - 1 + dj]
    j, i = np.nonzero(m.T)                # column-major scan order: columns outer
    return np.stack([j + 2, i + 2, c[i, j].astype(np.float64)], 1)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_maxima_match_brute_force(seed):
    rng = np.random.default_rng(seed)
    F = rng.standard_normal((33 + seed, 47)).astype(np.float32)
    got = oracle.find_maxima(F, -np.inf, 0)
    want = _maxima_brute(F)
    np.testing.assert_array_equal(got, want)
    # integer-valued frames: many ties, only strict maxima count
    G = rng.integers(0, 4, (30, 31)).astype(np.float32)
    np.testing.assert_array_equal(oracle.find_maxima(G, -np.inf, 0), _maxima_brute(G))


def test_plateau_has_no_maxima():
    F = np.zeros((12, 12), np.float32)
    F[4:7, 4:7] = 5.0                     # a 3x3 plateau: no pixel above all neighbours
    assert oracle.find_maxima(F, -1.0, 0).shape[0] == 0
    # a >= test would report the plateau's centre
    c = F[1:-1, 1:-1]
    ge = np.ones(c.shape, bool)
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            if di or dj:
                ge &= c >= F[1 + di:11 + di, 1 + dj:11 + dj]
    assert ge.any()


def test_caller_filters():
    rng = np.random.default_rng(5)
    F = rng.standard_normal((40, 50)).astype(np.float32)
    allm = _maxima_brute(F)
    for cutoff, dx, psf in ((0.5, 3, 0.0), (-np.inf, 6, 0.0), (1.0, 2, 27.0), (0.0, 1, 12.0)):
        got = oracle.find_maxima(F, cutoff, dx, psf)
        x, y, v = allm[:, 0], allm[:, 1], allm[:, 2]
        keep = (v > cutoff) & (x > dx) & (x <= 50 - dx) & (y > dx) & (y <= 40 - dx)
        if psf > 0:
            hp = psf / 2
            keep &= ~((x < hp) | (x > 40 - hp) | (y < hp) | (y > 40 - hp))     # nrows for both, as the GUI
        np.testing.assert_array_equal(got, allm[keep])


@pytest.mark.parametrize("mirror", [False, True])
@pytest.mark.parametrize("subtract_min", [False, True])
def test_find_rois_is_the_composition(mirror, subtract_min):
    frames, _ = gen.make_frames(3, 90, 77, 12, 11)
    h = np.asfortranarray(np.random.default_rng(2).uniform(-0.1, 0.3, (6, 5)).astype(np.float32))
    S, dx = 11, 5
    st, pk, val, n = oracle.find_rois(frames, S, h, cutoff=30.0, offset=100, conversion=0.5,
                                      subtract_min=subtract_min, mirror=mirror, first_frame=4, capacity=5000)
    want_st, want_pk, want_v = [], [], []
    for f in range(3):
        ph = (frames[:, :, f].astype(np.float32) - np.float32(100)) * np.float32(0.5)
        X = ph - ph[:, 0].min() if subtract_min else ph
        F = oracle.filter2(h, X)
        for x, y, v in _maxima_brute(F):
            x, y = int(x), int(y)
            if not (v > 30.0 and dx < x <= 77 - dx and dx < y <= 90 - dx):
                continue
            win = ph[y - 1 - dx:y + dx, x - 1 - dx:x + dx]
            want_st.append(win[:, ::-1] if mirror else win)
            want_pk.append([x, y, f + 4])
            want_v.append(v)
    assert n == len(want_pk) > 0
    np.testing.assert_array_equal(pk, np.array(want_pk, np.int32))
    np.testing.assert_array_equal(np.transpose(st, (2, 0, 1)), np.array(want_st))
    np.testing.assert_array_equal(val, np.array(want_v, np.float32))


def test_capacity_truncates_in_order():
    frames, _ = gen.make_frames(2, 64, 64, 10, 3)
    h = np.ones((3, 3), np.float32) / 9
    full = oracle.find_rois(frames, 7, h, cutoff=60.0, offset=100, conversion=0.5, capacity=1000)
    part = oracle.find_rois(frames, 7, h, cutoff=60.0, offset=100, conversion=0.5, capacity=5)
    assert part[3] == full[3] > 5
    np.testing.assert_array_equal(part[1], full[1][:5])
    np.testing.assert_array_equal(part[0], full[0][:, :, :5])


def test_planted_emitters_are_found():
    from bsfit import frames as FR
    frames, truth = gen.make_frames(4, 128, 128, 15, 21)
    h = FR.peak_filter("dog", 1.0)
    _, pk, _, n = oracle.find_rois(frames, 13, h, cutoff=15.0, offset=100, conversion=0.5, subtract_min=True,
                                   capacity=1000)
    found = 0
    for f, x, y in truth[:, :3]:
        m = pk[:, 2] == f + 1
        found += bool(m.any() and np.hypot(pk[m, 0] - 1 - x, pk[m, 1] - 1 - y).min() < 1.5)
    assert found >= 0.8 * len(truth), (found, len(truth))


# ------------------------------------------------------------- peak kernels
def _fspecial_gauss(n, sigma):
    c = (n - 1) / 2.0
    x, y = np.meshgrid(np.arange(n) - c, np.arange(n) - c)
    h = np.exp(-(x * x + y * y) / (2 * sigma * sigma))
    h[h < np.finfo(float).eps * h.max()] = 0
    return h / h.sum()


@pytest.mark.parametrize("sigma", [0.6, 1.0, 1.3, 2.0, 3.7])
def test_peak_kernels(sigma):
    from bsfit import frames as FR
    g = FR.peak_filter("gauss", sigma)
    n = max(3, math.ceil(3 * sigma + 1))
    assert g.shape == (n, n)
    np.testing.assert_allclose(g, _fspecial_gauss(n, sigma).astype(np.float32), rtol=1e-6, atol=1e-9)
    assert abs(g.astype(np.float64).sum() - 1.0) < 1e-6 and np.array_equal(g, g.T) and np.array_equal(g, g[::-1])
    d = FR.peak_filter("dog", sigma)
    m = max(math.ceil(6 * sigma + 1), 3)
    want = _fspecial_gauss(m, sigma) - _fspecial_gauss(m, max(1.0, 2.5 * sigma))
    assert d.shape == (m, m)
    np.testing.assert_allclose(d, want.astype(np.float32), rtol=1e-6, atol=1e-9)
    assert abs(d.astype(np.float64).sum()) < 1e-6


def test_psf_kernels():
    from bsfit import frames as FR
    rng = np.random.default_rng(8)
    psf = rng.uniform(0, 1, (27, 27, 9)).astype(np.float32)
    for roi in (13, 15, 9):
        lo, hi = (27 - roi) // 2, (27 + roi) // 2          # 1-based rsz = lo:hi
        k = psf[lo - 1:hi, lo - 1:hi, :].astype(np.float64).mean(2)
        k -= k.mean()
        k = 0.05 * k / k.max()
        np.testing.assert_allclose(FR.peak_filter_psf("window", psf, roi), k.astype(np.float32), rtol=1e-5,
                                   atol=1e-8)
    p64 = 5.0 * psf.astype(np.float64)
    mean_k = (p64 - p64.mean(0, keepdims=True)).mean(2)
    np.testing.assert_allclose(FR.peak_filter_psf("mean", psf, 13), mean_k.astype(np.float32), rtol=1e-5, atol=1e-6)


# ------------------------------------------------------------ aggregation
def _aggre_matlab(data, thr):
    """aggre (MultiD_fitter/simplefitter_spline.m:653-673) as MATLAB runs it:
    the distance matrix of the upper triangle with its zeros set to NaN, the
    first true entry of D < thr in column-major order, round (half away from
    zero) of the mean into row ii and row jj deleted."""
    data = np.array(data, np.float64).reshape(-1, 3)
    while True:
        n = len(data)
        D = np.zeros((max(n - 1, 0), n))
        for ii in range(n - 1):
            for jj in range(ii + 1, n):
                D[ii, jj] = np.sqrt((data[ii, 0] - data[jj, 0]) ** 2 + (data[ii, 1] - data[jj, 1]) ** 2)
        D[D == 0] = np.nan
        with np.errstate(invalid="ignore"):
            agind = D < thr
        hits = np.argwhere(agind.T)          # (jj, ii) in column-major order
        if len(hits) == 0:
            return data
        jj, ii = hits[0]
        m = (data[ii] + data[jj]) / 2
        data[ii] = np.sign(m) * np.floor(np.abs(m) + 0.5)
        data = np.delete(data, jj, 0)


def test_aggregate_known_answers():
    # distance 5: merged below a threshold above 5, kept at exactly 5 (D < thr)
    two = [[10, 10, 5.0], [13, 14, 8.0]]
    np.testing.assert_array_equal(oracle.aggregate(two, 5.5), [[12, 12, 7.0]])
    np.testing.assert_array_equal(oracle.aggregate(two, 5.0), two)
    # coincident maxima are never merged (D == 0 is NaN in the reference)
    same = [[10, 10, 1.0], [10, 10, 2.0]]
    np.testing.assert_array_equal(oracle.aggregate(same, 100.0), same)
    # the first pair in column-major order wins: A-B (column 2) before B-C
    # (column 3); the merged point (2) is then 4 from C
    row = [[1, 5, 1.0], [4, 5, 1.0], [7, 5, 1.0]]
    np.testing.assert_array_equal(oracle.aggregate(row, 3.5), [[3, 5, 1.0], [7, 5, 1.0]])
    # a merge that lands next to a third maximum merges again: A-C first
    # (column 3; A-B is 2 apart), round(1.5) = 2, then (2, 1) with B: round(2.5)
    tri = [[1, 1, 0.0], [3, 1, 0.0], [2, 1, 3.0]]
    np.testing.assert_array_equal(oracle.aggregate(tri, 1.5), [[3, 1, 1.0]])
    # half-way values round away from zero, negative ones too
    neg = [[1, 1, -3.0], [2, 1, -4.0]]
    np.testing.assert_array_equal(oracle.aggregate(neg, 2.0), [[2, 1, -4.0]])
    assert oracle.aggregate(np.zeros((0, 3)), 3.0).shape == (0, 3)
    np.testing.assert_array_equal(oracle.aggregate([[4, 4, 1.0]], 3.0), [[4, 4, 1.0]])


@pytest.mark.parametrize("seed", range(6))
def test_aggregate_matches_matlab_semantics(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 40))
    side = int(rng.integers(6, 40))
    data = np.stack([rng.integers(1, side, n), rng.integers(1, side, n),
                     rng.normal(20, 5, n).astype(np.float32).astype(np.float64)], 1)
    if seed % 2:
        data[n // 2] = data[0]                    # a coincident pair
    thr = float(rng.uniform(1.0, 8.0)) if seed != 5 else 3.0   # 3.0: distances equal to the threshold
    np.testing.assert_array_equal(oracle.aggregate(data, thr), _aggre_matlab(data, thr))


def test_find_rois_aggregates_before_the_psf_border_test():
    """find_rois with agg_dist is the composition in the reference's order:
    maxima (cutoff and dx borders), aggre, then the PSF border test, then the
    cut (simplefitter_spline.m:277-307)."""
    frames, _ = gen.make_frames(7, 96, 96, 40, 2)
    h = np.ones((3, 3), np.float32) / 9
    args = dict(cutoff=55.0, offset=100, conversion=0.5)
    roi, thr, psf = 9, 11.0, 30.0
    st, pk, val, n = oracle.find_rois(frames, roi, h, capacity=2000, psf_size=psf, agg_dist=thr, **args)
    want_pk, want_st, want_v = [], [], []
    for f in range(frames.shape[2]):
        ph = oracle.photons(frames[:, :, f], 100, 0.5)
        F = oracle.filter2(h, ph)
        m = _aggre_matlab(oracle.find_maxima(F, 55.0, roi // 2), thr)
        hp = psf / 2
        keep = ~((m[:, 0] < hp) | (m[:, 0] > 96 - hp) | (m[:, 1] < hp) | (m[:, 1] > 96 - hp))
        for x, y, v in m[keep]:
            x, y = int(x) - 1, int(y) - 1
            want_pk.append((x + 1, y + 1, f + 1))
            want_st.append(ph[y - roi // 2:y + roi // 2 + 1, x - roi // 2:x + roi // 2 + 1])
            want_v.append(v)
    assert n == len(want_pk) and n > 0
    merged = n < len(oracle.find_rois(frames, roi, h, capacity=2000, psf_size=psf, **args)[1])
    assert merged, "the case should merge some maxima"
    np.testing.assert_array_equal(pk, np.array(want_pk))
    np.testing.assert_array_equal(np.transpose(st, (2, 0, 1)), np.array(want_st))
    np.testing.assert_array_equal(val, np.array(want_v, np.float32))


def test_aggregate_planted_mistake_is_caught():
    """A search in row-major order (first ii, then jj) -- the plausible
    mistake -- gives another result on the three-in-a-row case."""
    def rowmajor(data, thr):
        data = [list(r) for r in data]
        while True:
            hit = None
            for ii in range(len(data)):
                for jj in range(ii + 1, len(data)):
                    d = np.hypot(data[ii][0] - data[jj][0], data[ii][1] - data[jj][1])
                    if 0 < d < thr:
                        hit = (ii, jj)
                        break
                if hit:
                    break
            if not hit:
                return np.array(data, np.float64)
            ii, jj = hit
            data[ii] = [np.floor((a + b) / 2 + 0.5) for a, b in zip(data[ii], data[jj])]
            del data[jj]
    # column-major merges 7-9 first (column 3) and then 3-5; row-major
    # merges 3-7 first and chains everything into one maximum
    case = [[3, 5, 1.0], [7, 5, 1.0], [9, 5, 1.0], [5, 5, 1.0]]
    np.testing.assert_array_equal(oracle.aggregate(case, 3.5), [[4, 5, 1.0], [8, 5, 1.0]])
    assert not np.array_equal(rowmajor(case, 3.5), oracle.aggregate(case, 3.5))
    np.testing.assert_array_equal(_aggre_matlab(case, 3.5), oracle.aggregate(case, 3.5))
