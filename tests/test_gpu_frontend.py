"""The GPU front-end (bsf_find_rois_*) against the oracle restatement of the
GUI fitter's frame -> ROI steps (oracle/bsf_frontend.c): the same FP32
operations in the same order, so the ROI count, the peak rows
(x, y, frame), the photon ROIs and the filtered values must be identical bit
for bit.  Covers both pixel types, Gaussian / DoG / PSF / random kernels
(odd, even, rectangular; the unrolled sizes up to 17 rows and the generic
kernel beyond), frame sizes that are not multiples of the 64 x 40 tile, the
DoG mode's first-column minimum, mirroring, the PSF-mode border test,
capacity truncation, empty results and a 2048 x 2048 frame; then frames ->
ROIs -> fits equals fitting the oracle's ROIs."""
import ctypes as C

import numpy as np
import pytest

import bsfit
import oracle
from bsfit import frames as FR
from bsfit import gen
from parity_util import Hip

pytestmark = pytest.mark.gpu


def _kernel(kind, rng):
    if kind[0] in ("dog", "gauss"):
        return FR.peak_filter(kind[0], kind[1])
    if kind[0] == "psf":
        psf = np.transpose(gen.psf_astig3d(), (0, 1, 2)).astype(np.float32)
        return FR.peak_filter_psf("window", psf, kind[1])
    return np.asfortranarray(rng.uniform(-0.05, 0.2, kind[1]).astype(np.float32))


CASES = [
    # (nrows, ncols, nframes, dtype, kernel, subtract_min, mirror, psf_size, roi, cutoff)
    (256, 200, 4, np.uint16, ("dog", 1.0), True, False, 0.0, 13, 15.0),
    (130, 97, 3, np.float32, ("gauss", 1.0), False, True, 0.0, 11, 40.0),
    (64, 64, 2, np.uint16, ("rand", (5, 8)), True, False, 27.0, 9, 20.0),
    (100, 301, 2, np.uint16, ("rand", (19, 19)), False, False, 0.0, 15, 60.0),
    (67, 45, 1, np.float32, ("rand", (1, 1)), False, False, 0.0, 7, 30.0),
    (300, 70, 1, np.uint16, ("rand", (31, 25)), True, True, 0.0, 13, 100.0),
    (200, 180, 2, np.uint16, ("psf", 13), True, False, 27.0, 13, 0.5),
    (129, 81, 2, np.float32, ("dog", 2.6), True, False, 0.0, 17, 5.0),
    (70, 66, 1, np.uint16, ("rand", (51, 3)), False, False, 0.0, 3, 10.0),
]


def _compare(got, want):
    st, pk, val, n = got
    st_o, pk_o, val_o, n_o = want
    assert n == n_o
    np.testing.assert_array_equal(pk, pk_o)
    np.testing.assert_array_equal(st, st_o)
    np.testing.assert_array_equal(val, val_o)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_find_rois_matches_oracle(case):
    nr, nc, nf, dt, kind, sub, mirror, psf, roi, cutoff = CASES[case]
    rng = np.random.default_rng(100 + case)
    frames, _ = gen.make_frames(nf, nr, nc, max(2, nr * nc // 2500), 500 + case, dtype=dt)
    h = _kernel(kind, rng)
    args = dict(cutoff=cutoff, offset=100.0, conversion=0.5, subtract_min=sub, mirror=mirror, psf_size=psf,
                first_frame=1, capacity=20000)
    got = FR.find_rois(frames, roi, h, **args)
    want = oracle.find_rois(frames, roi, h, **args)
    print("case %d: %dx%dx%d %s kernel %s: %d ROIs" % (case, nr, nc, nf, np.dtype(dt).name, h.shape, got[3]),
          flush=True)
    assert got[3] > 0
    _compare(got, want)


def test_capacity_truncation_and_empty_frames():
    frames, _ = gen.make_frames(3, 96, 96, 8, 77)
    h = FR.peak_filter("dog", 1.0)
    args = dict(cutoff=10.0, offset=100.0, conversion=0.5, subtract_min=True)
    full = FR.find_rois(frames, 11, h, capacity=5000, **args)
    part = FR.find_rois(frames, 11, h, capacity=7, **args)
    assert part[3] == full[3] > 7
    np.testing.assert_array_equal(part[1], full[1][:7])
    np.testing.assert_array_equal(part[0], full[0][:, :, :7])
    _compare(part, oracle.find_rois(frames, 11, h, capacity=7, **args))
    flat = np.full((80, 90, 2), 300, np.uint16)               # no strict maximum anywhere
    got = FR.find_rois(flat, 11, h, capacity=10, **args)
    assert got[3] == 0 and got[1].shape == (0, 3)
    none = FR.find_rois(np.zeros((80, 90, 0), np.uint16), 11, h, capacity=10, **args)
    assert none[3] == 0


def test_large_frame():
    frames, truth = gen.make_frames(1, 2048, 2048, 600, 9)
    h = FR.peak_filter("dog", 1.0)
    args = dict(cutoff=15.0, offset=100.0, conversion=0.5, subtract_min=True, capacity=20000)
    got = FR.find_rois(frames, 13, h, **args)
    _compare(got, oracle.find_rois(frames, 13, h, **args))
    assert got[3] >= 0.8 * len(truth)


def test_device_entry_point_matches_host():
    lib = FR._bind()
    frames, _ = gen.make_frames(3, 150, 130, 10, 31)
    h = FR.peak_filter("dog", 1.2)
    cap, roi = 400, 13
    host = FR.find_rois(frames, roi, h, cutoff=12.0, offset=100.0, conversion=0.5, subtract_min=True, capacity=cap)
    hip = Hip()
    d_fr, d_h = hip.put(frames), hip.put(h)
    q = FR.make_request(d_fr, FR.PIXEL_U16, frames.shape, roi, d_h, h.shape, 12.0, 100.0, 0.5, True, False, 0.0, 1,
                        cap)
    wb = lib.bsf_find_rois_work_bytes(C.byref(q))
    d_w, d_st, d_pk, d_v, d_n = (hip.alloc(wb), hip.alloc(roi * roi * cap * 4), hip.alloc(cap * 12),
                                 hip.alloc(cap * 4), hip.alloc(4))
    try:
        res = FR.RoiResult(d_st, d_pk, d_v, 0)
        rc = lib.bsf_find_rois_device(C.byref(q), C.byref(res), d_n, d_w, wb, None)
        assert rc == 0, lib.bsf_last_error()
        assert hip.h.hipDeviceSynchronize() == 0
        n = int(hip.get(d_n, 1, np.int32)[0])
        pk = hip.get(d_pk, cap * 3, np.int32).reshape(cap, 3, order="F")[:n]
        st = hip.get(d_st, roi * roi * cap, np.float32).reshape(roi, roi, cap, order="F")[:, :, :n]
        v = hip.get(d_v, cap, np.float32)[:n]
        # a work buffer that is too small is refused before any launch
        rc = lib.bsf_find_rois_device(C.byref(q), C.byref(res), d_n, d_w, wb - 8, None)
        assert rc == 1
    finally:
        hip.free(d_fr, d_h, d_w, d_st, d_pk, d_v, d_n)
    assert n == host[3]
    np.testing.assert_array_equal(pk, host[1])
    np.testing.assert_array_equal(st, host[0])
    np.testing.assert_array_equal(v, host[2])


def test_frames_to_fits():
    """The GUI's path on raw frames: ROIs from the GPU front-end fitted with
    fittype 6 give the same results as the oracle's ROIs fitted the same way
    (the ROIs are identical), and the fits land on the planted emitters."""
    coeff = gen.model_coeffs("astig3d")
    frames, truth = gen.make_frames(4, 192, 192, 20, 55, coeff=coeff)
    h = FR.peak_filter("dog", 1.0)
    args = dict(cutoff=15.0, offset=100.0, conversion=0.5, subtract_min=True, capacity=2000)
    st, pk, _, n = FR.find_rois(frames, 13, h, **args)
    st_o = oracle.find_rois(frames, 13, h, **args)[0]
    P, CR, LL = bsfit.fit(6, st, 50, coeff)
    Po, CRo, LLo = bsfit.fit(6, st_o, 50, coeff)
    np.testing.assert_array_equal(P, Po)
    # localisations in frame pixels: ROI corner (x - 1 - dx, y - 1 - dx) plus the fitted column/row position
    xs = pk[:, 0] - 1 - 6 + P[:, 0]
    ys = pk[:, 1] - 1 - 6 + P[:, 1]
    hit = 0
    for f, x, y in truth[:, :3]:
        m = pk[:, 2] == f + 1
        hit += bool(m.any() and np.hypot(xs[m] - x, ys[m] - y).min() < 0.3)
    assert hit >= 0.7 * len(truth), (hit, len(truth))


@pytest.mark.parametrize("ft", [6, 7])
def test_fit_frames_equals_find_then_fit(ft):
    """bsf_fit_frames_host (ROIs and fits in one call, the fit reading the
    ROI count on the device) equals bsf_find_rois_host followed by
    bsf_fit_host on its ROIs, bit for bit; a capacity below the count
    fits the first capacity ROIs."""
    coeff = gen.model_coeffs("astig3d" if ft == 6 else "color4d")
    frames, _ = gen.make_frames(3, 128, 120, 10, 40 + ft, coeff=coeff)
    h = FR.peak_filter("dog", 1.0)
    initP = None if ft == 6 else np.array([30.5, 1.5], np.float32)
    for cap in (1000, 9):
        args = dict(cutoff=15.0, offset=100.0, conversion=0.5, subtract_min=True, capacity=cap)
        st, pk, val, P, CR, LL, n = FR.fit_frames(frames, 13, h, fittype=ft, iterations=50, coeff=coeff,
                                                  initP=initP, **args)
        st2, pk2, val2, n2 = FR.find_rois(frames, 13, h, **args)
        assert n == n2 > 9
        np.testing.assert_array_equal(st, st2)
        np.testing.assert_array_equal(pk, pk2)
        P2, CR2, LL2 = bsfit.fit(ft, st2, 50, coeff, initP)
        np.testing.assert_array_equal(P, P2)
        np.testing.assert_array_equal(CR, CR2)
        np.testing.assert_array_equal(LL, LL2)


@pytest.mark.parametrize("seed", range(12))
def test_random_configurations_match_oracle(seed):
    """Randomly drawn configurations (frame shape, kernel shape 1..51 on each
    side, pixel type, first-column minimum, mirror, PSF border, ROI size,
    cutoff, capacity) against the oracle, bit for bit."""
    rng = np.random.default_rng(7000 + seed)
    nr, nc = int(rng.integers(3, 260)), int(rng.integers(3, 260))
    nf = int(rng.integers(1, 4))
    dt = np.uint16 if rng.random() < 0.6 else np.float32
    mr, mc = int(rng.integers(1, 52)), int(rng.integers(1, 52))
    roi = int(rng.choice(np.arange(3, 52, 2)))
    sub, mirror = bool(rng.random() < 0.5), bool(rng.random() < 0.3)
    psf = float(rng.choice([0.0, 0.0, 13.0, 27.0]))
    cap = int(rng.integers(0, 400))
    if nr > 2 * 11 + 2 and nc > 2 * 11 + 2:
        frames, _ = gen.make_frames(nf, nr, nc, max(1, nr * nc // 3000), 8000 + seed, dtype=dt)
    else:                                       # tiny frames: noise only
        frames = (rng.poisson(30, (nr, nc, nf)) / 0.5 + 100).astype(dt)
        frames = np.asfortranarray(frames)
    h = np.asfortranarray(rng.uniform(-0.02, 0.05, (mr, mc)).astype(np.float32))
    cutoff = float(rng.uniform(-5, 40))
    args = dict(cutoff=cutoff, offset=100.0, conversion=0.5, subtract_min=sub, mirror=mirror, psf_size=psf,
                first_frame=int(rng.integers(0, 5)), capacity=cap)
    print("seed %d: %dx%dx%d %s kernel %dx%d roi %d sub %d mirror %d psf %g cap %d" %
          (seed, nr, nc, nf, np.dtype(dt).name, mr, mc, roi, sub, mirror, psf, cap), flush=True)
    _compare(FR.find_rois(frames, roi, h, **args), oracle.find_rois(frames, roi, h, **args))


AGG_CASES = [
    # (nrows, ncols, nframes, dtype, kernel, subtract_min, mirror, psf_size, roi, cutoff, emitters, agg_dist)
    (128, 128, 3, np.uint16, ("dog", 1.0), True, False, 0.0, 13, 8.0, 60, 13.0),
    (160, 96, 2, np.float32, ("gauss", 1.0), False, True, 0.0, 11, 35.0, 50, 11.0),
    (96, 96, 3, np.uint16, ("rand", (5, 8)), True, False, 27.0, 9, 12.0, 40, 8.0),
    (70, 130, 2, np.uint16, ("dog", 1.0), True, False, 0.0, 7, 4.0, 30, 5.0),   # distance 5 exactly: kept
    (64, 64, 2, np.uint16, ("dog", 1.0), True, False, 0.0, 7, 2.0, 20, 1.0),    # below any distance: no merges
    (90, 90, 2, np.uint16, ("rand", (3, 3)), False, False, 0.0, 5, 0.0, 40, 500.0),   # one maximum per frame
]


@pytest.mark.parametrize("case", range(len(AGG_CASES)))
def test_aggregation_matches_oracle(case):
    """With agg_dist (the GUI's p.agg) the frames' maxima merge as aggre
    merges them, before the PSF border test: counts, peak rows, ROIs and the
    value column equal the oracle's bit for bit."""
    nr, nc, nf, dt, kind, sub, mirror, psf, roi, cutoff, em, thr = AGG_CASES[case]
    rng = np.random.default_rng(300 + case)
    frames, _ = gen.make_frames(nf, nr, nc, em, 900 + case, dtype=dt)
    h = _kernel(kind, rng)
    args = dict(cutoff=cutoff, offset=100.0, conversion=0.5, subtract_min=sub, mirror=mirror, psf_size=psf,
                first_frame=2, capacity=20000)
    plain = FR.find_rois(frames, roi, h, **args)
    got = FR.find_rois(frames, roi, h, agg_dist=thr, **args)
    want = oracle.find_rois(frames, roi, h, agg_dist=thr, **args)
    print("agg case %d: %d maxima, %d after aggregation at %g" % (case, plain[3], got[3], thr), flush=True)
    _compare(got, want)
    if thr <= 1.0:
        _compare(got, plain)              # strict maxima are at least 2 apart
    elif case < 3:
        assert got[3] < plain[3], "the case should merge some maxima"


def test_aggregation_capacity_device_and_fits():
    """Aggregation through the capacity limit, the device entry point (its
    larger workspace) and frames -> fits."""
    frames, _ = gen.make_frames(3, 128, 128, 70, 4242)
    h = FR.peak_filter("dog", 1.0)
    args = dict(cutoff=8.0, offset=100.0, conversion=0.5, subtract_min=True, agg_dist=13.0)
    full = FR.find_rois(frames, 13, h, capacity=5000, **args)
    part = FR.find_rois(frames, 13, h, capacity=9, **args)
    assert part[3] == full[3] > 9
    _compare(part, oracle.find_rois(frames, 13, h, capacity=9, **args))
    # device entry point: the workspace grows with aggregation
    lib = FR._bind()
    hip = Hip()
    cap, roi = 64, 13
    d_fr, d_h = hip.put(frames), hip.put(h)
    q0 = FR.make_request(d_fr, FR.PIXEL_U16, frames.shape, roi, d_h, h.shape, 8.0, 100.0, 0.5, True, False, 0.0, 1,
                         cap)
    q = FR.make_request(d_fr, FR.PIXEL_U16, frames.shape, roi, d_h, h.shape, 8.0, 100.0, 0.5, True, False, 0.0, 1,
                        cap, 13.0)
    w0, wb = lib.bsf_find_rois_work_bytes(C.byref(q0)), lib.bsf_find_rois_work_bytes(C.byref(q))
    assert wb >= w0 + 3 * 63 * 63 * 20
    d_w, d_st, d_pk, d_v, d_n = (hip.alloc(wb), hip.alloc(roi * roi * cap * 4), hip.alloc(cap * 12),
                                 hip.alloc(cap * 4), hip.alloc(4))
    try:
        res = FR.RoiResult(d_st, d_pk, d_v, 0)
        rc = lib.bsf_find_rois_device(C.byref(q), C.byref(res), d_n, d_w, wb - 8, None)
        assert rc == 1                     # refused before any launch
        rc = lib.bsf_find_rois_device(C.byref(q), C.byref(res), d_n, d_w, wb, None)
        assert rc == 0, lib.bsf_last_error()
        assert hip.h.hipDeviceSynchronize() == 0
        n = int(hip.get(d_n, 1, np.int32)[0])
        k = min(n, cap)
        pk = hip.get(d_pk, cap * 3, np.int32).reshape(cap, 3, order="F")[:k]
        st = hip.get(d_st, roi * roi * cap, np.float32).reshape(roi, roi, cap, order="F")[:, :, :k]
        v = hip.get(d_v, cap, np.float32)[:k]
        # the allocation-free entry point (its 16 MiB workspace holds these frames' maxima)
        lib.bsf_find_rois_device_noalloc.argtypes = [C.POINTER(FR.RoiRequest), C.POINTER(FR.RoiResult), C.c_void_p,
                                                     C.c_void_p]
        lib.bsf_find_rois_device_noalloc.restype = C.c_int
        rc = lib.bsf_find_rois_device_noalloc(C.byref(q), C.byref(res), d_n, None)
        assert rc == 0, lib.bsf_last_error()
        assert hip.h.hipDeviceSynchronize() == 0
        assert int(hip.get(d_n, 1, np.int32)[0]) == n
        np.testing.assert_array_equal(hip.get(d_pk, cap * 3, np.int32).reshape(cap, 3, order="F")[:k], pk)
        np.testing.assert_array_equal(hip.get(d_v, cap, np.float32)[:k], v)
    finally:
        hip.free(d_fr, d_h, d_w, d_st, d_pk, d_v, d_n)
    want = FR.find_rois(frames, roi, h, capacity=cap, **args)
    assert n == want[3] == full[3]
    np.testing.assert_array_equal(pk, want[1])
    np.testing.assert_array_equal(st, want[0])
    np.testing.assert_array_equal(v, want[2])
    # frames -> ROIs -> fits with aggregation equals fitting the aggregated ROIs
    w = gen.config(0, N=4)
    st, pk, val, P, CR, LL, n = FR.fit_frames(frames, 13, h, 8.0, 6, 50, w.coeff, offset=100.0, conversion=0.5,
                                              subtract_min=True, capacity=5000, agg_dist=13.0)
    assert n == full[3]
    np.testing.assert_array_equal(pk, full[1])
    P2, CR2, LL2 = bsfit.fit(6, full[0], 50, w.coeff)
    np.testing.assert_array_equal(P, P2)
    np.testing.assert_array_equal(CR, CR2)
    np.testing.assert_array_equal(LL, LL2)


@pytest.mark.parametrize("thr,psf", [(4.0, 0.0), (6.5, 27.0), (20.0, 0.0)])
def test_dense_aggregation_matches_oracle(thr, psf):
    """Every strict maximum of noisy frames (a cutoff far below the noise):
    ~670 per frame, merging in chains that carry merged maxima across the
    merge's cells (4, 6 and 19 pixels wide); counts, rows, ROIs and values
    equal the oracle's bit for bit."""
    frames, _ = gen.make_frames(2, 128, 144, 20, 7100, dtype=np.uint16)
    h = FR.peak_filter("dog", 1.0)
    args = dict(cutoff=-1e9, offset=100.0, conversion=0.5, subtract_min=True, psf_size=psf, first_frame=1,
                capacity=20000)
    plain = FR.find_rois(frames, 9, h, **args)
    got = FR.find_rois(frames, 9, h, agg_dist=thr, **args)
    want = oracle.find_rois(frames, 9, h, agg_dist=thr, **args)
    print("dense agg %g psf %g: %d maxima, %d after aggregation" % (thr, psf, plain[3], got[3]), flush=True)
    assert plain[3] > 500 and got[3] < plain[3]    # (the PSF border test leaves 820 of the 1,345)
    _compare(got, want)


def test_dense_aggregation_large_frames():
    """Two 1024 x 1024 frames with every strict maximum kept (tens of
    thousands per frame) and agg_dist 5: too many maxima for the oracle's
    literal search, so the result is checked for what aggre's loop leaves --
    no two maxima of a frame at a distance 0 < D < 5, each survivor within
    the frame, no more survivors than maxima -- and the merge, whose next
    pair is searched among the neighbouring cells only, ends in seconds."""
    import time
    frames, _ = gen.make_frames(2, 1024, 1024, 400, 7200, dtype=np.uint16)
    h = FR.peak_filter("dog", 1.0)
    args = dict(cutoff=-1e9, offset=100.0, conversion=0.5, subtract_min=True, first_frame=1, capacity=400000)
    plain = FR.find_rois(frames, 3, h, **args)
    t0 = time.perf_counter()
    got = FR.find_rois(frames, 3, h, agg_dist=5.0, **args)
    dt = time.perf_counter() - t0
    n = got[3]
    print("large dense agg: %d maxima, %d after aggregation, %.2f s" % (plain[3], n, dt), flush=True)
    assert plain[3] > 50000 and 0 < n < plain[3]
    assert dt < 20.0
    pk = got[1][:n]
    offs = [(a, b) for a in range(-4, 5) for b in range(-4, 5) if 0 < a * a + b * b <= 24]
    for f in (1, 2):
        q = pk[pk[:, 2] == f]
        x, y = q[:, 0], q[:, 1]
        assert x.min() >= 1 and x.max() <= 1024 and y.min() >= 1 and y.max() <= 1024
        occ = np.zeros((1024 + 10, 1024 + 10), bool)
        occ[x + 4, y + 4] = True
        for a, b in offs:
            assert not occ[x + 4 + a, y + 4 + b].any(), (f, a, b)


@pytest.mark.parametrize("seed", range(6))
def test_random_aggregation_matches_oracle(seed):
    """Random frames, kernels, ROI sizes, PSF borders and thresholds with
    aggregation, against the oracle bit for bit."""
    rng = np.random.default_rng(9100 + seed)
    nr, nc = int(rng.integers(40, 200)), int(rng.integers(40, 200))
    nf = int(rng.integers(1, 4))
    dt = np.uint16 if rng.random() < 0.6 else np.float32
    roi = int(rng.choice(np.arange(3, 16, 2)))
    psf = float(rng.choice([0.0, 13.0, 27.0]))
    frames, _ = gen.make_frames(nf, nr, nc, max(2, nr * nc // 400), 9500 + seed, dtype=dt)
    h = FR.peak_filter("dog", float(rng.uniform(0.8, 1.6)))
    thr = float(rng.choice([2.0, 2.5, np.sqrt(8.0), 7.0, roi * rng.uniform(0.5, 2.0)]))
    args = dict(cutoff=float(rng.uniform(2, 20)), offset=100.0, conversion=0.5, subtract_min=True,
                mirror=bool(rng.random() < 0.3), psf_size=psf, first_frame=1, capacity=int(rng.integers(1, 3000)),
                agg_dist=thr)
    print("seed %d: %dx%dx%d roi %d psf %g agg %g cap %d" % (seed, nr, nc, nf, roi, psf, thr, args["capacity"]),
          flush=True)
    _compare(FR.find_rois(frames, roi, h, **args), oracle.find_rois(frames, roi, h, **args))
