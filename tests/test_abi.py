"""The C-ABI library (libbsfit.so) loads and exports every function that
include/bsfit.h declares; host-only entry points answer without a GPU, and
compute entry points fail with BSF_ERR_DEVICE (never crash) where no gfx950
device is present."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "bsfit.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bsf_\w+)\s*\(", src)))


def test_exports_every_declared_symbol():
    import bsfit
    lib = bsfit.load()
    names = _declared()
    assert "bsf_fit_host" in names and "bsf_fit_4pi_host" in names
    for nm in names:
        assert hasattr(lib, nm), "libbsfit.so does not export %s" % nm


def test_host_only_entry_points():
    import bsfit
    assert [bsfit.num_params(f) for f in (6, 7, 8, 5)] == [5, 6, 8, -1]
    assert bsfit.num_params_4pi([1] * 6, 4) == 6
    assert bsfit.num_params_4pi([1, 1, 1, 0, 1, 1], 4) == 9


def test_errors_without_device_or_bad_args():
    import torch
    import bsfit
    from bsfit import gen
    w = gen.config(0, N=2)
    # argument errors are reported before any device work
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.fit(9, w.imstack, 50, w.coeff)
    assert e.value.code == 2
    if not torch.cuda.is_available():
        with pytest.raises(bsfit.BsfError) as e:
            bsfit.fit(6, w.imstack, 50, w.coeff)
        assert e.value.code == 6


def test_status_entry_points_without_device_or_bad_args():
    import torch
    import bsfit
    from bsfit import gen
    from bsfit.api import Request, Result, _request
    lib = bsfit.load()
    w = gen.config(0, N=2)
    im = np.asfortranarray(w.imstack, np.float32)
    c = np.asfortranarray(w.coeff, np.float32)
    P = np.zeros((2, 6), np.float32, order="F")
    CR = np.zeros((2, 5), np.float32, order="F")
    LL = np.zeros(2, np.float32)
    req = _request(6, im.shape[0], 2, 50, c.ctypes.data, c.shape, None, None, bsfit.PREC_MIX, im.ctypes.data)
    res = Result(P.ctypes.data, CR.ctypes.data, LL.ctypes.data)
    lib.bsf_fit_host_status.argtypes = [C.POINTER(Request), C.POINTER(Result), C.c_void_p]
    lib.bsf_fit_device_status.argtypes = [C.POINTER(Request), C.POINTER(Result), C.c_void_p, C.c_void_p]
    # a NULL status array is an argument error, reported before any device work
    assert lib.bsf_fit_host_status(C.byref(req), C.byref(res), None) == 1
    assert lib.bsf_fit_device_status(C.byref(req), C.byref(res), None, None) == 1
    assert "status" in lib.bsf_last_error().decode()
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.fit_status(9, w.imstack, 50, w.coeff)
    assert e.value.code == 2
    if not torch.cuda.is_available():
        with pytest.raises(bsfit.BsfError) as e:
            bsfit.fit_status(6, w.imstack, 50, w.coeff)
        assert e.value.code == 6


def test_session_errors_without_device_or_bad_args():
    import torch
    import bsfit
    # configuration errors are reported before any device work
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.Session(kernel_streams=17)
    assert e.value.code == 1
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.Session(hbm_budget=1 << 20, arena_bytes=2 << 20)
    assert e.value.code == 5
    if not torch.cuda.is_available():
        with pytest.raises(bsfit.BsfError) as e:
            bsfit.Session()
        assert e.value.code == 6


def test_coefficient_builder_argument_errors():
    import torch
    import bsfit
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.coeff_build(np.zeros((5, 1, 4), np.float32))       # an axis with one sample
    assert e.value.code == 4
    with pytest.raises(bsfit.BsfError) as e:
        bsfit.coeff_build(np.zeros((5, 4), np.float32), lam=[0.1, -1.0])
    assert e.value.code == 1
    if not torch.cuda.is_available():
        with pytest.raises(bsfit.BsfError) as e:
            bsfit.coeff_build(np.zeros((5, 4), np.float32))
        assert e.value.code == 6


def test_multistart_argument_errors():
    import torch
    import bsfit
    from bsfit import gen
    w = gen.config(1, N=2)
    with pytest.raises(bsfit.BsfError) as e:                      # fittype 6 has no start
        bsfit.fit_multistart(6, w.imstack, 50, w.coeff, np.zeros((2, 2), np.float32))
    assert e.value.code == 2
    with pytest.raises(bsfit.BsfError) as e:                      # fittype 7 starts are {z, w}
        bsfit.fit_multistart(7, w.imstack, 50, w.coeff, np.zeros((2, 1), np.float32))
    assert e.value.code == 1
    with pytest.raises(bsfit.BsfError) as e:                      # at most 64 starts
        bsfit.fit_multistart(7, w.imstack, 50, w.coeff, np.zeros((65, 2), np.float32))
    assert e.value.code == 1
    if not torch.cuda.is_available():
        with pytest.raises(bsfit.BsfError) as e:
            bsfit.fit_multistart(7, w.imstack, 50, w.coeff, np.ones((2, 2), np.float32))
        assert e.value.code == 6


def test_front_end_argument_errors_and_host_kernels():
    import torch
    import bsfit
    from bsfit import frames as FR
    h = FR.peak_filter("dog", 1.0)                                 # host code: no device needed
    assert h.shape == (7, 7)
    fr = np.zeros((32, 32, 1), np.uint16)
    for kw, code in ((dict(roi=12), 3), (dict(roi=53), 3), (dict(roi=7, filt=np.zeros((52, 3), np.float32)), 4),
                     (dict(roi=7, capacity=-1), 1), (dict(roi=7, agg_dist=-1.0), 1),
                     (dict(roi=7, agg_dist=float("nan")), 1), (dict(roi=7, agg_dist=float("inf")), 1)):
        args = dict(filt=h, cutoff=1.0, capacity=4)
        args.update(kw)
        roi = args.pop("roi")
        with pytest.raises(bsfit.BsfError) as e:
            FR.find_rois(fr, roi, **args)
        assert e.value.code == code, kw
    with pytest.raises(bsfit.BsfError) as e:
        FR.find_rois(np.zeros((2, 40, 1), np.uint16), 7, h, 1.0, capacity=4)   # frames below 3 x 3
    assert e.value.code == 4
    with pytest.raises(bsfit.BsfError):
        FR.find_rois(np.zeros((8, 8, 1), np.int32), 7, h, 1.0)                # pixel type
    if not torch.cuda.is_available():
        with pytest.raises(bsfit.BsfError) as e:
            FR.find_rois(fr, 7, h, 1.0, capacity=4)
        assert e.value.code == 6
    # aggregation's workspace: every maximum a frame can hold (strict maxima
    # are never adjacent: ceil((n-2)/2) per interior row and column), 12 B of
    # peak row, 8 B of value and a 4 B cell link each, per-frame counts, and
    # per frame the heads of the merge's cells: cs x cs pixels, cs the
    # integer square root of the largest d2 with sqrt(d2) < agg_dist, at
    # least 4 (agg_dist 13: d2 <= 168, cs 12; 2.5: d2 <= 6, cs 4)
    lib = FR._bind()
    for nr, nc, nf in ((32, 32, 1), (101, 67, 3), (1024, 1024, 2)):
        q0 = FR.make_request(0, FR.PIXEL_U16, (nr, nc, nf), 7, 0, h.shape, 1.0, 0.0, 1.0, False, False, 0.0, 1, 4)
        w0 = lib.bsf_find_rois_work_bytes(C.byref(q0))
        most = nf * ((nr - 1) // 2) * ((nc - 1) // 2)
        for dist, cs in ((13.0, 12), (2.5, 4)):
            q1 = FR.make_request(0, FR.PIXEL_U16, (nr, nc, nf), 7, 0, h.shape, 1.0, 0.0, 1.0, False, False, 0.0, 1,
                                 4, dist)
            w1 = lib.bsf_find_rois_work_bytes(C.byref(q1))
            heads = 4 * nf * (nc // cs + 1) * (nr // cs + 1)
            lo = 24 * most + heads
            assert lo <= w1 - w0 <= lo + 12 * nf + 2304, (nr, nc, nf, dist, w0, w1)


def test_oversized_coefficient_arrays_are_refused():
    """The kernels address coefficients with 32-bit byte offsets: arrays of
    2^31 bytes or more are refused (BSF_ERR_SHAPE) before any device work,
    by the 6/7/8 fit and by the 4Pi fit (no array that large is allocated:
    the checks read only the dimensions)."""
    import bsfit
    from bsfit.api import Request, Result, Request4Pi, Extras, _request
    lib = bsfit.load()
    buf = np.zeros(16, np.float32)
    P = np.zeros(16, np.float32)
    res = Result(P.ctypes.data, P.ctypes.data, P.ctypes.data)
    req = _request(6, 13, 1, 50, buf.ctypes.data, (1024, 1024, 600), None, None, bsfit.PREC_MIX, buf.ctypes.data)
    lib.bsf_fit_device.argtypes = [C.POINTER(Request), C.POINTER(Result), C.c_void_p]
    assert lib.bsf_fit_device(C.byref(req), C.byref(res), None) == 4
    assert "2^29" in lib.bsf_last_error().decode()
    lib.bsf_fit_device_ex.argtypes = [C.POINTER(Request), C.POINTER(Result), C.POINTER(Extras), C.c_void_p]
    x = Extras(None, None, None)
    assert lib.bsf_fit_device_ex(C.byref(req), C.byref(res), C.byref(x), None) == 4
    # just below the limit: the same check passes (then no device: BSF_ERR_DEVICE, or BSF_OK on a GPU box
    # would need real memory, so only the error code path is exercised where no GPU is present)
    import torch
    if not torch.cuda.is_available():
        req = _request(6, 13, 1, 50, buf.ctypes.data, (1024, 1024, 511), None, None, bsfit.PREC_MIX,
                       buf.ctypes.data)
        assert lib.bsf_fit_device(C.byref(req), C.byref(res), None) == 6
    sh = np.ones(6, np.uint32)
    ps = np.zeros(4, np.float32)
    r = Request4Pi()
    r.imstack, r.sz, r.nfits, r.nch = buf.ctypes.data, 13, 1, 4
    r.shared, r.iterations, r.coeff = sh.ctypes.data, 50, buf.ctypes.data
    for a, d in enumerate((1024, 1024, 64, 20)):
        r.coeff_dims[a] = d
    r.phaseShift, r.precision = ps.ctypes.data, bsfit.PREC_MIX
    lib.bsf_fit_4pi_device.argtypes = [C.POINTER(Request4Pi), C.POINTER(Result), C.c_void_p]
    assert lib.bsf_fit_4pi_device(C.byref(r), C.byref(res), None) == 4
    lib.bsf_fit_4pi_device_ex.argtypes = [C.POINTER(Request4Pi), C.POINTER(Result), C.POINTER(Extras), C.c_void_p]
    assert lib.bsf_fit_4pi_device_ex(C.byref(r), C.byref(res), C.byref(x), None) == 4


def test_roi_table_layout():
    """bsfit.roi_table packs bsf_roi records (16 bytes: off, cstride, sz)
    back to back by default."""
    import bsfit
    from bsfit.api import RoiDesc
    t = bsfit.roi_table([11, 13, 17])
    assert t.dtype.itemsize == C.sizeof(RoiDesc) == 16
    assert list(t["off"]) == [0, 121, 121 + 169]
    assert list(t["sz"]) == [11, 13, 17]


def test_tile_table_without_device():
    """The tile table (bsf_fit_tile / bsf_set_fit_tile) is host state:
    fittypes 6, 7 and FP32 fittype 8 take tiles 1 and 2, FP64 fittype 8
    only 1 (and no tile 2 above ROI size 32), tile 0 restores
    the default; bsf_tune and bsf_fit_device_ex check their arguments before
    any device work."""
    import bsfit
    from bsfit.api import Extras, Request, Result, _request
    lib = bsfit.load()
    for ft, prec in ((6, bsfit.PREC_MIX), (6, bsfit.PREC_F64), (7, bsfit.PREC_MIX), (7, bsfit.PREC_F64),
                     (8, bsfit.PREC_MIX)):
        d = bsfit.fit_tile(ft, prec, 13)
        assert d in (1, 2)
        try:
            bsfit.set_fit_tile(ft, prec, 13, 2)
            assert bsfit.fit_tile(ft, prec, 13) == 2
            bsfit.set_fit_tile(ft, prec, 13, 1)
            assert bsfit.fit_tile(ft, prec, 13) == 1
        finally:
            bsfit.set_fit_tile(ft, prec, 13, 0)
        assert bsfit.fit_tile(ft, prec, 13) == d
    assert bsfit.fit_tile(8, bsfit.PREC_MIX, 13) == 1 and bsfit.fit_tile(8, bsfit.PREC_F64, 13) == 1
    for bad in ((8, bsfit.PREC_F64, 13, 2), (8, bsfit.PREC_MIX, 33, 2), (6, bsfit.PREC_MIX, 13, 3),
                (6, bsfit.PREC_MIX, 0, 1), (6, 7, 13, 1), (7, bsfit.PREC_MIX, 65, 1)):
        with pytest.raises(bsfit.BsfError) as e:
            bsfit.set_fit_tile(*bad)
        assert e.value.code == 1
    buf = np.zeros(4096, np.float32)
    res = Result(buf.ctypes.data, buf.ctypes.data, buf.ctypes.data)
    req = _request(6, 13, 1, 50, buf.ctypes.data, (16, 16, 16), None, None, bsfit.PREC_MIX, buf.ctypes.data)
    t, ms = C.c_int(0), (C.c_float * 2)()
    lib.bsf_tune.argtypes = [C.POINTER(Request), C.POINTER(Result), C.c_int, C.POINTER(C.c_int),
                             C.POINTER(C.c_float), C.c_void_p]
    assert lib.bsf_tune(C.byref(req), C.byref(res), 0, C.byref(t), ms, None) == 1
    lib.bsf_fit_device_ex.argtypes = [C.POINTER(Request), C.POINTER(Result), C.POINTER(Extras), C.c_void_p]
    x = Extras(None, None, None, 3)
    assert lib.bsf_fit_device_ex(C.byref(req), C.byref(res), C.byref(x), None) == 1
    assert "tile" in lib.bsf_last_error().decode()


def test_session_graph_mode_argument_errors():
    import bsfit
    from bsfit.session import _bind
    lib = _bind(bsfit.load())
    assert lib.bsf_session_set_graphs(None, 32) == 1
    assert lib.bsf_session_graph_stats(None, None, None, None) == 1
    assert lib.bsf_session_set_hetero(None, 1) == 1
