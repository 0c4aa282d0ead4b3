"""The allocation-free front-end entry point (bsf_find_rois_device_noalloc)
through the C-ABI against bsf_find_rois_device with a caller-owned work
buffer: two calls in a row, the second with a smaller input (the library's
workspace keeps the first call's contents), each bit for bit equal in ROI
count, peak rows, ROI stack and filtered values; and a request larger than
the workspace is refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from parity_util import Hip

pytestmark = pytest.mark.gpu


def _run(lib, hip, frames, h, roi, cutoff, noalloc):
    from bsfit import frames as FR
    cap = frames.shape[2] * 200
    d_fr = hip.put(np.asfortranarray(frames))
    d_h = hip.put(np.asfortranarray(h, np.float32))
    d_st, d_pk, d_val = hip.alloc(roi * roi * cap * 4), hip.alloc(3 * cap * 4), hip.alloc(cap * 4)
    d_n = hip.put(np.full(1, -7, np.int32))
    q = FR.make_request(d_fr, FR.PIXEL_U16, frames.shape, roi, d_h, h.shape, cutoff, 100.0, 0.5, True, False, 0.0,
                        1, cap)
    res = FR.RoiResult(d_st, d_pk, d_val, 0)
    d_w = None
    try:
        if noalloc:
            rc = lib.bsf_find_rois_device_noalloc(C.byref(q), C.byref(res), d_n, None)
        else:
            wb = lib.bsf_find_rois_work_bytes(C.byref(q))
            d_w = hip.alloc(wb)
            rc = lib.bsf_find_rois_device(C.byref(q), C.byref(res), d_n, d_w, wb, None)
        assert rc == 0, lib.bsf_last_error().decode()
        assert hip.h.hipDeviceSynchronize() == 0
        n = int(hip.get(d_n, 1, np.int32)[0])
        k = min(n, cap)
        out = (n, hip.get(d_pk, 3 * cap, np.int32).reshape(3, cap)[:, :k].copy(),
               hip.get(d_st, roi * roi * cap, np.float32)[:roi * roi * k].copy(),
               hip.get(d_val, cap, np.float32)[:k].copy())
    finally:
        hip.free(d_fr, d_h, d_st, d_pk, d_val, d_n, d_w)
    return out


def test_noalloc_front_end_equals_work_buffer_entry_point():
    from bsfit import frames as FR
    from bsfit import gen
    lib = FR._bind()
    lib.bsf_find_rois_device_noalloc.argtypes = [C.POINTER(FR.RoiRequest), C.POINTER(FR.RoiResult), C.c_void_p,
                                                 C.c_void_p]
    lib.bsf_find_rois_device_noalloc.restype = C.c_int
    hip = Hip()
    coeff = gen.model_coeffs("astig3d")
    h = FR.peak_filter("dog", 1.0)
    # first call: 6 frames of 256 x 320; second: 2 frames of 128 x 192 (smaller)
    for nf, nr, nc, seed in ((6, 256, 320, 11), (2, 128, 192, 12)):
        frames, _ = gen.make_frames(nf, nr, nc, 40, seed, coeff=coeff)
        print("case noalloc frames=%d %dx%d" % (nf, nr, nc), flush=True)
        ref = _run(lib, hip, frames, h, 13, 15.0, noalloc=False)
        got = _run(lib, hip, frames, h, 13, 15.0, noalloc=True)
        print("ROIs found: %d (work buffer) / %d (no allocation)" % (ref[0], got[0]), flush=True)
        assert ref[0] > 10
        assert got[0] == ref[0]
        for a, b in zip(got[1:], ref[1:]):
            np.testing.assert_array_equal(a, b)


def test_noalloc_refuses_requests_beyond_its_workspace():
    """150 frames of 1024 x 1024 need more than the 16 MiB workspace: refused
    with BSF_ERR_NOMEM before anything is read from the (dummy) pointers."""
    from bsfit import frames as FR
    lib = FR._bind()
    lib.bsf_find_rois_device_noalloc.argtypes = [C.POINTER(FR.RoiRequest), C.POINTER(FR.RoiResult), C.c_void_p,
                                                 C.c_void_p]
    lib.bsf_find_rois_device_noalloc.restype = C.c_int
    h = np.ones((3, 3), np.float32)
    dummy = np.zeros(16, np.float32)
    q = FR.make_request(dummy.ctypes.data, FR.PIXEL_U16, (1024, 1024, 150), 13, h.ctypes.data, h.shape, 15.0, 100.0,
                        0.5, True, False, 0.0, 1, 1000)
    res = FR.RoiResult(dummy.ctypes.data, dummy.ctypes.data, None, 0)
    assert lib.bsf_find_rois_work_bytes(C.byref(q)) > 16 << 20
    assert lib.bsf_find_rois_device_noalloc(C.byref(q), C.byref(res), dummy.ctypes.data, None) == 5
    assert "workspace" in lib.bsf_last_error().decode()
