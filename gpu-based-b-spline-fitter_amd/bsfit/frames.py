"""Camera frames -> ROI stack: ctypes binding of bsf_find_rois_* and the peak
filters (include/bsfit.h), the GUI fitter's front-end
(MultiD_fitter/simplefitter_spline.m:197-307, MultiD_fitter/shared/
maximumfind2.c:18-65).

    h = peak_filter("dog", sigma)                    # or "gauss", or peak_filter_psf(...)
    imstack, peaks, value = find_rois(frames, roi=13, filt=h, cutoff=2.0,
                                      offset=100, conversion=0.5, subtract_min=True)

``frames`` is (nrows, ncols, nframes) in MATLAB order (Fortran order; a 2-D
array is one frame), uint16 or float32.  ``peaks`` rows are (x, y, frame) as
the GUI's peakcoordinates: x = column + 1, y = row + 1, frames counted from
``first_frame``.  All work runs in libbsfit.so on the GPU.
"""
import ctypes as C

import numpy as np

from .api import _NV, BsfError, Request, Result, _check, _f32, _request, load

PIXEL_U16 = 0
PIXEL_F32 = 1
PEAK_GAUSS = 0
PEAK_DOG = 1
PEAK_PSF_WINDOW = 2
PEAK_PSF_MEAN = 3
MAX_PEAK_FILTER = 51


class RoiRequest(C.Structure):
    _fields_ = [
        ("frames", C.c_void_p),
        ("pixel_type", C.c_int),
        ("nrows", C.c_int),
        ("ncols", C.c_int),
        ("nframes", C.c_int),
        ("first_frame", C.c_int),
        ("offset", C.c_float),
        ("conversion", C.c_float),
        ("filter", C.c_void_p),
        ("filter_rows", C.c_int),
        ("filter_cols", C.c_int),
        ("subtract_min", C.c_int),
        ("cutoff", C.c_double),
        ("roi", C.c_int),
        ("mirror", C.c_int),
        ("psf_size", C.c_double),
        ("capacity", C.c_int),
        ("agg_dist", C.c_double),
    ]


class RoiResult(C.Structure):
    _fields_ = [("imstack", C.c_void_p), ("peaks", C.c_void_p), ("value", C.c_void_p), ("count", C.c_int)]


def _bind():
    lib = load()
    if not getattr(lib, "_frames_bound", False):
        lib.bsf_find_rois_host.argtypes = [C.POINTER(RoiRequest), C.POINTER(RoiResult)]
        lib.bsf_find_rois_host.restype = C.c_int
        lib.bsf_find_rois_work_bytes.argtypes = [C.POINTER(RoiRequest)]
        lib.bsf_find_rois_work_bytes.restype = C.c_size_t
        lib.bsf_find_rois_device.argtypes = [C.POINTER(RoiRequest), C.POINTER(RoiResult), C.c_void_p, C.c_void_p,
                                             C.c_size_t, C.c_void_p]
        lib.bsf_find_rois_device.restype = C.c_int
        lib.bsf_fit_frames_host.argtypes = [C.POINTER(RoiRequest), C.POINTER(Request), C.POINTER(RoiResult),
                                            C.POINTER(Result)]
        lib.bsf_fit_frames_host.restype = C.c_int
        lib.bsf_fit_frames_device.argtypes = [C.POINTER(RoiRequest), C.POINTER(Request), C.POINTER(RoiResult),
                                              C.c_void_p, C.POINTER(Result), C.c_void_p, C.c_size_t, C.c_void_p]
        lib.bsf_fit_frames_device.restype = C.c_int
        lib.bsf_peak_filter.argtypes = [C.c_int, C.c_double, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        lib.bsf_peak_filter.restype = C.c_int
        lib.bsf_peak_filter_psf.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                            C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.bsf_peak_filter_psf.restype = C.c_int
        lib._frames_bound = True
    return lib


_KINDS = {"gauss": PEAK_GAUSS, "dog": PEAK_DOG, PEAK_GAUSS: PEAK_GAUSS, PEAK_DOG: PEAK_DOG}


def peak_filter(kind, sigma):
    """The GUI's Gaussian ('gauss', background mode 'none') or difference of
    Gaussians ('dog', background mode 'Gaussian') kernel as a float32
    (n, n) Fortran-order array."""
    lib = _bind()
    h = np.zeros(MAX_PEAK_FILTER * MAX_PEAK_FILTER, np.float32)
    n = C.c_int()
    _check(lib.bsf_peak_filter(_KINDS[kind], float(sigma), h.ctypes.data, h.size, C.byref(n)))
    return np.asfortranarray(h[:n.value * n.value].reshape(n.value, n.value, order="F"))


def peak_filter_psf(kind, psf, roi):
    """Kernels from a calibration PSF (n, n, nz): 'window' (the p.conv kernel)
    or 'mean' (the 'Conv PSF' background mode)."""
    lib = _bind()
    k = {"window": PEAK_PSF_WINDOW, "mean": PEAK_PSF_MEAN}.get(kind, kind)
    psf = np.asfortranarray(psf, np.float32)
    if psf.ndim == 2:
        psf = psf[:, :, None]
    dims = np.array(psf.shape, np.int32)
    h = np.zeros(MAX_PEAK_FILTER * MAX_PEAK_FILTER, np.float32)
    r, c = C.c_int(), C.c_int()
    _check(lib.bsf_peak_filter_psf(k, psf.ctypes.data, dims.ctypes.data, int(roi), h.ctypes.data, h.size,
                                   C.byref(r), C.byref(c)))
    return np.asfortranarray(h[:r.value * c.value].reshape(r.value, c.value, order="F"))


def make_request(frames_ptr, pixel_type, shape, roi, filt_ptr, filt_shape, cutoff, offset, conversion,
                 subtract_min, mirror, psf_size, first_frame, capacity, agg_dist=0.0):
    q = RoiRequest()
    q.frames = frames_ptr
    q.pixel_type = pixel_type
    q.nrows, q.ncols, q.nframes = (int(v) for v in shape)
    q.first_frame = int(first_frame)
    q.offset = float(offset)
    q.conversion = float(conversion)
    q.filter = filt_ptr
    q.filter_rows, q.filter_cols = (int(v) for v in filt_shape)
    q.subtract_min = int(bool(subtract_min))
    q.cutoff = float(cutoff)
    q.roi = int(roi)
    q.mirror = int(bool(mirror))
    q.psf_size = float(psf_size)
    q.capacity = int(capacity)
    q.agg_dist = float(agg_dist)
    return q


def find_rois(frames, roi, filt, cutoff, offset=0.0, conversion=1.0, subtract_min=False, mirror=False,
              psf_size=0.0, first_frame=1, capacity=None, with_value=True, agg_dist=0.0):
    """Run bsf_find_rois_host.  Returns (imstack (roi, roi, n) float32 F-order,
    peaks (n, 3) int32 [x, y, frame], value (n,) float32 or None, count);
    n = min(count, capacity).  agg_dist > 0 merges each frame's maxima closer
    than it, as the GUI's p.agg (agg_dist = p.roifit * p.aggsz)."""
    lib = _bind()
    if frames.dtype == np.uint16:
        pt = PIXEL_U16
    elif frames.dtype == np.float32:
        pt = PIXEL_F32
    else:
        raise BsfError(1, "frames must be uint16 or float32 (got %s)" % frames.dtype)
    frames = np.asfortranarray(frames)
    if frames.ndim == 2:
        frames = frames[:, :, None]
    if frames.ndim != 3:
        raise BsfError(1, "frames must be (nrows, ncols[, nframes])")
    filt = np.asfortranarray(filt, np.float32)
    if filt.ndim != 2:
        raise BsfError(1, "the filter kernel must be 2-D")
    if capacity is None:
        capacity = max(1, frames.shape[2] * ((frames.shape[0] + 1) // 2) * ((frames.shape[1] + 1) // 2) // 16)
    ncap = max(int(capacity), 0)          # a negative capacity is the library's error to report
    stack = np.zeros((max(int(roi), 0), max(int(roi), 0), ncap), np.float32, order="F")
    peaks = np.zeros((ncap, 3), np.int32, order="F")
    value = np.zeros(ncap, np.float32) if with_value else None
    q = make_request(frames.ctypes.data, pt, frames.shape, roi, filt.ctypes.data, filt.shape, cutoff, offset,
                     conversion, subtract_min, mirror, psf_size, first_frame, capacity, agg_dist)
    res = RoiResult(stack.ctypes.data, peaks.ctypes.data, value.ctypes.data if value is not None else None, 0)
    _check(lib.bsf_find_rois_host(C.byref(q), C.byref(res)))
    n = min(res.count, ncap)
    return (stack[:, :, :n], np.asfortranarray(peaks[:n]), value[:n] if value is not None else None, res.count)


def fit_frames(frames, roi, filt, cutoff, fittype, iterations, coeff, offset=0.0, conversion=1.0,
               subtract_min=False, mirror=False, psf_size=0.0, first_frame=1, capacity=None, initP=None,
               coeff3=None, precision=0, agg_dist=0.0):
    """Run bsf_fit_frames_host: the ROIs of find_rois and their fits in one
    call.  Returns (imstack, peaks, value, P, CRLB, LL, count) with
    n = min(count, capacity) ROIs; P (n, NV+1) and CRLB (n, NV) as fit()."""
    lib = _bind()
    if frames.dtype not in (np.uint16, np.float32):
        raise BsfError(1, "frames must be uint16 or float32 (got %s)" % frames.dtype)
    pt = PIXEL_U16 if frames.dtype == np.uint16 else PIXEL_F32
    frames = np.asfortranarray(frames)
    if frames.ndim == 2:
        frames = frames[:, :, None]
    filt = np.asfortranarray(filt, np.float32)
    if capacity is None:
        capacity = max(1, frames.shape[2] * ((frames.shape[0] + 1) // 2) * ((frames.shape[1] + 1) // 2) // 16)
    ncap = max(int(capacity), 0)
    NV = _NV.get(int(fittype), 1)
    stack = np.zeros((roi, roi, ncap), np.float32, order="F")
    peaks = np.zeros((ncap, 3), np.int32, order="F")
    value = np.zeros(ncap, np.float32)
    P = np.zeros((ncap, NV + 1), np.float32, order="F")
    CR = np.zeros((ncap, NV), np.float32, order="F")
    LL = np.zeros(ncap, np.float32)
    coeff = _f32(coeff)
    c3 = _f32(coeff3) if coeff3 is not None else None
    ip = np.ascontiguousarray(np.ravel(initP), np.float32) if initP is not None else None
    q = make_request(frames.ctypes.data, pt, frames.shape, roi, filt.ctypes.data, filt.shape, cutoff, offset,
                     conversion, subtract_min, mirror, psf_size, first_frame, capacity, agg_dist)
    f = _request(fittype, roi, ncap, iterations, coeff.ctypes.data, coeff.shape,
                 c3.ctypes.data if c3 is not None else None, ip, precision, None)
    r = RoiResult(stack.ctypes.data, peaks.ctypes.data, value.ctypes.data, 0)
    res = Result(P.ctypes.data, CR.ctypes.data, LL.ctypes.data)
    _check(lib.bsf_fit_frames_host(C.byref(q), C.byref(f), C.byref(r), C.byref(res)))
    n = min(r.count, ncap)
    return (stack[:, :, :n], np.asfortranarray(peaks[:n]), value[:n], np.asfortranarray(P[:n]),
            np.asfortranarray(CR[:n]), LL[:n], r.count)
