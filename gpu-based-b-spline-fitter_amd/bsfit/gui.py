"""The GUI fitter's localisation table and the colour demo's class mapping
(include/bsfit.h: bsf_gui_results, bsf_color_class;
MultiD_fitter/simplefitter_spline.m:550-630,
demo_color_classification/bspline_based_color_classification_v3.m:115-117)."""
import ctypes as C

import numpy as np

from .api import _check, load


def _bind():
    lib = load()
    if not getattr(lib, "_gui_bound", False):
        lib.bsf_gui_results_cols.argtypes = [C.c_int]
        lib.bsf_gui_results_cols.restype = C.c_int
        lib.bsf_gui_results.argtypes = [C.c_int, C.c_int] + [C.c_double] * 5 + [C.c_void_p] * 4 + [C.c_int,
                                                                                                   C.c_void_p]
        lib.bsf_gui_results.restype = C.c_int
        lib.bsf_color_class.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        lib.bsf_color_class.restype = C.c_int
        lib._gui_bound = True
    return lib


def gui_results(fittype, P, CRLB, LL, peaks, dx, z0, dz, du=1.0, dv=1.0, mirror=False):
    """The GUI's results table (N, 12 | 17 | 19), float64 Fortran order, from
    a fit block's P, CRLB, LL and its peak coordinates (N, 3) [x, y, frame]."""
    lib = _bind()
    P = np.asfortranarray(P, np.float32)
    CRLB = np.asfortranarray(CRLB, np.float32)
    LL = np.ascontiguousarray(np.ravel(LL), np.float32)
    peaks = np.asfortranarray(peaks, np.float64)
    N = P.shape[0]
    cols = lib.bsf_gui_results_cols(int(fittype))
    out = np.zeros((N, max(cols, 0)), np.float64, order="F")
    _check(lib.bsf_gui_results(int(fittype), int(bool(mirror)), float(dx), float(z0), float(dz), float(du),
                               float(dv), P.ctypes.data, CRLB.ctypes.data, LL.ctypes.data, peaks.ctypes.data, N,
                               out.ctypes.data))
    return out


def color_class(w):
    """(min(3, max(2, w)) - 2.5) * 2 in single precision."""
    lib = _bind()
    w = np.ascontiguousarray(np.ravel(w), np.float32)
    out = np.empty_like(w)
    _check(lib.bsf_color_class(w.ctypes.data, w.size, out.ctypes.data))
    return out
