"""Shared comparison of a GPU fit with the oracle (per-ROI agreement:
coordinates within 1e-4 px (spline units for z, w, u; absolute for g2),
photons / background within 1e-4 relative, CRLB within 1e-4 relative,
iteration counts equal), and device buffers for tests of the device entry
points."""
import ctypes as C

import numpy as np

COORD = {6: [0, 1, 4], 7: [0, 1, 4, 5], 8: [0, 1, 4, 5, 6, 7]}
REL = {6: [2, 3], 7: [2, 3], 8: [2, 3]}
# A CRLB (variance) above this, or infinite, means the parameter has no
# information (e.g. a coordinate on a mirror-padded bound of the spline, where
# its derivative vanishes): the kernel reports +inf there, the reference's
# arithmetic its rounding residue's 1e7-1e31.  Two such entries agree.
NO_INFO = 1e12


def _crlb_rel(CR, CRo):
    CR = np.asarray(CR, np.float64)
    CRo = np.asarray(CRo, np.float64)
    big, bigo = ~(CR <= NO_INFO), ~(CRo <= NO_INFO)
    with np.errstate(invalid="ignore", divide="ignore"):
        d = np.abs(CR - CRo) / np.maximum(np.abs(CRo), 1e-30)
    d = np.where(big & bigo, 0.0, d)
    return np.where(np.isnan(d), np.inf, d)


def compare(ft, P, CR, LL, Po, CRo, LLo, atol=1e-4, rtol=1e-4):
    NV = CR.shape[1]
    P = np.asarray(P, np.float64)
    Po = np.asarray(Po, np.float64)
    dcoord = np.abs(P[:, COORD[ft]] - Po[:, COORD[ft]]).max(1)
    drel = (np.abs(P[:, REL[ft]] - Po[:, REL[ft]]) / np.maximum(np.abs(Po[:, REL[ft]]), 1e-12)).max(1)
    dcr = _crlb_rel(CR, CRo).max(1)
    same_kk = P[:, NV] == Po[:, NV]
    ok = (dcoord <= atol) & (drel <= rtol) & (dcr <= rtol) & same_kk
    return dict(n=len(P), agree=float(ok.mean()), same_kk=float(same_kk.mean()),
                p_ok=float(((dcoord <= atol) & (drel <= rtol)).mean()),
                crlb_ok=float((dcr <= rtol).mean()),
                p99_dcoord=float(np.percentile(dcoord, 99)), max_dcoord=float(dcoord.max()),
                p99_drel=float(np.percentile(drel, 99)), p99_dcrlb=float(np.percentile(dcr, 99)),
                finite=bool(np.all(np.isfinite(P)) and np.all(np.isfinite(CR))))


def errors(ft, P, CR, Po, CRo, same_tol=1e-3):
    """Error of a fit against a reference fit of the same ROIs.

    coord: quantiles (50, 90, 99, max) over all ROIs of the largest
    coordinate error (px; spline units for z, w, u; absolute for g2), which
    includes ROIs whose LM trajectories ended at different points.
    same: the fraction of ROIs that reached the same optimum (coordinates
    within same_tol); rel / crlb: over those ROIs, quantiles of the largest
    relative N / bg error and of the largest relative CRLB error (entries
    without information on both sides count as equal) -- the precision of
    the fit at a given optimum."""
    P = np.asarray(P, np.float64)
    Po = np.asarray(Po, np.float64)
    dcoord = np.abs(P[:, COORD[ft]] - Po[:, COORD[ft]]).max(1)
    drel = (np.abs(P[:, REL[ft]] - Po[:, REL[ft]]) / np.maximum(np.abs(Po[:, REL[ft]]), 1e-12)).max(1)
    dcr = _crlb_rel(CR, CRo).max(1)
    same = dcoord <= same_tol

    def q(a):
        if len(a) == 0:
            return [np.nan] * 4
        return [float(np.percentile(a, 50)), float(np.percentile(a, 90)), float(np.percentile(a, 99)),
                float(a.max())]
    return dict(coord=q(dcoord), same=float(same.mean()), rel=q(drel[same]), crlb=q(dcr[same]))


def fmt_errors(e):
    return ("coord p50 %.1e p90 %.1e p99 %.1e max %.1e | same optimum %.3f: N,bg rel p50 %.1e p90 %.1e p99 %.1e"
            " | CRLB rel p50 %.1e p90 %.1e p99 %.1e" % (*e["coord"], e["same"], *e["rel"][:3], *e["crlb"][:3]))


def crlb_nan_free(CR):
    """No CRLB entry is NaN (an exactly singular Fisher matrix gives +inf for
    the parameter without information, not NaN everywhere)."""
    return not np.any(np.isnan(np.asarray(CR)))


class Hip:
    """Device buffers through the HIP runtime libbsfit.so links (not torch's
    bundled one: two runtimes in a process only get along when torch's
    initialises first)."""

    def __init__(self):
        import bsfit
        bsfit.load()
        self.h = C.CDLL("libamdhip64.so.7")
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
            getattr(self.h, f).restype = C.c_int
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def alloc(self, n):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(int(n), 4)) == 0
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a) if not a.flags.f_contiguous else a
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        if out.nbytes:
            assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self, *ps):
        for p in ps:
            if p:
                self.h.hipFree(p)


class Mex:
    """One MEX gateway of gpu-based-b-spline-fitter_amd/matlab/ compiled with
    the test MX API (tests/mex_api) into a library of its own and called as
    MATLAB would call it: numpy arrays in, numpy arrays out; a
    mexErrMsgIdAndTxt comes back as MexError(id, message)."""

    _CLS = {np.dtype(np.float64): 6, np.dtype(np.float32): 7, np.dtype(np.uint8): 9, np.dtype(np.uint16): 11, np.dtype(np.uint32): 13}
    _DT = {v: k for k, v in _CLS.items()}

    def __init__(self, name, outdir):
        import os
        import subprocess
        import bsfit
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        src = os.path.join(root, "gpu-based-b-spline-fitter_amd", "matlab", name + ".c")
        so = os.path.join(str(outdir), name + ".so")
        subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-shared", "-fPIC", src,
                        os.path.join(root, "tests", "mex_api", "mock_mx.c"), "-I", os.path.join(root, "include"),
                        "-I", os.path.join(root, "tests", "mex_api"), "-o", so], check=True)
        C.CDLL(bsfit.lib_path(), mode=C.RTLD_GLOBAL)      # the gateway's bsf_* symbols
        self.lib = C.CDLL(so)
        self.lib.mock_new.restype = C.c_void_p
        self.lib.mock_new.argtypes = [C.c_int, C.c_int, C.c_void_p]
        self.lib.mock_data.restype = C.c_void_p
        self.lib.mock_data.argtypes = [C.c_void_p]
        self.lib.mock_free.argtypes = [C.c_void_p]
        self.lib.mock_call.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        self.lib.mock_error_id.restype = C.c_char_p
        self.lib.mock_error_msg.restype = C.c_char_p
        for f in ("mxGetNumberOfDimensions", "mxGetClassID"):
            getattr(self.lib, f).argtypes = [C.c_void_p]
            getattr(self.lib, f).restype = C.c_size_t
        self.lib.mxGetDimensions.argtypes = [C.c_void_p]
        self.lib.mxGetDimensions.restype = C.POINTER(C.c_size_t)

    def _arr(self, a):
        a = np.asfortranarray(np.atleast_1d(a))
        dims = (C.c_size_t * max(a.ndim, 1))(*a.shape)
        p = self.lib.mock_new(self._CLS[a.dtype], a.ndim, dims)
        C.memmove(self.lib.mock_data(p), a.ctypes.data, a.nbytes)
        return p

    def _np(self, p):
        nd = self.lib.mxGetNumberOfDimensions(p)
        d = self.lib.mxGetDimensions(p)
        shape = tuple(d[k] for k in range(nd))
        dt = self._DT[int(self.lib.mxGetClassID(p))]
        n = int(np.prod(shape))
        out = np.empty(shape, dt, order="F")
        if n:
            C.memmove(out.ctypes.data, self.lib.mock_data(p), n * dt.itemsize)
        return out

    def __call__(self, nlhs, *args):
        ins = [self._arr(a) for a in args]
        prhs = (C.c_void_p * max(len(ins), 1))(*ins)
        plhs = (C.c_void_p * max(nlhs, 1))()
        try:
            rc = self.lib.mock_call(nlhs, plhs, len(ins), prhs)
            if rc:
                raise MexError(self.lib.mock_error_id().decode(), self.lib.mock_error_msg().decode())
            outs = [self._np(plhs[k]) for k in range(max(nlhs, 1))]
        finally:
            for p in ins:
                self.lib.mock_free(p)
        for k in range(max(nlhs, 1)):
            if plhs[k]:
                self.lib.mock_free(plhs[k])
        return outs


class MexError(RuntimeError):
    def __init__(self, ident, msg):
        super().__init__("%s: %s" % (ident, msg))
        self.ident = ident
