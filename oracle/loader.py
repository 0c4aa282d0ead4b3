"""ctypes bindings for the CPU oracle (test infrastructure only).

Array conventions follow the reference's MATLAB call surface
(``MultiD_bspline/visual code/CPUmleFit_LM/mexFunction.cpp:129-311``):
column-major (Fortran-order) float32 arrays.

* ``imstack``: shape (S, S, N); element (row ii, column jj) of ROI n.
* ``coeff``:   shape (nx+3, ny+3, nz+3[, nw+3[, nu+3]]), bsarray coefficients.
* outputs:     ``P`` (N, NV+1) with the iteration count in the last column,
               ``CRLB`` (N, NV) variances, ``LL`` (N,).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
REF_SPLINE = os.path.join(HERE, "_ref", "libref_bspline.so")
REF_MATINV = os.path.join(HERE, "_ref", "libref_matinv.so")

NV_OF = {6: 5, 7: 6, 8: 8}
D_OF = {6: 3, 7: 4, 8: 5}

_lib = None
_ref = None
_refinv = None

fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


def build(with_ref=None):
    """Compile liborc.so (and _ref/ when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "liborc.so"], check=True)
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference")
    if with_ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def available():
    return os.path.exists(LIB)


def _L():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(with_ref=False)
        _lib = C.CDLL(LIB)
        for sfx, rp in (("", fp), ("_f64", dp)):
            f = getattr(_lib, "orc_fit" + sfx)
            f.restype = C.c_int
            f.argtypes = [C.c_int, fp, C.c_int, C.c_int, fp, ip, fp, C.c_int, fp, rp, rp, rp, C.c_int]
            f = getattr(_lib, "orc_fit_status" + sfx)
            f.restype = C.c_int
            f.argtypes = [C.c_int, fp, C.c_int, C.c_int, fp, ip, fp, C.c_int, fp, rp, rp, rp, C.c_void_p, C.c_int]
            g = getattr(_lib, "orc_bspline" + sfx)
            g.restype = C.c_float if sfx == "" else C.c_double
            g.argtypes = [C.c_float if sfx == "" else C.c_double, C.c_int]
    return _lib


def _f32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float32))


def _ptr(a, t=fp):
    return a.ctypes.data_as(t) if a is not None else None


def _fit(sfx, rdt, fittype, imstack, iterations, coeff, initP=None, coeff3=None, nthreads=1, status=False):
    lib = _L()
    imstack = _f32(imstack)
    if imstack.ndim == 2:
        imstack = imstack[:, :, None]
    S, S2, N = imstack.shape
    assert S == S2, "ROIs must be square"
    coeff = _f32(coeff)
    NV = NV_OF[fittype]
    assert coeff.ndim == D_OF[fittype]
    cd = np.zeros(5, np.int32)
    cd[: coeff.ndim] = coeff.shape
    c3 = _f32(coeff3) if coeff3 is not None else None
    ip_ = _f32(np.ravel(initP)) if initP is not None else None
    P = np.zeros((N, NV + 1), rdt, order="F")
    CR = np.zeros((N, NV), rdt, order="F")
    LL = np.zeros(N, rdt)
    t = dp if rdt == np.float64 else fp
    if status:
        st = np.zeros(N, np.uint8)
        rc = getattr(lib, "orc_fit_status" + sfx)(
            fittype, _ptr(imstack), S, N, _ptr(coeff), cd.ctypes.data_as(ip), _ptr(c3),
            int(iterations), _ptr(ip_), _ptr(P, t), _ptr(CR, t), _ptr(LL, t), st.ctypes.data, int(nthreads))
    else:
        rc = getattr(lib, "orc_fit" + sfx)(
            fittype, _ptr(imstack), S, N, _ptr(coeff), cd.ctypes.data_as(ip), _ptr(c3),
            int(iterations), _ptr(ip_), _ptr(P, t), _ptr(CR, t), _ptr(LL, t), int(nthreads))
    if rc != 0:
        raise ValueError("oracle: unsupported fittype %r" % fittype)
    return (P, CR, LL, st) if status else (P, CR, LL)


def fit(fittype, imstack, iterations, coeff, initP=None, coeff3=None, nthreads=1, status=False):
    """Parity (float) restatement of the reference CPU fitter.  status=True
    also returns the per-ROI flags of the library's bsf_roi_flag (uint8),
    with bit 16 = the first damped solve failed."""
    return _fit("", np.float32, fittype, imstack, iterations, coeff, initP, coeff3, nthreads, status)


def fit_f64(fittype, imstack, iterations, coeff, initP=None, coeff3=None, nthreads=1, status=False):
    """The same fit with every quantity in double (FP64 oracle)."""
    return _fit("_f64", np.float64, fittype, imstack, iterations, coeff, initP, coeff3, nthreads, status)


class planted:
    """Context manager that plants mistake `k` in both oracle builds (see
    orc_plant in bsf_oracle.c) for tests showing that a pin catches it."""

    def __init__(self, k):
        self.k = int(k)

    def __enter__(self):
        lib = _L()
        for name in ("orc_plant", "orc_plant_f64"):
            C.c_int.in_dll(lib, name).value = self.k
        return self

    def __exit__(self, *exc):
        lib = _L()
        for name in ("orc_plant", "orc_plant_f64"):
            C.c_int.in_dll(lib, name).value = 0
        return False


# ---------------------------------------------------------------- primitives
def bspline(x, deg, f64=False):
    lib = _L()
    return (lib.orc_bspline_f64 if f64 else lib.orc_bspline)(x, deg)


def delta_tables(D, frac):
    lib = _L()
    fr = np.zeros(5, np.float32)
    fr[:D] = frac
    tab = np.zeros((D + 1) * 4 ** D, np.float32)
    lib.orc_delta_tables(C.c_int(D), _ptr(fr), _ptr(tab))
    return tab.reshape(D + 1, 4 ** D)


def deriv_eval(D, coord, nd, tab, coeff, theta):
    lib = _L()
    co = np.zeros(5, np.float32)
    co[:D] = coord
    n = np.zeros(5, np.int32)
    n[:D] = nd
    tab = np.ascontiguousarray(tab, np.float32).ravel()
    c = np.asarray(coeff, np.float32).ravel(order="F")
    th = np.zeros(9, np.float32)
    th[: len(theta)] = theta
    dudt = np.zeros(8, np.float32)
    model = np.zeros(1, np.float32)
    lib.orc_deriv_eval(C.c_int(D), _ptr(co), n.ctypes.data_as(ip), _ptr(tab), _ptr(c),
                       _ptr(th), _ptr(dudt), _ptr(model))
    return dudt[: D + 2], float(model[0])


def cholesky(A):
    lib = _L()
    n = A.shape[0]
    A = np.ascontiguousarray(A, np.float32).ravel()
    L = np.zeros(n * n, np.float32)
    U = np.zeros(n * n, np.float32)
    info = lib.orc_cholesky(_ptr(A), C.c_int(n), _ptr(L), _ptr(U))
    return info, L.reshape(n, n), U.reshape(n, n)


def lu_eval(L, U, b):
    lib = _L()
    n = len(b)
    x = np.zeros(n, np.float32)
    lib.orc_lu_eval(_ptr(np.ascontiguousarray(L, np.float32).ravel()),
                    _ptr(np.ascontiguousarray(U, np.float32).ravel()),
                    _ptr(np.ascontiguousarray(b, np.float32)), C.c_int(n), _ptr(x))
    return x


def matinv(M):
    lib = _L()
    n = M.shape[0]
    Mc = np.asfortranarray(M, np.float32).ravel(order="F").copy()
    Minv = np.zeros(n * n, np.float32)
    D = np.zeros(n, np.float32)
    lib.orc_matinv(_ptr(Mc), _ptr(Minv), _ptr(D), C.c_int(n))
    return D, Minv.reshape(n, n, order="F")


def com2d(roi):
    lib = _L()
    r = _f32(roi)
    x = C.c_float()
    y = C.c_float()
    lib.orc_com2d(C.c_int(r.shape[0]), _ptr(r), C.byref(x), C.byref(y))
    return x.value, y.value


def maxmin2d(roi):
    lib = _L()
    r = _f32(roi)
    a = C.c_float()
    b = C.c_float()
    lib.orc_maxmin2d(C.c_int(r.shape[0]), _ptr(r), C.byref(a), C.byref(b))
    return a.value, b.value


# ------------------------------------------------- the reference's own code
def ref_available():
    return os.path.exists(REF_SPLINE) and os.path.exists(REF_MATINV)


def _R():
    global _ref, _refinv
    if _ref is None:
        _ref = C.CDLL(REF_SPLINE)
        _refinv = C.CDLL(REF_MATINV)
    return _ref, _refinv


# mangled C++ names of the reference's functions (CPUsplineLib.h, MatInvLib.h)
_M = {
    "eval": "_Z18kernel_evalBSplinefiPf",
    "d3": "_Z29kernel_computeDelta3D_bSplinefffPfS_S_S_",
    "d4": "_Z29kernel_computeDelta4D_bSplineffffPfS_S_S_S_",
    "d5": "_Z29kernel_computeDelta5D_bSplinefffffPfS_S_S_S_S_",
    "v3": "_Z26kernel_Derivative_bSpline1fffiiiPfS_S_S_PKfS_S_S_",
    "v4": "_Z28kernel_Derivative_bSpline14DffffiiiiPfS_S_S_S_PKfS_S_S_",
    "v5": "_Z28kernel_Derivative_bSpline15DfffffiiiiiPfS_S_S_S_S_PKfS_S_S_",
    "chol": "_Z15kernel_choleskyPfiS_S_",
    "lu": "_Z17kernel_luEvaluatePfS_S_iS_",
    "lu5": "_Z19kernel_luEvaluate5DPfS_S_iS_",
    "inv": "_Z14kernel_MatInvNPfS_S_i",
}


def ref_bspline(x, deg):
    r, _ = _R()
    out = C.c_float()
    getattr(r, _M["eval"])(C.c_float(x), C.c_int(deg), C.byref(out))
    return out.value


def ref_delta_tables(D, frac):
    r, _ = _R()
    T = 4 ** D
    tabs = [np.zeros(T, np.float32) for _ in range(D + 1)]
    args = [C.c_float(float(v)) for v in frac[:D]] + [_ptr(t) for t in tabs]
    getattr(r, _M["d%d" % D])(*args)
    return np.stack(tabs)


def ref_deriv_eval(D, coord, nd, tab, coeff, theta):
    r, _ = _R()
    T = 4 ** D
    tab = np.ascontiguousarray(tab, np.float32).reshape(D + 1, T)
    tabs = [np.ascontiguousarray(tab[k]) for k in range(D + 1)]
    c = np.asarray(coeff, np.float32).ravel(order="F")
    th = np.zeros(9, np.float32)
    th[: len(theta)] = theta
    dudt = np.zeros(8, np.float32)
    model = C.c_float()
    args = ([C.c_float(float(v)) for v in coord[:D]] + [C.c_int(int(v)) for v in nd[:D]]
            + [_ptr(t) for t in tabs] + [_ptr(c), _ptr(th), _ptr(dudt), C.byref(model)])
    getattr(r, _M["v%d" % D])(*args)
    return dudt[: D + 2], model.value


def ref_cholesky(A):
    r, _ = _R()
    n = A.shape[0]
    A = np.ascontiguousarray(A, np.float32).ravel().copy()
    L = np.zeros(n * n, np.float32)
    U = np.zeros(n * n, np.float32)
    info = getattr(r, _M["chol"])(_ptr(A), C.c_int(n), _ptr(L), _ptr(U))
    return info, L.reshape(n, n), U.reshape(n, n)


def ref_lu_eval(L, U, b, five_d=False):
    r, _ = _R()
    n = len(b)
    x = np.zeros(n, np.float32)
    getattr(r, _M["lu5" if five_d else "lu"])(
        _ptr(np.ascontiguousarray(L, np.float32).ravel().copy()),
        _ptr(np.ascontiguousarray(U, np.float32).ravel().copy()),
        _ptr(np.ascontiguousarray(b, np.float32).copy()), C.c_int(n), _ptr(x))
    return x


def ref_matinv(M):
    _, ri = _R()
    n = M.shape[0]
    Mc = np.asfortranarray(M, np.float32).ravel(order="F").copy()
    Minv = np.zeros(n * n, np.float32)
    D = np.zeros(n, np.float32)
    getattr(ri, _M["inv"])(_ptr(Mc), _ptr(Minv), _ptr(D), C.c_int(n))
    return D, Minv.reshape(n, n, order="F")


def eval_pass(fittype, roi, theta, coeff, coeff3=None, f64=False, first=False):
    """One deviance / gradient / Hessian pass at theta (no LM step); first
    selects the first pass's quirks.  Returns (err, J, H, mu) with mu
    indexed [row, column]."""
    lib = _L()
    roi = _f32(roi)
    S = roi.shape[0]
    coeff = _f32(coeff)
    cd = np.zeros(5, np.int32)
    cd[: coeff.ndim] = coeff.shape
    c3 = _f32(coeff3) if coeff3 is not None else None
    NV = NV_OF[fittype]
    rdt = np.float64 if f64 else np.float32
    t = dp if f64 else fp
    th = np.zeros(9, rdt)
    th[:NV] = theta[:NV]
    J = np.zeros(NV, rdt)
    H = np.zeros(NV * NV, rdt)
    mu = np.zeros(S * S, rdt)
    f = lib.orc_eval_pass_f64 if f64 else lib.orc_eval_pass
    f.restype = C.c_double if f64 else C.c_float
    err = f(C.c_int(fittype), _ptr(roi), C.c_int(S), _ptr(coeff), cd.ctypes.data_as(ip),
            _ptr(c3), _ptr(th, t), _ptr(J, t), _ptr(H, t), _ptr(mu, t), C.c_int(1 if first else 0))
    return err, J, H.reshape(NV, NV), mu.reshape(S, S, order="F")


def nv_4pi(shared, nch=4):
    lib = _L()
    sh = np.ascontiguousarray(shared, np.int32)
    return lib.orc_nv_4pi(sh.ctypes.data_as(ip), C.c_int(nch))


def fit_4pi(imstack, shared, iterations, coeff, phaseShift, dT=None, initZ=None, initPhi=None,
            nthreads=1, f64=False):
    """4Pi multi-channel restatement (see bsf_oracle.c, A4).  imstack is
    (S, S, N, nch) Fortran-order; returns P (N, NV+1), CRLB (N, NV), LL (N,)."""
    lib = _L()
    im = _f32(imstack)
    S, _, N, nch = im.shape
    coeff = _f32(coeff)
    cd = np.zeros(5, np.int32)
    cd[:4] = coeff.shape
    sh = np.ascontiguousarray(shared, np.int32)
    NV = nv_4pi(sh, nch)
    rdt = np.float64 if f64 else np.float32
    t = dp if f64 else fp
    P = np.zeros((N, NV + 1), rdt, order="F")
    CR = np.zeros((N, NV), rdt, order="F")
    LL = np.zeros(N, rdt)
    dT_ = _f32(dT) if dT is not None else None
    iz = np.ascontiguousarray(initZ, np.float32) if initZ is not None else None
    iph = np.ascontiguousarray(initPhi, np.float32) if initPhi is not None else None
    ps = np.ascontiguousarray(phaseShift, np.float32)
    f = lib.orc_fit_4pi_f64 if f64 else lib.orc_fit_4pi
    rc = f(_ptr(im), C.c_int(S), C.c_int(N), C.c_int(nch), sh.ctypes.data_as(ip), C.c_int(int(iterations)),
           _ptr(coeff), cd.ctypes.data_as(ip), _ptr(dT_), _ptr(iz), _ptr(iph), _ptr(ps),
           _ptr(P, t), _ptr(CR, t), _ptr(LL, t), C.c_int(nthreads))
    if rc != 0:
        raise ValueError("oracle: bad 4Pi arguments")
    return P, CR, LL


# ------------------------------------------------------- frames -> ROI stack
def _frontend():
    lib = _L()
    if not getattr(lib, "_fe_bound", False):
        lib.orc_photons.argtypes = [C.c_void_p, C.c_int, C.c_long, C.c_float, C.c_float, C.c_void_p]
        lib.orc_photons.restype = None
        lib.orc_filter2.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        lib.orc_filter2.restype = None
        lib.orc_find_maxima.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_long]
        lib.orc_find_maxima.restype = C.c_long
        lib.orc_find_rois.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                      C.c_float, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                      C.c_int, C.c_double, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_double]
        lib.orc_find_rois.restype = C.c_long
        lib.orc_aggregate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_double]
        lib.orc_aggregate.restype = C.c_long
        lib._fe_bound = True
    return lib


def _pixels(frames):
    if frames.dtype == np.uint16:
        return 0
    if frames.dtype == np.float32:
        return 1
    raise ValueError("frames must be uint16 or float32")


def photons(raw, offset, conversion):
    """(single(raw) - offset) * conversion, elementwise (bsf_frontend.c)."""
    raw = np.asfortranarray(raw)
    out = np.empty(raw.shape, np.float32, order="F")
    _frontend().orc_photons(raw.ctypes.data, _pixels(raw), raw.size, float(offset), float(conversion),
                            out.ctypes.data)
    return out


def filter2(h, x):
    """filter2(h, x, 'same') in the front-end's FP32 fmaf order; x, h 2-D."""
    x = np.asfortranarray(x, np.float32)
    h = np.asfortranarray(h, np.float32)
    out = np.empty(x.shape, np.float32, order="F")
    _frontend().orc_filter2(x.ctypes.data, x.shape[0], x.shape[1], h.ctypes.data, h.shape[0], h.shape[1],
                            out.ctypes.data)
    return out


def find_maxima(F, cutoff, dx, psf_size=0.0):
    """maximumfind2's maxima of F plus the GUI's filters: (n, 3) [x, y, value]."""
    F = np.asfortranarray(F, np.float32)
    cap = F.size // 4 + 8
    xs = np.zeros(cap, np.int32)
    ys = np.zeros(cap, np.int32)
    vs = np.zeros(cap, np.float32)
    n = _frontend().orc_find_maxima(F.ctypes.data, F.shape[0], F.shape[1], float(cutoff), int(dx), float(psf_size),
                                    xs.ctypes.data, ys.ctypes.data, vs.ctypes.data, cap)
    return np.stack([xs[:n], ys[:n], vs[:n].astype(np.float64)], 1)


def aggregate(maxima, thr):
    """aggre on one frame's maxima (n, 3) [x, y, value] (bsf_frontend.c:
    orc_aggregate): the merged (m, 3) array."""
    m = np.asarray(maxima, np.float64).reshape(-1, 3)
    xs = np.ascontiguousarray(m[:, 0], np.int32)
    ys = np.ascontiguousarray(m[:, 1], np.int32)
    vs = np.ascontiguousarray(m[:, 2], np.float64)
    n = _frontend().orc_aggregate(xs.ctypes.data, ys.ctypes.data, vs.ctypes.data, len(m), float(thr))
    return np.stack([xs[:n].astype(np.float64), ys[:n].astype(np.float64), vs[:n]], 1)


def find_rois(frames, roi, filt, cutoff, offset=0.0, conversion=1.0, subtract_min=False, mirror=False,
              psf_size=0.0, first_frame=1, capacity=None, agg_dist=0.0):
    """The whole front-end (bsf_frontend.c: orc_find_rois); same arguments and
    returns as bsfit.frames.find_rois."""
    frames = np.asfortranarray(frames)
    if frames.ndim == 2:
        frames = frames[:, :, None]
    h = np.asfortranarray(filt, np.float32)
    nr, nc, nf = frames.shape
    if capacity is None:
        capacity = max(1, nf * ((nr + 1) // 2) * ((nc + 1) // 2) // 16)
    stack = np.zeros((roi, roi, capacity), np.float32, order="F")
    peaks = np.zeros((capacity, 3), np.int32, order="F")
    value = np.zeros(capacity, np.float32)
    n = _frontend().orc_find_rois(frames.ctypes.data, _pixels(frames), nr, nc, nf, int(first_frame), float(offset),
                                  float(conversion), h.ctypes.data, h.shape[0], h.shape[1], int(bool(subtract_min)),
                                  float(cutoff), int(roi), int(bool(mirror)), float(psf_size), capacity,
                                  stack.ctypes.data, peaks.ctypes.data, value.ctypes.data, float(agg_dist))
    k = min(n, capacity)
    return stack[:, :, :k], np.asfortranarray(peaks[:k]), value[:k], n
