"""ctypes binding of include/bsfit.h (libbsfit.so)."""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libbsfit.so")

PREC_MIX = 0
PREC_F64 = 1
_NV = {6: 5, 7: 6, 8: 8}
_D = {6: 3, 7: 4, 8: 5}

_STATUS = {0: "BSF_OK", 1: "BSF_ERR_ARG", 2: "BSF_ERR_FITTYPE", 3: "BSF_ERR_ROI_SIZE",
           4: "BSF_ERR_SHAPE", 5: "BSF_ERR_NOMEM", 6: "BSF_ERR_DEVICE"}


class BsfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (_STATUS.get(code, code), msg))
        self.code = code


class Request(C.Structure):
    _fields_ = [
        ("fittype", C.c_int),
        ("imstack", C.c_void_p),
        ("sz", C.c_int),
        ("nfits", C.c_int),
        ("iterations", C.c_int),
        ("coeff", C.c_void_p),
        ("coeff_ndim", C.c_int),
        ("coeff_dims", C.c_int * 5),
        ("coeff3", C.c_void_p),
        ("initP", C.c_void_p),
        ("n_initP", C.c_int),
        ("precision", C.c_int),
    ]


class Result(C.Structure):
    _fields_ = [("P", C.c_void_p), ("CRLB", C.c_void_p), ("LL", C.c_void_p)]


_lib = None


def lib_path():
    return LIB_PATH


def load():
    """Load libbsfit.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BsfError(6, "libbsfit.so is not built (run __graft_entry__.build() or make -C %s)" % PKG_DIR)
        lib = C.CDLL(LIB_PATH)
        lib.bsf_fit_host.argtypes = [C.POINTER(Request), C.POINTER(Result)]
        lib.bsf_fit_host.restype = C.c_int
        lib.bsf_fit_device.argtypes = [C.POINTER(Request), C.POINTER(Result), C.c_void_p]
        lib.bsf_fit_device.restype = C.c_int
        lib.bsf_last_error.restype = C.c_char_p
        lib.bsf_num_params.argtypes = [C.c_int]
        lib.bsf_num_params.restype = C.c_int
        lib.bsf_GPUmleFit_LM_multiD.restype = C.c_int
        _lib = lib
    return _lib


def num_params(fittype):
    return load().bsf_num_params(int(fittype))


def _check(rc):
    if rc != 0:
        raise BsfError(rc, load().bsf_last_error().decode())


def _f32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float32))


def _request(fittype, S, N, iterations, coeff_ptr, coeff_shape, coeff3_ptr, initP, precision, data_ptr):
    r = Request()
    r.fittype = int(fittype)
    r.imstack = data_ptr
    r.sz = int(S)
    r.nfits = int(N)
    r.iterations = int(iterations)
    r.coeff = coeff_ptr
    r.coeff_ndim = len(coeff_shape)
    for a, d in enumerate(coeff_shape):
        r.coeff_dims[a] = int(d)
    r.coeff3 = coeff3_ptr
    r.initP = initP.ctypes.data if initP is not None else None
    r.n_initP = 0 if initP is None else int(initP.size)
    r.precision = int(precision)
    return r


# per-ROI status flags (enum bsf_roi_flag)
ROI_MAX_ITER = 1
ROI_SOLVE_FAILED = 2
ROI_FISHER_SINGULAR = 4
ROI_NONFINITE = 8
ROI_FIRST_SOLVE_FAILED = 16


def fit(fittype, imstack, iterations, coeff, initP=None, coeff3=None, precision=PREC_MIX):
    """Fit an ROI stack (S, S, N) with host arrays; returns (P, CRLB, LL) with
    P (N, NV+1) and CRLB (N, NV) in Fortran order, as the MEX returns them."""
    return _fit(fittype, imstack, iterations, coeff, initP, coeff3, precision, False)


def fit_status(fittype, imstack, iterations, coeff, initP=None, coeff3=None, precision=PREC_MIX):
    """fit() plus the per-ROI status flags (bsf_fit_host_status): returns
    (P, CRLB, LL, status) with status (N,) uint8, bits ROI_MAX_ITER,
    ROI_SOLVE_FAILED, ROI_FISHER_SINGULAR, ROI_NONFINITE, ROI_FIRST_SOLVE_FAILED."""
    return _fit(fittype, imstack, iterations, coeff, initP, coeff3, precision, True)


def _fit(fittype, imstack, iterations, coeff, initP, coeff3, precision, with_status):
    lib = load()
    imstack = _f32(imstack)
    if imstack.ndim == 2:
        imstack = imstack[:, :, None]
    if imstack.ndim != 3:
        raise BsfError(1, "imstack must be 2-D or 3-D")
    S, S2, N = imstack.shape
    if S != S2:
        raise BsfError(4, "Fit Box must be square")
    coeff = _f32(coeff)
    c3 = _f32(coeff3) if coeff3 is not None else None
    ip = np.ascontiguousarray(np.ravel(initP), np.float32) if initP is not None else None
    NV = _NV.get(int(fittype), 1)
    P = np.zeros((N, NV + 1), np.float32, order="F")
    CR = np.zeros((N, NV), np.float32, order="F")
    LL = np.zeros(N, np.float32)
    req = _request(fittype, S, N, iterations, coeff.ctypes.data, coeff.shape,
                   c3.ctypes.data if c3 is not None else None, ip, precision, imstack.ctypes.data)
    res = Result(P.ctypes.data, CR.ctypes.data, LL.ctypes.data)
    if not with_status:
        _check(lib.bsf_fit_host(C.byref(req), C.byref(res)))
        return P, CR, LL
    st = np.zeros(N, np.uint8)
    lib.bsf_fit_host_status.argtypes = [C.POINTER(Request), C.POINTER(Result), C.c_void_p]
    lib.bsf_fit_host_status.restype = C.c_int
    _check(lib.bsf_fit_host_status(C.byref(req), C.byref(res), st.ctypes.data))
    return P, CR, LL, st


def fit_multistart(fittype, imstack, iterations, coeff, starts, coeff3=None, strict=False, precision=PREC_MIX,
                   status=False):
    """Multi-start fit (bsf_fit_multistart_host): starts (K, 2) for fittype 7
    or (K, 4) for fittype 8; a later start's result replaces the kept one
    where its log-likelihood is higher by >= 1e-4 (strict: > 1e-4), as the
    reference's callers combine their per-start MEX calls.  Returns
    (P, CRLB, LL) as fit(); with status=True also the kept results' per-ROI
    status flags (bsf_fit_multistart_host_status)."""
    lib = load()
    lib.bsf_fit_multistart_host.argtypes = [C.POINTER(Request), C.c_void_p, C.c_int, C.c_int, C.POINTER(Result)]
    lib.bsf_fit_multistart_host.restype = C.c_int
    imstack = _f32(imstack)
    if imstack.ndim == 2:
        imstack = imstack[:, :, None]
    S, _, N = imstack.shape
    coeff = _f32(coeff)
    c3 = _f32(coeff3) if coeff3 is not None else None
    st = np.ascontiguousarray(np.atleast_2d(starts), np.float32)
    NV = _NV.get(int(fittype), 1)
    P = np.zeros((N, NV + 1), np.float32, order="F")
    CR = np.zeros((N, NV), np.float32, order="F")
    LL = np.zeros(N, np.float32)
    req = _request(fittype, S, N, iterations, coeff.ctypes.data, coeff.shape,
                   c3.ctypes.data if c3 is not None else None, st[0], precision, imstack.ctypes.data)
    req.n_initP = st.shape[1]
    res = Result(P.ctypes.data, CR.ctypes.data, LL.ctypes.data)
    if not status:
        _check(lib.bsf_fit_multistart_host(C.byref(req), st.ctypes.data, st.shape[0], int(bool(strict)),
                                           C.byref(res)))
        return P, CR, LL
    flags = np.zeros(N, np.uint8)
    lib.bsf_fit_multistart_host_status.argtypes = [C.POINTER(Request), C.c_void_p, C.c_int, C.c_int,
                                                   C.POINTER(Result), C.c_void_p]
    lib.bsf_fit_multistart_host_status.restype = C.c_int
    _check(lib.bsf_fit_multistart_host_status(C.byref(req), st.ctypes.data, st.shape[0], int(bool(strict)),
                                              C.byref(res), flags.ctypes.data))
    return P, CR, LL, flags


class RoiDesc(C.Structure):
    """bsf_roi: one ROI of a launch that mixes ROI sizes."""
    _fields_ = [("off", C.c_longlong), ("cstride", C.c_int), ("sz", C.c_int)]


class Extras(C.Structure):
    """bsf_fit_extras: optional device-side inputs and outputs."""
    _fields_ = [("rois", C.c_void_p), ("status", C.c_void_p), ("evals", C.c_void_p), ("tile", C.c_int)]


def roi_table(sizes, offsets=None, cstride=0):
    """Host array of bsf_roi records (numpy structured, 16 bytes each) for
    ROI sizes `sizes` packed back to back (or at `offsets`, in floats)."""
    sizes = np.asarray(sizes, np.int64)
    if offsets is None:
        offsets = np.concatenate([[0], np.cumsum(sizes * sizes)[:-1]]) if len(sizes) else np.zeros(0, np.int64)
    t = np.zeros(len(sizes), dtype=np.dtype([("off", "<i8"), ("cstride", "<i4"), ("sz", "<i4")]))
    t["off"] = offsets
    t["cstride"] = cstride
    t["sz"] = sizes
    return t


def fit_device(fittype, data_ptr, S, N, iterations, coeff_ptr, coeff_shape, P_ptr, CRLB_ptr, LL_ptr,
               initP=None, coeff3_ptr=None, precision=PREC_MIX, stream=None, status_ptr=None, rois_ptr=None,
               evals_ptr=None, tile=0):
    """Queue a fit on device-resident buffers (raw device addresses, e.g.
    torch tensor .data_ptr()) on a HIP stream handle; returns immediately.
    status_ptr: a device array of N bytes for the per-ROI status flags
    (bsf_fit_device_status), or None.  rois_ptr: a device array of N bsf_roi
    records (roi_table) for ROIs of mixed sizes, S then being the largest;
    evals_ptr: a device int32 array of N for the evaluations each fit ran;
    tile: ROIs per wavefront (0: the recorded choice, 1 or 2) (the last three
    through bsf_fit_device_ex)."""
    lib = load()
    ip = np.ascontiguousarray(np.ravel(initP), np.float32) if initP is not None else None
    req = _request(fittype, S, N, iterations, coeff_ptr, coeff_shape, coeff3_ptr, ip, precision, data_ptr)
    res = Result(P_ptr, CRLB_ptr, LL_ptr)
    if rois_ptr is not None or evals_ptr is not None or tile:
        x = Extras(rois_ptr, status_ptr, evals_ptr, int(tile))
        lib.bsf_fit_device_ex.argtypes = [C.POINTER(Request), C.POINTER(Result), C.POINTER(Extras), C.c_void_p]
        lib.bsf_fit_device_ex.restype = C.c_int
        _check(lib.bsf_fit_device_ex(C.byref(req), C.byref(res), C.byref(x), stream))
        return
    if status_ptr is None:
        _check(lib.bsf_fit_device(C.byref(req), C.byref(res), stream))
        return
    lib.bsf_fit_device_status.argtypes = [C.POINTER(Request), C.POINTER(Result), C.c_void_p, C.c_void_p]
    lib.bsf_fit_device_status.restype = C.c_int
    _check(lib.bsf_fit_device_status(C.byref(req), C.byref(res), status_ptr, stream))


def tune(fittype, data_ptr, S, N, iterations, coeff_ptr, coeff_shape, P_ptr, CRLB_ptr, LL_ptr, initP=None,
         coeff3_ptr=None, precision=PREC_MIX, stream=None, reps=5):
    """bsf_tune: time each tile (ROIs per wavefront) on this fit (device
    buffers, as fit_device), record the fastest for (fittype, precision, S)
    and return (tile, [ms of tile 1, ms of tile 2]) (0 ms: not possible)."""
    lib = load()
    ip = np.ascontiguousarray(np.ravel(initP), np.float32) if initP is not None else None
    req = _request(fittype, S, N, iterations, coeff_ptr, coeff_shape, coeff3_ptr, ip, precision, data_ptr)
    res = Result(P_ptr, CRLB_ptr, LL_ptr)
    t = C.c_int(0)
    ms = (C.c_float * 2)()
    lib.bsf_tune.argtypes = [C.POINTER(Request), C.POINTER(Result), C.c_int, C.POINTER(C.c_int),
                             C.POINTER(C.c_float), C.c_void_p]
    lib.bsf_tune.restype = C.c_int
    _check(lib.bsf_tune(C.byref(req), C.byref(res), int(reps), C.byref(t), ms, stream))
    return t.value, [ms[0], ms[1]]


def fit_tile(fittype, precision=PREC_MIX, S=13):
    """The recorded tile (ROIs per wavefront) for (fittype, precision, S)."""
    lib = load()
    lib.bsf_fit_tile.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.bsf_fit_tile.restype = C.c_int
    return lib.bsf_fit_tile(int(fittype), int(precision), int(S))


def set_fit_tile(fittype, precision, S, tile):
    """Record a tile by hand (0: back to the measured default)."""
    lib = load()
    lib.bsf_set_fit_tile.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    lib.bsf_set_fit_tile.restype = C.c_int
    _check(lib.bsf_set_fit_tile(int(fittype), int(precision), int(S), int(tile)))


def GPUmleFit_LM_multiD(imstack, fittype, iterations=50, coeff=None, varmap=None, silent=1,
                        initP=None, coeff3=None):
    """The reference MEX's call surface (GPUmleFit_LM/mexFunction.cpp:14-55).
    varmap is ignored for fittypes 6-8 as in the MEX (:363-373)."""
    if coeff is None:
        raise BsfError(1, "coeff is required for fittypes 6-8")
    fittype = int(fittype)
    P, CR, LL = fit(fittype, imstack, int(iterations), coeff,
                    initP=initP if fittype > 6 else None, coeff3=coeff3)
    if not silent:
        print("subregion size: %d fittype: %d\nNumber of fits: %d" % (np.shape(imstack)[0], fittype, P.shape[0]))
    return P, CR, LL


def test_eval_pass(fittype, imstack, theta, coeff, coeff3=None, precision=PREC_MIX):
    """Test hook (bsf_test_eval_pass): one deviance / gradient / Hessian
    evaluation per ROI at theta (N, NV).  Returns err (N,), J (N, NV),
    H (N, NV, NV)."""
    lib = load()
    lib.bsf_test_eval_pass.argtypes = [C.POINTER(Request), C.c_void_p, C.c_void_p]
    lib.bsf_test_eval_pass.restype = C.c_int
    imstack = _f32(imstack)
    S, _, N = imstack.shape
    coeff = _f32(coeff)
    c3 = _f32(coeff3) if coeff3 is not None else None
    NV = _NV[int(fittype)]
    th = np.ascontiguousarray(theta, np.float32).reshape(N, NV)
    out = np.zeros((N, 1 + NV + NV * NV), np.float32)
    req = _request(fittype, S, N, 0, coeff.ctypes.data, coeff.shape,
                   c3.ctypes.data if c3 is not None else None, None, precision, imstack.ctypes.data)
    _check(lib.bsf_test_eval_pass(C.byref(req), th.ctypes.data, out.ctypes.data))
    return out[:, 0], out[:, 1:1 + NV], out[:, 1 + NV:].reshape(N, NV, NV)


class Request4Pi(C.Structure):
    _fields_ = [
        ("imstack", C.c_void_p),
        ("sz", C.c_int),
        ("nfits", C.c_int),
        ("nch", C.c_int),
        ("shared", C.c_void_p),
        ("iterations", C.c_int),
        ("coeff", C.c_void_p),
        ("coeff_dims", C.c_int * 4),
        ("dT", C.c_void_p),
        ("initZ", C.c_void_p),
        ("initPhi", C.c_void_p),
        ("phaseShift", C.c_void_p),
        ("precision", C.c_int),
    ]


def num_params_4pi(shared, nch=4):
    lib = load()
    sh = np.ascontiguousarray(shared, np.uint32)
    lib.bsf_num_params_4pi.argtypes = [C.c_void_p, C.c_int]
    return lib.bsf_num_params_4pi(sh.ctypes.data, int(nch))


def fit_4pi_device(data_ptr, S, N, nch, shared, iterations, coeff_ptr, coeff_shape, phaseShift, P_ptr, CRLB_ptr,
                   LL_ptr, initZ_ptr=None, initPhi_ptr=None, dT_ptr=None, precision=PREC_MIX, stream=None,
                   status_ptr=None, rois_ptr=None, evals_ptr=None):
    """Queue a 4Pi fit on device-resident buffers (bsf_fit_4pi_device_ex):
    data [S, S, N, nch], initZ / initPhi [N] or None, dT [6, nch, N] or None
    (device addresses); shared (6,) and phaseShift (nch,) are host arrays.
    Returns immediately; status / rois / evals as in fit_device."""
    lib = load()
    sh = np.ascontiguousarray(shared, np.uint32)
    ps = np.ascontiguousarray(phaseShift, np.float32)
    r = Request4Pi()
    r.imstack = data_ptr
    r.sz, r.nfits, r.nch = int(S), int(N), int(nch)
    r.shared = sh.ctypes.data
    r.iterations = int(iterations)
    r.coeff = coeff_ptr
    for a in range(4):
        r.coeff_dims[a] = int(coeff_shape[a])
    r.dT, r.initZ, r.initPhi = dT_ptr, initZ_ptr, initPhi_ptr
    r.phaseShift = ps.ctypes.data
    r.precision = int(precision)
    res = Result(P_ptr, CRLB_ptr, LL_ptr)
    x = Extras(rois_ptr, status_ptr, evals_ptr)
    lib.bsf_fit_4pi_device_ex.argtypes = [C.POINTER(Request4Pi), C.POINTER(Result), C.POINTER(Extras), C.c_void_p]
    lib.bsf_fit_4pi_device_ex.restype = C.c_int
    _check(lib.bsf_fit_4pi_device_ex(C.byref(r), C.byref(res), C.byref(x), stream))


def fit_4pi(imstack, shared, iterations, coeff, phaseShift, dT=None, initZ=None, initPhi=None,
            precision=PREC_MIX, status=False):
    """4Pi multi-channel fit on host arrays: imstack (S, S, N, nch) Fortran
    order, shared (6,), coeff 4D, phaseShift (nch,) in w units, optional
    dT (6, nch, N), initZ (N,), initPhi (N,).  Returns P (N, NV+1),
    CRLB (N, NV), LL (N,) -- GPUmleFit_LM_4Pi_bspline's outputs; with
    status=True also the per-ROI status flags (bsf_fit_4pi_host_status)."""
    lib = load()
    lib.bsf_fit_4pi_host.argtypes = [C.POINTER(Request4Pi), C.POINTER(Result)]
    lib.bsf_fit_4pi_host.restype = C.c_int
    im = _f32(imstack)
    if im.ndim != 4:
        raise BsfError(1, "4Pi imstack must be (S, S, N, nch)")
    S, S2, N, nch = im.shape
    if S != S2:
        raise BsfError(4, "Fit Box must be square")
    sh = np.ascontiguousarray(shared, np.uint32)
    coeff = _f32(coeff)
    ps = np.ascontiguousarray(phaseShift, np.float32)
    dT_ = _f32(dT) if dT is not None else None
    iz = np.ascontiguousarray(initZ, np.float32) if initZ is not None else None
    ip = np.ascontiguousarray(initPhi, np.float32) if initPhi is not None else None
    NV = num_params_4pi(sh, nch)
    if NV < 0:
        raise BsfError(1, "bad shared mask or channel count")
    P = np.zeros((N, NV + 1), np.float32, order="F")
    CR = np.zeros((N, NV), np.float32, order="F")
    LL = np.zeros(N, np.float32)
    r = Request4Pi()
    r.imstack = im.ctypes.data
    r.sz, r.nfits, r.nch = int(S), int(N), int(nch)
    r.shared = sh.ctypes.data
    r.iterations = int(iterations)
    r.coeff = coeff.ctypes.data
    for a in range(4):
        r.coeff_dims[a] = int(coeff.shape[a])
    r.dT = dT_.ctypes.data if dT_ is not None else None
    r.initZ = iz.ctypes.data if iz is not None else None
    r.initPhi = ip.ctypes.data if ip is not None else None
    r.phaseShift = ps.ctypes.data
    r.precision = int(precision)
    res = Result(P.ctypes.data, CR.ctypes.data, LL.ctypes.data)
    if not status:
        _check(lib.bsf_fit_4pi_host(C.byref(r), C.byref(res)))
        return P, CR, LL
    flags = np.zeros(N, np.uint8)
    lib.bsf_fit_4pi_host_status.argtypes = [C.POINTER(Request4Pi), C.POINTER(Result), C.c_void_p]
    lib.bsf_fit_4pi_host_status.restype = C.c_int
    _check(lib.bsf_fit_4pi_host_status(C.byref(r), C.byref(res), flags.ctypes.data))
    return P, CR, LL, flags


def GPUmleFit_LM_4Pi_bspline(imstack, shared, iterations, coeffs, dT=None, initZ=None, initPhi=None,
                             phaseShift=None):
    """The reference 4Pi demo's call surface (demo_4pi/bspline_based_4piloc_v4.m:196-206)."""
    return fit_4pi(imstack, shared, iterations, coeffs, phaseShift, dT=dT, initZ=initZ, initPhi=initPhi)


def coeff_build(samples, lam=None):
    """Cubic B-spline coefficients of a sample array on the GPU, the
    reference's bsarray(samples, 'lambda', lam).coeffs (degree 3): samples
    N-D (Fortran order), lam a scalar or one value per axis (smoothing), or
    None.  Returns float32 coefficients with every axis + 3, Fortran order."""
    lib = load()
    lib.bsf_coeff_build_host.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    lib.bsf_coeff_build_host.restype = C.c_int
    x = _f32(samples)
    nd = x.ndim
    dims = np.ascontiguousarray(x.shape, np.int32)
    lm = None if lam is None else np.ascontiguousarray(np.broadcast_to(np.asarray(lam, np.float64), (nd,)))
    out = np.zeros(tuple(d + 3 for d in x.shape), np.float32, order="F")
    _check(lib.bsf_coeff_build_host(x.ctypes.data, dims.ctypes.data, nd,
                                    lm.ctypes.data if lm is not None else None, out.ctypes.data))
    return out
