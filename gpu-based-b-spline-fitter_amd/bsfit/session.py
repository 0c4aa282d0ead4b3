"""ctypes binding of the serving-session part of include/bsfit.h.

    with Session(arena_bytes=64 << 20) as s:
        m = s.register_model(7, coeff)
        t = s.submit([Group(m, imstack_a, initP=initP), Group(m, imstack_b, initP=initP)])
        (P_a, CRLB_a, LL_a), (P_b, CRLB_b, LL_b) = s.poll(t, wait=True)

A group is one MEX call's worth of ROIs (one model, one ROI size); the
session fits all groups of a frame that share model, iterations and start in
one launch.  Outputs are returned per group in the MEX layout (P (N, NV+1),
CRLB (N, NV), LL (N,)).
"""
import ctypes as C

import numpy as np

from .api import BsfError, PREC_MIX, _check, _f32, load

FT_4PI = 104
BUSY = 7
PENDING = 8
SUBMIT_NONBLOCK = 1
_NV = {6: 5, 7: 6, 8: 8}


class SessionConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget", C.c_size_t), ("arena_bytes", C.c_size_t),
                ("kernel_streams", C.c_int), ("precision", C.c_int)]


class ModelDesc(C.Structure):
    _fields_ = [("fittype", C.c_int), ("coeff", C.c_void_p), ("coeff_ndim", C.c_int),
                ("coeff_dims", C.c_int * 5), ("coeff3", C.c_void_p), ("nch", C.c_int),
                ("shared", C.c_uint * 6), ("phaseShift", C.c_float * 8)]


class ModelSamples(C.Structure):
    _fields_ = [("fittype", C.c_int), ("samples", C.c_void_p), ("ndim", C.c_int), ("dims", C.c_int * 5),
                ("lambda_", C.c_void_p), ("samples3", C.c_void_p), ("lambda3", C.c_void_p), ("nch", C.c_int),
                ("shared", C.c_uint * 6), ("phaseShift", C.c_float * 8)]


class RoiGroup(C.Structure):
    _fields_ = [("model", C.c_int), ("sz", C.c_int), ("count", C.c_int), ("iterations", C.c_int),
                ("data", C.c_void_p), ("chan_stride", C.c_size_t), ("initP", C.c_void_p), ("n_initP", C.c_int),
                ("dT", C.c_void_p), ("initZ", C.c_void_p), ("initPhi", C.c_void_p),
                ("P", C.c_void_p), ("CRLB", C.c_void_p), ("LL", C.c_void_p),
                ("nstarts", C.c_int), ("strict", C.c_int)]


class Frame(C.Structure):
    _fields_ = [("groups", C.POINTER(RoiGroup)), ("ngroups", C.c_int)]


def _bind(lib):
    if getattr(lib, "_bsf_session_bound", False):
        return lib
    lib.bsf_session_create.argtypes = [C.POINTER(SessionConfig), C.POINTER(C.c_void_p)]
    lib.bsf_session_destroy.argtypes = [C.c_void_p]
    lib.bsf_model_register.argtypes = [C.c_void_p, C.POINTER(ModelDesc), C.POINTER(C.c_int)]
    lib.bsf_model_release.argtypes = [C.c_void_p, C.c_int]
    lib.bsf_model_register_samples.argtypes = [C.c_void_p, C.POINTER(ModelSamples), C.POINTER(C.c_int)]
    lib.bsf_session_set_precision.argtypes = [C.c_void_p, C.c_int]
    lib.bsf_submit.argtypes = [C.c_void_p, C.POINTER(Frame), C.c_int, C.POINTER(C.c_ulonglong)]
    lib.bsf_poll.argtypes = [C.c_void_p, C.c_ulonglong, C.c_int]
    lib.bsf_session_inflight.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]
    lib.bsf_session_set_graphs.argtypes = [C.c_void_p, C.c_int]
    lib.bsf_session_set_hetero.argtypes = [C.c_void_p, C.c_int]
    lib.bsf_session_graph_stats.argtypes = [C.c_void_p] + [C.POINTER(C.c_ulonglong)] * 3
    for f in ("bsf_session_create", "bsf_session_destroy", "bsf_model_register", "bsf_model_register_samples",
              "bsf_model_release", "bsf_session_set_graphs", "bsf_session_graph_stats", "bsf_session_set_hetero",
              "bsf_session_set_precision", "bsf_submit", "bsf_poll", "bsf_session_inflight"):
        getattr(lib, f).restype = C.c_int
    lib._bsf_session_bound = True
    return lib


def _roi_view(a, nch):
    """(array kept alive, pointer, channel stride in floats, S, count) of an
    ROI stack: (S, S, N) or, for 4Pi, (S, S, N, nch) -- a Fortran-ordered
    array or a slice of one along N (channels then sit chan_stride floats
    apart); anything else is copied to Fortran order first."""
    a = np.asarray(a)
    if a.dtype != np.float32:
        a = a.astype(np.float32)
    if nch > 1 or a.ndim == 4:
        if a.ndim != 4 or a.shape[3] != nch:
            raise BsfError(1, "4Pi groups take (S, S, N, %d) stacks" % nch)
        S, S2, N, _ = a.shape
        st = a.strides
        if not (st[0] == 4 and st[1] == 4 * S and st[2] == 4 * S * S and st[3] % 4 == 0 and st[3] >= 4 * N * S * S):
            a = np.asfortranarray(a)
            st = a.strides
        return a, a.ctypes.data, st[3] // 4, S, N
    if a.ndim == 2:
        a = a[:, :, None]
    if a.ndim != 3:
        raise BsfError(1, "imstack must be (S, S, N)")
    if not a.flags.f_contiguous:
        a = np.asfortranarray(a)
    return a, a.ctypes.data, 0, a.shape[0], a.shape[2]


class Group:
    """One group of a frame: ROIs of one model and one size.  starts (K, 2)
    for fittype 7 or (K, 4) for fittype 8 fits every ROI from K starts and
    keeps, start by start, the result whose log-likelihood is higher by at
    least 1e-4 (strict: more than 1e-4), as bsfit.fit_multistart does."""

    def __init__(self, model, imstack, iterations=50, initP=None, dT=None, initZ=None, initPhi=None,
                 starts=None, strict=False):
        self.model = model
        self.imstack = imstack
        self.iterations = int(iterations)
        self.nstarts, self.strict, self.row = 0, int(bool(strict)), 0
        if starts is not None:
            st = np.ascontiguousarray(np.atleast_2d(starts), np.float32)
            self.nstarts, self.row = st.shape[0], st.shape[1]
            initP = st
        self.initP = None if initP is None else np.ascontiguousarray(np.ravel(initP), np.float32)
        self.dT = None if dT is None else np.asfortranarray(dT, np.float32)
        self.initZ = None if initZ is None else np.ascontiguousarray(initZ, np.float32)
        self.initPhi = None if initPhi is None else np.ascontiguousarray(initPhi, np.float32)


class Session:
    """A device-resident fitting session (bsf_session_*)."""

    def __init__(self, device=0, hbm_budget=0, arena_bytes=0, kernel_streams=0, precision=PREC_MIX):
        self._lib = _bind(load())
        cfg = SessionConfig(int(device), int(hbm_budget), int(arena_bytes), int(kernel_streams), int(precision))
        h = C.c_void_p()
        _check(self._lib.bsf_session_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._models = {}
        self._frames = {}

    def close(self):
        if self._h:
            h, self._h = self._h, None
            _check(self._lib.bsf_session_destroy(h))
            self._frames.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def register_model(self, fittype, coeff, coeff3=None, nch=1, shared=None, phaseShift=None):
        coeff = _f32(coeff)
        c3 = _f32(coeff3) if coeff3 is not None else None
        d = ModelDesc()
        d.fittype = int(fittype)
        d.coeff = coeff.ctypes.data
        d.coeff_ndim = coeff.ndim
        for a, n in enumerate(coeff.shape[:5]):
            d.coeff_dims[a] = int(n)
        d.coeff3 = c3.ctypes.data if c3 is not None else None
        nv = _NV.get(int(fittype), 0)
        if int(fittype) == FT_4PI:
            d.nch = int(nch)
            sh = np.ravel(shared if shared is not None else np.ones(6))
            for p in range(6):
                d.shared[p] = int(sh[p] != 0)
            ps = np.ravel(phaseShift if phaseShift is not None else np.zeros(nch))
            for k in range(min(8, len(ps))):
                d.phaseShift[k] = float(ps[k])
            nv = int(sum(1 if sh[p] else nch for p in range(6)))
        mid = C.c_int(-1)
        _check(self._lib.bsf_model_register(self._h, C.byref(d), C.byref(mid)))
        self._models[mid.value] = (int(fittype), nv, int(nch) if int(fittype) == FT_4PI else 1)
        return mid.value

    def register_model_samples(self, fittype, samples, lam=None, samples3=None, lam3=None, nch=1, shared=None,
                               phaseShift=None):
        """Register a model from its PSF samples: the coefficients are built
        on the device (bsf_model_register_samples); lam / lam3 per-axis
        smoothing as bsarray's 'lambda'."""
        x = _f32(samples)
        x3 = _f32(samples3) if samples3 is not None else None
        d = ModelSamples()
        d.fittype = int(fittype)
        d.samples = x.ctypes.data
        d.ndim = x.ndim
        for a, n in enumerate(x.shape[:5]):
            d.dims[a] = int(n)
        lm = None if lam is None else np.ascontiguousarray(np.broadcast_to(np.asarray(lam, np.float64), (x.ndim,)))
        lm3 = None if lam3 is None else np.ascontiguousarray(np.broadcast_to(np.asarray(lam3, np.float64), (3,)))
        d.lambda_ = lm.ctypes.data if lm is not None else None
        d.samples3 = x3.ctypes.data if x3 is not None else None
        d.lambda3 = lm3.ctypes.data if lm3 is not None else None
        nv = _NV.get(int(fittype), 0)
        if int(fittype) == FT_4PI:
            d.nch = int(nch)
            sh = np.ravel(shared if shared is not None else np.ones(6))
            for p in range(6):
                d.shared[p] = int(sh[p] != 0)
            ps = np.ravel(phaseShift if phaseShift is not None else np.zeros(nch))
            for k in range(min(8, len(ps))):
                d.phaseShift[k] = float(ps[k])
            nv = int(sum(1 if sh[p] else nch for p in range(6)))
        mid = C.c_int(-1)
        _check(self._lib.bsf_model_register_samples(self._h, C.byref(d), C.byref(mid)))
        self._models[mid.value] = (int(fittype), nv, int(nch) if int(fittype) == FT_4PI else 1)
        return mid.value

    def release_model(self, model):
        _check(self._lib.bsf_model_release(self._h, int(model)))
        self._models.pop(int(model), None)

    def set_precision(self, precision):
        _check(self._lib.bsf_session_set_precision(self._h, int(precision)))

    def set_graphs(self, slots=16):
        """Graph mode (bsf_session_set_graphs) with `slots` arena slots:
        frames replay captured graphs of their launches; 0 (or False) returns
        to per-call launches.  Waits for the frames in flight."""
        _check(self._lib.bsf_session_set_graphs(self._h, int(slots)))

    def set_hetero(self, on=True):
        """Heterogeneous launches (bsf_session_set_hetero): a frame's
        single-start fittype 6-8 groups run as one launch."""
        _check(self._lib.bsf_session_set_hetero(self._h, 1 if on else 0))

    def graph_stats(self):
        """(graphs captured, frames replayed, frames launched call by call in graph mode)."""
        a, b, c = C.c_ulonglong(), C.c_ulonglong(), C.c_ulonglong()
        _check(self._lib.bsf_session_graph_stats(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def inflight(self):
        n, b = C.c_int(), C.c_size_t()
        _check(self._lib.bsf_session_inflight(self._h, C.byref(n), C.byref(b)))
        return n.value, b.value

    def submit(self, groups, block=True):
        """Queue a frame; returns its ticket, or None when block=False and the
        arena has no room now."""
        n = len(groups)
        arr = (RoiGroup * max(n, 1))()
        keep, outs = [], []
        for i, g in enumerate(groups):
            ft, nv, nch = self._models.get(int(g.model), (0, 0, 1))
            a, ptr, cst, S, N = _roi_view(g.imstack, nch)
            P = np.empty((N, nv + 1), np.float32, order="F")
            CR = np.empty((N, nv), np.float32, order="F")
            LL = np.empty(N, np.float32)
            keep.append((a, g))
            outs.append((P, CR, LL))
            r = arr[i]
            r.model, r.sz, r.count, r.iterations = int(g.model), int(S), int(N), g.iterations
            r.data, r.chan_stride = ptr, int(cst)
            if g.initP is not None:
                r.initP, r.n_initP = g.initP.ctypes.data, int(g.row if g.nstarts else g.initP.size)
                r.nstarts, r.strict = g.nstarts, g.strict
            r.dT = g.dT.ctypes.data if g.dT is not None else None
            r.initZ = g.initZ.ctypes.data if g.initZ is not None else None
            r.initPhi = g.initPhi.ctypes.data if g.initPhi is not None else None
            r.P, r.CRLB, r.LL = P.ctypes.data, CR.ctypes.data, LL.ctypes.data
        fr = Frame(arr, n)
        t = C.c_ulonglong(0)
        rc = self._lib.bsf_submit(self._h, C.byref(fr), 0 if block else SUBMIT_NONBLOCK, C.byref(t))
        if rc == BUSY and not block:
            return None
        _check(rc)
        self._frames[t.value] = outs
        return t.value

    def poll(self, ticket, wait=False):
        """The frame's per-group (P, CRLB, LL), or None while it runs."""
        rc = self._lib.bsf_poll(self._h, C.c_ulonglong(ticket), 1 if wait else 0)
        if rc == PENDING:
            return None
        _check(rc)
        if ticket not in self._frames:
            raise BsfError(1, "frame %d was already collected" % ticket)
        return self._frames.pop(ticket)
