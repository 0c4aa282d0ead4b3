"""Shared comparison of a GPU fit with the oracle (
coordinates within 1e-4 px (spline units for z, w, u; absolute for g2),
photons / background within 1e-4 relative, CRLB within 1e-4 relative,
iteration counts equal), and device buffers for tests of the device entry
points."""
import ctypes as C

import numpy as np

COORD = {6: [0, 1, 4], 7: [0, 1, 4, 5], 8: [0, 1, 4, 5, 6, 7]}
REL = {6: [2, 3], 7: [2, 3], 8: [2, 3]}


def compare(ft, P, CR, LL, Po, CRo, LLo, atol=1e-4, rtol=1e-4):
    NV = CR.shape[1]
    P = np.asarray(P, np.float64)
    Po = np.asarray(Po, np.float64)
    dcoord = np.abs(P[:, COORD[ft]] - Po[:, COORD[ft]]).max(1)
    drel = (np.abs(P[:, REL[ft]] - Po[:, REL[ft]]) / np.maximum(np.abs(Po[:, REL[ft]]), 1e-12)).max(1)
    dcr = (np.abs(CR - CRo) / np.maximum(np.abs(CRo), 1e-30)).max(1)
    same_kk = P[:, NV] == Po[:, NV]
    ok = (dcoord <= atol) & (drel <= rtol) & (dcr <= rtol) & same_kk
    return dict(n=len(P), agree=float(ok.mean()), same_kk=float(same_kk.mean()),
                p_ok=float(((dcoord <= atol) & (drel <= rtol)).mean()),
                crlb_ok=float((dcr <= rtol).mean()),
                p99_dcoord=float(np.percentile(dcoord, 99)), max_dcoord=float(dcoord.max()),
                p99_drel=float(np.percentile(drel, 99)), p99_dcrlb=float(np.percentile(dcr, 99)),
                finite=bool(np.all(np.isfinite(P)) and np.all(np.isfinite(CR))))


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
