"""The GPU coefficient builder (bsf_coeff_build_*, bsf_model_register_samples)
against the oracle restatement of the reference's bsarray (degree 3): FP64 on
both sides, so the only difference is the GPU result's rounding to FP32
(bound: 2e-7 of the array's largest coefficient, 1.7 float ulps there)."""
import ctypes as C

import numpy as np
import pytest

import bsfit
from bsfit import gen
from oracle import bsarray as B
from parity_util import Hip as _Hip

pytestmark = pytest.mark.gpu

CASES = [((7,), 0.0), ((12, 9), 0.0), ((12, 9), 0.2), ((8, 6, 5), 0.1), ((6, 5, 4, 7), [0.0, 0.0, 0.1, 0.2]),
         ((5, 4, 4, 3, 3), 1.0), ((2, 3, 2), 0.0), ((33, 17, 9), 5.0),
         # axes longer than 48 samples take the in-place global-memory path
         ((60, 5), 0.0), ((50, 3, 2), [0.3, 0.0, 0.1]), ((100,), 1.0), ((7, 64), [0.0, 2.0])]


def _close(gpu, ref):
    tol = 2e-7 * np.abs(ref).max()
    err = np.abs(gpu.astype(np.float64) - ref).max()
    assert err <= tol, (err, tol)
    return err / np.abs(ref).max()


@pytest.mark.parametrize("shape,lam", CASES)
def test_builder_matches_oracle(shape, lam):
    rng = np.random.default_rng(sum(shape))
    x = np.asfortranarray(rng.standard_normal(shape).astype(np.float32))
    g = bsfit.coeff_build(x, lam)
    r = B.bsarray(x.astype(np.float64), lam)
    assert g.shape == r.shape == tuple(d + 3 for d in shape)
    print("shape", shape, "lambda", lam, "max rel err %.2e" % _close(g, r), flush=True)


def test_builder_on_the_fit_models():
    """The 5D vortex model (21x21x21x17x9) and the colour model with the
    colour demo's smoothing: the arrays the fitter reads.  The builder takes
    FP32 samples, so the oracle gets the same FP32-rounded samples."""
    x5 = gen.psf_dipole5d().astype(np.float32)
    g5 = bsfit.coeff_build(x5)
    _close(g5, B.bsarray(x5.astype(np.float64)))
    # the generator behind the fit tests builds from the FP64 samples: the
    # samples' FP32 rounding, amplified by the prefilter, is all that differs
    assert np.abs(g5 - gen.model_coeffs("dipole5d")[0]).max() <= 3e-6 * np.abs(g5).max()
    xc = gen.psf_color4d().astype(np.float32)
    lam = [0.0, 0.0, 0.1, 0.2]
    _close(bsfit.coeff_build(xc, lam), B.bsarray(xc.astype(np.float64), lam))


def test_device_entry_point_matches_host():
    lib = bsfit.load()
    lib.bsf_coeff_work_bytes.argtypes = [C.c_void_p, C.c_int]
    lib.bsf_coeff_work_bytes.restype = C.c_size_t
    lib.bsf_coeff_build_device.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p]
    lib.bsf_coeff_build_device.restype = C.c_int
    x = np.asfortranarray(np.random.default_rng(9).standard_normal((10, 8, 6, 4)).astype(np.float32))
    lam = np.array([0.0, 0.05, 0.0, 0.3])
    dims = np.array(x.shape, np.int32)
    wb = lib.bsf_coeff_work_bytes(dims.ctypes.data, 4)
    assert wb == x.size * 16
    hip = _Hip()
    nout = int(np.prod([d + 3 for d in x.shape]))
    d_x = hip.put(x.ravel(order="F"))
    d_c, d_w = hip.alloc(nout * 4), hip.alloc(wb)
    try:
        rc = lib.bsf_coeff_build_device(d_x, dims.ctypes.data, 4, lam.ctypes.data, d_c, d_w, None)
        assert rc == 0, lib.bsf_last_error()
        assert hip.h.hipDeviceSynchronize() == 0
        got = hip.get(d_c, nout, np.float32).reshape(tuple(d + 3 for d in x.shape), order="F")
    finally:
        for p in (d_x, d_c, d_w):
            hip.h.hipFree(p)
    np.testing.assert_array_equal(got, bsfit.coeff_build(x, lam))


def test_session_registers_models_from_samples():
    """A model registered from samples fits exactly as one registered from
    coefficients built by bsf_coeff_build_host."""
    x5, x3 = gen.psf_dipole5d(), gen.psf_dipole3d()
    c5, c3 = bsfit.coeff_build(x5), bsfit.coeff_build(x3)
    w = gen.make_rois(8, 48, 13, 77, c5, c3)
    wc = gen.make_rois(7, 48, 15, 78, gen.model_coeffs("color4d"))
    xc = gen.psf_color4d()
    lam = [0.0, 0.0, 0.1, 0.2]
    with bsfit.Session(arena_bytes=32 << 20) as s:
        m_samples = s.register_model_samples(8, x5, samples3=x3)
        m_coeff = s.register_model(8, c5, c3)
        mc_samples = s.register_model_samples(7, xc, lam=lam)
        mc_coeff = s.register_model(7, bsfit.coeff_build(xc, lam))
        t = s.submit([bsfit.Group(m_samples, w.imstack, 50, initP=w.initP),
                      bsfit.Group(m_coeff, w.imstack, 50, initP=w.initP),
                      bsfit.Group(mc_samples, wc.imstack, 50, initP=wc.initP),
                      bsfit.Group(mc_coeff, wc.imstack, 50, initP=wc.initP)])
        r = s.poll(t, wait=True)
    for a, b in ((r[0], r[1]), (r[2], r[3])):
        for u, v in zip(a, b):
            np.testing.assert_array_equal(u, v)
    assert np.isfinite(r[0][0]).all()


@pytest.mark.parametrize("seed", range(8))
def test_random_shapes_match_oracle(seed):
    """Randomly drawn ranks (1..5), lengths (2..60: both the staged and the
    in-place paths) and per-axis lambda (0 or up to 5) against the oracle."""
    rng = np.random.default_rng(900 + seed)
    nd = int(rng.integers(1, 6))
    budget = 400000
    shape = []
    for a in range(nd):
        hi = max(3, min(61, int(budget ** (1.0 / (nd - a)))))
        d = int(rng.integers(2, hi))
        shape.append(d)
        budget = max(8, budget // d)
    lam = [0.0 if rng.random() < 0.5 else float(rng.uniform(0.01, 5.0)) for _ in shape]
    x = np.asfortranarray(rng.standard_normal(tuple(shape)).astype(np.float32))
    print("seed %d shape %s lambda %s" % (seed, tuple(shape), np.round(lam, 3).tolist()), flush=True)
    _close(bsfit.coeff_build(x, lam), B.bsarray(x.astype(np.float64), lam))
