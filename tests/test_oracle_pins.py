"""Pins of the CPU oracle (oracle/bsf_oracle.c).

* Against golden vectors written by the reference's own compiled
  CPUsplineLib.cpp / MatInvLib.h (tests/golden/make_ref_vectors.py): the
  basis, weight tables, gather-contract, Cholesky, solves and Fisher inverse
  must match bit for bit.
* Against mathematics for what those leave open: partition of unity,
  derivative identities, interpolation at knots, finite differences,
  library inverses, the centre of mass of a point image.
"""
import numpy as np
import pytest


def same(a, b):
    return np.asarray(a, np.float32).tobytes() == np.asarray(b, np.float32).tobytes()


# ---------------------------------------------------------- golden vectors
def test_bspline_basis_bitexact(orc, golden):
    xs = golden["bspline_x"]
    for d in (2, 3):
        got = np.array([orc.bspline(float(x), d) for x in xs], np.float32)
        assert same(got, golden["bspline_d%d" % d]), "basis deg %d differs from the reference" % d


@pytest.mark.parametrize("D", [3, 4, 5])
def test_weight_tables_bitexact(orc, golden, D):
    for fr, ref in zip(golden["D%d_frac" % D], golden["D%d_tabs" % D]):
        assert same(orc.delta_tables(D, fr), ref)


@pytest.mark.parametrize("D", [3, 4, 5])
def test_gather_contract_bitexact(orc, golden, D):
    c = golden["D%d_coeff" % D]
    nd = golden["D%d_nd" % D]
    for co, tab, th, du, m in zip(golden["D%d_coords" % D], golden["D%d_tabs" % D],
                                  golden["D%d_theta" % D], golden["D%d_dudt" % D],
                                  golden["D%d_model" % D]):
        d, mm = orc.deriv_eval(D, co, nd, tab, c, th)
        assert same(d, du[: D + 2]) and same(mm, m)


@pytest.mark.parametrize("n", [5, 6, 8])
def test_cholesky_solve_inverse_bitexact(orc, golden, n):
    for i, (A, b) in enumerate(zip(golden["chol%d_A" % n], golden["chol%d_b" % n])):
        info, L, U = orc.cholesky(A)
        assert info == golden["chol%d_info" % n][i]
        if info == 0:
            assert same(L, golden["chol%d_L" % n][i]) and same(U, golden["chol%d_U" % n][i])
            if n <= 6:
                assert same(orc.lu_eval(L, U, b), golden["chol%d_x" % n][i])
        if i > 0:
            d, M = orc.matinv(A)
            assert same(d, golden["inv%d_diag" % n][i]) and same(M, golden["inv%d_M" % n][i])


# ------------------------------------------------------------- mathematics
def test_basis_partition_of_unity_and_derivative(orc):
    rng = np.random.default_rng(0)
    for x in rng.uniform(0, 1, 200):
        s3 = sum(orc.bspline(float(x - k), 3, f64=True) for k in range(-2, 3))
        s2 = sum(orc.bspline(float(x - k), 2, f64=True) for k in range(-2, 3))
        assert abs(s3 - 1) < 1e-12 and abs(s2 - 1) < 1e-12
        # beta3'(x) = beta2(x + 1/2) - beta2(x - 1/2)
        for y in (x - 1.5, x - 0.5, x + 0.5):
            h = 1e-6
            fd = (orc.bspline(y + h, 3, f64=True) - orc.bspline(y - h, 3, f64=True)) / (2 * h)
            an = orc.bspline(y + 0.5, 2, f64=True) - orc.bspline(y - 0.5, 2, f64=True)
            assert abs(fd - an) < 1e-6
    assert orc.bspline(2.0, 3) == 0 and orc.bspline(-2.5, 2) == 0
    assert orc.bspline(0.3, 3) == orc.bspline(-0.3, 3)


def test_spline_interpolates_samples_and_derivatives_match_fd(orc):
    from bsfit import gen
    rng = np.random.default_rng(1)
    samples = rng.standard_normal((12, 12, 9))
    c = gen.bspline_coeffs(samples)
    nd = [12, 12, 9]
    th = [0, 0, 1.0, 0.0, 0]            # N = 1, bg = 0: model = f
    for (r, q, z) in [(3, 4, 2), (7, 1, 5), (10, 11, 8)]:
        # coordinate = sample index + 1 reproduces the sample
        co = np.array([q + 1, r + 1, z + 1], np.float32)
        tab = orc.delta_tables(3, co - np.floor(co))
        _, m = orc.deriv_eval(3, co, nd, tab, c, th)
        assert abs(m - samples[r, q, z]) < 2e-5
    # derivatives against central differences of f (double-checked in FP64)
    for _ in range(20):
        co = rng.uniform(2.2, 8.8, 3).astype(np.float32)
        tab = orc.delta_tables(3, co - np.floor(co))
        d, _ = orc.deriv_eval(3, co, nd, tab, c, th)
        for ax, (k, sign) in enumerate([(0, -1), (1, -1), (4, 1)]):
            h = 1e-2
            cp, cm = co.copy(), co.copy()
            cp[ax] += h
            cm[ax] -= h
            fp = orc.deriv_eval(3, cp, nd, orc.delta_tables(3, cp - np.floor(cp)), c, th)[1]
            fm = orc.deriv_eval(3, cm, nd, orc.delta_tables(3, cm - np.floor(cm)), c, th)[1]
            fd = (fp - fm) / (2 * h)
            # dudt[k] = -N * d/d(col) for k=0 (ax 0 = column), -N d/d(row) for k=1
            assert abs(sign * d[k] - fd) < 2e-3 * (1 + abs(fd))


def test_matinv_matches_library_inverse(orc):
    rng = np.random.default_rng(2)
    for n in (5, 6, 8):
        X = rng.standard_normal((n, 2 * n))
        A = (X @ X.T).astype(np.float32)
        d, M = orc.matinv(A)
        ref = np.linalg.inv(A.astype(np.float64))
        assert np.allclose(M, ref, rtol=2e-3, atol=1e-4)
        assert np.allclose(d, np.diag(ref), rtol=2e-3)


def test_cholesky_closed_form(orc):
    rng = np.random.default_rng(3)
    for n in (5, 6, 8):
        X = rng.standard_normal((n, 2 * n))
        A = (X @ X.T).astype(np.float32)
        info, L, U = orc.cholesky(A)
        assert info == 0
        Lnp = np.linalg.cholesky(A.astype(np.float64))
        assert np.allclose(np.tril(U), Lnp, rtol=1e-4, atol=1e-5)
        b = rng.standard_normal(n).astype(np.float32)
        assert np.allclose(orc.lu_eval(L, U, b), np.linalg.solve(A.astype(np.float64), b), rtol=1e-3, atol=1e-4)
        A2 = A.copy()
        A2[2, 2] = -abs(A2[2, 2])
        assert orc.cholesky(A2)[0] == 1


def test_initial_estimates_closed_form(orc):
    roi = np.zeros((9, 9), np.float32)
    roi[2, 6] = 10.0   # row 2, column 6
    x, y = orc.com2d(roi)
    assert (x, y) == (2.0, 6.0)          # the row centroid goes into theta0
    roi[4, 4] = -3.0
    mx, mn = orc.maxmin2d(roi)
    assert (mx, mn) == (10.0, -3.0)
    assert orc.maxmin2d(np.full((5, 5), -1, np.float32))[0] == 0.0   # max floored at 0


@pytest.mark.parametrize("ft", [6, 7])
def test_pass_gradient_and_hessian(orc, ft):
    """A9: deviance is zero at mu == data; J is half the deviance gradient
    (finite differences, FP64 build); H is the expected information."""
    from bsfit import gen
    w = gen.config(0 if ft == 6 else 1, N=4, noise=False)
    roi = np.asarray(w.imstack[:, :, 0], np.float64)
    th = w.truth[0].astype(np.float64)
    mu = orc.eval_pass(ft, roi.astype(np.float32), th, w.coeff, f64=True)[3]
    roi = mu.copy()          # data == model exactly (FP64)
    err, J, H, _ = orc.eval_pass(ft, roi, th, w.coeff, f64=True)
    assert abs(err) < 1e-6 and np.abs(J).max() < 1e-4   # data rounded to float32
    th2 = th.copy()
    th2[0] += 0.2
    th2[4] += 0.3
    err, J, H, mu = orc.eval_pass(ft, roi, th2, w.coeff, f64=True)
    for k in (0, 1, 2, 3, 4):
        h = 1e-4 * max(1.0, abs(th2[k]))
        tp, tm = th2.copy(), th2.copy()
        tp[k] += h
        tm[k] -= h
        fd = (orc.eval_pass(ft, roi, tp, w.coeff, f64=True)[0] - orc.eval_pass(ft, roi, tm, w.coeff, f64=True)[0]) / (2 * h)
        assert abs(J[k] - fd / 2) < 1e-4 * (1 + abs(fd))
    assert np.allclose(H, H.T) and np.all(np.linalg.eigvalsh(H) > -1e-6 * np.abs(H).max())


def test_ft7_first_hessian_singular(orc):
    """The reference's fittype-7 first pass sets dudt[5] = dudt[4]
    (CPUmleFit_LM.cpp:2896): the first Hessian has equal rows 4 and 5 and is
    singular, so the first Cholesky's success (and with it the whole
    trajectory) rests on rounding.  This is why fittype-7 and 4Pi fits agree
    across arithmetics only as rates."""
    from bsfit import gen
    w = gen.config(1, N=2)
    th = w.truth[0].astype(np.float64)
    err, J, H, _ = orc.eval_pass(7, w.imstack[:, :, 0], th, w.coeff, f64=True, first=True)
    assert np.array_equal(H[4], H[5]) and J[4] == J[5]
    assert abs(np.linalg.det(H)) <= 1e-9 * abs(np.prod(np.diag(H)))
    err2, J2, H2, _ = orc.eval_pass(7, w.imstack[:, :, 0], th, w.coeff, f64=True, first=False)
    assert not np.array_equal(H2[4], H2[5])
