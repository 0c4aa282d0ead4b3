"""Write golden vectors from the reference's own B-spline math library.

Runs only where the reference tree exists: it builds oracle/_ref from the
reference's CPUsplineLib.cpp and MatInvLib.h (oracle/Makefile 'ref' target),
calls those compiled functions on seeded inputs and stores inputs and
outputs in tests/golden/ref_primitives.npz.  Tests then pin the oracle's
restatement to these vectors on any machine.

    python tests/golden/make_ref_vectors.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import loader as L  # noqa: E402


def ref_deriv5_contig(coord, nd, tab, c, th):
    """Call the reference 5D contraction with its six tables in one buffer,
    the layout the reference's own locals have (its z term 27 reads past the
    z table for blocks 24..31)."""
    r, _ = L._R()
    buf = np.ascontiguousarray(tab, np.float32).ravel().copy()
    ptrs = [C.cast(buf.ctypes.data + 4 * 1024 * k, L.fp) for k in range(6)]
    cc = np.asarray(c, np.float32).ravel(order="F")
    t = np.zeros(9, np.float32)
    t[: len(th)] = th
    dudt = np.zeros(8, np.float32)
    m = C.c_float()
    args = ([C.c_float(float(v)) for v in coord[:5]] + [C.c_int(int(v)) for v in nd]
            + ptrs + [L._ptr(cc), L._ptr(t), L._ptr(dudt), C.byref(m)])
    getattr(r, L._M["v5"])(*args)
    return dudt[:7], m.value


def main():
    L.build(with_ref=True)
    rng = np.random.default_rng(20261203)
    out = {}
    # A6: basis at random points and at the piece boundaries
    xs = np.concatenate([rng.uniform(-2.6, 2.6, 4000), np.arange(-2.5, 2.51, 0.25)]).astype(np.float32)
    out["bspline_x"] = xs
    out["bspline_d2"] = np.array([L.ref_bspline(float(x), 2) for x in xs], np.float32)
    out["bspline_d3"] = np.array([L.ref_bspline(float(x), 3) for x in xs], np.float32)
    # A7 / A8: tables and contractions for D = 3, 4, 5
    for D, nd in ((3, [11, 11, 9]), (4, [9, 9, 7, 6]), (5, [7, 7, 5, 5, 5])):
        c = np.asfortranarray(rng.standard_normal([n + 3 for n in nd]).astype(np.float32))
        cases = 40 if D < 5 else 16
        fr = rng.uniform(0, 1, (cases, D)).astype(np.float32)
        fr[0] = 0
        coords = rng.uniform(0.5, np.array(nd, float) + 0.5, (cases, D)).astype(np.float32)
        # pixel coordinates share the ROI-level fractional part
        coords = (np.floor(coords) + fr).astype(np.float32)
        coords[1, 0] = nd[0]            # the xi == nData special case (column)
        coords[2, 1] = nd[1]            # (row)
        thetas = rng.uniform(1, 200, (cases, 5)).astype(np.float32)
        tabs = np.stack([L.ref_delta_tables(D, f) for f in fr])
        dudts, models = [], []
        for i in range(cases):
            if D == 5:
                du, m = ref_deriv5_contig(coords[i], nd, tabs[i], c, thetas[i])
            else:
                du, m = L.ref_deriv_eval(D, coords[i], nd, tabs[i], c, thetas[i])
            dudts.append(du)
            models.append(m)
        out["D%d_coeff" % D] = c
        out["D%d_nd" % D] = np.array(nd, np.int32)
        out["D%d_frac" % D] = fr
        out["D%d_tabs" % D] = tabs
        out["D%d_coords" % D] = coords
        out["D%d_theta" % D] = thetas
        out["D%d_dudt" % D] = np.stack(dudts)
        out["D%d_model" % D] = np.array(models, np.float32)
    # A10 / A12: Cholesky, triangular solves and the Fisher inverse on SPD
    # matrices (plus one with a negative pivot per size)
    for n in (5, 6, 8):
        As, bs, info, Ls, Us, xs_, invd, invm = [], [], [], [], [], [], [], []
        for i in range(30):
            X = rng.standard_normal((n, n + 3)).astype(np.float32)
            A = (X @ X.T).astype(np.float32) * np.float32(rng.uniform(0.1, 1e4))
            if i == 0:
                A[n - 1, n - 1] = -1
            b = rng.standard_normal(n).astype(np.float32)
            inf, Lm, Um = L.ref_cholesky(A)
            x = L.ref_lu_eval(Lm, Um, b, five_d=False) if (inf == 0 and n <= 6) else np.zeros(n, np.float32)
            d, m = L.ref_matinv(A) if i > 0 else (np.zeros(n, np.float32), np.zeros((n, n), np.float32))
            As.append(A); bs.append(b); info.append(inf); Ls.append(Lm); Us.append(Um)
            xs_.append(x); invd.append(d); invm.append(m)
        out["chol%d_A" % n] = np.stack(As)
        out["chol%d_b" % n] = np.stack(bs)
        out["chol%d_info" % n] = np.array(info, np.int32)
        out["chol%d_L" % n] = np.stack(Ls)
        out["chol%d_U" % n] = np.stack(Us)
        out["chol%d_x" % n] = np.stack(xs_)
        out["inv%d_diag" % n] = np.stack(invd)
        out["inv%d_M" % n] = np.stack(invm)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_primitives.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
