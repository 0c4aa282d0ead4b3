"""CPU pins of the oracle's 4Pi restatement (no reference source exists for
the 4Pi B-spline fitter, so what pins it is: the reduction to the reference's
fittype 7 with one channel; recovery of the truth from noise-free renders; the
NV formula of calculate_CRLB_YL_shared.m:13; and the shared-slot layout)."""
import numpy as np


def test_nv_layout(orc):
    assert orc.nv_4pi([1] * 6, 4) == 6
    assert orc.nv_4pi([1, 1, 1, 0, 1, 1], 4) == 9
    assert orc.nv_4pi([0] * 6, 4) == 24
    for sh in ([1, 1, 1, 0, 1, 1], [0, 1, 1, 1, 0, 1], [1, 0, 0, 1, 1, 1]):
        assert orc.nv_4pi(sh, 4) == 6 * 4 - sum(sh) * 3      # calculate_CRLB_YL_shared.m:13


def test_reduces_to_fittype7(orc):
    """One channel, zero phase offset, everything shared: the 4Pi fit is the
    reference's fittype-7 fit with initP = (initZ, initPhi), bit for bit,
    whenever the phase stays inside (1, nw-1] (no wrap) and inside fittype 7's
    bounds."""
    from bsfit import gen
    coeff = gen.model_coeffs("4pi")
    w = gen.make_rois_4pi(24, 13, 77, coeff, offsets=(0.0,))
    ip = np.clip(w.truth["w"], 4, 14).astype(np.float32)
    same = 0
    for n in range(w.N):
        a = orc.fit(7, np.asfortranarray(w.imstack[:, :, n:n + 1, 0]), 50, coeff, [w.initZ[n], ip[n]])
        b = orc.fit_4pi(np.asfortranarray(w.imstack[:, :, n:n + 1, :]), [1] * 6, 50, coeff, [0.0],
                        initZ=w.initZ[n:n + 1], initPhi=ip[n:n + 1])
        same += all(np.array_equal(x, y) for x, y in zip(a, b))
    assert same >= 0.9 * w.N


def test_noise_free_recovers_truth(orc):
    from bsfit import gen
    coeff = gen.model_coeffs("4pi")
    w = gen.make_rois_4pi(16, 13, 99, coeff, noise=False)
    P, CR, LL = orc.fit_4pi(w.imstack, w.shared, 100, coeff, w.phaseShift,
                            initZ=w.truth["z"].astype(np.float32) + 0.3,
                            initPhi=w.truth["w"].astype(np.float32) + 0.2, nthreads=4)
    t = w.truth
    dphi = (P[:, 5] - t["w"] + 8) % 16 - 8
    ok = (np.abs(P[:, 0] - t["x"]) < 1e-3) & (np.abs(P[:, 1] - t["y"]) < 1e-3) & \
         (np.abs(P[:, 4] - t["z"]) < 1e-2) & (np.abs(dphi) < 1e-2) & (np.abs(P[:, 2] / t["N"] - 1) < 1e-3)
    assert ok.mean() >= 0.8
