"""One GPU evaluation pass (model, derivatives, Poisson deviance, gradient,
Hessian) at given parameters against the oracle's pass (oracle.eval_pass)
on the same ROIs: the FP64 kernel must agree to ~1e-9, the FP32 kernel to
FP32 rounding of sums over S^2 pixels.  This pins the factorised
contraction (including the reference's two irregular 5D derivative terms)
independently of LM trajectories."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(cfg, S, N, rng):
    from bsfit import gen
    w = gen.config(cfg, N=N, S=S)
    w = w[0] if isinstance(w, list) else w
    th = w.truth.copy()
    # move away from the truth so gradients are non-zero
    th[:, 0] += rng.uniform(-0.6, 0.6, N)
    th[:, 1] += rng.uniform(-0.6, 0.6, N)
    th[:, 2] *= rng.uniform(0.7, 1.3, N)
    th[:, 4] += rng.uniform(-1.5, 1.5, N)
    if w.fittype == 8:
        th[:, 5] += rng.uniform(-0.7, 0.7, N)
        th[:, 6] = np.clip(th[:, 6] + rng.uniform(-0.7, 0.7, N), 1.05, 8.4)
        th[:, 7] = np.clip(th[:, 7] + rng.uniform(-0.2, 0.2, N), 0, 1)
    return w, th.astype(np.float32)


@pytest.mark.parametrize("cfg,S,N", [(0, None, 64), (1, None, 64), (3, 11, 32), (3, 17, 32)])
def test_eval_pass_matches_oracle(orc, cfg, S, N):
    import bsfit
    rng = np.random.default_rng(7 + cfg)
    w, th = _case(cfg, S, N, rng)
    ft = w.fittype
    print("case fittype=%d S=%d N=%d" % (ft, w.S, N), flush=True)
    for prec, tol in ((bsfit.PREC_F64, 3e-7), (bsfit.PREC_MIX, 2e-4)):
        e, J, H = bsfit.api.test_eval_pass(ft, w.imstack, th, w.coeff, w.coeff3, precision=prec)
        worst = np.zeros(3)
        for n in range(N):
            eo, Jo, Ho, _ = orc.eval_pass(ft, w.imstack[:, :, n], th[n].astype(np.float64), w.coeff, w.coeff3, f64=True)
            sJ = np.abs(Jo).max() + 1e-30
            sH = np.abs(Ho).max() + 1e-30
            worst = np.maximum(worst, [abs(e[n] - eo) / abs(eo), np.abs(J[n] - Jo).max() / sJ,
                                       np.abs(H[n] - Ho).max() / sH])
        print("precision", prec, "max rel err (err, J, H):", worst, flush=True)
        assert np.all(worst < tol), worst
