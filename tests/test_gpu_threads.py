"""Re-entrancy: bsf_fit_host / bsf_fit_4pi_host called from several host
threads at once (ctypes releases the GIL during the calls) give the same
results, bit for bit, as the same calls made one after another.  Each call
works on its own device buffers and the per-thread default stream, and the
error text is per thread."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import bsfit
from bsfit import gen

pytestmark = pytest.mark.gpu


def _jobs():
    w6 = gen.config(0, N=192)
    w7 = gen.config(1, N=96)
    w8 = gen.config(3, N=48, S=11)[0]
    w4 = gen.make_rois_4pi(32, 13, 77, gen.model_coeffs("4pi"))
    return [
        lambda: bsfit.fit(6, w6.imstack, 50, w6.coeff),
        lambda: bsfit.fit(7, w7.imstack, 50, w7.coeff, w7.initP),
        lambda: bsfit.fit(8, w8.imstack, 50, w8.coeff, w8.initP, w8.coeff3, precision=bsfit.PREC_F64),
        lambda: bsfit.fit_4pi(w4.imstack, w4.shared, 50, w4.coeff, w4.phaseShift, initZ=w4.initZ,
                              initPhi=w4.initPhi),
        lambda: bsfit.fit(6, np.zeros((53, 53, 2), np.float32), 50, w6.coeff),   # an error (size > 51)
    ]


def _run(job):
    try:
        return job()
    except bsfit.BsfError as e:
        return ("error", e.code)


def test_concurrent_host_fits_equal_sequential():
    jobs = _jobs()
    want = [_run(j) for j in jobs]
    assert isinstance(want[-1][0], str) and want[-1] == ("error", 3)
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        for rep in range(3):
            got = list(ex.map(_run, jobs * 2))
            for g, w in zip(got, want * 2):
                if isinstance(w[0], str):   # the expected error
                    assert g == w
                    continue
                for a, b in zip(g, w):
                    np.testing.assert_array_equal(a, b)
            print("round %d: %d concurrent calls equal the sequential results" % (rep, len(got)), flush=True)
