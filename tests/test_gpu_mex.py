"""The MATLAB MEX gateways on the GPU, called as MATLAB calls them (through
the test MX API): GPUmleFit_LM_multiD equals bsf_fit_host bit for bit for
fittypes 6, 7 (initP) and 8 (coeff3), with 2-D and 3-D data and fewer
outputs; bsf_find_rois equals bsf_find_rois_host; and the GUI's whole path
(frames -> bsf_find_rois -> GPUmleFit_LM_multiD -> bsf_gui_results) puts the
planted emitters where they are: x/y in frame pixels (1-based), z in nm with
the GUI's sign (MultiD_fitter/simplefitter_spline.m:598-630), the colour
coordinate in column 5 of the 4D table."""
import numpy as np
import pytest

import bsfit
from bsfit import gen
from parity_util import Mex

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mexdir(tmp_path_factory):
    return tmp_path_factory.mktemp("mexgpu")


def test_fitter_gateway_equals_library(mexdir):
    m = Mex("GPUmleFit_LM_multiD", mexdir)
    w6 = gen.config(0, N=64)
    got = m(3, w6.imstack, 6.0, 50.0, w6.coeff)
    for a, b in zip(got, bsfit.fit(6, w6.imstack, 50, w6.coeff)):
        np.testing.assert_array_equal(a.reshape(b.shape), b)
    w7 = gen.config(1, N=48)
    got = m(3, w7.imstack, 7.0, 50.0, w7.coeff, 0.0, 1.0, w7.initP)
    for a, b in zip(got, bsfit.fit(7, w7.imstack, 50, w7.coeff, w7.initP)):
        np.testing.assert_array_equal(a.reshape(b.shape), b)
    w8 = gen.config(3, N=40, S=13)[0]
    c5, c3 = gen.model_coeffs("dipole5d")
    got = m(1, w8.imstack, 8.0, 50.0, c5, 0.0, 1.0, w8.initP, c3)          # P only
    assert len(got) == 1
    np.testing.assert_array_equal(got[0], bsfit.fit(8, w8.imstack, 50, c5, w8.initP, c3)[0])
    one = np.asfortranarray(w6.imstack[:, :, 5])                          # 2-D data: one ROI
    got = m(3, one, 6.0, 50.0, w6.coeff)
    np.testing.assert_array_equal(got[0], bsfit.fit(6, one, 50, w6.coeff)[0])


def test_fitter_gateway_status_output(mexdir):
    """The optional fourth output: the per-ROI status flags (uint8, N x 1),
    equal to bsfit.fit_status; the first three outputs are unchanged by it."""
    m = Mex("GPUmleFit_LM_multiD", mexdir)
    w7 = gen.config(1, N=48)
    for iters in (50.0, 3.0):
        got = m(4, w7.imstack, 7.0, iters, w7.coeff, 0.0, 1.0, w7.initP)
        ref = bsfit.fit_status(7, w7.imstack, int(iters), w7.coeff, w7.initP)
        assert got[3].dtype == np.uint8 and got[3].shape == (48, 1)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a.reshape(b.shape), b)
    assert np.all(got[3].ravel() & bsfit.ROI_MAX_ITER)                  # 3 iterations: every ROI hit the limit


def test_roi_gateway_equals_library(mexdir):
    m = Mex("bsf_find_rois", mexdir)
    frames, _ = gen.make_frames(3, 120, 110, 10, 12)
    h = bsfit.peak_filter("dog", 1.0)
    st, pc, val = m(3, frames, 13.0, h, 15.0, 100.0, 0.5, 1.0, 0.0, 0.0, 1.0, 500.0)
    st_l, pk_l, val_l, n = bsfit.find_rois(frames, 13, h, 15.0, 100.0, 0.5, subtract_min=True, capacity=500)
    assert st.shape == (13, 13, n) and pc.shape == (n, 3) and n > 0
    np.testing.assert_array_equal(st, st_l)
    np.testing.assert_array_equal(pc, pk_l.astype(np.float64))
    np.testing.assert_array_equal(val.ravel(), val_l)


def _nearest(truth_xy, frame_of_truth, found_xy, frame_of_found):
    d = np.full(len(truth_xy), np.inf)
    idx = np.full(len(truth_xy), -1)
    for k, ((x, y), f) in enumerate(zip(truth_xy, frame_of_truth)):
        m = np.nonzero(frame_of_found == f)[0]
        if len(m):
            dd = np.hypot(found_xy[m, 0] - x, found_xy[m, 1] - y)
            j = int(np.argmin(dd))
            d[k], idx[k] = dd[j], m[j]
    return d, idx


def test_gui_path_on_colour_frames(mexdir):
    """The 4D GUI path: x = P(:,1) - dx + x_peak (the x/y exchange), y from
    P(:,2), z = -(P(:,5) - z0) dz, column 5 the colour coordinate."""
    find, fitm, table = (Mex(n, mexdir) for n in ("bsf_find_rois", "GPUmleFit_LM_multiD", "bsf_gui_results"))
    coeff = gen.model_coeffs("color4d")
    frames, truth = gen.make_frames(4, 160, 160, 12, 77, coeff=coeff)
    h = bsfit.peak_filter("dog", 1.0)
    roi, dx = 13, 6
    st, pc, _ = find(3, frames, float(roi), h, 15.0, 100.0, 0.5, 1.0)
    nz, nw = coeff.shape[2] - 3, coeff.shape[3] - 3
    P, CR, LL = fitm(3, st, 7.0, 50.0, coeff, 0.0, 1.0, np.array([nz / 2, nw / 2], np.float32))
    z0, dz = 31.0, 20.0                       # spline coordinate of z = 0 nm and nm per plane (generator)
    R = table(1, P, CR, LL.reshape(-1, 1), pc, 7.0, 0.0, float(dx), z0, dz)[0]
    assert R.shape == (len(pc), 17)
    d, idx = _nearest(truth[:, 1:3] + 1, truth[:, 0] + 1, R[:, 1:3], R[:, 0])
    ok = d < 0.3
    assert ok.mean() >= 0.7, ok.mean()
    z_nm = (truth[:, 3] - 31.0) * 20.0                                      # generator's z in nm
    dzerr = np.abs(R[idx[ok], 3] - (-z_nm[ok]))                             # the GUI's z is -(z - z0) dz
    assert np.median(dzerr) < 30.0, np.median(dzerr)
    np.testing.assert_array_equal(R[:, 4], P[:, 5])                         # column 5: the colour coordinate
    inside = ok & (truth[:, 6] < 2.4)                                       # the fitter bounds w <= nw - 0.5
    # the colour channels differ by 10 % in PSF width: a noisy estimate, but it follows the truth
    assert np.corrcoef(R[idx[inside], 4], truth[inside, 6])[0, 1] > 0.5


def test_gui_path_on_vortex_rois(mexdir):
    """The 5D table: x, y and z = (P(:,5) - z0) dz (no sign flip) for ROIs
    placed at known peaks."""
    fitm, table = Mex("GPUmleFit_LM_multiD", mexdir), Mex("bsf_gui_results", mexdir)
    w = gen.config(3, N=256, S=13)[0]
    c5, c3 = gen.model_coeffs("dipole5d")
    P, CR, LL = fitm(3, w.imstack, 8.0, 50.0, c5, 0.0, 1.0, w.initP, c3)
    rng = np.random.default_rng(4)
    dx = 6
    pk = np.stack([rng.integers(20, 400, w.N), rng.integers(20, 400, w.N), np.arange(w.N) + 1], 1).astype(float)
    z0, dz = 11.0, 60.0                       # the generator's 5D z axis: 60 nm planes, z = 0 at coordinate 11
    R = table(1, P, CR, LL.reshape(-1, 1), pk, 8.0, 0.0, float(dx), z0, dz, 1.0, 1.0)[0]
    assert R.shape == (w.N, 19)
    tx = w.truth[:, 0] - dx + pk[:, 0]
    ty = w.truth[:, 1] - dx + pk[:, 1]
    # single-start 5D fits land in local minima now and then: medians; on
    # the ROIs whose emitter sits > 2 px off the diagonal, the reading with
    # x and y swapped is far worse
    d = np.hypot(R[:, 1] - tx, R[:, 2] - ty)
    off = np.abs(w.truth[:, 0] - w.truth[:, 1]) > 2.0
    d_swapped = np.hypot(R[:, 1] - (w.truth[:, 1] - dx + pk[:, 0]), R[:, 2] - (w.truth[:, 0] - dx + pk[:, 1]))
    assert off.sum() >= 10
    assert np.median(d) < 0.5 and np.median(d_swapped[off]) > 2.0, (np.median(d), np.median(d_swapped[off]))
    z_nm = (w.truth[:, 4] - z0) * dz
    ez, ez_mirrored = np.median(np.abs(R[:, 3] - z_nm)), np.median(np.abs(R[:, 3] + z_nm))
    assert ez < 80.0 and ez_mirrored > 2 * ez, (ez, ez_mirrored)
