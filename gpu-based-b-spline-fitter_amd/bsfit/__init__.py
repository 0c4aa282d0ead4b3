"""bsfit -- MI355X B-spline PSF MLE fitter, Python side of the C-ABI.

``GPUmleFit_LM_multiD`` mirrors the reference's MATLAB entry point
(``MultiD_bspline/visual code/GPUmleFit_LM/mexFunction.cpp:14-55``):

    P, CRLB, LogL = GPUmleFit_LM_multiD(imstack, fittype, iterations, coeff,
                                        varmap=None, silent=1, initP=None, coeff3=None)

with ``imstack`` (S, S, N) and ``coeff`` in MATLAB (Fortran) order.  All work
runs in ``libbsfit.so`` (gfx950 HIP kernels); there is no CPU path: if the
library or a gfx950 device is missing, calls raise ``BsfError``.
"""
from .api import (  # noqa: F401
    BsfError, GPUmleFit_LM_multiD, GPUmleFit_LM_4Pi_bspline, coeff_build, fit, fit_4pi, fit_4pi_device,
    fit_device, fit_multistart, fit_tile, roi_table, set_fit_tile, tune,
    fit_status, ROI_MAX_ITER, ROI_SOLVE_FAILED, ROI_FISHER_SINGULAR, ROI_NONFINITE,
    ROI_FIRST_SOLVE_FAILED,
    lib_path, load, num_params, num_params_4pi, PREC_MIX, PREC_F64,
)
from .frames import find_rois, peak_filter, peak_filter_psf  # noqa: F401
from .gui import color_class, gui_results  # noqa: F401
from .session import FT_4PI, Group, Session  # noqa: F401
