"""CPU oracle for the B-spline PSF MLE fitter -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker: the shipped library in
``gpu-based-b-spline-fitter_amd/`` never calls it.

``liborc.so`` (built by ``oracle/Makefile`` from ``bsf_oracle.c``) holds two
restatements of the reference's CPU fitter
(``MultiD_bspline/visual code/CPUmleFit_LM/CPUmleFit_LM.cpp:2411-3620``):
the float "parity" build (``fit``) and the all-double build (``fit_f64``),
and the frame -> ROI front-end of the GUI fitter (``find_rois`` and its
steps ``photons``, ``filter2``, ``find_maxima``; ``bsf_frontend.c``).

``_ref/`` (built by ``make ref`` where the reference tree exists) holds the
reference's own ``CPUsplineLib.cpp`` and ``MatInvLib.h`` compiled from source;
``ref_*`` functions call them so the oracle's primitives can be pinned bit for
bit.  The reference's fitter translation unit includes MATLAB's ``mex.h`` and
is not buildable here.
"""
from .loader import (  # noqa: F401
    available, build, fit, fit_f64, bspline, delta_tables, deriv_eval,
    cholesky, lu_eval, matinv, com2d, maxmin2d, eval_pass, fit_4pi, nv_4pi,
    ref_available, ref_bspline, ref_delta_tables, ref_deriv_eval,
    ref_cholesky, ref_lu_eval, ref_matinv, NV_OF, D_OF,
    photons, filter2, find_maxima, find_rois, aggregate, planted,
)
