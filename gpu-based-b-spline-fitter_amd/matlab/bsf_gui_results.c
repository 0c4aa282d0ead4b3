/*
 * bsf_gui_results.c -- MATLAB MEX gateway of the GUI's localisation table
 * (MultiD_fitter/simplefitter_spline.m:550-630):
 *
 *   results = bsf_gui_results(P, CRLB, LogL, peakcoordinates, fittype, mirror,
 *                             dx, z0, dz, du, dv)
 *
 * P, CRLB, LogL: single, as GPUmleFit_LM_multiD returns them;
 * peakcoordinates: double [N, 3]; the rest are the GUI's p.* scalars
 * (du, dv used by fittype 8 only).  results: double [N, 12 | 17 | 19].
 *
 * Build (MATLAB):  mex -I<repo>/include bsf_gui_results.c \
 *                      -L<repo>/gpu-based-b-spline-fitter_amd -lbsfit
 */
#include "mex.h"
#include "bsfit.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    int fittype, cols, rc, a;
    size_t N;
    double s[5];
    (void)nlhs;
    if (nrhs < 9) mexErrMsgIdAndTxt("bsf_gui_results:WrongNoArgs",
                                    "Inputs: P, CRLB, LogL, peakcoordinates, fittype, mirror, dx, z0, dz[, du, dv]");
    if (mxGetClassID(prhs[0]) != mxSINGLE_CLASS || mxGetClassID(prhs[1]) != mxSINGLE_CLASS ||
        mxGetClassID(prhs[2]) != mxSINGLE_CLASS || mxGetClassID(prhs[3]) != mxDOUBLE_CLASS)
        mexErrMsgIdAndTxt("bsf_gui_results:WrongDataType", "P, CRLB, LogL single; peakcoordinates double");
    fittype = (int)mxGetScalar(prhs[4]);
    cols = bsf_gui_results_cols(fittype);
    if (cols < 0) mexErrMsgIdAndTxt("bsf_gui_results:FitType", "fittype must be 6, 7 or 8");
    N = mxGetM(prhs[0]);
    if (mxGetM(prhs[3]) != N || mxGetN(prhs[3]) != 3 || mxGetN(prhs[0]) != (size_t)bsf_num_params(fittype) + 1 ||
        mxGetNumberOfElements(prhs[2]) != N || mxGetM(prhs[1]) != N)
        mexErrMsgIdAndTxt("bsf_gui_results:WrongDimensions", "inconsistent sizes");
    for (a = 0; a < 5; a++) s[a] = nrhs > 6 + a ? mxGetScalar(prhs[6 + a]) : 1.0;
    plhs[0] = mxCreateNumericMatrix(N, (mwSize)cols, mxDOUBLE_CLASS, mxREAL);
    rc = bsf_gui_results(fittype, (int)mxGetScalar(prhs[5]), s[0], s[1], s[2], s[3], s[4],
                         (const float *)mxGetData(prhs[0]), (const float *)mxGetData(prhs[1]),
                         (const float *)mxGetData(prhs[2]), (const double *)mxGetData(prhs[3]), (int)N,
                         (double *)mxGetData(plhs[0]));
    if (rc) mexErrMsgIdAndTxt("bsf_gui_results:Run", "%s", bsf_last_error());
}
