/*
 * GPUmleFit_LM_multiD.c -- MATLAB MEX gateway of the MI355X fitter with the
 * reference MEX's name and call surface
 * (MultiD_bspline/visual code/GPUmleFit_LM/mexFunction.cpp:14-55, 208):
 *
 *   [P, CRLB, LogL] = GPUmleFit_LM_multiD(imstack, fittype, iterations, coeff,
 *                                         varmap, silent, initP, coeff3)
 *
 * for the B-spline fittypes 6, 7 and 8.  An optional fourth output, which the
 * reference MEX does not have, returns the per-ROI status flags (uint8, N x 1,
 * bits of enum bsf_roi_flag).  Argument checks follow the
 * reference MEX (:247-385): single-precision square data, ROI size <= 51,
 * iterations default 50, varmap ignored for these fittypes, initP and coeff3
 * read when given.  Everything else is bsf_GPUmleFit_LM_multiD.
 *
 * Build (MATLAB):  mex -I<repo>/include GPUmleFit_LM_multiD.c \
 *                      -L<repo>/gpu-based-b-spline-fitter_amd -lbsfit
 */
#include "mex.h"
#include "bsfit.h"

static const float *single_arg(const mxArray *a, const char *what)
{
    if (mxGetClassID(a) != mxSINGLE_CLASS)
        mexErrMsgIdAndTxt("GPUmleFit_LM:WrongDataType", "%s must be comprised of single floats!", what);
    return (const float *)mxGetData(a);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    size_t imdims[3], cdims[5], N, ninit = 0;
    int imnd, cnd, fittype, iterations, silent, NV, rc, a;
    const mwSize *d;
    const float *data, *coeff, *initP = NULL, *coeff3 = NULL;
    mxArray *out[4];
    if (nlhs > 4) mexErrMsgIdAndTxt("GPUmleFit_LM:WrongNoArgs", "Outputs: P, CRLB, LogL[, status]");
    if (nrhs < 4)
        mexErrMsgIdAndTxt("GPUmleFit_LM:WrongNoArgs",
                          "Inputs: imstack, fittype, iterations, coeff[, varmap, silent, initP, coeff3]");
    data = single_arg(prhs[0], "Data");
    coeff = single_arg(prhs[3], "coeff");
    imnd = (int)mxGetNumberOfDimensions(prhs[0]);
    if (imnd < 2 || imnd > 3)
        mexErrMsgIdAndTxt("GPUmleFit_LM:WrongDimensions", "Data must be a 2-D or 3-D array");
    d = mxGetDimensions(prhs[0]);
    imdims[0] = d[0];
    imdims[1] = d[1];
    imdims[2] = imnd > 2 ? d[2] : 1;
    fittype = (int)mxGetScalar(prhs[1]);
    iterations = mxGetNumberOfElements(prhs[2]) ? (int)mxGetScalar(prhs[2]) : 50;
    cnd = (int)mxGetNumberOfDimensions(prhs[3]);
    if (cnd > 5) mexErrMsgIdAndTxt("GPUmleFit_LM:WrongDimensions", "coeff has more than 5 dimensions");
    d = mxGetDimensions(prhs[3]);
    for (a = 0; a < cnd; a++) cdims[a] = d[a];
    silent = nrhs > 5 && mxGetNumberOfElements(prhs[5]) ? (int)mxGetScalar(prhs[5]) : 0;
    if (nrhs > 6 && mxGetNumberOfElements(prhs[6])) {
        initP = single_arg(prhs[6], "initP");
        ninit = mxGetNumberOfElements(prhs[6]);
    }
    if (nrhs > 7 && mxGetNumberOfElements(prhs[7])) coeff3 = single_arg(prhs[7], "coeff3");
    NV = bsf_num_params(fittype);
    if (NV < 0) mexErrMsgIdAndTxt("GPUmleFit_LM:FitType", "fittype %d: this build fits 6, 7 and 8", fittype);
    N = imdims[2];
    out[0] = mxCreateNumericMatrix(N, (mwSize)NV + 1, mxSINGLE_CLASS, mxREAL);
    out[1] = mxCreateNumericMatrix(N, (mwSize)NV, mxSINGLE_CLASS, mxREAL);
    out[2] = mxCreateNumericMatrix(N, 1, mxSINGLE_CLASS, mxREAL);
    out[3] = nlhs > 3 ? mxCreateNumericMatrix(N, 1, mxUINT8_CLASS, mxREAL) : NULL;
    if (out[3])
        rc = bsf_GPUmleFit_LM_multiD_status(data, imdims, imnd > 2 ? 3 : 2, fittype, iterations, coeff, cdims, cnd,
                                            NULL, 0, silent, initP, ninit, coeff3, (float *)mxGetData(out[0]),
                                            (float *)mxGetData(out[1]), (float *)mxGetData(out[2]),
                                            (unsigned char *)mxGetData(out[3]));
    else
        rc = bsf_GPUmleFit_LM_multiD(data, imdims, imnd > 2 ? 3 : 2, fittype, iterations, coeff, cdims, cnd, NULL, 0,
                                     silent, initP, ninit, coeff3, (float *)mxGetData(out[0]),
                                     (float *)mxGetData(out[1]), (float *)mxGetData(out[2]));
    if (rc) {
        for (a = 0; a < 4; a++)
            if (out[a]) mxDestroyArray(out[a]);
        mexErrMsgIdAndTxt("GPUmleFit_LM:Fit", "%s", bsf_last_error());
    }
    /* plhs holds max(nlhs, 1) outputs */
    for (a = 0; a < 4; a++) {
        if (!out[a]) continue;
        if (a < (nlhs > 0 ? nlhs : 1))
            plhs[a] = out[a];
        else
            mxDestroyArray(out[a]);
    }
}
