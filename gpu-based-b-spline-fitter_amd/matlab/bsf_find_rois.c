/*
 * bsf_find_rois.c -- MATLAB MEX gateway of the GPU frame -> ROI front-end,
 * the GUI fitter's per-frame loop (MultiD_fitter/simplefitter_spline.m:
 * 197-307: photon conversion, filter2, maximumfind2, cutoff and border
 * tests, ROI cut):
 *
 *   [imstack, peakcoordinates, value] = bsf_find_rois(frames, roifit, h,
 *       peakcutoff, offset, conversion, subtract_min, mirror, psf_size,
 *       first_frame, capacity, agg_dist)
 *
 * frames: uint16 or single [nrows, ncols, nframes]; h: single filter kernel
 * (fspecial-style, e.g. the GUI's hgauss / hdog); subtract_min: 1 for the
 * DoG and PSF background modes; psf_size: size(cal.PSF, 1) in the p.conv
 * mode, else 0; agg_dist: p.roifit * p.aggsz when p.agg (maxima closer
 * than it are merged as aggre merges them), else 0.  Outputs: imstack single [roifit, roifit, n], peakcoordinates
 * double [n, 3] (x, y, frame), value single [n, 1]; n = min(found, capacity).
 *
 * Build (MATLAB):  mex -I<repo>/include bsf_find_rois.c \
 *                      -L<repo>/gpu-based-b-spline-fitter_amd -lbsfit
 */
#include <stdlib.h>
#include <string.h>

#include "mex.h"
#include "bsfit.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    bsf_roi_request q;
    bsf_roi_result r;
    const mwSize *d;
    mwSize nd, dims[3];
    int cap, n, k, c, rc;
    int *peaks;
    float *stack, *value;
    double *pc;
    if (nrhs < 6)
        mexErrMsgIdAndTxt("bsf_find_rois:WrongNoArgs",
                          "Inputs: frames, roifit, h, peakcutoff, offset, conversion[, subtract_min, mirror, "
                          "psf_size, first_frame, capacity, agg_dist]");
    memset(&q, 0, sizeof q);
    if (mxGetClassID(prhs[0]) == mxUINT16_CLASS)
        q.pixel_type = BSF_PIXEL_U16;
    else if (mxGetClassID(prhs[0]) == mxSINGLE_CLASS)
        q.pixel_type = BSF_PIXEL_F32;
    else
        mexErrMsgIdAndTxt("bsf_find_rois:WrongDataType", "frames must be uint16 or single");
    if (mxGetClassID(prhs[2]) != mxSINGLE_CLASS || mxGetNumberOfDimensions(prhs[2]) != 2)
        mexErrMsgIdAndTxt("bsf_find_rois:WrongDataType", "h must be a single 2-D kernel");
    nd = mxGetNumberOfDimensions(prhs[0]);
    d = mxGetDimensions(prhs[0]);
    q.frames = mxGetData(prhs[0]);
    q.nrows = (int)d[0];
    q.ncols = (int)d[1];
    q.nframes = nd > 2 ? (int)d[2] : 1;
    q.roi = (int)mxGetScalar(prhs[1]);
    q.filter = (const float *)mxGetData(prhs[2]);
    q.filter_rows = (int)mxGetDimensions(prhs[2])[0];
    q.filter_cols = (int)mxGetDimensions(prhs[2])[1];
    q.cutoff = mxGetScalar(prhs[3]);
    q.offset = (float)mxGetScalar(prhs[4]);
    q.conversion = (float)mxGetScalar(prhs[5]);
    q.subtract_min = nrhs > 6 ? (int)mxGetScalar(prhs[6]) : 0;
    q.mirror = nrhs > 7 ? (int)mxGetScalar(prhs[7]) : 0;
    q.psf_size = nrhs > 8 ? mxGetScalar(prhs[8]) : 0.0;
    q.agg_dist = nrhs > 11 ? mxGetScalar(prhs[11]) : 0.0;
    q.first_frame = nrhs > 9 ? (int)mxGetScalar(prhs[9]) : 1;
    cap = nrhs > 10 ? (int)mxGetScalar(prhs[10]) : q.nframes * 2000;
    q.capacity = cap;
    if (q.roi < 1 || cap < 0) mexErrMsgIdAndTxt("bsf_find_rois:Args", "bad roifit or capacity");
    stack = (float *)malloc(sizeof(float) * (size_t)q.roi * q.roi * (size_t)(cap ? cap : 1));
    peaks = (int *)malloc(sizeof(int) * 3 * (size_t)(cap ? cap : 1));
    value = (float *)malloc(sizeof(float) * (size_t)(cap ? cap : 1));
    if (!stack || !peaks || !value) mexErrMsgIdAndTxt("bsf_find_rois:NoMemory", "out of host memory");
    r.imstack = stack;
    r.peaks = peaks;
    r.value = value;
    r.count = 0;
    rc = bsf_find_rois_host(&q, &r);
    if (rc) {
        free(stack);
        free(peaks);
        free(value);
        mexErrMsgIdAndTxt("bsf_find_rois:Run", "%s", bsf_last_error());
    }
    n = r.count < cap ? r.count : cap;
    dims[0] = (mwSize)q.roi;
    dims[1] = (mwSize)q.roi;
    dims[2] = (mwSize)n;
    /* plhs holds max(nlhs, 1) outputs */
    plhs[0] = mxCreateNumericArray(3, dims, mxSINGLE_CLASS, mxREAL);
    memcpy(mxGetData(plhs[0]), stack, sizeof(float) * (size_t)q.roi * q.roi * (size_t)n);
    if (nlhs > 1) {
        plhs[1] = mxCreateNumericMatrix((mwSize)n, 3, mxDOUBLE_CLASS, mxREAL);
        pc = (double *)mxGetData(plhs[1]);
        for (c = 0; c < 3; c++)
            for (k = 0; k < n; k++) pc[(size_t)c * n + k] = peaks[(size_t)c * cap + k];
    }
    if (nlhs > 2) {
        plhs[2] = mxCreateNumericMatrix((mwSize)n, 1, mxSINGLE_CLASS, mxREAL);
        memcpy(mxGetData(plhs[2]), value, sizeof(float) * (size_t)n);
    }
    free(stack);
    free(peaks);
    free(value);
}
