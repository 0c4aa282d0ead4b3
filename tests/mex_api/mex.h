/*
 * Test infrastructure: declarations of the subset of MATLAB's documented C
 * MEX API (mex.h / matrix.h) that the gateways in
 * gpu-based-b-spline-fitter_amd/matlab/ use, so they can be compiled and
 * called without MATLAB (mock_mx.c implements them over plain memory).
 * Class IDs carry MATLAB's documented values.
 */
#ifndef BSF_TEST_MEX_API_H
#define BSF_TEST_MEX_API_H

#include <stddef.h>

typedef size_t mwSize;
typedef struct mxArray_tag mxArray;

typedef enum {
    mxUNKNOWN_CLASS = 0,
    mxDOUBLE_CLASS = 6,
    mxSINGLE_CLASS = 7,
    mxUINT8_CLASS = 9,
    mxUINT16_CLASS = 11,
    mxUINT32_CLASS = 13
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

mxClassID mxGetClassID(const mxArray *pm);
mwSize mxGetNumberOfDimensions(const mxArray *pm);
const mwSize *mxGetDimensions(const mxArray *pm);
size_t mxGetNumberOfElements(const mxArray *pm);
size_t mxGetM(const mxArray *pm);
size_t mxGetN(const mxArray *pm);
double mxGetScalar(const mxArray *pm);
void *mxGetData(const mxArray *pm);
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID classid, mxComplexity flag);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID classid, mxComplexity flag);
void mxDestroyArray(mxArray *pm);
void mexErrMsgIdAndTxt(const char *errorid, const char *errormsg, ...);
int mexPrintf(const char *message, ...);

#endif
