/*
 * Test infrastructure: the MX API of mex.h over plain memory, linked with
 * one MEX gateway into a test library.  mexErrMsgIdAndTxt unwinds to
 * mock_call (as MATLAB unwinds to the prompt) and keeps the id and message.
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

struct mxArray_tag {
    mxClassID cls;
    mwSize ndim;
    mwSize dims[8];
    void *data;
};

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

static jmp_buf g_env;
static char g_id[256], g_msg[1024];

static size_t esize(mxClassID c)
{
    switch (c) {
    case mxDOUBLE_CLASS: return 8;
    case mxSINGLE_CLASS: return 4;
    case mxUINT16_CLASS: return 2;
    case mxUINT32_CLASS: return 4;
    default: return 1;
    }
}

mxArray *mock_new(int cls, int ndim, const size_t *dims)
{
    mxArray *a = (mxArray *)calloc(1, sizeof *a);
    size_t n = 1;
    a->cls = (mxClassID)cls;
    a->ndim = (mwSize)(ndim < 2 ? 2 : ndim);
    a->dims[0] = a->dims[1] = 1;
    for (int k = 0; k < ndim; k++) {
        a->dims[k] = dims[k];
        n *= dims[k];
    }
    a->data = calloc(n ? n : 1, esize(a->cls));
    return a;
}

void *mock_data(mxArray *a) { return a->data; }
void mock_free(mxArray *a)
{
    if (a) {
        free(a->data);
        free(a);
    }
}
const char *mock_error_id(void) { return g_id; }
const char *mock_error_msg(void) { return g_msg; }

int mock_call(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs)
{
    g_id[0] = g_msg[0] = 0;
    if (setjmp(g_env)) return 1;
    mexFunction(nlhs, plhs, nrhs, prhs);
    return 0;
}

mxClassID mxGetClassID(const mxArray *a) { return a->cls; }
mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->ndim; }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims; }
size_t mxGetNumberOfElements(const mxArray *a)
{
    size_t n = 1;
    for (mwSize k = 0; k < a->ndim; k++) n *= a->dims[k];
    return n;
}
size_t mxGetM(const mxArray *a) { return a->dims[0]; }
size_t mxGetN(const mxArray *a) { return mxGetNumberOfElements(a) / (a->dims[0] ? a->dims[0] : 1); }
void *mxGetData(const mxArray *a) { return a->data; }

double mxGetScalar(const mxArray *a)
{
    if (!mxGetNumberOfElements(a)) return 0.0;
    switch (a->cls) {
    case mxDOUBLE_CLASS: return ((double *)a->data)[0];
    case mxSINGLE_CLASS: return ((float *)a->data)[0];
    case mxUINT8_CLASS: return ((uint8_t *)a->data)[0];
    case mxUINT16_CLASS: return ((uint16_t *)a->data)[0];
    case mxUINT32_CLASS: return ((uint32_t *)a->data)[0];
    default: return 0.0;
    }
}

mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID c, mxComplexity f)
{
    size_t d[2] = {m, n};
    (void)f;
    return mock_new(c, 2, d);
}

mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID c, mxComplexity f)
{
    (void)f;
    return mock_new(c, (int)ndim, dims);
}

void mxDestroyArray(mxArray *a) { mock_free(a); }

void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...)
{
    va_list ap;
    snprintf(g_id, sizeof g_id, "%s", id);
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof g_msg, fmt, ap);
    va_end(ap);
    longjmp(g_env, 1);
}

int mexPrintf(const char *fmt, ...)
{
    va_list ap;
    int n;
    va_start(ap, fmt);
    n = vprintf(fmt, ap);
    va_end(ap);
    return n;
}
