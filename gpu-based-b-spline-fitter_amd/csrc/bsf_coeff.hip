// bsf_coeff.hip -- cubic B-spline coefficients of an N-D sample array on
// gfx950: the reference's bsarray(a, 'lambda', lambda) with degree 3 on every
// axis (MultiD_bspline/bsarray/bsarray.m:58-173, directFilter.m:1-297).
//
// Per axis, per pole z of that axis (getBSplineFiltCoeffs.m: sqrt(3)-2 for
// lambda = 0; for lambda > 0 the two roots inside the unit circle of
// z^4 + a z^3 + b z^2 + a z + 1, a = -4 + 1/(6 lambda), b = 6 + 2/(3 lambda)),
// every 1-D line along the axis runs the symmetric exponential filter of
// symExpFilt.m:26-45 (K0-term causal start over the whole-sample mirror
// extension, causal pass, anticausal start, anticausal pass); after the
// axis's poles the real part is scaled by the gain (6, or 1/lambda).  A last
// pass mirror-pads every axis from M to M+3 samples and rounds to FP32.
// Everything runs in FP64 (the reference filters in double), complex where
// a pole is.  One launch per axis runs all of its poles and the gain, one
// thread per line (lines along axis d with stride prod(dims[0..d)),
// consecutive threads on consecutive lines); the first also converts the
// FP32 samples, and one more launch pads and rounds: ndim + 1 launches.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>

#include "bsf_kernels.h"

namespace bsf {

namespace {

struct cplx {
    double re, im;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }

// whole-sample reflection of index k into [0, M)
__device__ __forceinline__ long reflect(long k, long M)
{
    const long period = 2 * M - 2;
    k %= period;
    if (k < 0) k += period;
    return k < M ? k : period - k;
}

// One axis (M >= 2 samples): every pole of the axis in turn along one line,
// then the real part times the gain.  Line l = (outer, inner), elements at
// base + k * stride, base = outer * M * stride + inner; the first axis reads
// the FP32 samples itself (x != NULL).
struct AxisPoles {
    int n;
    cplx z[2], c0[2];
    int K0[2];
    double gain;
};

__global__ void axis_kernel(const float *__restrict__ x, cplx *__restrict__ w, long M, long stride, long nlines,
                            AxisPoles P)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    const long outer = l / stride, inner = l - outer * stride;
    const long base = outer * M * stride + inner;
    cplx *p = w + base;
    if (x) {
        for (long k = 0; k < M; k++) p[k * stride] = {(double)x[base + k * stride], 0.0};
    }
    for (int i = 0; i < P.n; i++) {
        const cplx z = P.z[i];
        // causal start: sum_k z^k x[reflect(k)]
        cplx acc = {0.0, 0.0}, zk = {1.0, 0.0};
        for (int k = 0; k < P.K0[i]; k++) {
            acc = cadd(acc, cmul(zk, p[reflect(k, M) * stride]));
            zk = cmul(zk, z);
        }
        const cplx xlast = p[(M - 1) * stride];   // the anticausal start needs it
        p[0] = acc;
        cplx prev = acc;
        for (long k = 1; k < M; k++) {
            const cplx y = cadd(p[k * stride], cmul(z, prev));
            p[k * stride] = y;
            prev = y;
        }
        cplx y = cmul(csub(cadd(prev, prev), xlast), P.c0[i]);   // (2 y[M-1] - x[M-1]) c0
        p[(M - 1) * stride] = y;
        for (long k = M - 2; k >= 0; k--) {
            y = cmul(csub(y, p[k * stride]), z);
            p[k * stride] = y;
        }
    }
    for (long k = 0; k < M; k++) p[k * stride] = {p[k * stride].re * P.gain, 0.0};
}

struct PadDims {
    int ndim;
    long in[kMaxCoeffDims];    // sample dims
    long out[kMaxCoeffDims];   // padded dims (+3)
};

__global__ void pad_kernel(const cplx *__restrict__ w, float *__restrict__ c, long nout, PadDims D)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nout) return;
    long r = i, src = 0, sstride = 1;
    for (int a = 0; a < D.ndim; a++) {
        const long q = r / D.out[a], k = r - q * D.out[a];
        r = q;
        src += reflect(k - 1, D.in[a]) * sstride;
        sstride *= D.in[a];
    }
    c[i] = (float)w[src].re;
}

}  // namespace

// Poles inside the unit circle and gain of the cubic (smoothing) prefilter.
int coeff_poles(double lambda, std::complex<double> *z, double *gain)
{
    if (lambda == 0.0) {
        z[0] = std::sqrt(3.0) - 2.0;
        *gain = 6.0;
        return 1;
    }
    // p(z)/z^2 = (w^2 - 2) + a w + b with w = z + 1/z
    const double a = -4.0 + 1.0 / (6.0 * lambda), b = 6.0 + 2.0 / (3.0 * lambda);
    const std::complex<double> disc = std::sqrt(std::complex<double>(a * a - 4.0 * (b - 2.0), 0.0));
    int n = 0;
    for (int s = -1; s <= 1; s += 2) {
        const std::complex<double> w = (-a + (double)s * disc) / 2.0;
        const std::complex<double> r = std::sqrt(w * w - 4.0);
        std::complex<double> z1 = (w + r) / 2.0, z2 = (w - r) / 2.0;
        z[n++] = std::abs(z1) <= 1.0 ? z1 : z2;
    }
    *gain = 1.0 / lambda;
    return n;
}

size_t coeff_work_bytes(const int *dims, int ndim)
{
    size_t n = 1;
    for (int a = 0; a < ndim; a++) n *= (size_t)dims[a];
    return n * sizeof(cplx);
}

hipError_t coeff_build(const float *samples, const int *dims, int ndim, const double *lambda, float *coeff,
                       void *work, hipStream_t st)
{
    cplx *w = (cplx *)work;
    long n = 1, nout = 1;
    PadDims D;
    D.ndim = ndim;
    for (int a = 0; a < ndim; a++) {
        n *= dims[a];
        nout *= dims[a] + 3;
        D.in[a] = dims[a];
        D.out[a] = dims[a] + 3;
    }
    const int TB = 256;
    hipError_t e = hipSuccess;
    long stride = 1;
    for (int a = 0; a < ndim && e == hipSuccess; a++) {
        const long M = dims[a], nlines = n / M;
        std::complex<double> zs[2];
        AxisPoles P;
        P.n = coeff_poles(lambda ? lambda[a] : 0.0, zs, &P.gain);
        for (int i = 0; i < P.n; i++) {
            const std::complex<double> z = zs[i], c0 = -z / (1.0 - z * z);
            P.z[i] = {z.real(), z.imag()};
            P.c0[i] = {c0.real(), c0.imag()};
            P.K0[i] = (int)std::ceil(std::log(2.220446049250313e-16) / std::log(std::abs(z)));
        }
        hipLaunchKernelGGL(axis_kernel, dim3((unsigned)((nlines + TB - 1) / TB)), dim3(TB), 0, st,
                           a == 0 ? samples : nullptr, w, M, stride, nlines, P);
        e = hipGetLastError();
        stride *= M;
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(pad_kernel, dim3((unsigned)((nout + TB - 1) / TB)), dim3(TB), 0, st, w, coeff, nout, D);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace bsf
