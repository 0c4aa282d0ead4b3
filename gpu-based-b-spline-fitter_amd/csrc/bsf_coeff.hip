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
// a pole is.  One launch per axis runs all of its poles and the gain: lines
// of up to 48 samples are staged 64 at a time in one wave's LDS (read once,
// filtered there, the real results written once); longer lines are filtered
// in place in global memory, one thread per line.  The first axis converts
// the FP32 samples; between axes the work array holds real FP64 values.
// The last axis, when staged in LDS, writes the mirror-padded FP32 output
// itself (ndim launches); otherwise one more launch pads and rounds.
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

struct AxisPoles {
    int n;
    bool cplx_poles;         // a complex pair (lambda > 0 with strong smoothing)
    cplx z[2], c0[2];
    int K0[2];
    double gain;
};

constexpr int kLB = 64;      // lines per workgroup (one wave) of the LDS path
constexpr int kLdsM = 48;    // lines up to this length are staged in LDS

// Every pole of the axis along one line held by an accessor (get/set of
// element k), then the real part times the gain: the causal start over the
// K0-term mirror extension, the causal pass, the anticausal start and pass.
// The next element is read before the current one is stored, so the loads
// of the recursion do not wait on its stores.
template <bool CPLX, class Line>
__device__ __forceinline__ void filter_line(Line &L, long M, const AxisPoles &P)
{
    for (int i = 0; i < P.n; i++) {
        const cplx z = P.z[i];
        cplx acc = {0.0, 0.0}, zk = {1.0, 0.0};
        for (int k = 0; k < P.K0[i]; k++) {
            const cplx v = L.get(reflect(k, M));
            acc = CPLX ? cadd(acc, cmul(zk, v)) : cplx{acc.re + zk.re * v.re, 0.0};
            zk = CPLX ? cmul(zk, z) : cplx{zk.re * z.re, 0.0};
        }
        const cplx xlast = L.get(M - 1);
        cplx xk = L.get(1);
        L.set(0, acc);
        cplx prev = acc;
        for (long k = 1; k < M; k++) {
            const cplx xn = L.get(k + 1 < M ? k + 1 : k);
            const cplx y = CPLX ? cadd(xk, cmul(z, prev)) : cplx{xk.re + z.re * prev.re, 0.0};
            L.set(k, y);
            prev = y;
            xk = xn;
        }
        cplx y = CPLX ? cmul(csub(cadd(prev, prev), xlast), P.c0[i])
                      : cplx{((prev.re + prev.re) - xlast.re) * P.c0[i].re, 0.0};
        L.set(M - 1, y);
        xk = L.get(M - 2);
        for (long k = M - 2; k >= 0; k--) {
            const cplx xn = L.get(k > 0 ? k - 1 : 0);
            y = CPLX ? cmul(csub(y, xk), z) : cplx{(y.re - xk.re) * z.re, 0.0};
            L.set(k, y);
            xk = xn;
        }
    }
    for (long k = 0; k < M; k++) L.set(k, {L.get(k).re * P.gain, 0.0});
}

template <bool CPLX>
struct LdsLine {
    double *s;               // element k at s[k * kLB * W]
    __device__ cplx get(long k) const
    {
        return CPLX ? cplx{s[k * kLB * 2], s[k * kLB * 2 + 1]} : cplx{s[k * kLB], 0.0};
    }
    __device__ void set(long k, cplx v) const
    {
        if (CPLX) {
            s[k * kLB * 2] = v.re;
            s[k * kLB * 2 + 1] = v.im;
        } else {
            s[k * kLB] = v.re;
        }
    }
};

template <bool CPLX>
struct GlobalLine {
    double *re, *im;         // split complex, element k at [k * stride]
    long stride;
    __device__ cplx get(long k) const { return {re[k * stride], CPLX ? im[k * stride] : 0.0}; }
    __device__ void set(long k, cplx v) const
    {
        re[k * stride] = v.re;
        if (CPLX) im[k * stride] = v.im;
    }
};

struct PadDims {
    int ndim;
    long in[kMaxCoeffDims];    // sample dims
    long out[kMaxCoeffDims];   // padded dims (+3)
};

constexpr int kStage = 24;   // global loads a thread keeps in flight while staging

// Lines of up to kLdsM samples: a workgroup stages its kLB lines in LDS
// (element k of its line t at [k][t]), filters them there and writes the
// real results back.  Lines along the first axis (stride 1) lie one after
// another, so the workgroup's block is read and written as one contiguous
// run; along the other axes element k of neighbouring lines is contiguous.
// The first axis reads the FP32 samples (x32), the others the previous
// axis's result in R.  Loads are issued kStage at a time from clamped
// addresses, so they are in flight together.  On the last axis (coeff !=
// NULL) the results go straight to the padded FP32 output instead of R:
// each line along the last axis, at sample indices (i_0, .., i_{d-2}) of
// the other axes, is written, mirror-padded to M+3, to every padded line
// whose indices reflect onto it (j_a = i_a + 1, and 0, M_a + 1 or M_a + 2
// where the reflection of -1, M_a or M_a + 1 is i_a).
template <bool CPLX>
__global__ __launch_bounds__(kLB) void axis_lds_kernel(const float *__restrict__ x32, double *__restrict__ R, long M,
                                                       long stride, long nlines, AxisPoles P, float *__restrict__ coeff,
                                                       PadDims D)
{
    extern __shared__ double sh[];
    constexpr int W = CPLX ? 2 : 1;
    const long l0 = (long)blockIdx.x * kLB;
    const int t = threadIdx.x;
    const int nl = (int)(nlines - l0 < kLB ? nlines - l0 : kLB);
    const int Mi = (int)M;
    if (stride == 1) {
        const long g0 = l0 * M;
        const int tot = nl * Mi;
        int ln = t / Mi, k = t - (t / Mi) * Mi;                 // (line, element) of run index t
        const int dl = kLB / Mi, dk = kLB - dl * Mi;
        for (int e0 = t; e0 < tot; e0 += kLB * kStage) {
            double v[kStage];
            int at[kStage];
#pragma unroll
            for (int q = 0; q < kStage; q++) {
                const int e = e0 + q * kLB;
                const long g = g0 + (e < tot ? e : tot - 1);
                v[q] = x32 ? (double)x32[g] : R[g];
                at[q] = e < tot ? (k * kLB + ln) * W : -1;
                ln += dl;
                k += dk;
                if (k >= Mi) {
                    k -= Mi;
                    ln++;
                }
            }
#pragma unroll
            for (int q = 0; q < kStage; q++) {
                if (at[q] < 0) continue;
                sh[at[q]] = v[q];
                if (CPLX) sh[at[q] + 1] = 0.0;
            }
        }
    } else if (t < nl) {
        const long l = l0 + t, outer = l / stride, inner = l - outer * stride, base = outer * M * stride + inner;
        for (int k0 = 0; k0 < Mi; k0 += kStage) {
            double v[kStage];
#pragma unroll
            for (int q = 0; q < kStage; q++) {
                const int k = k0 + q < Mi ? k0 + q : Mi - 1;
                const long g = base + k * stride;
                v[q] = x32 ? (double)x32[g] : R[g];
            }
#pragma unroll
            for (int q = 0; q < kStage; q++) {
                if (k0 + q >= Mi) break;
                sh[((k0 + q) * kLB + t) * W] = v[q];
                if (CPLX) sh[((k0 + q) * kLB + t) * W + 1] = 0.0;
            }
        }
    }
    __syncthreads();
    if (t < nl) {
        LdsLine<CPLX> L{sh + t * W};
        filter_line<CPLX>(L, M, P);
    }
    if (coeff) {                                              // last axis: padded FP32 output
        if (t >= nl) return;
        const int d = D.ndim;
        // sample indices of this line on the other axes (the line index is
        // their flattened value: outer is 0 on the last axis); per axis the
        // padded indices that reflect onto it: i + 1, then those of 0, M+1,
        // M+2 whose reflection is i.  All loops run over compile-time axis
        // numbers so the per-axis values stay in registers.
        int jv[kMaxCoeffDims - 1][4], nj[kMaxCoeffDims - 1], dig[kMaxCoeffDims - 1];
        int rem = (int)(l0 + t);
#pragma unroll
        for (int a = 0; a < kMaxCoeffDims - 1; a++) {
            nj[a] = 1;
            dig[a] = 0;
            jv[a][0] = jv[a][1] = jv[a][2] = jv[a][3] = 0;
            if (a + 1 < d) {
                const int Ma = (int)D.in[a];
                const int i = rem % Ma;
                rem /= Ma;
                jv[a][0] = i + 1;
                const int cand[3] = {0, Ma + 1, Ma + 2};
                const bool hit[3] = {(int)reflect(-1, Ma) == i, (int)reflect(Ma, Ma) == i,
                                     (int)reflect(Ma + 1, Ma) == i};
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    // slot nj[a] by selects: no run-time index into the array
                    const int n = nj[a];
                    if (hit[c]) {
                        jv[a][1] = n == 1 ? cand[c] : jv[a][1];
                        jv[a][2] = n == 2 ? cand[c] : jv[a][2];
                        jv[a][3] = n == 3 ? cand[c] : jv[a][3];
                        nj[a] = n + 1;
                    }
                }
            }
        }
        // the last axis: padded j reads sample j - 1, reflected at the ends
        const int Mi3 = Mi + 3;
        const int r0 = (int)reflect(-1, M), rM = (int)reflect(M, M), rM1 = (int)reflect(M + 1, M);
        const double *line = sh + t * W;
        for (;;) {
            long base = 0, os = 1;
#pragma unroll
            for (int a = 0; a < kMaxCoeffDims - 1; a++) {
                if (a + 1 < d) {
                    const int q = dig[a];
                    const int j = q == 0 ? jv[a][0] : q == 1 ? jv[a][1] : q == 2 ? jv[a][2] : jv[a][3];
                    base += (long)j * os;
                    os *= D.out[a];
                }
            }
            for (int j = 0; j < Mi3; j++) {
                const int src = j == 0 ? r0 : j <= Mi ? j - 1 : j == Mi + 1 ? rM : rM1;
                coeff[base + (long)j * os] = (float)line[src * kLB * W];
            }
            // next combination (odometer over the axes)
            bool carry = true;
#pragma unroll
            for (int a = 0; a < kMaxCoeffDims - 1; a++) {
                if (carry && a + 1 < d) {
                    if (++dig[a] == nj[a])
                        dig[a] = 0;
                    else
                        carry = false;
                }
            }
            if (carry) break;
        }
        return;
    }
    __syncthreads();
    if (stride == 1) {
        const long g0 = l0 * M;
        const int tot = nl * Mi;
        int ln = t / Mi, k = t - (t / Mi) * Mi;
        const int dl = kLB / Mi, dk = kLB - dl * Mi;
        for (int e = t; e < tot; e += kLB) {
            R[g0 + e] = sh[(k * kLB + ln) * W];
            ln += dl;
            k += dk;
            if (k >= Mi) {
                k -= Mi;
                ln++;
            }
        }
    } else if (t < nl) {
        const long l = l0 + t, outer = l / stride, inner = l - outer * stride, base = outer * M * stride + inner;
        for (int k = 0; k < Mi; k++) R[base + k * stride] = sh[(k * kLB + t) * W];
    }
}

// Longer lines: one thread per line, in place in global memory (R holds the
// real parts, I the imaginary parts of a complex pole's intermediates).
template <bool CPLX>
__global__ __launch_bounds__(256) void axis_global_kernel(const float *__restrict__ x32, double *__restrict__ R,
                                                          double *__restrict__ I, long M, long stride, long nlines,
                                                          AxisPoles P)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    const long outer = l / stride, inner = l - outer * stride, base = outer * M * stride + inner;
    for (long k = 0; k < M; k++) {
        const long g = base + k * stride;
        if (x32) R[g] = (double)x32[g];
        if (CPLX) I[g] = 0.0;
    }
    GlobalLine<CPLX> L{R + base, I + base, stride};
    filter_line<CPLX>(L, M, P);
}

__global__ void pad_kernel(const double *__restrict__ R, float *__restrict__ c, long nout, PadDims D)
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
    c[i] = (float)R[src];
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
    long n = 1, nout = 1;
    PadDims D;
    D.ndim = ndim;
    for (int a = 0; a < ndim; a++) {
        n *= dims[a];
        nout *= dims[a] + 3;
        D.in[a] = dims[a];
        D.out[a] = dims[a] + 3;
    }
    double *R = (double *)work, *I = R + n;       // work: 2 n doubles
    hipError_t e = hipSuccess;
    bool padded = false;
    long stride = 1;
    for (int a = 0; a < ndim && e == hipSuccess; a++) {
        const long M = dims[a], nlines = n / M;
        std::complex<double> zs[2];
        AxisPoles P;
        P.n = coeff_poles(lambda ? lambda[a] : 0.0, zs, &P.gain);
        P.cplx_poles = false;
        for (int i = 0; i < P.n; i++) {
            const std::complex<double> z = zs[i], c0 = -z / (1.0 - z * z);
            P.z[i] = {z.real(), z.imag()};
            P.c0[i] = {c0.real(), c0.imag()};
            P.K0[i] = (int)std::ceil(std::log(2.220446049250313e-16) / std::log(std::abs(z)));
            if (z.imag() != 0.0) P.cplx_poles = true;
        }
        const float *x = a == 0 ? samples : nullptr;
        const bool last = a == ndim - 1;
        if (M <= kLdsM) {
            // the last axis writes the padded output itself
            float *out = last ? coeff : nullptr;
            const unsigned blocks = (unsigned)((nlines + kLB - 1) / kLB);
            const size_t lds = sizeof(double) * (size_t)M * kLB * (P.cplx_poles ? 2 : 1);
            if (P.cplx_poles)
                hipLaunchKernelGGL(axis_lds_kernel<true>, dim3(blocks), dim3(kLB), lds, st, x, R, M, stride, nlines, P,
                                   out, D);
            else
                hipLaunchKernelGGL(axis_lds_kernel<false>, dim3(blocks), dim3(kLB), lds, st, x, R, M, stride, nlines,
                                   P, out, D);
            if (last) padded = true;
        } else {
            const unsigned blocks = (unsigned)((nlines + 255) / 256);
            if (P.cplx_poles)
                hipLaunchKernelGGL(axis_global_kernel<true>, dim3(blocks), dim3(256), 0, st, x, R, I, M, stride, nlines,
                                   P);
            else
                hipLaunchKernelGGL(axis_global_kernel<false>, dim3(blocks), dim3(256), 0, st, x, R, I, M, stride,
                                   nlines, P);
        }
        e = hipGetLastError();
        stride *= M;
    }
    if (e == hipSuccess && !padded) {
        hipLaunchKernelGGL(pad_kernel, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, R, coeff, nout, D);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace bsf
