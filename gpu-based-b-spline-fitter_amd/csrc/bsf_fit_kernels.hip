// bsf_fit_kernels.hip -- gfx950 kernels of the multi-D B-spline PSF
// Poisson-MLE Levenberg-Marquardt fitter (fittypes 6: 3D, 7: 4D, 8: 5D+3D).
//
// One wavefront fits one ROI (workgroup = one wave64).  Per ROI evaluation:
//   stage A  lanes walk the (rows x cols) support window of the ROI and
//            contract the non-xy axes (z[, w[, u]]) of the coefficient array
//            with separable 4-tap value / derivative stencils; the results
//            ("planes": value, d/dz, d/dw, d/du, ...) go to LDS;
//   stage B  lanes walk pixels and apply the 4x4 xy stencil to the planes:
//            model, d/dcol, d/drow and the non-xy derivatives;
//   stage C  per-pixel Poisson deviance, gradient and expected-information
//            Hessian terms, reduced across the wave through LDS.
// The LM control flow (accept / reject / lambda / limiter / bounds / wrap)
// runs lane-uniform and mirrors the reference CPU fitter
// (MultiD_bspline/visual code/CPUmleFit_LM/CPUmleFit_LM.cpp:2411-3620);
// the damped normal-equation solve and the Fisher inverse run in FP64.
//
// Precision: E = float  -> FP32 spline evaluation and pixel sums (MIX);
//            E = double -> everything in FP64 (F64).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bsf_kernels.h"

namespace bsf {

template <int FT>
struct Ft;
template <>
struct Ft<6> { static constexpr int D = 3, K = 1, NV = 5, NPL = 2; };
template <>
struct Ft<7> { static constexpr int D = 4, K = 2, NV = 6, NPL = 3; };
// planes: value, dz, dw, du, z-irregular, u-irregular, 3D value, 3D dz
template <>
struct Ft<8> { static constexpr int D = 5, K = 3, NV = 8, NPL = 8; };

// ---------------------------------------------------------------- basis
// Centred cubic / quadratic B-spline, the reference's piecewise formulas in
// the same operation order (CPUsplineLib.cpp:1289-1326).
template <typename E>
__device__ __forceinline__ E bsp3(E x)
{
    const E c43 = E(4) / E(3), c23 = E(2) / E(3), c16 = E(1) / E(6);
    E x2 = x * x, x3 = x2 * x, b = E(0);
    if (x >= E(1) && x < E(2)) b = c43 - E(2) * x + x2 - c16 * x3;
    else if (x >= E(0) && x < E(1)) b = c23 - x2 + E(0.5) * x3;
    else if (x >= E(-1) && x < E(0)) b = c23 - x2 - E(0.5) * x3;
    else if (x >= E(-2) && x < E(-1)) b = c43 + E(2) * x + x2 + c16 * x3;
    return b;
}

template <typename E>
__device__ __forceinline__ E bsp2(E x)
{
    E x2 = x * x, b = E(0);
    if (x >= E(0.5) && x < E(1.5)) b = E(1.125) - E(1.5) * x + E(0.5) * x2;
    else if (x >= E(-0.5) && x < E(0.5)) b = E(0.75) - x2;
    else if (x >= E(-1.5) && x < E(-0.5)) b = (E)((double)E(1.125) + 1.5 * (double)x + (double)(E(0.5) * x2));
    return b;
}

// Per-axis factors for fractional part d (the reference's B1/B2/dB1/dB2 at
// idx 1, 2) and the 4-tap stencils over taps k-2..k+1:
//   W  = value weights, Dv = derivative weights (forward differences with
//   quadratic weights rewritten as a stencil on the coefficients).
template <typename E>
struct Axis {
    E B1[3], B2[3], dB1[3], dB2[3];
    E W[4], Dv[4];
};

template <typename E>
__device__ __forceinline__ void axis_factors(E d, Axis<E> &a)
{
    E dd = d - E(1);
#pragma unroll
    for (int idx = 1; idx <= 2; idx++) {
        a.B1[idx] = bsp3<E>(d + (E)idx - E(1));
        a.B2[idx] = bsp3<E>(d - (E)idx);
        a.dB1[idx] = bsp2<E>(dd + (E)idx - E(0.5));
        a.dB2[idx] = bsp2<E>(dd - (E)idx + E(0.5));
    }
    a.W[0] = a.B1[2];
    a.W[1] = a.B1[1];
    a.W[2] = a.B2[1];
    a.W[3] = a.B2[2];
    // sum_t (c[t+1]-c[t]) dW[t], dW = (dB1[2], dB1[1], dB2[1], dB2[2]);
    // dB2[2] = beta2(d - 2.5) = 0 for d in [0,1), so tap k+2 drops out.
    a.Dv[0] = -a.dB1[2];
    a.Dv[1] = a.dB1[2] - a.dB1[1];
    a.Dv[2] = a.dB1[1] - a.dB2[1];
    a.Dv[3] = a.dB2[1] - a.dB2[2];
}

// Clamped base tap index for one coordinate (CPUsplineLib.cpp:1203-1225):
// floor, clamp to [1, n-1], coordinate == n selects n; returns k = kc + 1.
template <typename E>
__device__ __forceinline__ int base_tap(E c, int n)
{
    int kc = (int)floor(c);
    kc = kc < 1 ? 1 : kc;
    kc = kc > n - 1 ? n - 1 : kc;
    if (c == (E)n) kc = n;
    return kc + 1;
}

// -------------------------------------------------------------- geometry
// Everything one evaluation needs that depends on the ROI-level spline
// coordinates (and so on theta): per-axis factors, per-pixel taps, windows.
template <typename E, int FT>
struct Geo {
    Axis<E> ax[5];     // 0: row (dim 0), 1: column (dim 1), 2..: z, w, u
    int knx[3];        // base taps of the non-xy axes
    int row0, col0;    // window origin (0-based coefficient index)
    int PR, PC;        // window extent
};

struct Dev {
    const float *data, *coeff, *coeff3;
    float *P, *CRLB, *LL;
    int N, S, iterations, roi_offset;
    int nd[5];
    long s1, s2, s3, s4;   // coefficient strides of dims 1..4
    int cbytes, c3bytes;   // coefficient array sizes (buffer record counts)
    const float *theta_in; // test hook: evaluate once at these parameters
    float *dbg;            // test hook output: err, J[NV], H[NV*NV] per ROI
    float initP[4];
    int has_initP;
};

template <typename E, int FT>
__device__ __forceinline__ void make_geo(const Dev &A, const E *cth, Geo<E, FT> &g)
{
    constexpr int D = Ft<FT>::D;
    const int S = A.S;
    const E off = (E)(A.nd[0] - S) / E(2);
    const E colc = (E)((S + 1) / 2) - cth[0];
    const E rowc = (E)((S + 1) / 2) - cth[1];
    axis_factors<E>(rowc - floor(rowc), g.ax[0]);
    axis_factors<E>(colc - floor(colc), g.ax[1]);
#pragma unroll
    for (int a = 2; a < D; a++) {
        E c = cth[2 + a];
        axis_factors<E>(c - floor(c), g.ax[a]);
        g.knx[a - 2] = base_tap<E>(c, A.nd[a]);
    }
    int rmin = 1 << 30, rmax = -1, cmin = 1 << 30, cmax = -1;
    for (int i = 0; i < S; i++) {
        int kr = base_tap<E>(rowc + (E)i + off, A.nd[1]);
        int kq = base_tap<E>(colc + (E)i + off, A.nd[0]);
        rmin = min(rmin, kr);
        rmax = max(rmax, kr);
        cmin = min(cmin, kq);
        cmax = max(cmax, kq);
    }
    g.row0 = rmin - 2;
    g.col0 = cmin - 2;
    g.PR = rmax - rmin + 4;
    g.PC = cmax - cmin + 4;
}

// ------------------------------------------------------------ stage A
// Coefficient gathers go through buffer loads: one per-lane byte offset for
// the window point (VGPR) plus a wave-uniform tap offset (SGPR soffset), so
// the 4^K taps of a point cost no per-tap address registers; the descriptor's
// record count also makes any out-of-range tap read 0.
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// Contract the non-xy axes over the window; planes[p][b*PR + a].
template <typename E, int FT>
__device__ __forceinline__ void stage_a(const Dev &A, const Geo<E, FT> &g, E *planes, int plane_stride, int lane)
{
    constexpr int K = Ft<FT>::K;
    const int PR = g.PR, npts = g.PR * g.PC;
    const __amdgpu_buffer_rsrc_t rc = __builtin_amdgcn_make_buffer_rsrc((void *)A.coeff, (short)0, A.cbytes, 0x00020000);
    const int s2 = (int)A.s2 * 4, s3 = (int)A.s3 * 4, s4 = (int)A.s4 * 4;
    int tapb = (g.knx[0] - 2) * s2;
    if constexpr (K >= 2) tapb += (g.knx[1] - 2) * s3;
    if constexpr (K >= 3) tapb += (g.knx[2] - 2) * s4;
    tapb = __builtin_amdgcn_readfirstlane(tapb);
    for (int q = lane; q < npts; q += 64) {
        const int a = q % PR, b = q / PR;
        const int vo = ((g.row0 + a) + (g.col0 + b) * (int)A.s1) * 4;
        if constexpr (K == 1) {
            const Axis<E> &z = g.ax[2];
            E cv[4];
#pragma unroll
            for (int t = 0; t < 4; t++) cv[t] = (E)bload(rc, vo, tapb + t * s2);
            E v = E(0), dz = E(0);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                v += z.W[t] * cv[t];
                dz += z.Dv[t] * cv[t];
            }
            planes[q] = v;
            planes[plane_stride + q] = dz;
        } else if constexpr (K == 2) {
            const Axis<E> &z = g.ax[2], &w = g.ax[3];
            E v0[4], v1[4];
#pragma unroll
            for (int tz = 0; tz < 4; tz++) {
                E cv[4];
#pragma unroll
                for (int tw = 0; tw < 4; tw++) cv[tw] = (E)bload(rc, vo, tapb + tz * s2 + tw * s3);
                E s0 = E(0), s1 = E(0);
#pragma unroll
                for (int tw = 0; tw < 4; tw++) {
                    s0 += w.W[tw] * cv[tw];
                    s1 += w.Dv[tw] * cv[tw];
                }
                v0[tz] = s0;
                v1[tz] = s1;
            }
            E v = E(0), dz = E(0), dw = E(0);
#pragma unroll
            for (int tz = 0; tz < 4; tz++) {
                v += z.W[tz] * v0[tz];
                dz += z.Dv[tz] * v0[tz];
                dw += z.W[tz] * v1[tz];
            }
            planes[q] = v;
            planes[plane_stride + q] = dz;
            planes[2 * plane_stride + q] = dw;
        } else {
            const Axis<E> &z = g.ax[2], &w = g.ax[3], &u = g.ax[4];
            E cv[4][4][4];
#pragma unroll
            for (int tz = 0; tz < 4; tz++)
#pragma unroll
                for (int tw = 0; tw < 4; tw++)
#pragma unroll
                    for (int tu = 0; tu < 4; tu++) cv[tz][tw][tu] = (E)bload(rc, vo, tapb + tz * s2 + tw * s3 + tu * s4);
            E v = E(0), dz = E(0), dw = E(0), du = E(0);
#pragma unroll
            for (int tz = 0; tz < 4; tz++) {
                E a00 = E(0), a01 = E(0), a10 = E(0);
#pragma unroll
                for (int tw = 0; tw < 4; tw++) {
                    E u0 = E(0), u1 = E(0);
#pragma unroll
                    for (int tu = 0; tu < 4; tu++) {
                        u0 += u.W[tu] * cv[tz][tw][tu];
                        u1 += u.Dv[tu] * cv[tz][tw][tu];
                    }
                    a00 += w.W[tw] * u0;
                    a01 += w.Dv[tw] * u0;
                    a10 += w.W[tw] * u1;
                }
                v += z.W[tz] * a00;
                dz += z.Dv[tz] * a00;
                dw += z.W[tz] * a01;
                du += z.W[tz] * a10;
            }
            // The reference's 5D contraction has two irregular terms in
            // both its CPU and GPU code (CPUsplineLib.cpp:313 / 420-421,
            // GPUsplineLib.cuh:573 / 681); they are reproduced as two extra
            // planes applied later with the upper / lower 2x2 xy stencils.
            //  z: term 27 of block (q,r,s) weighs the z-difference at tap
            //     kz-q (w tap kw+r-1, u tap ku+s-1) with a weight taken from
            //     block +8, term 15 (past the z table: the w table).
            E zb = E(0);
#pragma unroll
            for (int q2 = 1; q2 <= 2; q2++)
#pragma unroll
                for (int r = 1; r <= 2; r++)
#pragma unroll
                    for (int s = 1; s <= 2; s++) {
                        E bug;
                        if (r == 1) bug = z.dB2[q2] * w.B2[2] * u.B1[s];
                        else if (s == 1) bug = z.dB2[q2] * w.B2[1] * u.B1[2];
                        else bug = z.B2[q2] * w.dB2[1] * u.B1[1];
                        E ok = z.dB1[q2] * w.B2[r] * u.B2[s];
                        const int tz = 2 - q2, tw = 1 + r, tu = 1 + s;
                        zb += (bug - ok) * (cv[tz + 1][tw][tu] - cv[tz][tw][tu]);
                    }
            //  u: term 16 differences u tap ku-s instead of ku+s-1; with
            //     dB2[2] = 0 only s = 1 survives: weight * (2c[ku]-c[ku-1]-c[ku+1]).
            E ub = E(0);
#pragma unroll
            for (int q2 = 1; q2 <= 2; q2++)
#pragma unroll
                for (int r = 1; r <= 2; r++) {
                    const int tz = 2 - q2, tw = 2 - r;
                    ub += z.B1[q2] * w.B1[r] * (E(2) * cv[tz][tw][2] - cv[tz][tw][1] - cv[tz][tw][3]);
                }
            ub *= u.dB2[1];
            planes[q] = v;
            planes[plane_stride + q] = dz;
            planes[2 * plane_stride + q] = dw;
            planes[3 * plane_stride + q] = du;
            planes[4 * plane_stride + q] = zb;
            planes[5 * plane_stride + q] = ub;
            // the 3D model (coeff3) shares the xy window and the z taps
            const __amdgpu_buffer_rsrc_t r3 =
                __builtin_amdgcn_make_buffer_rsrc((void *)A.coeff3, (short)0, A.c3bytes, 0x00020000);
            const int tap3 = __builtin_amdgcn_readfirstlane((g.knx[0] - 2) * s2);
            E v3 = E(0), dz3 = E(0);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                E cc = (E)bload(r3, vo, tap3 + t * s2);
                v3 += z.W[t] * cc;
                dz3 += z.Dv[t] * cc;
            }
            planes[6 * plane_stride + q] = v3;
            planes[7 * plane_stride + q] = dz3;
        }
    }
}

// --------------------------------------------------------- stage B + C
// Model and dudt at one pixel from the planes.
template <typename E, int FT>
__device__ __forceinline__ void pixel_model(const Dev &A, const Geo<E, FT> &g, const E *planes, int plane_stride,
                                            int ii, int jj, E cth0, E cth1, E Nph, E bg, E g2, E *dudt, E &model)
{
    constexpr int D = Ft<FT>::D;
    const int S = A.S;
    const E off = (E)(A.nd[0] - S) / E(2);
    const E colc = (E)((S + 1) / 2) - cth0;
    const E rowc = (E)((S + 1) / 2) - cth1;
    const int r = base_tap<E>(rowc + (E)ii + off, A.nd[1]) - 2 - g.row0;
    const int c = base_tap<E>(colc + (E)jj + off, A.nd[0]) - 2 - g.col0;
    const int PR = g.PR;
    const Axis<E> &ay = g.ax[0], &ax = g.ax[1];
    const E *p0 = planes + c * PR + r;
    // value plane: f, d/dcol, d/drow
    E f = E(0), dcol = E(0), drow = E(0);
#pragma unroll
    for (int ty = 0; ty < 4; ty++) {
        E sw = E(0), sd = E(0);
#pragma unroll
        for (int tx = 0; tx < 4; tx++) {
            E gv = p0[tx * PR + ty];
            sw += ax.W[tx] * gv;
            sd += ax.Dv[tx] * gv;
        }
        f += ay.W[ty] * sw;
        dcol += ay.W[ty] * sd;
        drow += ay.Dv[ty] * sw;
    }
    E dn[3] = {E(0), E(0), E(0)};
#pragma unroll
    for (int k = 0; k < D - 2; k++) {
        const E *pk = p0 + (k + 1) * plane_stride;
        E s = E(0);
#pragma unroll
        for (int ty = 0; ty < 4; ty++) {
            E sw = E(0);
#pragma unroll
            for (int tx = 0; tx < 4; tx++) sw += ax.W[tx] * pk[tx * PR + ty];
            s += ay.W[ty] * sw;
        }
        dn[k] = s;
    }
    if constexpr (FT == 8) {
        const E *pz = p0 + 4 * plane_stride, *pu = p0 + 5 * plane_stride;
        E zb = E(0), ub = E(0);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                zb += ay.W[2 + i] * ax.W[2 + j] * pz[(2 + j) * PR + 2 + i];
                ub += ay.W[i] * ax.W[j] * pu[j * PR + i];
            }
        dn[0] += zb;
        dn[2] += ub;
    }
    dudt[0] = E(-1) * Nph * dcol;
    dudt[1] = E(-1) * Nph * drow;
    dudt[2] = f;
    dudt[3] = E(1);
#pragma unroll
    for (int k = 0; k < D - 2; k++) dudt[4 + k] = Nph * dn[k];
    model = bg + Nph * f;
    if constexpr (FT == 8) {
        const E *q0 = p0 + 6 * plane_stride, *q1 = p0 + 7 * plane_stride;
        E f3 = E(0), dc3 = E(0), dr3 = E(0), dz3 = E(0);
#pragma unroll
        for (int ty = 0; ty < 4; ty++) {
            E sw = E(0), sd = E(0), sz = E(0);
#pragma unroll
            for (int tx = 0; tx < 4; tx++) {
                E gv = q0[tx * PR + ty];
                sw += ax.W[tx] * gv;
                sd += ax.Dv[tx] * gv;
                sz += ax.W[tx] * q1[tx * PR + ty];
            }
            f3 += ay.W[ty] * sw;
            dc3 += ay.W[ty] * sd;
            dr3 += ay.Dv[ty] * sw;
            dz3 += ay.W[ty] * sz;
        }
        E d3[5];
        d3[0] = E(-1) * Nph * dc3;
        d3[1] = E(-1) * Nph * dr3;
        d3[2] = f3;
        d3[3] = E(1);
        d3[4] = Nph * dz3;
        E m3 = bg + Nph * f3;
        E m5 = model;
#pragma unroll
        for (int l = 0; l < 5; l++) dudt[l] = dudt[l] * g2 + (E(1) - g2) * d3[l];
        model = m3 * (E(1) - g2) + m5 * g2;
        dudt[7] = m5 - m3;
    }
}

// ------------------------------------------------------------- LM solve
// The NV x NV systems are spread over the wave: lane i holds row i, pivots
// and columns travel by v_readlane (uniform lane index), so the solve costs a
// few registers per lane instead of NV^2 doubles.
__device__ __forceinline__ double readlane_d(double v, int l)
{
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ float readlane_f(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

template <typename E>
__device__ __forceinline__ E readlane_e(E v, int l)
{
    if constexpr (sizeof(E) == 8) return readlane_d(v, l);
    else return readlane_f(v, l);
}

// Damped normal equations H x = J in FP64 by Cholesky (right-looking), the
// two triangular solves; lane i holds H row i (a[]) and J_i (b).  Fails, as
// the reference's kernel_cholesky (CPUsplineLib.cpp:1363-1388), when a pivot
// radicand is negative.  On success every lane returns x_lane in xi.
template <int NV>
__device__ __forceinline__ int chol_solve_rows(double (&a)[NV], double b, int lane, double &xi)
{
    double colL[NV], rinv[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) colL[j] = 0.0;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const double akk = readlane_d(a[k], k);
        if (!(akk >= 0.0)) return 1;
        const double lkk = sqrt(akk);
        rinv[k] = 1.0 / lkk;
        const double lik = lane > k ? a[k] * rinv[k] : (lane == k ? lkk : 0.0);
        a[k] = lik;
#pragma unroll
        for (int j = k + 1; j < NV; j++) {
            const double ljk = readlane_d(lik, j);
            a[j] -= lik * ljk;
            colL[j] = (lane == k) ? ljk : colL[j];   // lane k keeps column k of L
        }
    }
    double acc = 0.0, yv = 0.0;
#pragma unroll
    for (int k = 0; k < NV; k++) {                   // L y = b
        const double yk = readlane_d((b - acc) * rinv[k], k);
        yv = (lane == k) ? yk : yv;
        acc += a[k] * yk;
    }
    double acc2 = 0.0;
    xi = 0.0;
#pragma unroll
    for (int k = NV - 1; k >= 0; k--) {              // L^T x = y
        const double xk = readlane_d((yv - acc2) * rinv[k], k);
        xi = (lane == k) ? xk : xi;
        acc2 += colL[k] * xk;
    }
    return 0;
}

// Diagonal of the inverse Fisher matrix by unpivoted Gauss-Jordan elimination
// over rows held one per lane, in FP64: the reference's kernel_MatInvN
// (MatInvLib.h:17-77) is the same unpivoted elimination written as LU.
template <int NV>
__device__ __forceinline__ double inverse_diag_rows(double (&m)[NV], int lane)
{
    double v[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) v[j] = (j == lane) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const double mkk = readlane_d(m[k], k);
        const double r = 1.0 / mkk;
        const double fac = m[k] * r;
#pragma unroll
        for (int j = 0; j < NV; j++) {
            const double mk = readlane_d(m[j], k), vk = readlane_d(v[j], k);
            m[j] = (lane == k) ? mk * r : m[j] - fac * mk;
            v[j] = (lane == k) ? vk * r : v[j] - fac * vk;
        }
    }
    double d = 0.0;
#pragma unroll
    for (int j = 0; j < NV; j++) d = (j == lane) ? v[j] : d;
    return d;
}

// Index of (row i, column j) in the packed upper triangle used by the sums.
template <int NV>
__device__ __forceinline__ int packed_index(int i, int j)
{
    const int a = i < j ? i : j, b = i < j ? j : i;
    return a * NV - (a * (a - 1)) / 2 + (b - a);
}

// Sum V per-lane values across the wave; the totals stay in LDS at
// scratch[V*65 + v] for every lane to read.
template <typename E, int V>
__device__ __forceinline__ void wave_sum_lds(const E (&v)[V], E *scratch, int lane)
{
#pragma unroll
    for (int k = 0; k < V; k++) scratch[k * 65 + lane] = v[k];
    __syncthreads();
    if (lane < V) {
        E s0 = E(0), s1 = E(0), s2 = E(0), s3 = E(0);
        const E *row = scratch + lane * 65;
#pragma unroll 4
        for (int j = 0; j < 64; j += 4) {
            s0 += row[j];
            s1 += row[j + 1];
            s2 += row[j + 2];
            s3 += row[j + 3];
        }
        scratch[V * 65 + lane] = (s0 + s1) + (s2 + s3);
    }
    __syncthreads();
}

template <typename E>
__device__ __forceinline__ E dmaxr(E a, double b) { return ((double)a > b) ? a : (E)b; }
template <typename E>
__device__ __forceinline__ E dminr(E a, double b) { return ((double)a < b) ? a : (E)b; }

// ----------------------------------------------------------------- kernel
template <int FT, typename E>
__global__ __launch_bounds__(64) void fit_kernel(Dev A)
{
    constexpr int NV = Ft<FT>::NV;
    constexpr int NPL = Ft<FT>::NPL;
    constexpr int NH = NV * (NV + 1) / 2;
    constexpr int NRED = 1 + NV + NH;
    extern __shared__ __align__(16) unsigned char smem[];

    const int n = blockIdx.x + A.roi_offset;
    if (n >= A.N) return;
    const int lane = threadIdx.x;
    const int S = A.S, S2 = S * S;
    const int W = S + 4, plane_stride = W * W;
    float *s_data = (float *)smem;
    E *s_pl = (E *)(smem + (((size_t)S2 * 4 + 15) & ~(size_t)15));
    E *s_red = s_pl;   // reductions reuse the plane storage (never live together)

    const float *dptr = A.data + (size_t)n * S2;
    for (int p = lane; p < S2; p += 64) s_data[p] = dptr[p];
    __syncthreads();

    // ---- initial values: centre of mass, max / min (CPUgaussLib.cpp:172-223)
    // summed in the reference's sequential order so they match it bit for bit.
    E th[NV], oth[NV], sgn[NV], osgn[NV], clampv[NV], oclamp[NV], fade[NV], upd[NV];
    (void)NPL;
    {
        E tx = E(0), ty = E(0), ts = E(0), mx = E(0), mn = E(10e10);
        for (int ii = 0; ii < S; ii++)
            for (int jj = 0; jj < S; jj++) {
                E d = (E)s_data[S * jj + ii];
                tx += d * (E)ii;
                ty += d * (E)jj;
                ts += d;
            }
        for (int p = 0; p < S2; p++) {
            E d = (E)s_data[p];
            mx = (mx > d) ? mx : d;
            mn = (mn < d) ? mn : d;
        }
        th[0] = tx / ts;
        th[1] = ty / ts;
        double t = (double)(((mx - mn) * E(2)) * (E)3.141592) * 1.5 * 1.5;   // PI, definitions.h:15
        th[2] = (E)((0.0 > t) ? 0.0 : t);
        th[3] = dmaxr<E>(mn, 0.01);
        if constexpr (FT == 6) {
            th[4] = (E)A.nd[2] / E(2);
        } else if constexpr (FT == 7) {
            if (A.has_initP && A.initP[0] != 0.f) {
                th[4] = (E)A.initP[0];
                th[5] = (E)A.initP[1];
            } else {
                th[4] = (E)A.nd[2] / E(2);
                th[5] = (E)A.nd[3] / E(2);
            }
        } else {
            if (A.has_initP) {
#pragma unroll
                for (int i = 4; i < 8; i++) th[i] = (E)A.initP[i - 4];
            } else {
                th[4] = (E)A.nd[2] / E(2);
                th[5] = (E)A.nd[3] / E(2);
                th[6] = (E)A.nd[4] / E(2);
                th[7] = (E)0.9;
            }
        }
    }
    {
        const float c6[5] = {2.0f, 2.0f, 10000.f, 20.f, 20.f};
        const float c7[6] = {1.0f, 1.0f, 10000.f, 20.f, 2.f, 0.00001f};
        const float f7[6] = {0.5f, 0.5f, 0.4f, 0.5f, 0.7f, 0.8f};
        const float c8[8] = {1.0f, 1.0f, 10000.f, 20.f, 2.f, 2.f, 2.f, 0.1f};
#pragma unroll
        for (int i = 0; i < NV; i++) {
            clampv[i] = (E)(FT == 6 ? c6[i] : FT == 7 ? c7[i] : c8[i]);
            oclamp[i] = clampv[i];
            fade[i] = (E)(FT == 7 ? f7[i] : 0.5f);
            sgn[i] = E(0);
            osgn[i] = E(0);
            upd[i] = E(0);
            oth[i] = th[i];
        }
    }

    // coordinate-bearing parameters of the last evaluation (theta 0,1,4..7)
    E cth[9];
    Geo<E, FT> g;
    E g2 = E(0);
    // lane i (< NV) holds row i of the Hessian and J_i; the reference keeps
    // the Hessian between iterations (a rejected step reuses it, a failed
    // Cholesky leaves its damped diagonal), so does this.
    E hrow[NV], jv = E(0);
    const int myrow = lane < NV ? lane : NV - 1;

    // one ROI evaluation at th: fills hrow / jv and returns the deviance
    auto evaluate = [&](bool first) __attribute__((always_inline)) -> E {
#pragma unroll
        for (int i = 0; i < NV; i++) cth[i] = th[i];
        if constexpr (FT == 8) g2 = th[7];
        make_geo<E, FT>(A, cth, g);
        __syncthreads();   // the scratch of the last reduction is read out
        stage_a<E, FT>(A, g, s_pl, plane_stride, lane);
        __syncthreads();
        E acc[NRED];
#pragma unroll
        for (int k = 0; k < NRED; k++) acc[k] = E(0);
        for (int p = lane; p < S2; p += 64) {
            const int pjj = p / S, pii = p - pjj * S;
            E dudt[8], model;
            pixel_model<E, FT>(A, g, s_pl, plane_stride, pii, pjj, cth[0], cth[1], th[2], th[3], g2, dudt,
                               model);
            if constexpr (FT == 7) {
                if (first) dudt[5] = dudt[4];
            }
            E data = (E)s_data[p];
            if (data > E(0)) {
                acc[0] += E(2) * ((model - data) - data * log(model / data));
            } else {
                acc[0] += E(2) * model;
                data = E(0);
            }
            E t1 = E(1) - data / model;
            E t2 = data / (model * model);
#pragma unroll
            for (int l = 0; l < NV; l++) acc[1 + l] += t1 * dudt[l];
            int h = 1 + NV;
#pragma unroll
            for (int l = 0; l < NV; l++)
#pragma unroll
                for (int m = l; m < NV; m++) acc[h++] += t2 * dudt[l] * dudt[m];
        }
        __syncthreads();   // planes are dead; the scratch reuses them
        wave_sum_lds<E, NRED>(acc, s_red, lane);
        const E *tot = s_red + NRED * 65;
        jv = tot[1 + myrow];
#pragma unroll
        for (int j = 0; j < NV; j++) hrow[j] = tot[1 + NV + packed_index<NV>(myrow, j)];
        return tot[0];
    };

    if (A.theta_in) {
        // test hook: one evaluation at a given theta; write err, J, H
#pragma unroll
        for (int i = 0; i < NV; i++) th[i] = (E)A.theta_in[(size_t)n * NV + i];
        const E e = evaluate(false);
        float *o = A.dbg + (size_t)n * (1 + NV + NV * NV);
        if (lane == 0) o[0] = (float)e;
        if (lane < NV) {
            o[1 + lane] = (float)jv;
#pragma unroll
            for (int j = 0; j < NV; j++) o[1 + NV + lane * NV + j] = (float)hrow[j];
        }
        return;
    }
    E newErr = evaluate(true);
    E oldErr = (E)1e13;
    E newLambda = E(1), oldLambda = E(1), sumUpdate = (E)1e10;
    int turn = 1, kk;
    for (kk = 0; kk < A.iterations; kk++) {
        if (fabs((newErr - oldErr) / newErr) < (E)1e-6 || (FT == 8 && sumUpdate < (E)1e-5)) break;   // TOLERANCE, TOLERANCE1
        if ((double)newErr > 1.5 * (double)oldErr) {
#pragma unroll
            for (int i = 0; i < NV; i++) {
                sgn[i] = osgn[i];
                clampv[i] = oclamp[i];
                th[i] = oth[i];
            }
            newLambda = oldLambda;
            newErr = oldErr;
            newLambda = E(10) * newLambda;
        } else if (newErr < oldErr) {
            if (newLambda > E(1)) newLambda = (E)((double)newLambda * 0.8);
            else if (newLambda < E(1)) newLambda = E(1);
        }
        double arow[NV], xi;
#pragma unroll
        for (int j = 0; j < NV; j++) {
            hrow[j] = (j == lane) ? hrow[j] * newLambda : hrow[j];
            arow[j] = (double)hrow[j];
        }
        int info = chol_solve_rows<NV>(arow, (double)jv, lane, xi);
        if (info == 0) {
#pragma unroll
            for (int l = 0; l < NV; l++) upd[l] = (E)readlane_d(xi, l);
            if constexpr (FT != 6) {
#pragma unroll
                for (int l = 0; l < NV; l++)
                    if (isnan(upd[l])) info = 1;
            }
        }
        if (info != 0) {
            newLambda = E(10) * newLambda;
            continue;
        }
        if constexpr (FT == 7) upd[5] = upd[5] * E(10);
#pragma unroll
        for (int i = 0; i < NV; i++) {
            osgn[i] = sgn[i];
            oclamp[i] = clampv[i];
            oth[i] = th[i];
        }
        oldLambda = newLambda;
        oldErr = newErr;
        if constexpr (FT == 7) {
            if ((double)fabs(upd[4]) < 0.005 && turn) {
                clampv[5] = (E)A.nd[3] / E(10);
                oclamp[5] = (E)A.nd[3] / E(10);
                turn = 0;
                clampv[4] = E(2);
                oclamp[4] = E(2);
            }
        }
        if constexpr (FT == 8) sumUpdate = E(0);
#pragma unroll
        for (int l = 0; l < NV; l++) {
            if (sgn[l] != E(0)) {
                if ((sgn[l] == E(1) && upd[l] < E(0)) || (sgn[l] == E(-1) && upd[l] > E(0)))
                    clampv[l] = clampv[l] * fade[l];
            }
            sgn[l] = (upd[l] > E(0)) ? E(1) : E(-1);
            if constexpr (FT == 8) {
                upd[l] = upd[l] / (E(1) + fabs(upd[l] / clampv[l]));
                th[l] = th[l] - upd[l];
                sumUpdate += fabs(upd[l]);
            } else {
                th[l] = th[l] - upd[l] / (E(1) + fabs(upd[l] / clampv[l]));
            }
        }
        if constexpr (FT == 8) {
            const double lo = (double)(((E)S - E(1)) / E(2)) - S / 4.0;
            const double hi = (double)(((E)S - E(1)) / E(2)) + S / 4.0;
            th[0] = dmaxr<E>(th[0], lo);
            th[0] = dminr<E>(th[0], hi);
            th[1] = dmaxr<E>(th[1], lo);
            th[1] = dminr<E>(th[1], hi);
        }
        th[2] = dmaxr<E>(th[2], 1.0);
        th[3] = dmaxr<E>(th[3], 0.01);
        th[4] = dmaxr<E>(th[4], 0.0);
        if constexpr (FT == 6) {
            E zb = (E)(A.nd[2] - 4);
            th[4] = (th[4] < zb) ? th[4] : zb;
        } else {
            th[4] = dminr<E>(th[4], (double)(E)A.nd[2] - 0.5);
        }
        if constexpr (FT == 7) {
            th[5] = dmaxr<E>(th[5], 0.0);
            th[5] = dminr<E>(th[5], (double)(E)A.nd[3] - 0.5);
        }
        if constexpr (FT == 8) {
            if (th[5] <= E(1)) th[5] = (E)(A.nd[3] - 2) + th[5];
            else if (th[5] > (E)(A.nd[3] - 1)) th[5] = (E(2) + th[5]) - (E)A.nd[3];
            th[6] = (th[6] > E(1)) ? th[6] : E(1);
            th[6] = dminr<E>(th[6], (double)(E)A.nd[4] - 0.5);
            th[7] = (th[7] < E(1)) ? th[7] : E(1);
            th[7] = (th[7] > E(0)) ? th[7] : E(0);
        }
        newErr = evaluate(false);
#ifdef BSF_TRACE
        if (lane == 0 && A.dbg) {
            float *o = A.dbg + ((size_t)n * 64 + kk) * (3 + 2 * NV);
            o[0] = (float)kk;
            o[1] = (float)newErr;
            o[2] = (float)newLambda;
            for (int i = 0; i < NV; i++) o[3 + i] = (float)th[i];
            for (int i = 0; i < NV; i++) o[3 + NV + i] = (float)upd[i];
        }
#endif
    }

    // ---- Fisher information, CRLB, log-likelihood.  Fittypes 6/7 evaluate
    // at the final theta; fittype 8 keeps the last evaluation's coordinates
    // and g2 with the final N, bg (CPUmleFit_LM.cpp:3566-3583).
    if constexpr (FT != 8) {
#pragma unroll
        for (int i = 0; i < NV; i++) cth[i] = th[i];
        make_geo<E, FT>(A, cth, g);
    }
    __syncthreads();
    stage_a<E, FT>(A, g, s_pl, plane_stride, lane);
    __syncthreads();
    constexpr int NF = 1 + NH;
    E acc[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) acc[k] = E(0);
    for (int p = lane; p < S2; p += 64) {
        const int pjj = p / S, pii = p - pjj * S;
        E dudt[8], model;
        pixel_model<E, FT>(A, g, s_pl, plane_stride, pii, pjj, cth[0], cth[1], th[2], th[3], g2, dudt, model);
        E data = (E)s_data[p];
        int h = 1;
#pragma unroll
        for (int a = 0; a < NV; a++)
#pragma unroll
            for (int b = a; b < NV; b++) acc[h++] += dudt[b] * dudt[a] / model;
        if (model > E(0)) {
            if (data > E(0)) acc[0] += data * log(model) - model - data * log(data) + data;
            else acc[0] += -model;
        }
    }
    __syncthreads();
    wave_sum_lds<E, NF>(acc, s_red, lane);
    const E *tot = s_red + NF * 65;
    double mrow[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) mrow[j] = (double)tot[1 + packed_index<NV>(myrow, j)];
    const double crlb = inverse_diag_rows<NV>(mrow, lane);
    if (lane < NV) A.CRLB[(size_t)A.N * lane + n] = (float)crlb;
    if (lane == 0) {
#pragma unroll
        for (int a = 0; a < NV; a++) A.P[(size_t)A.N * a + n] = (float)th[a];
        A.P[(size_t)A.N * NV + n] = (float)kk;
        A.LL[n] = (float)tot[0];
    }
}

// ------------------------------------------------------------ launchers
template <int FT, typename E>
static hipError_t launch_t(const Dev &A, int nblocks, hipStream_t st)
{
    const int S = A.S;
    const int W = S + 4;
    constexpr int NV = Ft<FT>::NV;
    constexpr int NRED = 1 + NV + NV * (NV + 1) / 2;
    size_t pl = (size_t)Ft<FT>::NPL * W * W;
    size_t red = (size_t)(NRED + 1) * 65;
    size_t lds = (((size_t)S * S * 4 + 15) & ~(size_t)15) + sizeof(E) * (pl > red ? pl : red);
    hipLaunchKernelGGL((fit_kernel<FT, E>), dim3(nblocks), dim3(64), lds, st, A);
    return hipGetLastError();
}

size_t lds_bytes(int fittype, int f64, int S)
{
    int NV = fittype == 6 ? 5 : fittype == 7 ? 6 : 8;
    int NPL = fittype == 6 ? 2 : fittype == 7 ? 3 : 8;
    int W = S + 4;
    size_t pl = (size_t)NPL * W * W, red = (size_t)(1 + NV + NV * (NV + 1) / 2 + 1) * 65;
    return (((size_t)S * S * 4 + 15) & ~(size_t)15) + (f64 ? 8 : 4) * (pl > red ? pl : red);
}

template <int FT, typename E>
static hipError_t launch_ft(const Dev &A, int nblocks, hipStream_t st)
{
    return launch_t<FT, E>(A, nblocks, st);
}

hipError_t launch_fit(const FitLaunch &L, hipStream_t st)
{
    Dev A;
    A.data = L.data;
    A.coeff = L.coeff;
    A.coeff3 = L.coeff3;
    A.P = L.P;
    A.CRLB = L.CRLB;
    A.LL = L.LL;
    A.N = L.N;
    A.S = L.S;
    A.iterations = L.iterations;
    A.roi_offset = L.roi_offset;
    for (int a = 0; a < 5; a++) A.nd[a] = L.cdims[a] - 3;
    A.s1 = L.cdims[0];
    A.s2 = A.s1 * L.cdims[1];
    A.s3 = A.s2 * (L.ndim > 2 ? L.cdims[2] : 1);
    A.s4 = A.s3 * (L.ndim > 3 ? L.cdims[3] : 1);
    {
        long nc = 1;
        for (int a = 0; a < L.ndim; a++) nc *= L.cdims[a];
        A.cbytes = (int)(nc * 4);
        A.c3bytes = (int)(A.s3 * 4);
    }
    for (int i = 0; i < 4; i++) A.initP[i] = L.initP[i];
    A.has_initP = L.has_initP;
    A.theta_in = L.theta_in;
    A.dbg = L.dbg;
    const int nb = L.count;
    switch (L.fittype * 2 + (L.f64 ? 1 : 0)) {
    case 12: return launch_ft<6, float>(A, nb, st);
    case 13: return launch_ft<6, double>(A, nb, st);
    case 14: return launch_ft<7, float>(A, nb, st);
    case 15: return launch_ft<7, double>(A, nb, st);
    case 16: return launch_ft<8, float>(A, nb, st);
    case 17: return launch_ft<8, double>(A, nb, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace bsf
