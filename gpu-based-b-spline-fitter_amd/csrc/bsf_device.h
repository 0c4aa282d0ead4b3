// bsf_device.h -- gfx950 device building blocks shared by the fit kernels:
// B-spline basis and per-axis stencils, the ROI window geometry, the
// non-xy contraction (stage A), the per-pixel model (stage B), wave
// reductions and the lane-distributed FP64 solves.
#pragma once
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

// ----------------------------------------------------------- lane reads
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

// Waves that fit R ROIs, one per 64 / R lanes: a value of lane hb + i of
// the caller's ROI (hb its first lane, i the same for all of its lanes).
// R = 1 is v_readlane (a scalar); otherwise ds_bpermute (per lane).
template <int R>
__device__ __forceinline__ int roi_read_i(int v, int i, int hb)
{
    if constexpr (R == 1) return __builtin_amdgcn_readlane(v, i);
    else return __builtin_amdgcn_ds_bpermute((hb + i) << 2, v);
}

template <int R, typename E>
__device__ __forceinline__ E roi_read(E v, int i, int hb)
{
    if constexpr (R == 1) return readlane_e<E>(v, i);
    else if constexpr (sizeof(E) == 8) {
        const long long b = __double_as_longlong(v);
        const int lo = roi_read_i<R>((int)(b & 0xffffffffLL), i, hb), hi = roi_read_i<R>((int)(b >> 32), i, hb);
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    } else return __int_as_float(roi_read_i<R>(__float_as_int(v), i, hb));
}

// Ballot over the caller's ROI's lanes: bit i = lane hb + i.
template <int R>
__device__ __forceinline__ unsigned long long roi_ballot(bool c, int hb)
{
    const unsigned long long b = __ballot(c);
    if constexpr (R == 1) return b;
    else return (b >> hb) & ((1ull << (64 / R)) - 1);
}

// ---------------------------------------------------------------- basis
// Centred cubic / quadratic B-splines (the reference's kernel_evalBSpline,
// CPUsplineLib.cpp:1289-1326, writes them as four / three polynomial pieces
// in x) in closed symmetric form without branches:
//   beta3(x) = 2/3 - x^2 + |x|^3/2 (|x| < 1), (2 - |x|)^3 / 6 (|x| < 2), else 0
//   beta2(x) = 3/4 - x^2 (|x| < 1/2), (3/2 - |x|)^2 / 2 (|x| < 3/2), else 0
// Equal to the pieces up to rounding, and more accurate near the ends of the
// support; every lane of make_geo evaluates a different piece, so branches
// would run all of them.
template <typename E>
__device__ __forceinline__ E bsp3s(E x)
{
    const E a = fabs(x), a2 = a * a, t = E(2) - a;
    const E inner = (E(2) / E(3) - a2) + E(0.5) * a2 * a;
    const E outer = t * t * t * (E(1) / E(6));
    return a < E(1) ? inner : (a < E(2) ? outer : E(0));
}

template <typename E>
__device__ __forceinline__ E bsp2s(E x)
{
    const E a = fabs(x), t = E(1.5) - a;
    const E inner = E(0.75) - a * a;
    const E outer = E(0.5) * t * t;
    return a < E(0.5) ? inner : (a < E(1.5) ? outer : E(0));
}

// Per-axis factors for fractional part d (the reference's B1/B2/dB1/dB2 at
// idx 1, 2) and the 4-tap stencils over taps k-2..k+1:
//   W  = value weights, Dv = derivative weights (forward differences with
//   quadratic weights rewritten as a stencil on the coefficients).
// (Filled by make_geo.)
template <typename E>
struct Axis {
    E B1[3], B2[3], dB1[3], dB2[3];
    E W[4], Dv[4];
};

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
    int CS;            // plane column stride (PR, or S + 4 in size-specialised kernels)
};

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8) and each XCD has its own L2.  This gives each XCD a contiguous range
// of ROIs, so the column-major output stores of neighbouring ROIs meet in one
// L2 and are written back as full lines.  A bijection of [0, G).
__device__ __forceinline__ int xcd_block(int b, int G)
{
    const int x = b & 7, q = b >> 3, base = G >> 3, rem = G & 7;
    return x * base + (x < rem ? x : rem) + q;
}

struct Dev {
    const float *data, *coeff, *coeff3;
    const RoiDesc *rois;   // per-ROI placement, or NULL: uniform S, ROI n at data + n*S*S
    float *P, *CRLB, *LL;
    int N, S, iterations, roi_offset;
    const int *ncount;     // device: ROIs to fit (fewer than N), or NULL: all N
    int nd[5];
    long s1, s2, s3, s4;   // coefficient strides of dims 1..4
    int cbytes, c3bytes;   // coefficient array sizes (buffer record counts)
    const float *theta_in; // test hook: evaluate once at these parameters
    float *dbg;            // test hook output: err, J[NV], H[NV*NV] per ROI
    unsigned char *status; // per-ROI status flags (bsf_roi_flag), or NULL
    int *evals;            // per-ROI count of evaluations run (the final Fisher pass included), or NULL
    float initP[4];
    int has_initP;
    int keep;              // 0: write the results; 1 / 2: write them only where the log-likelihood
                           // beats the one in LL by >= / > 1e-4 (a later start of a multi-start fit)
};

// LDS layout of a fit workgroup (one wave): ROI data (nch channels), pixel
// (row, column) table, per-row / per-column tap offsets, axis factors, then
// the planes (reused as reduction scratch) and kernel-specific extras.
struct LdsLayout {
    size_t rdat, rtap, ctap, fac, hj, pl, total;
};

__host__ __device__ inline size_t lds_align16(size_t x) { return (x + 15) & ~(size_t)15; }

__host__ __device__ inline LdsLayout lds_layout(int S, int nch, size_t esz, size_t pl_elems, size_t hj_elems = 0)
{
    LdsLayout L;
    size_t o = lds_align16((size_t)nch * S * S * 4);
    L.rdat = o;                                  // 1/data per pixel (FP32 evaluation only)
    o = lds_align16(o + (esz == 4 ? (size_t)nch * S * S * 4 : 0));
    L.rtap = o;
    o = lds_align16(o + (size_t)S * 4);
    L.ctap = o;
    o = lds_align16(o + (size_t)S * 4);
    L.fac = o;
    o = lds_align16(o + 40 * esz);
    L.hj = o;                                    // J and packed H of the last evaluation
    o = lds_align16(o + hj_elems * esz);
    L.pl = o;
    L.total = o + esz * pl_elems;
    return L;
}

// LDS layout of the fit kernel (fittypes 6-8, one ROI per wave): [ROI data,]
// 1/data per pixel (FP32 evaluation only), tap tables, J / H, then the
// planes.  Fittypes 6 / 7 keep the ROI data in global memory (each pixel pass
// reads its pixel; they stay cached while the wave fits the ROI), and the
// per-axis factors and the ROI copy the initial values are computed from live
// in the plane region (both are read back before stage A writes the planes):
// for fittype 7 at S = 15 this is 10.2 KB, so 16 waves fit a CU (11.0 KB: 14).
// Fittype 8 runs 2 waves per SIMD (registers), so LDS does not bind it and
// it reads the data from LDS (global reads measured 2 % slower).
__host__ __device__ constexpr bool fit_data_in_lds(int ft) { return ft == 8; }

__host__ __device__ inline LdsLayout fit_lds_layout(int S, size_t esz, size_t pl_elems, size_t hj_elems,
                                                    bool data_in_lds)
{
    LdsLayout L;
    size_t o = data_in_lds ? lds_align16((size_t)S * S * 4) : 0;
    L.rdat = o;
    o = lds_align16(o + (esz == 4 ? (size_t)S * S * 4 : 0));
    L.rtap = o;
    o = lds_align16(o + (size_t)S * 4);
    L.ctap = o;
    o = lds_align16(o + (size_t)S * 4);
    L.hj = o;
    o = lds_align16(o + hj_elems * esz);
    L.pl = o;
    L.fac = o;
    size_t need = pl_elems * esz;
    need = need > 40 * esz ? need : 40 * esz;
    need = need > (size_t)S * S * 4 ? need : (size_t)S * S * 4;
    L.total = o + need;
    return L;
}

// Per-wave LDS workspace of one evaluation's geometry.
template <typename E>
struct GeoLds {
    E *fac;      // [8*D] per-axis factors B1[1],B1[2],B2[1],B2[2],dB1[1],dB1[2],dB2[1],dB2[2]
    int *rtap;   // [S] window row offset of tap k-2, per ROI row
    int *ctap;   // [S] window column offset of tap k-2, per ROI column
};

// Geometry of one evaluation, computed across the wave: lane 8a+f evaluates
// factor f of axis a (the reference's per-axis B-spline values, in its
// operation order), lane i evaluates the clamped taps of ROI row / column i
// (CPUsplineLib.cpp:1203-1225, per pixel exactly as the reference), and the
// window spans the taps of all rows / columns.  Every lane must call it.
template <typename E, int FT, int R = 1>
__device__ __forceinline__ void make_geo(const Dev &A, const E *cth, Geo<E, FT> &g, int lane, const GeoLds<E> &L,
                                         bool fixed_cs = false, int hb = 0)
{
    constexpr int D = Ft<FT>::D;
    const int S = A.S;
    const E off = (E)(A.nd[0] - S) / E(2);
    const E colc = (E)((S + 1) / 2) - cth[0];
    const E rowc = (E)((S + 1) / 2) - cth[1];
    // 8 factors per axis: lanes 0 .. 8D-1 of the ROI (fittype 8 in a
    // two-ROI wave has 32 lanes for its 40: the first 8 take two)
#pragma unroll
    for (int fl = lane; fl < 8 * D; fl += 64 / R) {
        const int a = fl >> 3, f = fl & 7, idx = (f & 1) + 1;
        E c = rowc;
        c = (a == 1) ? colc : c;
#pragma unroll
        for (int i = 2; i < D; i++) c = (a == i) ? cth[2 + i] : c;
        const E d = c - floor(c), dd = d - E(1);
        // f 0,1: beta3(d + idx - 1); 2,3: beta3(d - idx);
        //   4,5: beta2(dd + idx - 1/2); 6,7: beta2(dd - idx + 1/2)
        const E x3 = (f < 2) ? d + (E)(idx - 1) : d - (E)idx;
        const E x2 = (f < 6) ? dd + ((E)idx - E(0.5)) : dd - ((E)idx - E(0.5));
        L.fac[fl] = (f < 4) ? bsp3s<E>(x3) : bsp2s<E>(x2);
    }
    // per row / column taps; the sequence is non-decreasing except where a
    // coordinate equals n exactly (tap n+1, the maximum; all later ones are n)
    const int li = lane < S ? lane : S - 1;
    const E rc = rowc + (E)li + off, cc = colc + (E)li + off;
    const int n1 = A.nd[1], n0 = A.nd[0];
    const int kr = base_tap<E>(rc, n1), kq = base_tap<E>(cc, n0);
    // (lane is the lane within the ROI's lanes, hb their first lane in the wave)
    const unsigned long long sr = roi_ballot<R>(lane < S && rc == (E)n1, hb);
    const unsigned long long sq = roi_ballot<R>(lane < S && cc == (E)n0, hb);
    const int kr0 = roi_read_i<R>(kr, 0, hb), krl = roi_read_i<R>(kr, S - 1, hb);
    const int kq0 = roi_read_i<R>(kq, 0, hb), kql = roi_read_i<R>(kq, S - 1, hb);
    int rmin = kr0, rmax = krl, cmin = kq0, cmax = kql;
    if (sr) {
        rmax = n1 + 1;
        rmin = (sr & 1ull) ? (S > 1 ? n1 : n1 + 1) : kr0;
    }
    if (sq) {
        cmax = n0 + 1;
        cmin = (sq & 1ull) ? (S > 1 ? n0 : n0 + 1) : kq0;
    }
    g.row0 = rmin - 2;
    g.col0 = cmin - 2;
    g.PR = rmax - rmin + 4;
    // planes column stride: the window height, or with a compile-time ROI
    // size its largest value S + 4, so that plane offsets are constants
    g.CS = fixed_cs ? S + 4 : g.PR;
    g.PC = cmax - cmin + 4;
    if (lane < S) {
        L.rtap[lane] = kr - rmin;
        L.ctap[lane] = kq - cmin;
    }
#pragma unroll
    for (int a = 2; a < D; a++) g.knx[a - 2] = base_tap<E>(cth[2 + a], A.nd[a]);
    __syncthreads();
#pragma unroll
    for (int a = 0; a < D; a++) {
        Axis<E> &x = g.ax[a];
        const E *f = L.fac + 8 * a;
        x.B1[1] = f[0];
        x.B1[2] = f[1];
        x.B2[1] = f[2];
        x.B2[2] = f[3];
        x.dB1[1] = f[4];
        x.dB1[2] = f[5];
        x.dB2[1] = f[6];
        x.dB2[2] = f[7];
        x.W[0] = x.B1[2];
        x.W[1] = x.B1[1];
        x.W[2] = x.B2[1];
        x.W[3] = x.B2[2];
        // sum_t (c[t+1]-c[t]) dW[t], dW = (dB1[2], dB1[1], dB2[1], dB2[2]);
        // dB2[2] = beta2(d - 2.5) = 0 for d in [0,1), so tap k+2 drops out.
        x.Dv[0] = -x.dB1[2];
        x.Dv[1] = x.dB1[2] - x.dB1[1];
        x.Dv[2] = x.dB1[1] - x.dB2[1];
        x.Dv[3] = x.dB2[1] - x.dB2[2];
    }
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

// Contract the non-xy axes over the window; planes[p][b*CS + a].
// R ROIs per wave: lane is the lane within the ROI, the window points are
// dealt over the ROI's 64 / R lanes, and the ROI's tap offset (uniform for
// the whole wave when R = 1, then an SGPR) joins the per-lane offset.
template <typename E, int FT, int R = 1>
__device__ __forceinline__ void stage_a(const Dev &A, const Geo<E, FT> &g, E *planes, int plane_stride, int lane)
{
    constexpr int K = Ft<FT>::K;
    const int PR = g.PR, npts = g.PR * g.PC;
    const __amdgpu_buffer_rsrc_t rc = __builtin_amdgcn_make_buffer_rsrc((void *)A.coeff, (short)0, A.cbytes, 0x00020000);
    const int s2 = (int)A.s2 * 4, s3 = (int)A.s3 * 4, s4 = (int)A.s4 * 4;
    int tapb = (g.knx[0] - 2) * s2;
    if constexpr (K >= 2) tapb += (g.knx[1] - 2) * s3;
    if constexpr (K >= 3) tapb += (g.knx[2] - 2) * s4;
    if constexpr (R == 1) tapb = __builtin_amdgcn_readfirstlane(tapb);
    const int vtap = R == 1 ? 0 : tapb;   // per-lane part of the tap offset
    if constexpr (R != 1) tapb = 0;
    // window point q = b*PR + a without an integer divide: for q < 2^12 the
    // FP32 quotient (q + 0.5) / PR stays within 1e-4 of its exact value,
    // which is at least 0.5/PR away from an integer, so the floor is exact
    const float invPR = 1.0f / (float)PR;
#pragma nounroll
    for (int q = lane; q < npts; q += 64 / R) {
        const int b = (int)(((float)q + 0.5f) * invPR), a = q - b * PR;
        const int pq = b * g.CS + a;
        const int vo = ((g.row0 + a) + (g.col0 + b) * (int)A.s1) * 4 + vtap;
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
            planes[pq] = v;
            planes[plane_stride + pq] = dz;
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
            planes[pq] = v;
            planes[plane_stride + pq] = dz;
            planes[2 * plane_stride + pq] = dw;
        } else {
            // Slice by slice along z (16 taps at a time) to bound live
            // registers.  The reference's 5D contraction has two irregular
            // terms in both its CPU and GPU code (CPUsplineLib.cpp:313 /
            // 420-421, GPUsplineLib.cuh:573 / 681), reproduced as two extra
            // planes applied later with the upper / lower 2x2 xy stencils:
            //  z: term 27 of block (q,r,s) weighs the z-difference at tap
            //     kz-q (w tap kw+r-1, u tap ku+s-1) with a weight taken from
            //     block +8, term 15 (past the z table: the w table);
            //  u: term 16 differences u tap ku-s instead of ku+s-1; with
            //     dB2[2] = 0 only s = 1 survives: weight * (2c[ku]-c[ku-1]-c[ku+1]).
            const Axis<E> &z = g.ax[2], &w = g.ax[3], &u = g.ax[4];
            E Vz[3][3][3], Vu[3][3];
#pragma unroll
            for (int q2 = 1; q2 <= 2; q2++)
#pragma unroll
                for (int r = 1; r <= 2; r++) {
#pragma unroll
                    for (int s = 1; s <= 2; s++) {
                        E bug;
                        if (r == 1) bug = z.dB2[q2] * w.B2[2] * u.B1[s];
                        else if (s == 1) bug = z.dB2[q2] * w.B2[1] * u.B1[2];
                        else bug = z.B2[q2] * w.dB2[1] * u.B1[1];
                        Vz[q2][r][s] = bug - z.dB1[q2] * w.B2[r] * u.B2[s];
                    }
                    Vu[q2][r] = z.B1[q2] * w.B1[r] * u.dB2[1];
                }
            E v = E(0), dz = E(0), dw = E(0), du = E(0), zb = E(0), ub = E(0);
            E prev[2][2];
            // the u tap goes into the per-lane offset (4 VGPRs), the (z, w)
            // taps into the wave-uniform one (16 SGPRs instead of 64)
            const int vu[4] = {vo, vo + s4, vo + 2 * s4, vo + 3 * s4};
#pragma unroll
            for (int tz = 0; tz < 4; tz++) {
                E cs[4][4];
#pragma unroll
                for (int tw = 0; tw < 4; tw++)
#pragma unroll
                    for (int tu = 0; tu < 4; tu++) cs[tw][tu] = (E)bload(rc, vu[tu], tapb + tz * s2 + tw * s3);
                E a00 = E(0), a01 = E(0), a10 = E(0);
#pragma unroll
                for (int tw = 0; tw < 4; tw++) {
                    E u0 = E(0), u1 = E(0);
#pragma unroll
                    for (int tu = 0; tu < 4; tu++) {
                        u0 += u.W[tu] * cs[tw][tu];
                        u1 += u.Dv[tu] * cs[tw][tu];
                    }
                    a00 += w.W[tw] * u0;
                    a01 += w.Dv[tw] * u0;
                    a10 += w.W[tw] * u1;
                }
                v += z.W[tz] * a00;
                dz += z.Dv[tz] * a00;
                dw += z.W[tz] * a01;
                du += z.W[tz] * a10;
                if (tz == 1 || tz == 2) {
                    const int q2 = 3 - tz;      // difference between slices tz-1 and tz
#pragma unroll
                    for (int r = 1; r <= 2; r++)
#pragma unroll
                        for (int s = 1; s <= 2; s++) zb += Vz[q2][r][s] * (cs[1 + r][1 + s] - prev[r - 1][s - 1]);
                }
                if (tz <= 1) {
                    const int q2 = 2 - tz;
#pragma unroll
                    for (int r = 1; r <= 2; r++) {
                        const int tw = 2 - r;
                        ub += Vu[q2][r] * (E(2) * cs[tw][2] - cs[tw][1] - cs[tw][3]);
                    }
                }
#pragma unroll
                for (int r = 0; r < 2; r++)
#pragma unroll
                    for (int s = 0; s < 2; s++) prev[r][s] = cs[2 + r][2 + s];
            }
            planes[pq] = v;
            planes[plane_stride + pq] = dz;
            planes[2 * plane_stride + pq] = dw;
            planes[3 * plane_stride + pq] = du;
            planes[4 * plane_stride + pq] = zb;
            planes[5 * plane_stride + pq] = ub;
            // the 3D model (coeff3) shares the xy window and the z taps (the
            // ROI's z tap: wave-uniform, or per lane in a two-ROI wave)
            const __amdgpu_buffer_rsrc_t r3 =
                __builtin_amdgcn_make_buffer_rsrc((void *)A.coeff3, (short)0, A.c3bytes, 0x00020000);
            int tap3 = (g.knx[0] - 2) * s2, vo3 = vo - vtap;
            if constexpr (R == 1) tap3 = __builtin_amdgcn_readfirstlane(tap3);
            else vo3 += tap3, tap3 = 0;
            E v3 = E(0), dz3 = E(0);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                E cc = (E)bload(r3, vo3, tap3 + t * s2);
                v3 += z.W[t] * cc;
                dz3 += z.Dv[t] * cc;
            }
            planes[6 * plane_stride + pq] = v3;
            planes[7 * plane_stride + pq] = dz3;
        }
    }
}

// --------------------------------------------------------- stage B + C
// Fittype 8: blend the 5D model (dudt, model) with the 3D one (f3 and its
// column, row and z derivatives) by g2 (CPUmleFit_LM.cpp:3318-3322), every
// product rounded on its own as the reference computes it -- so the one- and
// the two-pass stage B, whose code leaves the compiler different fusion
// choices, give the same bits.
template <typename E>
__device__ __forceinline__ void blend_5n3(E Nph, E bg, E g2, E f3, E dc3, E dr3, E dz3, E *dudt, E &model)
{
#pragma clang fp contract(off)
    E d3[5];
    d3[0] = E(-1) * Nph * dc3;
    d3[1] = E(-1) * Nph * dr3;
    d3[2] = f3;
    d3[3] = E(1);
    d3[4] = Nph * dz3;
    const E m3 = bg + Nph * f3;
    const E m5 = model;
#pragma unroll
    for (int l = 0; l < 5; l++) dudt[l] = dudt[l] * g2 + (E(1) - g2) * d3[l];
    model = m3 * (E(1) - g2) + m5 * g2;
    dudt[7] = m5 - m3;
}

// Model and dudt at one pixel from the planes.
template <typename E, int FT>
__device__ __forceinline__ void pixel_model(const Geo<E, FT> &g, const E *planes, int plane_stride, int r, int c,
                                            E Nph, E bg, E g2, E *dudt, E &model)
{
    constexpr int D = Ft<FT>::D;
    const int CS = g.CS;   // plane column stride
    const Axis<E> &ay = g.ax[0], &ax = g.ax[1];
    const E *p0 = planes + c * CS + r;
    // value plane: f, d/dcol, d/drow
    E f = E(0), dcol = E(0), drow = E(0);
#pragma unroll
    for (int ty = 0; ty < 4; ty++) {
        E sw = E(0), sd = E(0);
#pragma unroll
        for (int tx = 0; tx < 4; tx++) {
            E gv = p0[tx * CS + ty];
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
            for (int tx = 0; tx < 4; tx++) sw += ax.W[tx] * pk[tx * CS + ty];
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
                zb += ay.W[2 + i] * ax.W[2 + j] * pz[(2 + j) * CS + 2 + i];
                ub += ay.W[i] * ax.W[j] * pu[j * CS + i];
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
                E gv = q0[tx * CS + ty];
                sw += ax.W[tx] * gv;
                sd += ax.Dv[tx] * gv;
                sz += ax.W[tx] * q1[tx * CS + ty];
            }
            f3 += ay.W[ty] * sw;
            dc3 += ay.W[ty] * sd;
            dr3 += ay.Dv[ty] * sw;
            dz3 += ay.W[ty] * sz;
        }
        blend_5n3<E>(Nph, bg, g2, f3, dc3, dr3, dz3, dudt, model);
    }
}

// Stage B in two passes.  The xy stencil of pixel (i, j) sums over window
// columns (tx) and then window rows (ty); the column sums depend only on the
// pixel column j and the window row, so the first pass forms them once per
// (j, row) -- shared by the S pixels of column j -- and the second pass
// applies the row weights per pixel: 16 + 4 multiply-adds per value instead
// of 16 + 4 per pixel and row tap.  Both passes keep pixel_model's
// operation order, so every value is the same as there.
// Column sums of item (j, row) at h[((j * Wd) + row) * NHP + m], Wd = S + 4,
// in LDS after the planes (NHP values per item, 16-byte groups):
//   ft6/7 (NHP 4): W.p0, Dv.p0, W.pk (k = 1..K)      [ft6: one pad]
//   ft8   (NHP 8): W.p0, Dv.p0, W.p1, W.p2, W.p3, W.p6, Dv.p6, W.p7
// (W. = sum_tx W[tx] x, Dv. likewise; the irregular 2x2 planes 4, 5 of ft8
// stay per pixel.)
// Where it is used: FP32 fittypes 6/7 always (a kernel that also holds the
// one-pass form needs more than 128 VGPRs for fittype 7); FP64 fittypes 6/7
// up to S = 15 (2 waves per SIMD: the extra LDS leaves occupancy as it is);
// FP64 fittype 8 nowhere (kSplit8MaxF64), though its kernel holds the form
// for the registers it saves.
// FP32 fittype 8 keeps one pass: two gained nothing at S <= 13 and the
// kernel holding both lost 3 % at S >= 15.  Elsewhere pixel_model does one
// pass.  (Keeping the sums in registers and writing them over the planes
// instead needs ~20 more VGPRs for fittype 7 at S = 15: a wave per SIMD less.)
template <int FT>
struct ColSums { static constexpr int NHP = FT == 8 ? 8 : 4; };

// FP64 fittype 8 takes the two-pass form up to this ROI size: none.  Its
// kernel still holds that form (the size is a run-time value): with it the
// compiler allocates 306 instead of 382 registers, and 256 with 49 VGPRs
// spilled instead of 155 in the two-wave variant.  Measured with the
// two-wave registers (40,000 ROIs, ms per launch, one pass / two passes,
// same bits): S = 9 20.2 / 20.0, 11 23.0 / 29.3, 12 27.5 / 36.4, 13 23.7 /
// 35.3.
constexpr int kSplit8MaxF64 = 0;

__host__ __device__ inline bool stage_b_split(int ft, bool f64, int S)
{
    if (ft == 8) return f64 && S <= kSplit8MaxF64;
    return !f64 || S <= 15;
}

// LDS elements of the column sums for ROI size S
__host__ __device__ inline size_t col_sum_elems(int ft, int S) { return (size_t)(ft == 8 ? 8 : 4) * S * (S + 4); }

template <typename E, int FT>
__device__ __forceinline__ void col_sums(const Geo<E, FT> &g, const E *planes, int plane_stride, const int *ctap,
                                         int Wd, float invWd, int q, E *hv)
{
    constexpr int K = Ft<FT>::K;
    const int PR = g.PR;
    const Axis<E> &ax = g.ax[1];
    const int j = (int)(((float)q + 0.5f) * invWd);
    int a = q - j * Wd;
    a = a < PR ? a : PR - 1;   // rows past the window: a copy of the last (never read)
    const int CS = g.CS;
    const E *p0 = planes + ctap[j] * CS + a;
    auto wd = [&](const E *pp, E &sw, E &sd) __attribute__((always_inline)) {
        sw = E(0);
        sd = E(0);
#pragma unroll
        for (int tx = 0; tx < 4; tx++) {
            E gv = pp[tx * CS];
            sw += ax.W[tx] * gv;
            sd += ax.Dv[tx] * gv;
        }
    };
    auto w = [&](const E *pp) __attribute__((always_inline)) -> E {
        E sw = E(0);
#pragma unroll
        for (int tx = 0; tx < 4; tx++) sw += ax.W[tx] * pp[tx * CS];
        return sw;
    };
    wd(p0, hv[0], hv[1]);
#pragma unroll
    for (int k = 0; k < K; k++) hv[2 + k] = w(p0 + (k + 1) * plane_stride);
    if constexpr (FT == 8) {
        wd(p0 + 6 * plane_stride, hv[5], hv[6]);
        hv[7] = w(p0 + 7 * plane_stride);
    } else if constexpr (K == 1) {
        hv[3] = E(0);
    }
}

template <typename E>
struct alignas(16) Vec4 { E v[4]; };

// First pass.  Every lane must call it; it ends with the sums visible to
// the wave.
template <typename E, int FT, int R = 1>
__device__ __forceinline__ void stage_b_cols(const Geo<E, FT> &g, const E *planes, int plane_stride, const int *ctap,
                                             int S, float invWd, E *h, int lane)
{
    constexpr int NHP = ColSums<FT>::NHP;
    const int Wd = S + 4, items = S * Wd;
#pragma nounroll
    for (int q = lane; q < items; q += 64 / R) {
        E hv[NHP];
        col_sums<E, FT>(g, planes, plane_stride, ctap, Wd, invWd, q, hv);
        Vec4<E> *dst = (Vec4<E> *)(h + (size_t)q * NHP);
#pragma unroll
        for (int m = 0; m < NHP / 4; m++) dst[m] = Vec4<E>{{hv[4 * m], hv[4 * m + 1], hv[4 * m + 2], hv[4 * m + 3]}};
    }
    __syncthreads();
}

// Second pass: model and dudt at pixel (row r's taps, column j) from the
// column sums (hj = h + j * Wd * NHP), and for fittype 8 the irregular 2x2
// planes (pz, pu at the pixel's window origin, row stride 1, column stride CS).
template <typename E, int FT>
__device__ __forceinline__ void pixel_model_h(const Geo<E, FT> &g, const E *hj, int r, const E *pz, const E *pu,
                                              E Nph, E bg, E g2, E *dudt, E &model)
{
    constexpr int K = Ft<FT>::K, NHP = ColSums<FT>::NHP;
    const Axis<E> &ay = g.ax[0];
    E hv[4][NHP];
#pragma unroll
    for (int ty = 0; ty < 4; ty++) {
        const Vec4<E> *src = (const Vec4<E> *)(hj + (size_t)(r + ty) * NHP);
#pragma unroll
        for (int m = 0; m < NHP / 4; m++) {
            const Vec4<E> t = src[m];
#pragma unroll
            for (int i = 0; i < 4; i++) hv[ty][4 * m + i] = t.v[i];
        }
    }
    E f = E(0), dcol = E(0), drow = E(0);
#pragma unroll
    for (int ty = 0; ty < 4; ty++) {
        f += ay.W[ty] * hv[ty][0];
        dcol += ay.W[ty] * hv[ty][1];
        drow += ay.Dv[ty] * hv[ty][0];
    }
    E dn[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        E s = E(0);
#pragma unroll
        for (int ty = 0; ty < 4; ty++) s += ay.W[ty] * hv[ty][2 + k];
        dn[k] = s;
    }
    if constexpr (FT == 8) {
        const int CS = g.CS;   // plane column stride
        const Axis<E> &ax = g.ax[1];
        E zb = E(0), ub = E(0);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                zb += ay.W[2 + i] * ax.W[2 + j] * pz[(2 + j) * CS + 2 + i];
                ub += ay.W[i] * ax.W[j] * pu[j * CS + i];
            }
        dn[0] += zb;
        dn[2] += ub;
    }
    dudt[0] = E(-1) * Nph * dcol;
    dudt[1] = E(-1) * Nph * drow;
    dudt[2] = f;
    dudt[3] = E(1);
#pragma unroll
    for (int k = 0; k < K; k++) dudt[4 + k] = Nph * dn[k];
    model = bg + Nph * f;
    if constexpr (FT == 8) {
        E f3 = E(0), dc3 = E(0), dr3 = E(0), dz3 = E(0);
#pragma unroll
        for (int ty = 0; ty < 4; ty++) {
            f3 += ay.W[ty] * hv[ty][5];
            dc3 += ay.W[ty] * hv[ty][6];
            dr3 += ay.Dv[ty] * hv[ty][5];
            dz3 += ay.W[ty] * hv[ty][7];
        }
        blend_5n3<E>(Nph, bg, g2, f3, dc3, dr3, dz3, dudt, model);
    }
}

// Poisson deviance term of one pixel, 2[(mu - d) - d ln(mu/d)] for d > 0 and
// 2 mu otherwise (d <= 0 counts as 0, as the reference), without a branch.
// FP32: mu * (1/d) with 1/d precomputed per pixel and the hardware log2 (the
// ratio is never denormal for a fit's models); FP64: the exact expression.
template <typename E>
__device__ __forceinline__ E deviance_term(E model, E data, float rdat, E &d0)
{
    const bool pos = data > E(0);
    d0 = pos ? data : E(0);
    E lq;
    if constexpr (sizeof(E) == 8) lq = log(pos ? model / data : E(1));
    else lq = __builtin_amdgcn_logf(pos ? model * rdat : 1.0f) * 0.69314718055994531f;
    return E(2) * ((model - d0) - d0 * lq);
}

// ------------------------------------------------------------- LM solve
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
// A pivot that is exactly zero belongs to a parameter without information
// (every dudt of it is exactly 0: a coordinate on a mirror-padded bound of
// the spline, where the derivative vanishes by symmetry): its row and column
// are zero, so it decouples; its CRLB is +inf and the others come from the
// rest of the matrix.  (The reference divides by its rounding residue there
// and reports 1e7-1e31 for that parameter, the others as here.)
template <int NV, int R = 1>
__device__ __forceinline__ double inverse_diag_rows(double (&m)[NV], int lane, int hb = 0)
{
    double v[NV];
    bool none = false;   // this lane's parameter has no information
#pragma unroll
    for (int j = 0; j < NV; j++) v[j] = (j == lane) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const double mkk = roi_read<R, double>(m[k], k, hb);
        const bool zp = mkk == 0.0;
        none = none || (zp && lane == k);
        const double r = zp ? 0.0 : 1.0 / mkk;
        const double fac = zp ? 0.0 : m[k] * r;
#pragma unroll
        for (int j = 0; j < NV; j++) {
            const double mk = roi_read<R, double>(m[j], k, hb), vk = roi_read<R, double>(v[j], k, hb);
            m[j] = (lane == k) ? (zp ? m[j] : mk * r) : m[j] - fac * mk;
            v[j] = (lane == k) ? (zp ? v[j] : vk * r) : v[j] - fac * vk;
        }
    }
    double d = 0.0;
#pragma unroll
    for (int j = 0; j < NV; j++) d = (j == lane) ? v[j] : d;
    return none ? __builtin_inf() : d;
}

// Index of (row i, column j) in the packed upper triangle used by the sums.
template <int NV>
__device__ __forceinline__ int packed_index(int i, int j)
{
    const int a = i < j ? i : j, b = i < j ? j : i;
    return a * NV - (a * (a - 1)) / 2 + (b - a);
}

// Damped normal equations H x = J in FP64, the same in every lane: J and the
// packed upper triangle of H come from LDS (jh[0..NV) = J, jh[NV..) = H, the
// evaluation's sums held exactly in FP64).
// Square-root-free LDL^T, right-looking.  Its pivots are the reference
// Cholesky's radicands (CPUsplineLib.cpp:1363-1388): the solve fails, as
// there, when one is negative.  Returns 0 and x on success.
// 1/d in FP64: v_rcp_f64 and two Newton steps (within an ulp for the
// normal pivots here); 1/0 = +inf as a divide gives.
__device__ __forceinline__ double rcp_f64(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    return d == 0.0 ? __builtin_inf() : r;
}

// A parameter whose row of the system (and gradient entry) is exactly zero
// has no information in this evaluation: every dudt of it is 0, which
// happens exactly on the mirror-padded first sample of a spline axis, where
// the fits' lower bounds clamp z (0), the colour w (0) and the polar u (1).
// The reference's arithmetic leaves rounding residue there (|J| ~ 1e-16,
// H ~ 1e-30), so its Newton step for that parameter is a ratio of residues
// that the limiter caps at the parameter's clamp, and the bound keeps only
// the steps leaving it: its fits walk off the bound.  Here such a row is
// solved out (the rest of the system as it is) and the parameter gets
// kEscapeStep, which the limiter turns into a step of its clamp upward, off
// the bound; without it every later solve would fail and the fit freeze.
constexpr double kEscapeStep = -1e30;

// Rows of the damped system without information, as a bit mask the same in
// every lane.  H_kk is a sum of non-negative terms t2 dudt_k^2, so H_kk == 0
// means dudt_k vanishes at every pixel with weight and the whole row is zero;
// J_k == 0 as well (otherwise the pixels without weight still see the
// parameter).  jh: J[NV] then the packed H, as ldlt_solve takes them.
template <int NV, int R = 1>
__device__ __forceinline__ unsigned zero_rows(const double *jh)
{
    unsigned zr = 0;
#pragma unroll
    for (int k = 0; k < NV; k++)
        zr |= (jh[NV + packed_index<NV>(k, k)] == 0.0 && jh[k] == 0.0) ? 1u << k : 0u;
    if constexpr (R == 1) return __builtin_amdgcn_readfirstlane(zr);
    else return zr;   // the same in the ROI's lanes, different between ROIs
}

// zr: rows without information (zero_rows) -- their diagonal is taken as 1,
// so that they decouple and their unknown comes out 0.
template <int NV>
__device__ __forceinline__ int ldlt_solve(const double *jh, double (&x)[NV], unsigned zr = 0)
{
    double a[NV][NV];
#pragma unroll
    for (int i = 0; i < NV; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) a[i][j] = jh[NV + packed_index<NV>(j, i)];
#pragma unroll
    for (int i = 0; i < NV; i++) a[i][i] = ((zr >> i) & 1u) ? 1.0 : a[i][i];
    double r[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const double d = a[k][k];
        if (!(d >= 0.0)) return 1;
        r[k] = rcp_f64(d);
        double t[NV];
#pragma unroll
        for (int i = k + 1; i < NV; i++) {
            t[i] = a[i][k];              // l_ik d_k
            a[i][k] = t[i] * r[k];       // l_ik
        }
#pragma unroll
        for (int i = k + 1; i < NV; i++)
#pragma unroll
            for (int j = k + 1; j <= i; j++) a[i][j] -= a[i][k] * t[j];
    }
#pragma unroll
    for (int i = 0; i < NV; i++) {       // L y = J, then z = D^-1 y
        double y = jh[i];
#pragma unroll
        for (int k = 0; k < i; k++) y -= a[i][k] * x[k];
        x[i] = y;
    }
#pragma unroll
    for (int i = 0; i < NV; i++) x[i] *= r[i];
#pragma unroll
    for (int i = NV - 1; i >= 0; i--) {  // L^T x = z
#pragma unroll
        for (int k = i + 1; k < NV; k++) x[i] -= a[k][i] * x[k];
    }
    return 0;
}

// Per-ROI status flags (enum bsf_roi_flag of bsfit.h) of a finished fit:
// 1 the iteration limit ended it, 2 its last damped solve failed, 4 a CRLB
// entry is not a finite positive number, 8 a parameter or the log-likelihood
// is not finite, 16 the damped solve of its first iteration failed.  Every lane of the wave calls it; `own` marks the lanes that
// hold a parameter (th) and its CRLB.
__device__ __forceinline__ void write_roi_status(unsigned char *status, int n, bool own, float th, float crlb,
                                                 float ll, bool max_iter, bool solve_failed, bool first_failed,
                                                 int lane)
{
    const unsigned long long badc = __ballot(own && !(crlb > 0.f && isfinite(crlb)));
    const unsigned long long badp = __ballot(own && !isfinite(th));
    if (lane == 0) {
        unsigned f = 0;
        if (max_iter) f |= 1u;
        if (solve_failed) f |= 2u;
        if (first_failed) f |= 16u;
        if (badc) f |= 4u;
        if (badp || !isfinite(ll)) f |= 8u;
        status[n] = (unsigned char)f;
    }
}

// ROI pixel p (column-major, p = jj*S + ii) to row ii and column jj without
// an integer divide: exact for p < 2^12, as the window decode of stage_a.
__device__ __forceinline__ void pixel_rc(int p, float invS, int S, int &ii, int &jj)
{
    jj = (int)(((float)p + 0.5f) * invS);
    ii = p - jj * S;
}

// Lane-permuting moves within quads (DPP quad_perm), FP32 and FP64.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
constexpr int kQuadXor1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;   // quad_perm [2,3,0,1]

// Reduction scratch: 16 values x 65 (transposed, padded) plus the totals.
constexpr int kRedChunk = 16;
__host__ __device__ constexpr int red_elems(int V) { return kRedChunk * 65 + V; }


// LDS bytes of one ROI of a fit workgroup (fittypes 6-8) whose largest ROI
// has size S, a multiple of 16: the planes, the column sums of the two-pass
// stage B after them and the reduction scratch over both, with the ROI data,
// tap tables and J / H before them (fit_lds_layout).  A wave that fits R
// ROIs holds R such regions back to back.
__host__ __device__ inline size_t fit_lds_bytes(int fittype, int f64, int S)
{
    const int NV = fittype == 6 ? 5 : fittype == 7 ? 6 : 8;
    const int NPL = fittype == 6 ? 2 : fittype == 7 ? 3 : 8;
    const size_t W = S + 4;
    const size_t pl = (size_t)NPL * W * W;
    const size_t h = col_sum_elems(fittype, S) + 3;   // after the planes, 16-byte aligned
    const size_t red = (size_t)red_elems(1 + NV + NV * (NV + 1) / 2);
    const size_t e0 = (fittype != 8 && !f64) || stage_b_split(fittype, f64, S) ? pl + h : pl;
    const size_t e = e0 > red ? e0 : red;
    return lds_align16(fit_lds_layout(S, f64 ? 8 : 4, e, (NV + NV * (NV + 1) / 2) * (f64 ? 1 : 2),
                                      fit_data_in_lds(fittype))
                           .total);
}

// Sum V per-lane values across the wave, 16 values at a time: the chunk is
// written transposed to LDS, lane 4v+q adds entries 16q..16q+15 of value v,
// and the four quarters combine by quad permutes (the same total in all four
// lanes).  The totals land in scratch[16*65 + k] for every lane to read.
template <typename E, int V>
__device__ __forceinline__ void wave_sum_lds(const E (&v)[V], E *scratch, int lane)
{
    E *tot = scratch + kRedChunk * 65;
    const int vi = lane >> 2, q4 = lane & 3;
#pragma unroll
    for (int k0 = 0; k0 < V; k0 += kRedChunk) {
#pragma unroll
        for (int k = k0; k < V && k < k0 + kRedChunk; k++) scratch[(k - k0) * 65 + lane] = v[k];
        __syncthreads();
        const int C = (V - k0) < kRedChunk ? (V - k0) : kRedChunk;
        E s0 = E(0), s1 = E(0);
        if (vi < C) {
            const E *row = scratch + vi * 65 + q4 * 16;
#pragma unroll
            for (int j = 0; j < 16; j += 2) {
                s0 += row[j];
                s1 += row[j + 1];
            }
        }
        E t = s0 + s1;
        t += dpp_move<kQuadXor1>(t);
        t += dpp_move<kQuadXor2>(t);
        if (vi < C && q4 == 0) tot[k0 + vi] = t;
        __syncthreads();
    }
}

// wave_sum_lds for a wave that fits two ROIs, one per 32 lanes: lane l of a
// half holds the partial sums of the one-ROI kernel's lanes l (v0) and
// l + 32 (v1), and each half reduces its 64 virtual lanes in wave_sum_lds's
// order, so that the totals are the same bits.  scratch is the half's own.
template <typename E, int V>
__device__ __forceinline__ void wave_sum_lds2(const E (&v0)[V], const E (&v1)[V], E *scratch, int hl)
{
    E *tot = scratch + kRedChunk * 65;
    const int vi = hl >> 2, q4 = hl & 3;
#pragma unroll
    for (int k0 = 0; k0 < V; k0 += kRedChunk) {
#pragma unroll
        for (int k = k0; k < V && k < k0 + kRedChunk; k++) {
            scratch[(k - k0) * 65 + hl] = v0[k];
            scratch[(k - k0) * 65 + hl + 32] = v1[k];
        }
        __syncthreads();
        const int C = (V - k0) < kRedChunk ? (V - k0) : kRedChunk;
#pragma unroll
        for (int half = 0; half < 2; half++) {
            const int vv = vi + 8 * half;
            E s0 = E(0), s1 = E(0);
            if (vv < C) {
                const E *row = scratch + vv * 65 + q4 * 16;
#pragma unroll
                for (int j = 0; j < 16; j += 2) {
                    s0 += row[j];
                    s1 += row[j + 1];
                }
            }
            E t = s0 + s1;
            t += dpp_move<kQuadXor1>(t);
            t += dpp_move<kQuadXor2>(t);
            if (vv < C && q4 == 0) tot[k0 + vv] = t;
        }
        __syncthreads();
    }
}

// wave_sum_lds for a wave that fits two ROIs with one sum per lane (fittype
// 8: lane l of a half holds the sums of the half's pixels l, l + 32, ...):
// each half reduces its 32 lanes, 8 values at a time (the chunk transposed
// to LDS, lane 4v+q adds entries 8q..8q+7 of value v, the quads combine).
// The totals land in scratch[16*65 + k], as wave_sum_lds leaves them.
template <typename E, int V>
__device__ __forceinline__ void wave_sum_lds_half(const E (&v)[V], E *scratch, int hl)
{
    E *tot = scratch + kRedChunk * 65;
    const int vi = hl >> 2, q4 = hl & 3;
#pragma unroll
    for (int k0 = 0; k0 < V; k0 += 8) {
#pragma unroll
        for (int k = k0; k < V && k < k0 + 8; k++) scratch[(k - k0) * 33 + hl] = v[k];
        __syncthreads();
        const int C = (V - k0) < 8 ? (V - k0) : 8;
        E s0 = E(0), s1 = E(0);
        if (vi < C) {
            const E *row = scratch + vi * 33 + q4 * 8;
#pragma unroll
            for (int j = 0; j < 8; j += 2) {
                s0 += row[j];
                s1 += row[j + 1];
            }
        }
        E t = s0 + s1;
        t += dpp_move<kQuadXor1>(t);
        t += dpp_move<kQuadXor2>(t);
        if (vi < C && q4 == 0) tot[k0 + vi] = t;
        __syncthreads();
    }
}

// 1/x for the per-pixel weights: the hardware reciprocal (1 ulp) in FP32, a
// correctly rounded divide in FP64.
template <typename E>
__device__ __forceinline__ E recip(E x)
{
    if constexpr (sizeof(E) == 8) return 1.0 / x;
    else return __builtin_amdgcn_rcpf(x);
}

template <typename E>
__device__ __forceinline__ E dmaxr(E a, double b) { return ((double)a > b) ? a : (E)b; }
template <typename E>
__device__ __forceinline__ E dminr(E a, double b) { return ((double)a < b) ? a : (E)b; }

// Runtime-size variants (nv <= NVMAX) of the lane-distributed solves, for
// fitters whose parameter count depends on the call (the 4Pi shared mask).
template <int NVMAX>
__device__ __forceinline__ int chol_solve_rows_rt(double (&a)[NVMAX], double b, int lane, int nv, double &xi)
{
    double colL[NVMAX], rinv[NVMAX];
    // rows (and gradient entries) that are exactly zero: see kEscapeStep
    bool mine = lane < nv && b == 0.0;
#pragma unroll
    for (int j = 0; j < NVMAX; j++) mine = mine && (j >= nv || a[j] == 0.0);
    const unsigned long long zrows = __ballot(mine);
#pragma unroll
    for (int j = 0; j < NVMAX; j++) colL[j] = 0.0, rinv[j] = 0.0;
#pragma unroll
    for (int k = 0; k < NVMAX; k++) {
        if (k < nv) {
            const double akk = readlane_d(a[k], k);
            if (!(akk >= 0.0)) return 1;
            const double lkk = sqrt(akk);
            rinv[k] = ((zrows >> k) & 1ull) ? 0.0 : 1.0 / lkk;
            const double lik = lane > k ? a[k] * rinv[k] : (lane == k ? lkk : 0.0);
            a[k] = lik;
#pragma unroll
            for (int j = k + 1; j < NVMAX; j++) {
                if (j < nv) {
                    const double ljk = readlane_d(lik, j);
                    a[j] -= lik * ljk;
                    colL[j] = (lane == k) ? ljk : colL[j];
                }
            }
        }
    }
    double acc = 0.0, yv = 0.0;
#pragma unroll
    for (int k = 0; k < NVMAX; k++) {
        if (k < nv) {
            const double yk = readlane_d((b - acc) * rinv[k], k);
            yv = (lane == k) ? yk : yv;
            acc += a[k] * yk;
        }
    }
    double acc2 = 0.0;
    xi = 0.0;
#pragma unroll
    for (int k = NVMAX - 1; k >= 0; k--) {
        if (k < nv) {
            const double xk = readlane_d((yv - acc2) * rinv[k], k);
            xi = (lane == k) ? xk : xi;
            acc2 += colL[k] * xk;
        }
    }
    if ((zrows >> lane) & 1ull) xi = kEscapeStep;
    return 0;
}

template <int NVMAX>
__device__ __forceinline__ double inverse_diag_rows_rt(double (&m)[NVMAX], int lane, int nv)
{
    double v[NVMAX];
    bool none = false;   // a zero pivot: as in inverse_diag_rows
#pragma unroll
    for (int j = 0; j < NVMAX; j++) v[j] = (j == lane) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < NVMAX; k++) {
        if (k < nv) {
            const double mkk = readlane_d(m[k], k);
            const bool zp = mkk == 0.0;
            none = none || (zp && lane == k);
            const double r = zp ? 0.0 : 1.0 / mkk;
            const double fac = zp ? 0.0 : m[k] * r;
#pragma unroll
            for (int j = 0; j < NVMAX; j++) {
                if (j < nv) {
                    const double mk = readlane_d(m[j], k), vk = readlane_d(v[j], k);
                    m[j] = (lane == k) ? (zp ? m[j] : mk * r) : m[j] - fac * mk;
                    v[j] = (lane == k) ? (zp ? v[j] : vk * r) : v[j] - fac * vk;
                }
            }
        }
    }
    double d = 0.0;
#pragma unroll
    for (int j = 0; j < NVMAX; j++) d = (j == lane) ? v[j] : d;
    return none ? __builtin_inf() : d;
}

}  // namespace bsf
