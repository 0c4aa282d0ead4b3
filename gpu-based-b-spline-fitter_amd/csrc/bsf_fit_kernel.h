// bsf_fit_kernel.h -- the fit kernel template (fittypes 6: 3D, 7: 4D,
// 8: 5D+3D) and its launcher, shared by the translation unit of the generic
// kernels (bsf_fit_kernels.hip) and those of the size-specialised ones
// (bsf_fit_sized6.hip, bsf_fit_sized7.hip), which compile in parallel.
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
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bsf_kernels.h"

#include "bsf_device.h"

namespace bsf {

// Initial x, y, N, bg of one ROI (CPUgaussLib.cpp:172-223, CPUmleFit_LM.cpp:
// 2478-2482): centre of mass and max / min, summed in the reference's
// sequential order, each product and sum rounded on its own.
template <typename E>
__device__ __forceinline__ void roi_init4(const float *d, int S, E *v)
{
#pragma clang fp contract(off)
    E tx = E(0), ty = E(0), ts = E(0), mx = E(0), mn = E(10e10);
    // partly unrolled only: fully unrolled (a compile-time S) these loops
    // hold every loaded pixel in a register
#pragma unroll 2
    for (int ii = 0; ii < S; ii++)
        for (int jj = 0; jj < S; jj++) {
            E dd = (E)d[S * jj + ii];
            tx += dd * (E)ii;
            ty += dd * (E)jj;
            ts += dd;
        }
#pragma unroll 8
    for (int p = 0; p < S * S; p++) {
        E dd = (E)d[p];
        mx = (mx > dd) ? mx : dd;
        mn = (mn < dd) ? mn : dd;
    }
    v[0] = tx / ts;
    v[1] = ty / ts;
    double t = (double)(((mx - mn) * E(2)) * (E)3.141592) * 1.5 * 1.5;   // PI, definitions.h:15
    v[2] = (E)((0.0 > t) ? 0.0 : t);
    v[3] = dmaxr<E>(mn, 0.01);
}

// ----------------------------------------------------------------- kernel
// SC > 0: every ROI of the launch has size SC, so the index arithmetic and
// LDS offsets that depend on the size fold to constants; SC = 0: any size.
// R ROIs per wave: R = 1, one wave per ROI; R = 2, each half of the wave
// (32 lanes) fits its own ROI with its own LDS region and LM state, and the
// per-ROI work that does not spread over lanes (geometry, the damped solve,
// the LM control, the reductions' tails) runs once for both.  A half's lane
// holds the pixel sums of two lanes of the one-ROI kernel and reduces them in
// its order, so both forms give the same bits -- except FP32 fittype 8,
// whose lanes keep one set of sums over the half's pixels (HALF_SUMS below).
// The fit of the ROIs of workgroup `block` of `nblocks`; fit_kernel runs it
// for its own workgroup, the heterogeneous kernel (bsf_fit_hetero.hip) for
// the workgroup's place within its launch entry.
template <int FT, typename E, int SC, int R = 1>
__device__ __forceinline__ void fit_roi(Dev A, const int block, const int nblocks)
{
    static_assert(R == 1 || (R == 2 && (FT != 8 || sizeof(E) == 4)), "two ROIs per wave: fittypes 6, 7, FP32 8");
    // fittype 8 in a two-ROI wave keeps one set of pixel sums per lane (the
    // half's pixels l, l + 32, ...): a second set, for the one-ROI order,
    // does not fit its registers, so its sums are not the one-ROI kernel's bits
    constexpr bool HALF_SUMS = R == 2 && FT == 8;
    constexpr int NV = Ft<FT>::NV;
    constexpr int NH = NV * (NV + 1) / 2;
    constexpr int NRED = 1 + NV + NH;
    constexpr int L = 64 / R;   // lanes per ROI
    extern __shared__ __align__(16) unsigned char smem_all[];

    // with a device-side ROI count, the blocks below it are spread over the
    // XCDs as the whole grid otherwise is (the rest exit at once)
    int G = nblocks;
    if (R == 1 && A.ncount) {   // (R = 2 launches take no device-side count)
        const int c = *A.ncount - A.roi_offset;
        G = c < 0 ? 0 : c < G ? c : G;
        if (block >= G) return;
    }
    // uniform launches: each XCD takes a contiguous range of ROIs (their
    // column-major output stores meet in one L2); launches of mixed sizes,
    // whose tables list the largest ROIs first: block b takes ROI b, so that
    // every XCD (blocks are dealt round-robin over the 8) gets the same mix
    // of sizes instead of a range of one size
    const int lane = threadIdx.x;
    const int hl = R == 1 ? lane : (lane & (L - 1));   // lane within the ROI
    const int hb = lane - hl;                          // the ROI's first lane
    int n = ((A.rois ? block : xcd_block(block, G)) * R + hb / L) + A.roi_offset;
    bool live = true;   // R = 2: the last block's second half may have no ROI of its own
    if (n >= A.N) {
        if constexpr (R == 1) return;
        live = false;
        n = A.N - 1;    // fits a copy of the last ROI (barriers and LDS stay per half) and writes nothing
    }
    // the launch's largest ROI size, which the LDS allocation is sized for
    const int Smax = SC ? SC : A.S;
    unsigned char *smem = smem_all;
    if constexpr (R != 1) smem += (hb / L) * fit_lds_bytes(FT, sizeof(E) == 8, Smax);
    const float *dptr = A.data + (size_t)n * Smax * Smax;
    if (A.rois) {
        const RoiDesc d = A.rois[n];
        A.S = d.S < 1 ? 1 : d.S > Smax ? Smax : d.S;   // within the LDS sized for the largest
        dptr = A.data + d.off;
    }
    if (SC) A.S = SC;
    const int S = A.S, S2 = S * S;
    const int W = S + 4, plane_stride = W * W;
    constexpr bool DL = fit_data_in_lds(FT);
    const LdsLayout LL = fit_lds_layout(S, sizeof(E), 0, (NV + NH) * 8 / sizeof(E), DL);
    // ROI data in LDS (fittype 8), or a copy for the initial values only
    float *s_data = DL ? (float *)smem : (float *)(smem + LL.pl);
    // the pixel passes read the data through a buffer descriptor (32-bit
    // offsets, no 64-bit address arithmetic per load)
    const __amdgpu_buffer_rsrc_t rsd =
        __builtin_amdgcn_make_buffer_rsrc((void *)dptr, (short)0, S2 * 4, 0x00020000);
    float *s_rdat = (float *)(smem + LL.rdat);
    const GeoLds<E> gl = {(E *)(smem + LL.fac), (int *)(smem + LL.rtap), (int *)(smem + LL.ctap)};
    E *s_pl = (E *)(smem + LL.pl);
    // J and packed H of the last evaluation: the reference keeps the Hessian
    // between iterations (a rejected step reuses it, a failed Cholesky leaves
    // its damped diagonal), so they live in LDS across iterations
    double *s_jh = (double *)(smem + LL.hj);
    E *s_red = s_pl;   // reductions reuse the plane storage (never live together)
    // column sums of the two-pass stage B (FP32 fittypes 6 / 7; FP64 where
    // stage_b_split allows it for the launch's largest size), after the planes
    const bool split = (FT != 8 && sizeof(E) == 4) || stage_b_split(FT, sizeof(E) == 8, Smax);
    E *s_h = s_pl + ((Ft<FT>::NPL * plane_stride + 3) & ~3);   // 16-byte groups
    constexpr int NHP = ColSums<FT>::NHP;
    const float invS = 1.0f / (float)S, invW = 1.0f / (float)W;

    // stage the ROI (fittypes 6 / 7: only for the initial values) and 1/data
    for (int p = hl; p < S2; p += L) {
        const float d = dptr[p];
        s_data[p] = d;
        if constexpr (sizeof(E) == 4) s_rdat[p] = d > 0.f ? 1.0f / d : 0.f;
    }
    __syncthreads();

    // ---- initial values: centre of mass, max / min (CPUgaussLib.cpp:172-223)
    // summed in the reference's sequential order so they match it bit for bit.
    // Lane l < NV owns parameter l: its value and LM state (sign, clamp,
    // update and the copies restored on a rejected step) are per-lane scalars.
    const int myl = hl < NV ? hl : NV - 1;
    E th;
    {
        E init[8];
        roi_init4<E>(s_data, S, init);
        init[4] = init[5] = init[6] = init[7] = E(0);
        if constexpr (FT == 6) {
            init[4] = (E)A.nd[2] / E(2);
        } else if constexpr (FT == 7) {
            if (A.has_initP && A.initP[0] != 0.f) {
                init[4] = (E)A.initP[0];
                init[5] = (E)A.initP[1];
            } else {
                init[4] = (E)A.nd[2] / E(2);
                init[5] = (E)A.nd[3] / E(2);
            }
        } else {
            if (A.has_initP) {
#pragma unroll
                for (int i = 4; i < 8; i++) init[i] = (E)A.initP[i - 4];
            } else {
                init[4] = (E)A.nd[2] / E(2);
                init[5] = (E)A.nd[3] / E(2);
                init[6] = (E)A.nd[4] / E(2);
                init[7] = (E)0.9;
            }
        }
        th = init[0];
#pragma unroll
        for (int i = 1; i < NV; i++) th = (myl == i) ? init[i] : th;
    }
    E clampv, fade;
    {
        const float c6[5] = {2.0f, 2.0f, 10000.f, 20.f, 20.f};
        const float c7[6] = {1.0f, 1.0f, 10000.f, 20.f, 2.f, 0.00001f};
        const float f7[6] = {0.5f, 0.5f, 0.4f, 0.5f, 0.7f, 0.8f};
        const float c8[8] = {1.0f, 1.0f, 10000.f, 20.f, 2.f, 2.f, 2.f, 0.1f};
        clampv = (E)(FT == 6 ? c6[0] : FT == 7 ? c7[0] : c8[0]);
        fade = (E)(FT == 7 ? f7[0] : 0.5f);
#pragma unroll
        for (int i = 1; i < NV; i++) {
            clampv = (myl == i) ? (E)(FT == 6 ? c6[i] : FT == 7 ? c7[i] : c8[i]) : clampv;
            fade = (myl == i) ? (E)(FT == 7 ? f7[i] : 0.5f) : fade;
        }
    }
    E oclamp = clampv, oth = th, sgn = E(0), osgn = E(0), upd = E(0);

    // coordinate-bearing parameters of the last evaluation (uniform)
    E cth[NV];
    Geo<E, FT> g;
    E g2 = E(0);
    const int myrow = myl;

    // one ROI evaluation at the current parameters: fills s_jh and returns
    // the deviance
    // the per-pixel terms of one evaluation (deviance, J, H) added to acc
    auto pixel_terms = [&](int p, bool first, E (&acc)[NRED]) __attribute__((always_inline)) {
        int ii, jj;
        pixel_rc(p, invS, S, ii, jj);
        E dudt[8], model;
        if (split) {
            const int r = gl.rtap[ii];
            const E *pp = s_pl + gl.ctap[jj] * g.CS + r;
            pixel_model_h<E, FT>(g, s_h + jj * W * NHP, r, pp + 4 * plane_stride, pp + 5 * plane_stride, cth[2],
                                 cth[3], g2, dudt, model);
        }
        else
            pixel_model<E, FT>(g, s_pl, plane_stride, gl.rtap[ii], gl.ctap[jj], cth[2], cth[3], g2, dudt, model);
        if constexpr (FT == 7) {
            if (first) dudt[5] = dudt[4];
        }
        E data;
        acc[0] += deviance_term<E>(model, DL ? (E)s_data[p] : (E)bload(rsd, p * 4, 0), s_rdat[p], data);
        const E rm = recip<E>(model);
        const E dr = data * rm;
        E t1 = E(1) - dr;
        E t2 = dr * rm;
#pragma unroll
        for (int l = 0; l < NV; l++) acc[1 + l] += t1 * dudt[l];
        int h = 1 + NV;
#pragma unroll
        for (int l = 0; l < NV; l++)
#pragma unroll
            for (int m = l; m < NV; m++) acc[h++] += t2 * dudt[l] * dudt[m];
    };

    // one ROI evaluation at the current parameters: fills s_jh and returns
    // the deviance
    auto evaluate = [&](bool first) __attribute__((always_inline)) -> E {
#pragma unroll
        for (int i = 0; i < NV; i++) cth[i] = roi_read<R, E>(th, i, hb);
        if constexpr (FT == 8) g2 = cth[7];
        __syncthreads();   // the scratch of the last reduction is read out
        make_geo<E, FT, R>(A, cth, g, hl, gl, SC != 0, hb);
        stage_a<E, FT, R>(A, g, s_pl, plane_stride, hl);
        __syncthreads();
        if (split) stage_b_cols<E, FT, R>(g, s_pl, plane_stride, gl.ctap, S, invW, s_h, hl);
        E acc[NRED];
#pragma unroll
        for (int k = 0; k < NRED; k++) acc[k] = E(0);
        if constexpr (R == 1) {
            // with a compile-time size every lane has a first pixel (S * S > 64), so
            // the loop needs no entry test (with one, the compiler zeroes the
            // accumulators a second time under the entry mask); not unrolled, which
            // with a fixed size would cost registers (a wave per SIMD)
            int p = lane;
            if (SC || p < S2) {
#pragma nounroll
                do {
                    pixel_terms(p, first, acc);
                    p += 64;
                } while (p < S2);
            }
            __syncthreads();   // planes are dead; the scratch reuses them
            wave_sum_lds<E, NRED>(acc, s_red, lane);
        } else if constexpr (HALF_SUMS) {
#pragma nounroll
            for (int p = hl; p < S2; p += L) pixel_terms(p, first, acc);
            __syncthreads();   // planes are dead; the scratch reuses them
            wave_sum_lds_half<E, NRED>(acc, s_red, hl);
        } else {
            // pixels of the one-ROI kernel's lanes hl (acc) and hl + 32 (acc1)
            E acc1[NRED];
#pragma unroll
            for (int k = 0; k < NRED; k++) acc1[k] = E(0);
#pragma nounroll
            for (int p = hl; p < S2; p += 64) {
                pixel_terms(p, first, acc);
                if (p + 32 < S2) pixel_terms(p + 32, first, acc1);
            }
            __syncthreads();   // planes are dead; the scratch reuses them
            wave_sum_lds2<E, NRED>(acc, acc1, s_red, hl);
        }
        const E *tot = s_red + kRedChunk * 65;
        // J and H (NV + NH values: 44 for fittype 8, more than a half-wave's lanes)
#pragma unroll
        for (int k = hl; k < NV + NH; k += L) s_jh[k] = (double)tot[1 + k];
        return tot[0];
    };

    if (R == 1 && A.theta_in) {
        // test hook: one evaluation at a given theta; write err, J, H
        th = (E)A.theta_in[(size_t)n * NV + myl];
        const E e = evaluate(false);
        float *o = A.dbg + (size_t)n * (1 + NV + NV * NV);
        __syncthreads();
        if (lane == 0) o[0] = (float)e;
        if (lane < NV) {
            o[1 + lane] = (float)s_jh[lane];
#pragma unroll
            for (int j = 0; j < NV; j++) o[1 + NV + lane * NV + j] = (float)s_jh[NV + packed_index<NV>(lane, j)];
        }
        return;
    }
    E newErr = evaluate(true);
    int nev = 1;   // evaluations run (for the evals output)
    E oldErr = (E)1e13;
    E newLambda = E(1), oldLambda = E(1), sumUpdate = (E)1e10;
    int turn = 1, kk, lastinfo = 0, firstinfo = 0;
    for (kk = 0; kk < A.iterations; kk++) {
        if (fabs((newErr - oldErr) / newErr) < (E)1e-6 || (FT == 8 && sumUpdate < (E)1e-5)) break;   // TOLERANCE, TOLERANCE1
        if ((double)newErr > 1.5 * (double)oldErr) {
            sgn = osgn;
            clampv = oclamp;
            th = oth;
            newLambda = oldLambda;
            newErr = oldErr;
            newLambda = E(10) * newLambda;
        } else if (newErr < oldErr) {
            if (newLambda > E(1)) newLambda = (E)((double)newLambda * 0.8);
            else if (newLambda < E(1)) newLambda = E(1);
        }
        // damp the diagonal in place (it persists, as the reference's)
        __syncthreads();
        if (hl < NV) {
            double *dg = s_jh + NV + packed_index<NV>(hl, hl);
            *dg = (double)((E)*dg * newLambda);
        }
        __syncthreads();
        // rows without information (see kEscapeStep): solved with a 1 on their
        // diagonal, which decouples them, and given the escape step
        const unsigned zr = zero_rows<NV, R>(s_jh);
        double xs[NV];
        int info = ldlt_solve<NV>(s_jh, xs, zr);
        if (info == 0) {
            E u = (E)xs[0];
#pragma unroll
            for (int i = 1; i < NV; i++) u = (myl == i) ? (E)xs[i] : u;
            upd = ((zr >> myl) & 1u) ? (E)kEscapeStep : u;
            if constexpr (FT != 6) {
#pragma unroll
                for (int i = 0; i < NV; i++)
                    if (isnan((E)xs[i])) info = 1;
            }
        }
        lastinfo = info;
        if (kk == 0) firstinfo = info;
        if (info != 0) {
            newLambda = E(10) * newLambda;
            continue;
        }
        if constexpr (FT == 7) upd = (myl == 5) ? upd * E(10) : upd;
        osgn = sgn;
        oclamp = clampv;
        oth = th;
        oldLambda = newLambda;
        oldErr = newErr;
        if constexpr (FT == 7) {
            const E u4 = roi_read<R, E>(upd, 4, hb);
            if ((double)fabs(u4) < 0.005 && turn) {
                if (myl == 5) clampv = oclamp = (E)A.nd[3] / E(10);
                if (myl == 4) clampv = oclamp = E(2);
                turn = 0;
            }
        }
        if (sgn != E(0)) {
            if ((sgn == E(1) && upd < E(0)) || (sgn == E(-1) && upd > E(0))) clampv = clampv * fade;
        }
        sgn = (upd > E(0)) ? E(1) : E(-1);
        if constexpr (FT == 8) {
            upd = upd / (E(1) + fabs(upd / clampv));
            th = th - upd;
            sumUpdate = E(0);
#pragma unroll
            for (int l = 0; l < NV; l++) sumUpdate += fabs(roi_read<R, E>(upd, l, hb));
        } else {
            th = th - upd / (E(1) + fabs(upd / clampv));
        }
        // bounds (CPUmleFit_LM.cpp:2631-2634, 3021-3026, 3449-3472), one per lane
        if constexpr (FT == 8) {
            if (myl <= 1) {
                const double lo = (double)(((E)S - E(1)) / E(2)) - S / 4.0;
                const double hi = (double)(((E)S - E(1)) / E(2)) + S / 4.0;
                th = dminr<E>(dmaxr<E>(th, lo), hi);
            }
        }
        if (myl == 2) th = dmaxr<E>(th, 1.0);
        if (myl == 3) th = dmaxr<E>(th, 0.01);
        if (myl == 4) {
            th = dmaxr<E>(th, 0.0);
            if constexpr (FT == 6) {
                const E zb = (E)(A.nd[2] - 4);
                th = (th < zb) ? th : zb;
            } else {
                th = dminr<E>(th, (double)(E)A.nd[2] - 0.5);
            }
        }
        if constexpr (FT == 7) {
            if (myl == 5) th = dminr<E>(dmaxr<E>(th, 0.0), (double)(E)A.nd[3] - 0.5);
        }
        if constexpr (FT == 8) {
            if (myl == 5) {
                if (th <= E(1)) th = (E)(A.nd[3] - 2) + th;
                else if (th > (E)(A.nd[3] - 1)) th = (E(2) + th) - (E)A.nd[3];
            }
            if (myl == 6) {
                th = (th > E(1)) ? th : E(1);
                th = dminr<E>(th, (double)(E)A.nd[4] - 0.5);
            }
            if (myl == 7) {
                th = (th < E(1)) ? th : E(1);
                th = (th > E(0)) ? th : E(0);
            }
        }
        newErr = evaluate(false);
        nev++;
    }

    // ---- Fisher information, CRLB, log-likelihood.  Fittypes 6/7 evaluate
    // at the final theta; fittype 8 keeps the last evaluation's coordinates
    // and g2 with the final N, bg (CPUmleFit_LM.cpp:3566-3583).
    if constexpr (FT != 8) {
#pragma unroll
        for (int i = 0; i < NV; i++) cth[i] = roi_read<R, E>(th, i, hb);
        __syncthreads();
        make_geo<E, FT, R>(A, cth, g, hl, gl, SC != 0, hb);
    } else {
        // the last evaluation's geometry again (x, y, z, w, u unchanged in
        // cth; recomputing it is cheaper than keeping it live in registers)
        cth[2] = roi_read<R, E>(th, 2, hb);
        cth[3] = roi_read<R, E>(th, 3, hb);
        __syncthreads();
        make_geo<E, FT, R>(A, cth, g, hl, gl, SC != 0, hb);
    }
    __syncthreads();
    stage_a<E, FT, R>(A, g, s_pl, plane_stride, hl);
    __syncthreads();
    if (split) stage_b_cols<E, FT, R>(g, s_pl, plane_stride, gl.ctap, S, invW, s_h, hl);
    constexpr int NF = 1 + NH;
    // the per-pixel Fisher terms and log-likelihood added to acc
    auto fisher_terms = [&](int p, E (&acc)[NF]) __attribute__((always_inline)) {
        int ii, jj;
        pixel_rc(p, invS, S, ii, jj);
        E dudt[8], model;
        if (split) {
            const int r = gl.rtap[ii];
            const E *pp = s_pl + gl.ctap[jj] * g.CS + r;
            pixel_model_h<E, FT>(g, s_h + jj * W * NHP, r, pp + 4 * plane_stride, pp + 5 * plane_stride, cth[2],
                                 cth[3], g2, dudt, model);
        }
        else
            pixel_model<E, FT>(g, s_pl, plane_stride, gl.rtap[ii], gl.ctap[jj], cth[2], cth[3], g2, dudt, model);
        E data = DL ? (E)s_data[p] : (E)bload(rsd, p * 4, 0);
        const E rmf = recip<E>(model);   // Fisher terms dudt_a dudt_b / model
        int h = 1;
#pragma unroll
        for (int a = 0; a < NV; a++)
#pragma unroll
            for (int b = a; b < NV; b++) acc[h++] += dudt[b] * dudt[a] * rmf;
        if (model > E(0)) {
            if (data > E(0)) acc[0] += data * log(model) - model - data * log(data) + data;
            else acc[0] += -model;
        }
    };
    E acc[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) acc[k] = E(0);
    if constexpr (R == 1) {
        // with a compile-time size every lane has a first pixel (S * S > 64), so
        // the loop needs no entry test (with one, the compiler zeroes the
        // accumulators a second time under the entry mask); not unrolled, which
        // with a fixed size would cost registers (a wave per SIMD)
        int p = lane;
        if (SC || p < S2) {
#pragma nounroll
            do {
                fisher_terms(p, acc);
                p += 64;
            } while (p < S2);
        }
        __syncthreads();
        wave_sum_lds<E, NF>(acc, s_red, lane);
    } else if constexpr (HALF_SUMS) {
#pragma nounroll
        for (int p = hl; p < S2; p += L) fisher_terms(p, acc);
        __syncthreads();
        wave_sum_lds_half<E, NF>(acc, s_red, hl);
    } else {
        E acc1[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) acc1[k] = E(0);
#pragma nounroll
        for (int p = hl; p < S2; p += 64) {
            fisher_terms(p, acc);
            if (p + 32 < S2) fisher_terms(p + 32, acc1);
        }
        __syncthreads();
        wave_sum_lds2<E, NF>(acc, acc1, s_red, hl);
    }
    const E *tot = s_red + kRedChunk * 65;
    double mrow[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) mrow[j] = (double)tot[1 + packed_index<NV>(myrow, j)];
    const double crlb = inverse_diag_rows<NV, R>(mrow, hl, hb);
    const float llf = (float)tot[0];
    if (!live) return;
    if (A.evals && hl == 0) A.evals[n] = nev + 1;
    // A later start of a multi-start fit replaces the kept result only where
    // its log-likelihood beats it by 1e-4, compared in FP32 as the
    // reference's callers compare the MEX outputs
    // (demo_orientation/store/fitb5n3.m:38-45: >=; simplefitter_spline.m:488: >).
    if (A.keep) {
        const float d = llf - A.LL[n];
        if (!(A.keep == 2 ? d > 1e-4f : d >= 1e-4f)) return;
    }
    if (hl < NV) {
        A.P[(size_t)A.N * hl + n] = (float)th;
        A.CRLB[(size_t)A.N * hl + n] = (float)crlb;
    }
    if (hl == 0) {
        A.P[(size_t)A.N * NV + n] = (float)kk;
        A.LL[n] = llf;
    }
    if (A.status) {
        // per-ROI status flags, as write_roi_status writes them; kept inline
        // here, where the call form costs this kernel SGPR spills
        const float cf = (float)crlb, tf = (float)th;
        const unsigned long long badc = roi_ballot<R>(hl < NV && !(cf > 0.f && isfinite(cf)), hb);
        const unsigned long long badp = roi_ballot<R>(hl < NV && !isfinite(tf), hb);
        if (hl == 0) {
            unsigned f = 0;
            if (kk >= A.iterations) f |= 1u;
            if (lastinfo != 0) f |= 2u;
            if (firstinfo != 0) f |= 16u;
            if (badc) f |= 4u;
            if (badp || !isfinite(llf)) f |= 8u;
            A.status[n] = (unsigned char)f;
        }
    }
}

// W2: registers for two waves per SIMD (256) where the kernel would
// otherwise take more and run one (FP64 fittype 8: 306 registers, or 256
// with 49 VGPRs spilled, most of them outside the pixel loops; the launcher
// picks by ROI size); the fittype-8 two-ROI tile always takes it.
template <int FT, typename E, int SC, int R = 1, int W2 = 0>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu((FT == 8 && R == 2) || W2 ? 2 : 1)))
void fit_kernel(Dev A)
{
    fit_roi<FT, E, SC, R>(A, (int)blockIdx.x, (int)gridDim.x);
}

// ------------------------------------------------------------ launcher
// Kernel arguments of a launch.
inline Dev fit_dev(const FitLaunch &L)
{
    Dev A;
    A.data = L.data;
    A.coeff = L.coeff;
    A.coeff3 = L.coeff3;
    A.rois = L.rois;
    A.P = L.P;
    A.CRLB = L.CRLB;
    A.LL = L.LL;
    A.N = L.N;
    A.S = L.S;
    A.iterations = L.iterations;
    A.roi_offset = L.roi_offset;
    A.ncount = L.ncount;
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
    A.keep = L.keep;
    A.theta_in = L.theta_in;
    A.dbg = L.dbg;
    A.status = L.status;
    A.evals = L.evals;
    return A;
}

// nrois ROIs in R-ROI waves (one wave per workgroup, R LDS regions)
template <int FT, typename E, int SC = 0, int R = 1, int W2 = 0>
inline hipError_t launch_t(const Dev &A, int nrois, hipStream_t st)
{
    const size_t lds = R * lds_bytes(FT, sizeof(E) == 8, A.S);   // A.S = largest ROI size of the launch
    hipLaunchKernelGGL((fit_kernel<FT, E, SC, R, W2>), dim3((nrois + R - 1) / R), dim3(64), lds, st, A);
    return hipGetLastError();
}

// FP32 fittypes 6 / 7 where every ROI of a launch has one of the common
// sizes: a kernel specialised on it, whose size-dependent index arithmetic
// and LDS offsets fold to constants (4.6 % faster on 100k 15x15 fittype-7
// ROIs, 7-11 % on fittype 6).  (The initial-value loops are only partly
// unrolled: fully unrolled at a compile-time S they held every pixel in a
// register, up to 234 VGPRs.)
inline bool sized_kernel(int, int S) { return S == 11 || S == 13 || S == 15 || S == 17; }
// Launch the specialised kernel for S (sized_kernel(ft, S) must hold);
// in bsf_fit_sized6.hip / bsf_fit_sized7.hip.
hipError_t launch_sized6(const Dev &A, int nrois, hipStream_t st, int S, int R = 1);
hipError_t launch_sized7(const Dev &A, int nrois, hipStream_t st, int S, int R = 1);
// FP64: fittype 6 at every one of these sizes, fittype 7 up to S = 15
inline bool sized_kernel_f64(int ft, int S)
{
    return ft == 6 ? sized_kernel(6, S) : ft == 7 && (S == 11 || S == 13 || S == 15);
}
hipError_t launch_sized6d(const Dev &A, int nblocks, hipStream_t st, int S);
hipError_t launch_sized7d(const Dev &A, int nblocks, hipStream_t st, int S);

}  // namespace bsf
