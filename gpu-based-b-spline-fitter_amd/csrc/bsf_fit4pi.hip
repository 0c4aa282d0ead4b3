// bsf_fit4pi.hip -- gfx950 kernel of the 4Pi multi-channel B-spline fit.
//
// The reference calls GPUmleFit_LM_4Pi_bspline (demo_4pi/bspline_based_4piloc_v4.m:
// 196-206) but ships no source for it.  This fitter is specified from that
// call site, the channel model of demo_4pi/store/simBSplinePSF.m:38-53 (channel
// k evaluates the 4D (x, y, z, phi) spline at its phase shifted by the
// channel offset) and the shared-parameter layout of
// demo_4pi/store/calculate_CRLB_YL_shared.m:11-13, 98-124 (base parameters x, y,
// N, bg, z, phi; a shared one is one slot, an unshared one nch slots).  The
// LM control is fittype 7's (CPUmleFit_LM.cpp:2763-3158) applied slot by
// slot, with phase slots wrapped (period nw-2) instead of clamped.
//
// One wave fits one ROI of nch channels.  Lane l < NV owns parameter slot l
// (its value, sign, clamp, update and their saved copies) and row l of the
// Hessian; the per-channel contraction and pixel sums reuse the fittype-7
// stages, and the channel blocks are scattered into the NV x NV system in LDS.
#include <hip/hip_runtime.h>

#include <cstring>

#include "bsf_device.h"
#include "bsf_kernels.h"

namespace bsf {

struct Dev4 {
    Dev b;
    int nch, nv;
    int slot[6][8];
    int base_of[24], chan_of[24];   // slot -> base parameter, channel (-1: shared)
    float shift[8];                 // channel phase offsets, w units
    const float *dT;                // [6, nch, N] x/y offsets, or NULL
    const float *initZ, *initPhi;   // [N] starts, or NULL
};

template <typename E>
__device__ __forceinline__ E wrap_phase(E w, int nw)
{
    if (w <= E(1)) w = (E)(nw - 2) + w;
    else if (w > (E)(nw - 1)) w = (E(2) + w) - (E)nw;
    return w;
}

__device__ __forceinline__ int packed6(int p, int q)
{
    const int a = p < q ? p : q, b = p < q ? q : p;
    return a * 6 - (a * (a - 1)) / 2 + (b - a);
}

__host__ __device__ inline bool split_4pi(bool f64, int S) { return f64 && stage_b_split(7, true, S); }

template <typename E, int NVMAX>
__global__ __launch_bounds__(64) void fit4pi_kernel(Dev4 A4)
{
    Dev A = A4.b;
    // a device-side ROI count (launches with a ROI table only): blocks past it exit
    if (A.ncount && (int)blockIdx.x + A.roi_offset >= *A.ncount) return;
    // mixed sizes: block b takes ROI b (every XCD the same size mix, as fit_kernel)
    const int n = (A.rois ? (int)blockIdx.x : xcd_block(blockIdx.x, gridDim.x)) + A.roi_offset;
    if (n >= A.N) return;
    const int lane = threadIdx.x;
    size_t dbase = (size_t)n * A.S * A.S, cstride = (size_t)A.N * A.S * A.S;
    // the launch's largest ROI size, which the LDS allocation is sized for
    const int Smax = A.S;
    if (A.rois) {
        const RoiDesc d = A.rois[n];
        A.S = d.S < 1 ? 1 : d.S > A.S ? A.S : d.S;   // within the LDS sized for the largest
        dbase = (size_t)d.off;
        cstride = (size_t)d.cstride;
    }
    const int S = A.S, S2 = S * S, nch = A4.nch, nv = A4.nv;
    const int W = S + 4, plane_stride = W * W;
    extern __shared__ __align__(16) unsigned char smem[];
    const LdsLayout LL = lds_layout(S, nch, sizeof(E), 0);
    float *s_data = (float *)smem;
    float *s_rdat = (float *)(smem + LL.rdat);
    const GeoLds<E> gl = {(E *)(smem + LL.fac), (int *)(smem + LL.rtap), (int *)(smem + LL.ctap)};
    E *s_pl = (E *)(smem + LL.pl);
    const size_t pl = (size_t)3 * plane_stride;
    // column sums of the two-pass stage B after the planes, the reduction
    // scratch over both (FP64 only: with FP32 the two-pass form measured 6 %
    // slower here, 3.02 vs 2.84 M fits/s on 200k 4x13x13 ROIs; FP64 8 % faster)
    const bool split = split_4pi(sizeof(E) == 8, Smax);
    const size_t pls = split ? pl + col_sum_elems(7, S) + 3 : pl;
    const size_t plred = pls > (size_t)red_elems(28) ? pls : (size_t)red_elems(28);
    const float invS = 1.0f / (float)S, invW = 1.0f / (float)W;
    E *s_red = s_pl;
    E *s_h = s_pl + ((pl + 3) & ~(size_t)3);   // 16-byte groups
    E *s_H = s_pl + plred;          // nv * nv
    E *s_J = s_H + nv * nv;         // nv
    E *s_err = s_J + nv;            // 1

    for (int p = lane; p < nch * S2; p += 64) {
        const int k = p / S2, q = p - k * S2;
        const float d = A.data[dbase + k * cstride + q];
        s_data[p] = d;
        if constexpr (sizeof(E) == 4) s_rdat[p] = d > 0.f ? 1.0f / d : 0.f;
    }
    __syncthreads();

    const int myrow = lane < nv ? lane : nv - 1;
    const int myb = A4.base_of[myrow], myc = A4.chan_of[myrow];

    // ---- initial values: fittype 7's per channel, channel mean for a shared
    // slot (CPUmleFit_LM.cpp:2831-2845, CPUgaussLib.cpp:172-223)
    E th;
    {
        E sum = E(0);
        for (int k = 0; k < nch; k++) {
            const float *d = s_data + k * S2;
            E v;
            if (myb <= 1) {
                E tx = E(0), ty = E(0), ts = E(0);
                for (int ii = 0; ii < S; ii++)
                    for (int jj = 0; jj < S; jj++) {
                        E dd = (E)d[S * jj + ii];
                        tx += dd * (E)ii;
                        ty += dd * (E)jj;
                        ts += dd;
                    }
                v = (myb == 0) ? tx / ts : ty / ts;
            } else if (myb <= 3) {
                E mx = E(0), mn = E(10e10);
                for (int p = 0; p < S2; p++) {
                    E dd = (E)d[p];
                    mx = (mx > dd) ? mx : dd;
                    mn = (mn < dd) ? mn : dd;
                }
                if (myb == 2) {
                    double t = (double)(((mx - mn) * E(2)) * (E)3.141592) * 1.5 * 1.5;
                    v = (E)((0.0 > t) ? 0.0 : t);
                } else {
                    v = dmaxr<E>(mn, 0.01);
                }
            } else if (myb == 4) {
                v = A4.initZ ? (E)A4.initZ[n] : (E)A.nd[2] / E(2);
            } else {
                v = A4.initPhi ? (E)A4.initPhi[n] : (E)A.nd[3] / E(2);
            }
            if (myc < 0 || myc == k) sum += v;
        }
        th = (myc < 0) ? sum / (E)nch : sum;
    }
    const float c7[6] = {1.0f, 1.0f, 10000.f, 20.f, 2.f, 0.00001f};
    const float f7[6] = {0.5f, 0.5f, 0.4f, 0.5f, 0.7f, 0.8f};
    E clampv = (E)c7[myb], oclamp = clampv, fade = (E)f7[myb];
    E oth = th, sgn = E(0), osgn = E(0), upd = E(0);

    Geo<E, 7> g;
    E hrow[NVMAX], jv = E(0);

    auto channel_params = [&](int k, E *cth) __attribute__((always_inline)) {
#pragma unroll
        for (int p = 0; p < 6; p++) cth[p] = readlane_e<E>(th, A4.slot[p][k]);
        if (A4.dT) {
            cth[0] += (E)A4.dT[((size_t)n * nch + k) * 6 + 0];
            cth[1] += (E)A4.dT[((size_t)n * nch + k) * 6 + 1];
        }
        cth[5] = wrap_phase<E>(cth[5] + (E)A4.shift[k], A.nd[3]);
    };

    // with every parameter shared (nv = 6) each channel's 6 x 6 block is the
    // whole Hessian: the lanes' sums run on over the channels and are reduced
    // once, after the last channel (one reduction and two barriers per
    // evaluation instead of one per channel: +12 % MIX, +5 % FP64 on 200k
    // four-channel 13 x 13 ROIs)
    const bool fold = nv == 6;

    // one evaluation over all channels: the NV x NV Hessian, J and deviance
    // assembled in LDS from the per-channel 6 x 6 blocks
    auto evaluate = [&](bool first) __attribute__((always_inline)) -> E {
        __syncthreads();
        for (int i = lane; i < nv * nv + nv + 1; i += 64) s_H[i] = E(0);
        E acc[28];
#pragma unroll
        for (int q = 0; q < 28; q++) acc[q] = E(0);
        for (int k = 0; k < nch; k++) {
            E cth[6];
            channel_params(k, cth);
            __syncthreads();
            make_geo<E, 7>(A, cth, g, lane, gl);
            stage_a<E, 7>(A, g, s_pl, plane_stride, lane);
            __syncthreads();
            if (split) stage_b_cols<E, 7>(g, s_pl, plane_stride, gl.ctap, S, invW, s_h, lane);
            for (int p = lane; p < S2; p += 64) {
                int ii, jj;
                pixel_rc(p, invS, S, ii, jj);
                E dudt[8], model;
                if (split)
                    pixel_model_h<E, 7>(g, s_h + jj * W * 4, gl.rtap[ii], nullptr, nullptr, cth[2], cth[3], E(0),
                                        dudt, model);
                else
                    pixel_model<E, 7>(g, s_pl, plane_stride, gl.rtap[ii], gl.ctap[jj], cth[2], cth[3], E(0), dudt,
                                      model);
                if (first) dudt[5] = dudt[4];
                E data;
                acc[0] += deviance_term<E>(model, (E)s_data[k * S2 + p], s_rdat[k * S2 + p], data);
                const E rm = recip<E>(model);
                const E dr = data * rm;
                E t1 = E(1) - dr;
                E t2 = dr * rm;
#pragma unroll
                for (int l = 0; l < 6; l++) acc[1 + l] += t1 * dudt[l];
                int h = 7;
#pragma unroll
                for (int l = 0; l < 6; l++)
#pragma unroll
                    for (int m = l; m < 6; m++) acc[h++] += t2 * dudt[l] * dudt[m];
            }
            if (fold && k + 1 < nch) continue;
            __syncthreads();
            wave_sum_lds<E, 28>(acc, s_red, lane);
            const E *tot = s_red + kRedChunk * 65;
            if (lane < 36) {
                const int p = lane / 6, q = lane - p * 6;
                s_H[A4.slot[p][k] * nv + A4.slot[q][k]] += tot[7 + packed6(p, q)];
            } else if (lane < 42) {
                s_J[A4.slot[lane - 36][k]] += tot[1 + lane - 36];
            } else if (lane == 42) {
                s_err[0] += tot[0];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 28; q++) acc[q] = E(0);
        }
        jv = s_J[myrow];
#pragma unroll
        for (int j = 0; j < NVMAX; j++) hrow[j] = j < nv ? s_H[myrow * nv + j] : E(0);
        return s_err[0];
    };

    E newErr = evaluate(true);
    int nev = 1;   // evaluations run (for the evals output)
    E oldErr = (E)1e13, newLambda = E(1), oldLambda = E(1);
    int turn = 1, kk, lastinfo = 0, firstinfo = 0;
    const int zslot = A4.slot[4][0];
    for (kk = 0; kk < A.iterations; kk++) {
        if (fabs((newErr - oldErr) / newErr) < (E)1e-6) break;
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
        double arow[NVMAX], xi;
#pragma unroll
        for (int j = 0; j < NVMAX; j++) {
            hrow[j] = (j == lane) ? hrow[j] * newLambda : hrow[j];
            arow[j] = (double)hrow[j];
        }
        int info = chol_solve_rows_rt<NVMAX>(arow, (double)jv, lane, nv, xi);
        if (info == 0) {
            upd = (E)xi;
            if (__any(lane < nv && isnan(upd))) info = 1;
        }
        lastinfo = info;
        if (kk == 0) firstinfo = info;
        if (info != 0) {
            newLambda = E(10) * newLambda;
            continue;
        }
        if (myb == 5) upd = upd * E(10);
        osgn = sgn;
        oclamp = clampv;
        oth = th;
        oldLambda = newLambda;
        oldErr = newErr;
        const E uz = readlane_e<E>(upd, zslot);
        if ((double)fabs(uz) < 0.005 && turn) {
            if (myb == 5) clampv = oclamp = (E)A.nd[3] / E(10);
            if (myb == 4) clampv = oclamp = E(2);
            turn = 0;
        }
        if (sgn != E(0)) {
            if ((sgn == E(1) && upd < E(0)) || (sgn == E(-1) && upd > E(0))) clampv = clampv * fade;
        }
        sgn = (upd > E(0)) ? E(1) : E(-1);
        th = th - upd / (E(1) + fabs(upd / clampv));
        if (myb == 2) th = dmaxr<E>(th, 1.0);
        else if (myb == 3) th = dmaxr<E>(th, 0.01);
        else if (myb == 4) th = dminr<E>(dmaxr<E>(th, 0.0), (double)(E)A.nd[2] - 0.5);
        else if (myb == 5) th = wrap_phase<E>(th, A.nd[3]);
        newErr = evaluate(false);
        nev++;
    }

    // ---- Fisher information over all channels, CRLB, log-likelihood
    __syncthreads();
    for (int i = lane; i < nv * nv + nv + 1; i += 64) s_H[i] = E(0);
    E facc[22];
#pragma unroll
    for (int q = 0; q < 22; q++) facc[q] = E(0);
    for (int k = 0; k < nch; k++) {
        E cth[6];
        channel_params(k, cth);
        __syncthreads();
        make_geo<E, 7>(A, cth, g, lane, gl);
        stage_a<E, 7>(A, g, s_pl, plane_stride, lane);
        __syncthreads();
        if (split) stage_b_cols<E, 7>(g, s_pl, plane_stride, gl.ctap, S, invW, s_h, lane);
        E (&acc)[22] = facc;
        for (int p = lane; p < S2; p += 64) {
            int ii, jj;
            pixel_rc(p, invS, S, ii, jj);
            E dudt[8], model;
            if (split)
                pixel_model_h<E, 7>(g, s_h + jj * W * 4, gl.rtap[ii], nullptr, nullptr, cth[2], cth[3], E(0), dudt,
                                    model);
            else
                pixel_model<E, 7>(g, s_pl, plane_stride, gl.rtap[ii], gl.ctap[jj], cth[2], cth[3], E(0), dudt, model);
            E data = (E)s_data[k * S2 + p];
            const E rmf = recip<E>(model);   // Fisher terms dudt_a dudt_b / model
            int h = 1;
#pragma unroll
            for (int a = 0; a < 6; a++)
#pragma unroll
                for (int b = a; b < 6; b++) acc[h++] += dudt[b] * dudt[a] * rmf;
            if (model > E(0)) {
                if (data > E(0)) acc[0] += data * log(model) - model - data * log(data) + data;
                else acc[0] += -model;
            }
        }
        if (fold && k + 1 < nch) continue;
        __syncthreads();
        wave_sum_lds<E, 22>(acc, s_red, lane);
        const E *tot = s_red + kRedChunk * 65;
        if (lane < 36) {
            const int p = lane / 6, q = lane - p * 6;
            s_H[A4.slot[p][k] * nv + A4.slot[q][k]] += tot[1 + packed6(p, q)];
        } else if (lane == 42) {
            s_err[0] += tot[0];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 22; q++) acc[q] = E(0);
    }
    double mrow[NVMAX];
#pragma unroll
    for (int j = 0; j < NVMAX; j++) mrow[j] = j < nv ? (double)s_H[myrow * nv + j] : 0.0;
    const double crlb = inverse_diag_rows_rt<NVMAX>(mrow, lane, nv);
    if (lane < nv) {
        A.P[(size_t)A.N * lane + n] = (float)th;
        A.CRLB[(size_t)A.N * lane + n] = (float)crlb;
    }
    if (lane == 0) {
        A.P[(size_t)A.N * nv + n] = (float)kk;
        A.LL[n] = (float)s_err[0];
        if (A.evals) A.evals[n] = nev + 1;
    }
    if (A.status)
        write_roi_status(A.status, n, lane < nv, (float)th, (float)crlb, (float)s_err[0], kk >= A.iterations,
                         lastinfo != 0, firstinfo != 0, lane);
}

// ---------------------------------------------------------------------------
// The same fit with one wave per channel: a workgroup of nch waves per ROI.
// Wave k evaluates channel k (geometry, planes, pixel sums, its reduction in
// its own LDS region) and stores its 28 totals; wave 0 then adds the channel
// blocks into the NV x NV system in channel order, exactly as the one-wave
// kernel does, so both give identical results.  Every wave keeps its own copy
// of the parameter slots and runs the same lane-uniform LM step on the same
// LDS data, so all waves take the same branches and meet the same barriers.
// Work that is per ROI rather than per channel is not repeated: wave k
// computes channel k's initial values, and wave 0 alone solves each damped
// system and passes the step on through LDS.
// A ROI's evaluation then takes about one channel's time instead of nch,
// for launches too small to fill the GPU.  Up to 4 channels (256 threads).
struct Lds4W {
    size_t rdat, blk, hje, lm, wave0, wstride, rtap, ctap, fac, pl, total;
};

__host__ __device__ inline Lds4W lds_4pi_wpc(int S, int nch, size_t esz, int nv)
{
    Lds4W L;
    const size_t W = S + 4, pl = 3 * W * W;
    const size_t plred = pl > (size_t)red_elems(28) ? pl : (size_t)red_elems(28);
    size_t o = lds_align16((size_t)nch * S * S * 4);
    L.rdat = o;
    o = lds_align16(o + (esz == 4 ? (size_t)nch * S * S * 4 : 0));
    L.blk = o;                                   // [nch][28] channel totals
    o = lds_align16(o + (size_t)nch * 28 * esz);
    L.hje = o;                                   // H [nv][nv], J [nv], err
    o = lds_align16(o + ((size_t)nv * nv + nv + 1) * esz);
    L.lm = o;                                    // step [24] (FP64), solve status, initial values [8][6]
    o = lds_align16(o + 24 * 8 + 16 + (size_t)nch * 6 * esz);
    L.wave0 = o;                                 // per-wave regions follow
    size_t w = 0;
    L.rtap = w;
    w = lds_align16(w + (size_t)S * 4);
    L.ctap = w;
    w = lds_align16(w + (size_t)S * 4);
    L.fac = w;
    w = lds_align16(w + 40 * esz);
    L.pl = w;
    w = lds_align16(w + plred * esz);
    L.wstride = w;
    L.total = o + (size_t)nch * w;
    return L;
}

// orders wave 0's own LDS updates (one wave: no workgroup barrier allowed there)
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename E, int NVMAX>
__global__ __launch_bounds__(256) void fit4pi_wpc_kernel(Dev4 A4)
{
    Dev A = A4.b;
    // a device-side ROI count (launches with a ROI table only): blocks past it exit
    if (A.ncount && (int)blockIdx.x + A.roi_offset >= *A.ncount) return;
    // mixed sizes: block b takes ROI b (every XCD the same size mix, as fit_kernel)
    const int n = (A.rois ? (int)blockIdx.x : xcd_block(blockIdx.x, gridDim.x)) + A.roi_offset;
    if (n >= A.N) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;      // wave = channel
    size_t dbase = (size_t)n * A.S * A.S, cstride = (size_t)A.N * A.S * A.S;
    if (A.rois) {
        const RoiDesc d = A.rois[n];
        A.S = d.S < 1 ? 1 : d.S > A.S ? A.S : d.S;   // within the LDS sized for the largest
        dbase = (size_t)d.off;
        cstride = (size_t)d.cstride;
    }
    const int S = A.S, S2 = S * S, nch = A4.nch, nv = A4.nv;
    const int W = S + 4, plane_stride = W * W;
    extern __shared__ __align__(16) unsigned char smem[];
    const Lds4W LW = lds_4pi_wpc(S, nch, sizeof(E), nv);
    float *s_data = (float *)smem;
    float *s_rdat = (float *)(smem + LW.rdat);
    E *s_blk = (E *)(smem + LW.blk);
    E *s_H = (E *)(smem + LW.hje);
    E *s_J = s_H + nv * nv;
    E *s_err = s_J + nv;
    double *s_xi = (double *)(smem + LW.lm);
    int *s_info = (int *)(smem + LW.lm + 24 * 8);
    E *s_init = (E *)(smem + LW.lm + 24 * 8 + 16);
    unsigned char *wb = smem + LW.wave0 + (size_t)wave * LW.wstride;
    const GeoLds<E> gl = {(E *)(wb + LW.fac), (int *)(wb + LW.rtap), (int *)(wb + LW.ctap)};
    E *s_pl = (E *)(wb + LW.pl);
    E *s_red = s_pl;
    const float invS = 1.0f / (float)S;

    for (int p = threadIdx.x; p < nch * S2; p += (int)blockDim.x) {
        const int k = p / S2, q = p - k * S2;
        const float d = A.data[dbase + k * cstride + q];
        s_data[p] = d;
        if constexpr (sizeof(E) == 4) s_rdat[p] = d > 0.f ? 1.0f / d : 0.f;
    }
    __syncthreads();

    const int myrow = lane < nv ? lane : nv - 1;
    const int myb = A4.base_of[myrow], myc = A4.chan_of[myrow];

    // initial values: wave k works out channel k's six (lane b: base
    // parameter b, as the one-wave kernel does per channel), then every lane
    // combines them over the channels in channel order
    E th;
    {
        const int k = wave;
        if (lane < 6) {
            const int b = lane;
            const float *d = s_data + k * S2;
            E v;
            if (b <= 1) {
                E tx = E(0), ty = E(0), ts = E(0);
                for (int ii = 0; ii < S; ii++)
                    for (int jj = 0; jj < S; jj++) {
                        E dd = (E)d[S * jj + ii];
                        tx += dd * (E)ii;
                        ty += dd * (E)jj;
                        ts += dd;
                    }
                v = (b == 0) ? tx / ts : ty / ts;
            } else if (b <= 3) {
                E mx = E(0), mn = E(10e10);
                for (int p = 0; p < S2; p++) {
                    E dd = (E)d[p];
                    mx = (mx > dd) ? mx : dd;
                    mn = (mn < dd) ? mn : dd;
                }
                if (b == 2) {
                    double t = (double)(((mx - mn) * E(2)) * (E)3.141592) * 1.5 * 1.5;
                    v = (E)((0.0 > t) ? 0.0 : t);
                } else {
                    v = dmaxr<E>(mn, 0.01);
                }
            } else if (b == 4) {
                v = A4.initZ ? (E)A4.initZ[n] : (E)A.nd[2] / E(2);
            } else {
                v = A4.initPhi ? (E)A4.initPhi[n] : (E)A.nd[3] / E(2);
            }
            s_init[k * 6 + b] = v;
        }
        __syncthreads();
        E sum = E(0);
        for (int c = 0; c < nch; c++)
            if (myc < 0 || myc == c) sum += s_init[c * 6 + myb];
        th = (myc < 0) ? sum / (E)nch : sum;
    }
    const float c7[6] = {1.0f, 1.0f, 10000.f, 20.f, 2.f, 0.00001f};
    const float f7[6] = {0.5f, 0.5f, 0.4f, 0.5f, 0.7f, 0.8f};
    E clampv = (E)c7[myb], oclamp = clampv, fade = (E)f7[myb];
    E oth = th, sgn = E(0), osgn = E(0), upd = E(0);

    Geo<E, 7> g;
    E hrow[NVMAX], jv = E(0);

    auto channel_params = [&](int k, E *cth) __attribute__((always_inline)) {
#pragma unroll
        for (int p = 0; p < 6; p++) cth[p] = readlane_e<E>(th, A4.slot[p][k]);
        if (A4.dT) {
            cth[0] += (E)A4.dT[((size_t)n * nch + k) * 6 + 0];
            cth[1] += (E)A4.dT[((size_t)n * nch + k) * 6 + 1];
        }
        cth[5] = wrap_phase<E>(cth[5] + (E)A4.shift[k], A.nd[3]);
    };

    // wave 0 adds the channel totals (V per channel: err, J[6], packed H)
    // into the NV x NV system, channel by channel
    auto assemble = [&](int hoff, bool with_j) __attribute__((always_inline)) {
        if (wave != 0) return;
        for (int i = lane; i < nv * nv + nv + 1; i += 64) s_H[i] = E(0);
        wave_lds_sync();
        for (int k = 0; k < nch; k++) {
            const E *tot = s_blk + k * 28;
            if (lane < 36) {
                const int p = lane / 6, q = lane - p * 6;
                s_H[A4.slot[p][k] * nv + A4.slot[q][k]] += tot[hoff + packed6(p, q)];
            } else if (lane < 42) {
                if (with_j) s_J[A4.slot[lane - 36][k]] += tot[1 + lane - 36];
            } else if (lane == 42) {
                s_err[0] += tot[0];
            }
            wave_lds_sync();
        }
    };

    auto evaluate = [&](bool first) __attribute__((always_inline)) -> E {
        __syncthreads();
        const int k = wave;
        E cth[6];
        channel_params(k, cth);
        make_geo<E, 7>(A, cth, g, lane, gl);
        stage_a<E, 7>(A, g, s_pl, plane_stride, lane);
        __syncthreads();
        E acc[28];
#pragma unroll
        for (int q = 0; q < 28; q++) acc[q] = E(0);
        for (int p = lane; p < S2; p += 64) {
            int ii, jj;
            pixel_rc(p, invS, S, ii, jj);
            E dudt[8], model;
            pixel_model<E, 7>(g, s_pl, plane_stride, gl.rtap[ii], gl.ctap[jj], cth[2], cth[3], E(0), dudt, model);
            if (first) dudt[5] = dudt[4];
            E data;
            acc[0] += deviance_term<E>(model, (E)s_data[k * S2 + p], s_rdat[k * S2 + p], data);
            const E rm = recip<E>(model);
            const E dr = data * rm;
            E t1 = E(1) - dr;
            E t2 = dr * rm;
#pragma unroll
            for (int l = 0; l < 6; l++) acc[1 + l] += t1 * dudt[l];
            int h = 7;
#pragma unroll
            for (int l = 0; l < 6; l++)
#pragma unroll
                for (int m = l; m < 6; m++) acc[h++] += t2 * dudt[l] * dudt[m];
        }
        __syncthreads();
        wave_sum_lds<E, 28>(acc, s_red, lane);
        if (lane < 28) s_blk[k * 28 + lane] = s_red[kRedChunk * 65 + lane];
        __syncthreads();
        assemble(7, true);
        __syncthreads();
        jv = s_J[myrow];
#pragma unroll
        for (int j = 0; j < NVMAX; j++) hrow[j] = j < nv ? s_H[myrow * nv + j] : E(0);
        return s_err[0];
    };

    E newErr = evaluate(true);
    int nev = 1;   // evaluations run (for the evals output)
    E oldErr = (E)1e13, newLambda = E(1), oldLambda = E(1);
    int turn = 1, kk, lastinfo = 0, firstinfo = 0;
    const int zslot = A4.slot[4][0];
    for (kk = 0; kk < A.iterations; kk++) {
        if (fabs((newErr - oldErr) / newErr) < (E)1e-6) break;
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
        double arow[NVMAX], xi;
#pragma unroll
        for (int j = 0; j < NVMAX; j++) {
            hrow[j] = (j == lane) ? hrow[j] * newLambda : hrow[j];
            arow[j] = (double)hrow[j];
        }
        // wave 0 solves; every wave takes the step from LDS
        if (wave == 0) {
            const int inf0 = chol_solve_rows_rt<NVMAX>(arow, (double)jv, lane, nv, xi);
            if (lane < nv) s_xi[lane] = xi;
            if (lane == 0) s_info[0] = inf0;
        }
        __syncthreads();
        int info = s_info[0];
        xi = lane < nv ? s_xi[lane] : 0.0;
        __syncthreads();                      // the step is read before the next solve writes it
        if (info == 0) {
            upd = (E)xi;
            if (__any(lane < nv && isnan(upd))) info = 1;
        }
        lastinfo = info;
        if (kk == 0) firstinfo = info;
        if (info != 0) {
            newLambda = E(10) * newLambda;
            continue;
        }
        if (myb == 5) upd = upd * E(10);
        osgn = sgn;
        oclamp = clampv;
        oth = th;
        oldLambda = newLambda;
        oldErr = newErr;
        const E uz = readlane_e<E>(upd, zslot);
        if ((double)fabs(uz) < 0.005 && turn) {
            if (myb == 5) clampv = oclamp = (E)A.nd[3] / E(10);
            if (myb == 4) clampv = oclamp = E(2);
            turn = 0;
        }
        if (sgn != E(0)) {
            if ((sgn == E(1) && upd < E(0)) || (sgn == E(-1) && upd > E(0))) clampv = clampv * fade;
        }
        sgn = (upd > E(0)) ? E(1) : E(-1);
        th = th - upd / (E(1) + fabs(upd / clampv));
        if (myb == 2) th = dmaxr<E>(th, 1.0);
        else if (myb == 3) th = dmaxr<E>(th, 0.01);
        else if (myb == 4) th = dminr<E>(dmaxr<E>(th, 0.0), (double)(E)A.nd[2] - 0.5);
        else if (myb == 5) th = wrap_phase<E>(th, A.nd[3]);
        newErr = evaluate(false);
        nev++;
    }

    // ---- Fisher information, CRLB, log-likelihood: channel k on wave k
    __syncthreads();
    {
        const int k = wave;
        E cth[6];
        channel_params(k, cth);
        make_geo<E, 7>(A, cth, g, lane, gl);
        stage_a<E, 7>(A, g, s_pl, plane_stride, lane);
        __syncthreads();
        E acc[22];
#pragma unroll
        for (int q = 0; q < 22; q++) acc[q] = E(0);
        for (int p = lane; p < S2; p += 64) {
            int ii, jj;
            pixel_rc(p, invS, S, ii, jj);
            E dudt[8], model;
            pixel_model<E, 7>(g, s_pl, plane_stride, gl.rtap[ii], gl.ctap[jj], cth[2], cth[3], E(0), dudt, model);
            E data = (E)s_data[k * S2 + p];
            const E rmf = recip<E>(model);
            int h = 1;
#pragma unroll
            for (int a = 0; a < 6; a++)
#pragma unroll
                for (int b = a; b < 6; b++) acc[h++] += dudt[b] * dudt[a] * rmf;
            if (model > E(0)) {
                if (data > E(0)) acc[0] += data * log(model) - model - data * log(data) + data;
                else acc[0] += -model;
            }
        }
        __syncthreads();
        wave_sum_lds<E, 22>(acc, s_red, lane);
        if (lane < 22) s_blk[k * 28 + lane] = s_red[kRedChunk * 65 + lane];
        __syncthreads();
        assemble(1, false);
        __syncthreads();
    }
    if (wave != 0) return;
    double mrow[NVMAX];
#pragma unroll
    for (int j = 0; j < NVMAX; j++) mrow[j] = j < nv ? (double)s_H[myrow * nv + j] : 0.0;
    const double crlb = inverse_diag_rows_rt<NVMAX>(mrow, lane, nv);
    if (lane < nv) {
        A.P[(size_t)A.N * lane + n] = (float)th;
        A.CRLB[(size_t)A.N * lane + n] = (float)crlb;
    }
    if (lane == 0) {
        A.P[(size_t)A.N * nv + n] = (float)kk;
        A.LL[n] = (float)s_err[0];
        if (A.evals) A.evals[n] = nev + 1;
    }
    if (A.status)
        write_roi_status(A.status, n, lane < nv, (float)th, (float)crlb, (float)s_err[0], kk >= A.iterations,
                         lastinfo != 0, firstinfo != 0, lane);
}

size_t lds_bytes_4pi_wpc(int f64, int S, int nch, int nv) { return lds_4pi_wpc(S, nch, f64 ? 8 : 4, nv).total; }

size_t lds_bytes_4pi(int f64, int S, int nch, int nv)
{
    const size_t W = S + 4;
    const size_t pl = 3 * W * W;
    const size_t pls = split_4pi(f64, S) ? pl + col_sum_elems(7, S) + 3 : pl;
    const size_t plred = pls > (size_t)red_elems(28) ? pls : (size_t)red_elems(28);
    return lds_layout(S, nch, f64 ? 8 : 4, plred + (size_t)nv * nv + nv + 1).total;
}

int g_4pi_kernel = -1;

int use_wpc_4pi(int prefer, int f64, int S, int nch, int nv, size_t lds_limit)
{
    const int want = g_4pi_kernel >= 0 ? g_4pi_kernel : prefer;
    // up to 4 channels: a 256-thread workgroup keeps the one-wave kernel's
    // register budget.  Fits with every parameter shared (nv = 6) always take
    // the one-wave kernel: its sums run on over the channels before one
    // reduction, an order the per-channel waves cannot give, and both forms
    // must give the same bits whatever the launch size
    return want && nv != 6 && nch <= 4 && lds_bytes_4pi_wpc(f64, S, nch, nv) <= lds_limit;
}

template <typename E, int NVMAX>
static hipError_t launch4(const Dev4 &A4, int nblocks, hipStream_t st, size_t lds, int wpc)
{
    if (wpc)
        hipLaunchKernelGGL((fit4pi_wpc_kernel<E, NVMAX>), dim3(nblocks), dim3(64 * A4.nch), lds, st, A4);
    else
        hipLaunchKernelGGL((fit4pi_kernel<E, NVMAX>), dim3(nblocks), dim3(64), lds, st, A4);
    return hipGetLastError();
}

hipError_t launch_fit_4pi(const Fit4PiLaunch &L, hipStream_t st)
{
    Dev4 A4;
    std::memset(&A4, 0, sizeof A4);
    Dev &A = A4.b;
    A.data = L.data;
    A.coeff = L.coeff;
    A.coeff3 = nullptr;
    A.rois = L.rois;
    A.P = L.P;
    A.CRLB = L.CRLB;
    A.LL = L.LL;
    A.N = L.N;
    A.S = L.S;
    A.iterations = L.iterations;
    A.roi_offset = L.roi_offset;
    for (int a = 0; a < 4; a++) A.nd[a] = L.cdims[a] - 3;
    A.nd[4] = 0;
    A.s1 = L.cdims[0];
    A.s2 = A.s1 * L.cdims[1];
    A.s3 = A.s2 * L.cdims[2];
    A.s4 = A.s3 * L.cdims[3];
    A.cbytes = (int)(A.s4 * 4);
    A.c3bytes = 0;
    A.has_initP = 0;
    A.theta_in = nullptr;
    A.dbg = nullptr;
    A.status = L.status;
    A.evals = L.evals;
    if (L.ncount && !L.rois) return hipErrorInvalidValue;   // a device count needs the identity block mapping
    A.ncount = L.ncount;
    A4.nch = L.nch;
    int n = 0;
    for (int p = 0; p < 6; p++) {
        for (int k = 0; k < L.nch; k++) A4.slot[p][k] = L.shared[p] ? n : n + k;
        for (int k = L.nch; k < 8; k++) A4.slot[p][k] = 0;
        for (int k = 0; k < (L.shared[p] ? 1 : L.nch); k++) {
            A4.base_of[n + k] = p;
            A4.chan_of[n + k] = L.shared[p] ? -1 : k;
        }
        n += L.shared[p] ? 1 : L.nch;
    }
    A4.nv = n;
    for (int i = n; i < 24; i++) A4.base_of[i] = 0, A4.chan_of[i] = -1;
    for (int k = 0; k < 8; k++) A4.shift[k] = k < L.nch ? L.phaseShift[k] : 0.f;
    A4.dT = L.dT;
    A4.initZ = L.initZ;
    A4.initPhi = L.initPhi;
    const int nb = L.count;
    const int w = L.wpc;
    const size_t lds = w ? lds_bytes_4pi_wpc(L.f64, L.S, L.nch, n) : lds_bytes_4pi(L.f64, L.S, L.nch, n);
    if (L.f64) {
        if (n <= 8) return launch4<double, 8>(A4, nb, st, lds, w);
        if (n <= 12) return launch4<double, 12>(A4, nb, st, lds, w);
        return launch4<double, 24>(A4, nb, st, lds, w);
    }
    if (n <= 8) return launch4<float, 8>(A4, nb, st, lds, w);
    if (n <= 12) return launch4<float, 12>(A4, nb, st, lds, w);
    return launch4<float, 24>(A4, nb, st, lds, w);
}

}  // namespace bsf
