// bsf_frontend.hip -- camera frames to the ROI stack the fitter takes, on
// gfx950: the GUI fitter's per-frame front-end
// (MultiD_fitter/simplefitter_spline.m:259-307, MultiD_fitter/shared/
// maximumfind2.c:18-65).
//
// Per frame: photons = (single(raw) - offset) * conversion; optionally minus
// the minimum photon count of the frame's first column; filter2 'same'
// correlation with an mr x mc kernel (zero padding, tap (u, v) reads row
// i + u - floor((mr-1)/2), column j + v - floor((mc-1)/2), summed as
// acc = fmaf(h[v*mr+u], x, acc) from 0, v outer, u inner); strict
// 8-neighbour maxima of the filtered frame on interior pixels with value >
// cutoff and dx < x <= ncols - dx, dx < y <= nrows - dx (x = column + 1,
// y = row + 1, dx = S / 2); optionally the PSF-convolution mode's border
// test; then for each maximum, in column-major scan order over the frames,
// the S x S window of photons around it (columns reversed when mirrored).
//
// Kernels, all on the caller's stream:
//   column_min_kernel  one workgroup per frame: the first column's minimum.
//   detect_kernel      one workgroup per 64-row x 40-column tile (tiles dealt
//                      out in contiguous runs per XCD): raw pixels
//                      of the tile plus its filter and NMS halo -> photons in
//                      LDS -> filtered values of the tile plus a one-pixel
//                      ring in LDS -> maxima test; each wave writes one
//                      64-bit ballot per tile column, so bit b of word
//                      (frame, column, q) marks row 64q + b and the words in
//                      index order list the maxima in scan order.  Nothing
//                      else of the frame is written.
//   count/scan/emit    order-preserving compaction of the mask words into
//                      (x, y, frame) rows.
//   aggregate_kernel   with aggregation (the GUI's p.agg): one workgroup per
//                      frame merges the frame's maxima as aggre does and
//                      applies the PSF-convolution border test after it;
//                      gather_kernel then packs the frames' survivors.
//   cut_kernel         one wave per ROI (grid-stride over the device count):
//                      the window re-read from the raw frame and converted,
//                      the filtered value recomputed in the same order.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdint.h>

#include <cmath>
#include <utility>

#include "bsf_device.h"
#include "bsf_kernels.h"

// photons and differences are rounded separately, as MATLAB rounds each
// single-precision operation: no contraction into fused multiply-adds
#pragma clang fp contract(off)

namespace bsf {

namespace {

constexpr int kThreads = 256;
constexpr int kTR = 64;           // tile rows: one mask word per column
constexpr int kTC = 40;           // tile columns
constexpr int kFR = kTR + 2;      // filtered rows incl. the NMS ring
constexpr int kFC = kTC + 2;      // filtered columns incl. the ring
constexpr int kFLD = kFR + 1;     // LDS leading dimension of the filtered tile
constexpr int kR = 11;            // filtered rows per thread
constexpr int kRG = kFR / kR;     // row groups (6)
static_assert(kRG * kR == kFR, "row groups must tile the filtered rows");
static_assert(kRG * kFC <= kThreads, "one filter thread per (row group, column)");
constexpr int kMaxTemplR = 17;    // kernels with mr <= this get an unrolled filter
constexpr int kLoadBatch = 2;     // lower-halo loads a thread keeps in flight
constexpr int kColBatch = 12;     // region columns a wave stages per round (loads in flight)
static_assert(kTC % 4 == 0, "tile columns split evenly over the four waves");

constexpr int kCW = 8;                    // mask words per thread in the compaction
constexpr int kCB = kThreads * kCW;       // mask words per compaction block

template <typename T>
__device__ __forceinline__ T load_px(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ uint16_t load_px<uint16_t>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
}
template <>
__device__ __forceinline__ float load_px<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

template <typename T>
__device__ __forceinline__ float photon(T v, float offset, float conv)
{
    const float d = (float)v - offset;
    return d * conv;
}

struct DetectArgs {
    const void *raw;
    int nr, nc;
    float offset, conv;
    const float *h;       // device [mr, mc]
    int mr, mc;
    const float *sub;     // per-frame value subtracted from the photons, or NULL
    float cutoff_f;       // the cutoff rounded down to FP32
    int dx;
    double psf_half;      // > 0: border test of the PSF-convolution mode
    uint64_t *mask;
    int wpc;              // mask words per column
    int tiles_x;          // column tiles per frame
    int f0;               // first frame of this launch
};

template <typename T>
__global__ __launch_bounds__(kThreads) void column_min_kernel(const T *__restrict__ raw, int nr, int nc, float offset,
                                                              float conv, float *__restrict__ sub)
{
    __shared__ float red[kThreads / 64];
    const T *col = raw + (size_t)blockIdx.x * nr * nc;
    float m = __builtin_nanf("");
    for (int i = threadIdx.x; i < nr; i += kThreads) m = fminf(m, photon(col[i], offset, conv));
    for (int o = 32; o; o >>= 1) m = fminf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = red[0];
        for (int w = 1; w < kThreads / 64; w++) r = fminf(r, red[w]);
        sub[blockIdx.x] = r;
    }
}

template <typename T, int MR>
__global__ __launch_bounds__(kThreads) void detect_kernel(DetectArgs A)
{
    extern __shared__ float lds[];
    const int mr = MR > 0 ? MR : A.mr;
    const int mc = A.mc;
    const int ar = (mr - 1) / 2, ac = (mc - 1) / 2;
    const int XR = kFR + mr - 1, XC = kFC + mc - 1;
    const int LDX = XR | 1;                       // odd: lanes on neighbouring columns hit different banks
    float *X = lds;                               // photons [XC][LDX]
    float *Fl = lds + XC * LDX;                   // filtered [kFC][kFLD]

    // tile of this workgroup: the flat grid is dealt out so that each XCD
    // (blocks go round-robin over the 8) gets a contiguous run of tiles,
    // column tiles fastest, and neighbouring tiles share halos and mask
    // lines in one L2
    const int tq = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int tx = tq % A.tiles_x, tr = tq / A.tiles_x;
    const int ty = tr % A.wpc, f = A.f0 + tr / A.wpc;
    const int i0 = ty * kTR, j0 = tx * kTC;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const T *fr = (const T *)A.raw + (size_t)f * A.nr * A.nc;
    const float sub = A.sub ? A.sub[f] : 0.0f;

    // photons of the tile, the filter halo and the NMS ring; zero outside the
    // frame.  Region rows [0, 64): lane = row, wave w takes columns w, w+4,
    // ..., kColBatch of them per round with their loads in flight together;
    // the column tests are wave-uniform and the rows' are fixed per lane.
    // Rows [64, XR) (the lower halo): a flat list over all threads.  Buffer
    // loads return 0 outside the frame's bytes; every load address is clamped
    // into the frame and the padding is applied by a select, because
    // predicated loads would each be waited for on their own.
    const int gr0 = i0 - 1 - ar, gc0 = j0 - 1 - ac;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)fr, (short)0, (int)(sizeof(T) * (size_t)A.nr * A.nc), 0x00020000);
    {
        const int gr = gr0 + lane;
        const bool row_in = gr >= 0 && gr < A.nr;
        const int voff = (gr < 0 ? 0 : gr >= A.nr ? A.nr - 1 : gr) * (int)sizeof(T);
        const int w = __builtin_amdgcn_readfirstlane(wave);
        for (int c0 = w; c0 < XC; c0 += 4 * kColBatch) {
            T v[kColBatch];
            // columns past the region are clamped to its last one, which is
            // then written again with the same value: no branch, so the
            // compiler keeps the loads together
#pragma unroll
            for (int k = 0; k < kColBatch; k++) {
                const int c = min(c0 + 4 * k, XC - 1), gc = gc0 + c;
                const int cgc = gc < 0 ? 0 : gc >= A.nc ? A.nc - 1 : gc;
                v[k] = load_px<T>(rs, voff, cgc * A.nr * (int)sizeof(T));
            }
#pragma unroll
            for (int k = 0; k < kColBatch; k++) {
                const int c = min(c0 + 4 * k, XC - 1), gc = gc0 + c;
                float x = photon(v[k], A.offset, A.conv);
                if (A.sub) x = x - sub;
                X[c * LDX + lane] = row_in && gc >= 0 && gc < A.nc ? x : 0.0f;
            }
        }
    }
    {
        const int nB = XR - 64;                                       // >= 2
        const int nQ = nB * XC;
        const float inv = 1.0f / (float)nB;
        for (int q0 = threadIdx.x; q0 < nQ; q0 += kThreads * kLoadBatch) {
            T v[kLoadBatch];
            int at[kLoadBatch];
            bool in[kLoadBatch];
#pragma unroll
            for (int k = 0; k < kLoadBatch; k++) {
                const int q = q0 + k * kThreads;
                const int c = (int)(((float)q + 0.5f) * inv), r = 64 + q - c * nB;   // exact: q < 2^14, nB <= 52
                const int gc = gc0 + c, gr = gr0 + r;
                at[k] = q < nQ ? c * LDX + r : -1;
                in[k] = q < nQ && gc >= 0 && gc < A.nc && gr >= 0 && gr < A.nr;
                const int cgc = gc < 0 ? 0 : gc >= A.nc ? A.nc - 1 : gc;
                const int cgr = gr < 0 ? 0 : gr >= A.nr ? A.nr - 1 : gr;
                v[k] = load_px<T>(rs, (cgc * A.nr + cgr) * (int)sizeof(T), 0);
            }
#pragma unroll
            for (int k = 0; k < kLoadBatch; k++) {
                if (at[k] < 0) continue;
                float x = photon(v[k], A.offset, A.conv);
                if (A.sub) x = x - sub;
                X[at[k]] = in[k] ? x : 0.0f;
            }
        }
    }
    __syncthreads();

    // filtered values of the tile and its ring: thread (row group g, column c)
    // accumulates kR rows of one column
    if (threadIdx.x < kRG * kFC) {
        const int c = threadIdx.x % kFC, g = threadIdx.x / kFC;
        float acc[kR];
#pragma unroll
        for (int r = 0; r < kR; r++) acc[r] = 0.0f;
        for (int v = 0; v < mc; v++) {
            const float *xc = X + (c + v) * LDX + g * kR;
            const float *hv = A.h + v * mr;
            if constexpr (MR > 0) {
                // input row t feeds output row r through tap u = t - r
#pragma unroll
                for (int t = 0; t < kR + MR - 1; t++) {
                    const float x = xc[t];
#pragma unroll
                    for (int r = 0; r < kR; r++) {
                        const int u = t - r;
                        if (u >= 0 && u < MR) acc[r] = fmaf(hv[u], x, acc[r]);
                    }
                }
            } else {
                for (int u = 0; u < mr; u++) {
                    const float hu = hv[u];
#pragma unroll
                    for (int r = 0; r < kR; r++) acc[r] = fmaf(hu, xc[r + u], acc[r]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kR; r++) Fl[c * kFLD + g * kR + r] = acc[r];
    }
    __syncthreads();

    // maxima: lane = row of the tile, wave w takes tile columns
    // [w * kTC/4, (w+1) * kTC/4) with a sliding window of three filtered
    // columns (rows lane-1, lane, lane+1 of each).  Row tests are fixed per
    // lane, column tests are wave-uniform; m > cutoff_f in FP32 is the same
    // test as (double)m > cutoff (cutoff_f = cutoff rounded down to FP32).
    const int gi = i0 + lane, y = gi + 1;
    bool row_ok = gi >= 1 && gi <= A.nr - 2 && y > A.dx && y <= A.nr - A.dx;
    if (A.psf_half > 0.0) row_ok = row_ok && !((double)y < A.psf_half || (double)y > A.nr - A.psf_half);
    constexpr int kCPW = kTC / (kThreads / 64);                      // tile columns per wave
    const int cb = __builtin_amdgcn_readfirstlane(wave) * kCPW;
    const float *pl = Fl + lane;                                      // row lane-1 of the tile, column 0 of Fl
    float l0 = pl[cb * kFLD], l1 = pl[cb * kFLD + 1], l2 = pl[cb * kFLD + 2];
    float c0 = pl[(cb + 1) * kFLD], c1 = pl[(cb + 1) * kFLD + 1], c2 = pl[(cb + 1) * kFLD + 2];
    for (int cc = cb; cc < cb + kCPW; cc++) {
        const int gj = j0 + cc;
        if (gj >= A.nc) break;                                        // wave-uniform
        const float *pr = pl + (cc + 2) * kFLD;
        const float r0 = pr[0], r1 = pr[1], r2 = pr[2];
        const int x = gj + 1;
        bool col_ok = gj >= 1 && gj <= A.nc - 2 && x > A.dx && x <= A.nc - A.dx;
        if (A.psf_half > 0.0) col_ok = col_ok && !((double)x < A.psf_half || (double)x > A.nr - A.psf_half);
        const float m = c1;
        const bool keep = col_ok && row_ok && m > A.cutoff_f && m > c0 && m > c2 && m > l0 && m > l1 && m > l2 &&
                          m > r0 && m > r1 && m > r2;
        const uint64_t bits = __ballot(keep);
        if (lane == 0) A.mask[((size_t)f * A.nc + gj) * A.wpc + ty] = bits;
        l0 = c0, l1 = c1, l2 = c2;
        c0 = r0, c1 = r1, c2 = r2;
    }
}

// exclusive prefix sum over the workgroup of one value per thread; *total
// receives the sum
__device__ __forceinline__ int block_exclusive_scan(int v, int *total)
{
    __shared__ int wsum[kThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int base = 0, all = 0;
    for (int w = 0; w < kThreads / 64; w++) {
        if (w < wave) base += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return base + inc - v;
}

__global__ __launch_bounds__(kThreads) void count_kernel(const uint64_t *__restrict__ mask, long W, int *__restrict__ bsum)
{
    const long w0 = (long)blockIdx.x * kCB + (long)threadIdx.x * kCW;
    int n = 0;
#pragma unroll
    for (int k = 0; k < kCW; k++)
        if (w0 + k < W) n += __popcll(mask[w0 + k]);
    int total;
    block_exclusive_scan(n, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// one workgroup: exclusive scan of the block sums in place; the grand total
// goes to *count
__global__ __launch_bounds__(kThreads) void scan_kernel(int *__restrict__ bsum, int nb, int *__restrict__ count)
{
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kThreads) {
        const int b = b0 + threadIdx.x;
        const int v = b < nb ? bsum[b] : 0;
        int total;
        const int ex = block_exclusive_scan(v, &total);
        if (b < nb) bsum[b] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *count = carry;
}

__global__ __launch_bounds__(kThreads) void emit_kernel(const uint64_t *__restrict__ mask, long W,
                                                        const int *__restrict__ boff, int wpc, int nc, int first_frame,
                                                        int cap, int *__restrict__ peaks)
{
    const long w0 = (long)blockIdx.x * kCB + (long)threadIdx.x * kCW;
    uint64_t m[kCW];
    int n = 0;
#pragma unroll
    for (int k = 0; k < kCW; k++) {
        m[k] = w0 + k < W ? mask[w0 + k] : 0;
        n += __popcll(m[k]);
    }
    int total;
    int idx = boff[blockIdx.x] + block_exclusive_scan(n, &total);
#pragma unroll
    for (int k = 0; k < kCW; k++) {
        uint64_t b = m[k];
        if (!b) continue;
        const long w = w0 + k;
        const long fc = w / wpc;
        const int q = (int)(w - fc * wpc);
        const int f = (int)(fc / nc), j = (int)(fc - (long)f * nc);
        while (b) {
            const int bit = __builtin_ctzll(b);
            b &= b - 1;
            if (idx < cap) {
                peaks[idx] = j + 1;
                peaks[cap + idx] = q * 64 + bit + 1;
                peaks[2 * cap + idx] = first_frame + f;
            }
            idx++;
        }
    }
}

struct CutArgs {
    const void *raw;
    int nr, nc;
    float offset, conv;
    const int *peaks;
    const int *count;
    int cap, first_frame, S, dx, mirror;
    float *stack;
    float *value;         // filtered value at each maximum, or NULL
    const float *h;
    int mr, mc;
    const float *sub;
};

// The filtered value at pixel (row y, column x) of frame fr: the detect
// kernel's sum for that pixel, term for term.
template <typename T>
__device__ float filtered_at(const T *fr, int nr, int nc, float offset, float conv, const float *h, int mr, int mc,
                             const float *subp, int x, int y)
{
    const int ar = (mr - 1) / 2, ac = (mc - 1) / 2;
    const float sub = subp ? *subp : 0.0f;
    float acc = 0.0f;
    for (int v = 0; v < mc; v++) {
        const int gc = x - ac + v;
        for (int u = 0; u < mr; u++) {
            const int gr = y - ar + u;
            float xv = 0.0f;
            if (gc >= 0 && gc < nc && gr >= 0 && gr < nr) {
                xv = photon(fr[(size_t)gc * nr + gr], offset, conv);
                if (subp) xv = xv - sub;
            }
            acc = fmaf(h[v * mr + u], xv, acc);
        }
    }
    return acc;
}

template <typename T>
__global__ __launch_bounds__(64) void cut_kernel(CutArgs A)
{
    const int n = min(*A.count, A.cap);
    const int SS = A.S * A.S;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int x = A.peaks[k] - 1, y = A.peaks[A.cap + k] - 1, f = A.peaks[2 * A.cap + k] - A.first_frame;
        const T *fr = (const T *)A.raw + (size_t)f * A.nr * A.nc;
        float *dst = A.stack + (size_t)k * SS;
        for (int e = threadIdx.x; e < SS; e += 64) {
            const int c = e / A.S, r = e - c * A.S;
            const int gc = A.mirror ? x + A.dx - c : x - A.dx + c;
            dst[e] = photon(fr[(size_t)gc * A.nr + (y - A.dx + r)], A.offset, A.conv);
        }
        if (A.value && threadIdx.x == 0)
            A.value[k] = filtered_at<T>(fr, A.nr, A.nc, A.offset, A.conv, A.h, A.mr, A.mc, A.sub ? A.sub + f : nullptr,
                                        x, y);
    }
}

// ---------------------------------------------------------- aggregation
// aggre (MultiD_fitter/simplefitter_spline.m:281-283, 653-673) on each
// frame's maxima [x, y, value] in scan order: while two maxima ii < jj lie
// at a distance D with 0 < D < agg_dist, the first such pair in MATLAB's
// column-major find order (smallest jj, then smallest ii) is replaced by
// round(mean) of the two rows (x, y and value; round half away from zero) in
// row ii, and row jj is deleted.  D == 0 never merges (the reference sets
// those entries to NaN).  Coordinates are integers, so D < agg_dist is
// d2 = dx^2 + dy^2 <= d2max with d2max the largest integer whose correctly
// rounded square root is below agg_dist (agg_d2max, on the host).  After a
// merge into ii the first close pair can only lie in column ii or later (the
// pairs of earlier columns are unchanged), so the search resumes there.
// Rows keep their first index for the whole merge (a deleted row is marked,
// x = -1, and dropped when the survivors are packed), so "ii < jj" compares
// indices; each frame's maxima sit in lists of cs x cs cells (cs >= the
// largest |dx| or |dy| of a close pair), so the first ii of column jj is the
// smallest index in the 3 x 3 cells around it instead of a scan of every
// earlier row.
struct AggArgs {
    const void *raw;
    int nr, nc;
    float offset, conv;
    const float *h;
    int mr, mc;
    const float *sub;
    int *cand;            // [3, ccap]: x, y, frame number of the maxima, frames in order
    double *cval;         // [ccap] values (when requested), or NULL
    const int *ccount;    // maxima found (at most ccap)
    int ccap, first_frame;
    long long d2max;
    double psf_half;      // > 0: the PSF-convolution mode's border test, after the merge
    int *fkept, *foff;    // per frame: maxima kept (both; foff is then scanned to offsets)
    int *fseg;            // per frame: first index in cand
    int *cnext;           // [ccap]: the next maximum in the same cell, -1 ends the list
    int *chead;           // [frames, gx, gy]: the first maximum of each cell, -1: none
    int cs, gx, gy;       // cell size; cells along x (columns) and y (rows)
};

// The cell heads are built with atomics, which act at L2, so every access
// to them goes to L2 (agent-scope atomic loads and stores) and none can see
// a stale line in the CU's vector cache.
__device__ __forceinline__ int head_load(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void head_store(int *p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Smallest index below `bound` among the maxima of the 3 x 3 cells around
// (x, y) whose squared distance d2 from it is in (0, d2max]; bound if none.
__device__ __forceinline__ int first_close(const AggArgs &A, const int *xs, const int *ys, const int *cnext,
                                           const int *chead, int x, int y, int bound)
{
    const int cx = x / A.cs, cy = y / A.cs;
    const int a0 = cx > 0 ? cx - 1 : 0, a1 = cx + 1 < A.gx ? cx + 1 : A.gx - 1;
    const int b0 = cy > 0 ? cy - 1 : 0, b1 = cy + 1 < A.gy ? cy + 1 : A.gy - 1;
    for (int a = a0; a <= a1; a++)
        for (int b = b0; b <= b1; b++)
            for (int k = head_load(&chead[a * A.gy + b]); k >= 0; k = cnext[k]) {
                if (k >= bound) continue;
                const long long dxv = xs[k] - x, dyv = ys[k] - y;
                const long long d2 = dxv * dxv + dyv * dyv;
                if (d2 > 0 && d2 <= A.d2max) bound = k;
            }
    return bound;
}

// Take maximum k out of the list of cell c (one thread).
__device__ __forceinline__ void cell_unlink(int *cnext, int *chead, int c, int k)
{
    int q = head_load(&chead[c]);
    if (q == k) {
        head_store(&chead[c], cnext[k]);
        return;
    }
    while (q >= 0 && cnext[q] != k) q = cnext[q];   // (k is always in the list)
    if (q >= 0) cnext[q] = cnext[k];
}

template <typename T>
__global__ __launch_bounds__(kThreads) void aggregate_kernel(AggArgs A)
{
    const int f = blockIdx.x, tid = threadIdx.x;
    const int total = min(*A.ccount, A.ccap);
    // the frame's run of maxima (the frame column is non-decreasing)
    const int *fr = A.cand + 2 * (size_t)A.ccap;
    const int key = A.first_frame + f;
    int lo = 0, hi = total;
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (fr[m] < key) lo = m + 1;
        else hi = m;
    }
    int end = lo, top = total;
    while (end < top) {
        const int m = (end + top) >> 1;
        if (fr[m] <= key) end = m + 1;
        else top = m;
    }
    int *xs = A.cand + lo, *ys = A.cand + A.ccap + lo;
    double *vs = A.cval ? A.cval + lo : nullptr;
    int n = end - lo;
    if (vs) {
        const T *frm = (const T *)A.raw + (size_t)f * A.nr * A.nc;
        for (int k = tid; k < n; k += kThreads)
            vs[k] = (double)filtered_at<T>(frm, A.nr, A.nc, A.offset, A.conv, A.h, A.mr, A.mc,
                                           A.sub ? A.sub + f : nullptr, xs[k] - 1, ys[k] - 1);
    }
    // every maximum into the list of its cell (no merge possible: d2max < 1)
    const bool merge = A.d2max >= 1;
    int *cnext = A.cnext + lo;
    int *chead = A.chead + (size_t)f * A.gx * A.gy;
    auto cell = [&](int x, int y) { return (x / A.cs) * A.gy + y / A.cs; };
    if (merge) {
        for (int c = tid; c < A.gx * A.gy; c += kThreads) head_store(&chead[c], -1);
        __syncthreads();
        for (int k = tid; k < n; k += kThreads) cnext[k] = atomicExch(&chead[cell(xs[k], ys[k])], k);
        __syncthreads();
    }
    __shared__ unsigned long long best;
    int start = 1;
    while (merge) {
        if (tid == 0) best = ~0ull;
        __syncthreads();
        // each thread takes columns jj = start + tid, + 256, ... and stops at
        // its first close pair (or once a smaller column has one)
        for (int jj = start + tid; jj < n; jj += kThreads) {
            if (((unsigned long long)jj << 32) > *(volatile unsigned long long *)&best) break;
            const int xj = xs[jj];
            if (xj < 0) continue;   // merged away
            const int hit = first_close(A, xs, ys, cnext, chead, xj, ys[jj], jj);
            if (hit < jj) {
                atomicMin(&best, ((unsigned long long)jj << 32) | (unsigned)hit);
                break;
            }
        }
        __syncthreads();
        const unsigned long long b = best;
        if (b == ~0ull) break;
        const int jj = (int)(b >> 32), ii = (int)(b & 0xffffffffu);
        if (tid == 0) {
            const int xo = xs[ii], yo = ys[ii];
            const int xn = (int)round(((double)xo + (double)xs[jj]) / 2.0);
            const int yn = (int)round(((double)yo + (double)ys[jj]) / 2.0);
            if (vs) vs[ii] = round((vs[ii] + vs[jj]) / 2.0);
            cell_unlink(cnext, chead, cell(xs[jj], ys[jj]), jj);
            xs[jj] = -1;
            const int co = cell(xo, yo), cn = cell(xn, yn);
            if (co != cn) {
                cell_unlink(cnext, chead, co, ii);
                cnext[ii] = head_load(&chead[cn]);
                head_store(&chead[cn], ii);
            }
            xs[ii] = xn;
            ys[ii] = yn;
        }
        __syncthreads();
        start = ii > 1 ? ii : 1;
    }
    // the survivors packed in order, with the PSF-convolution mode's border
    // test (simplefitter_spline.m:284-295, rows used for both axes) after the
    // merge
    int kept = n;
    if (merge || A.psf_half > 0.0) {
        const double hp = A.psf_half, lim = (double)A.nr - hp;
        const bool psf = hp > 0.0;
        int base = 0;
        for (int k0 = 0; k0 < n; k0 += kThreads) {
            const int k = k0 + tid;
            int tx = 0, ty = 0;
            double tv = 0.0;
            bool keep = false;
            if (k < n) {
                tx = xs[k];
                ty = ys[k];
                if (vs) tv = vs[k];
                keep = tx >= 0 &&
                       !(psf && ((double)tx < hp || (double)tx > lim || (double)ty < hp || (double)ty > lim));
            }
            int cnt;
            const int pos = base + block_exclusive_scan(keep ? 1 : 0, &cnt);
            if (keep) {
                xs[pos] = tx;
                ys[pos] = ty;
                if (vs) vs[pos] = tv;
            }
            base += cnt;
            __syncthreads();
        }
        kept = base;
    }
    if (tid == 0) {
        A.fkept[f] = kept;
        A.foff[f] = kept;
        A.fseg[f] = lo;
    }
}

// The frames' kept maxima, in frame order, into the outputs: peaks [cap, 3]
// and value (float) at the frame's offset (fcount scanned to offsets).
__global__ __launch_bounds__(kThreads) void gather_kernel(const int *__restrict__ cand, const double *__restrict__ cval,
                                                          int ccap, const int *__restrict__ fkept,
                                                          const int *__restrict__ foff, const int *__restrict__ fseg,
                                                          int cap, int *__restrict__ peaks, float *__restrict__ value)
{
    const int f = blockIdx.x;
    const int n = fkept[f], off = foff[f], lo = fseg[f];
    for (int k = threadIdx.x; k < n; k += kThreads) {
        const int o = off + k;
        if (o >= cap) break;
        peaks[o] = cand[lo + k];
        peaks[cap + o] = cand[ccap + lo + k];
        peaks[2 * cap + o] = cand[2 * (size_t)ccap + lo + k];
        if (value) value[o] = (float)cval[lo + k];
    }
}

__global__ void zero_count_kernel(int *count) { *count = 0; }

// The largest integer d2 with sqrt(d2) < thr (correctly rounded, as MATLAB
// computes the distance of two integer points), or -1 when thr <= 0; capped
// above the largest squared distance within an nr x nc frame.
long long agg_d2max(double thr, int nr, int nc)
{
    if (!(thr > 0.0)) return -1;
    const double lim = (double)nr * nr + (double)nc * nc;
    const double t2 = thr * thr;
    long long k = t2 > lim ? (long long)lim + 1 : (long long)t2 + 2;
    while (k > 0 && !(std::sqrt((double)k) < thr)) k--;
    return k;
}

// The merge's cells: cs x cs pixels, cs at least the largest |dx| of a
// close pair (the integer square root of d2max) and at least 4 (the lists'
// heads then take at most a quarter of a byte per pixel).
struct AggGrid {
    long long d2max;
    int cs, gx, gy;
};

AggGrid agg_grid(double thr, int nr, int nc)
{
    AggGrid g;
    g.d2max = agg_d2max(thr, nr, nc);
    long long r = g.d2max > 0 ? (long long)std::sqrt((double)g.d2max) : 0;
    while (r * r > g.d2max) r--;
    while ((r + 1) * (r + 1) <= g.d2max) r++;
    g.cs = (int)(r > 4 ? r : 4);
    g.gx = nc / g.cs + 1;
    g.gy = nr / g.cs + 1;
    return g;
}

struct WorkLayout {
    size_t sub, mask, bsum, total;
    // with aggregation: every maximum of every frame (ccap: the most a frame
    // of nr x nc can hold, strict maxima being never adjacent), their values,
    // their count, the per-frame kept counts, offsets and first indices, and
    // the merge's cell lists (a link per maximum, gx x gy heads per frame)
    size_t cand, cval, ccount, fkept, foff, fseg, cnext, chead;
    long W;
    int nb, wpc, ccap;
    AggGrid grid;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

WorkLayout work_layout(int nr, int nc, int nframes, double agg_dist)
{
    const bool agg = agg_dist > 0.0;
    WorkLayout w;
    w.wpc = (nr + kTR - 1) / kTR;
    w.W = (long)nframes * nc * w.wpc;
    w.nb = (int)((w.W + kCB - 1) / kCB);
    w.sub = 0;
    w.mask = align256(sizeof(float) * (size_t)nframes);
    w.bsum = w.mask + align256(sizeof(uint64_t) * (size_t)w.W);
    w.total = w.bsum + align256(sizeof(int) * (size_t)(w.nb > 0 ? w.nb : 1));
    w.ccap = 0;
    w.cand = w.cval = w.ccount = w.fkept = w.foff = w.fseg = w.cnext = w.chead = w.total;
    w.grid = AggGrid{-1, 4, 1, 1};
    if (agg) {
        w.grid = agg_grid(agg_dist, nr, nc);
        const size_t per = (size_t)((nr - 1) / 2) * (size_t)((nc - 1) / 2);
        const size_t cc = per * (size_t)(nframes > 0 ? nframes : 0);
        w.ccap = (int)cc;   // below 2^31 (the entry points check the bound)
        const size_t nf = (size_t)(nframes > 0 ? nframes : 1);
        w.cand = w.total;
        w.cval = w.cand + align256(sizeof(int) * 3 * (cc ? cc : 1));
        w.ccount = w.cval + align256(sizeof(double) * (cc ? cc : 1));
        w.fkept = w.ccount + 256;
        w.foff = w.fkept + align256(sizeof(int) * nf);
        w.fseg = w.foff + align256(sizeof(int) * nf);
        w.cnext = w.fseg + align256(sizeof(int) * nf);
        w.chead = w.cnext + align256(sizeof(int) * (cc ? cc : 1));
        w.total = w.chead + align256(sizeof(int) * nf * (size_t)w.grid.gx * (size_t)w.grid.gy);
    }
    return w;
}

using DetectFn = void (*)(DetectArgs);

template <typename T, size_t... I>
DetectFn detect_fn(int mr, std::index_sequence<I...>)
{
    // mr == 0 selects the generic kernel; mr in 1..kMaxTemplR the unrolled ones
    static const DetectFn table[] = {detect_kernel<T, (int)I>...};
    return mr <= kMaxTemplR ? table[mr] : table[0];
}

template <typename T>
hipError_t launch_typed(const FrontLaunch &L, hipStream_t st)
{
    const bool agg = L.agg_dist > 0.0;
    const WorkLayout w = work_layout(L.nr, L.nc, L.nframes, L.agg_dist);
    char *work = (char *)L.work;
    float *sub = L.subtract_min ? (float *)(work + w.sub) : nullptr;
    uint64_t *mask = (uint64_t *)(work + w.mask);
    int *bsum = (int *)(work + w.bsum);
    const T *raw = (const T *)L.frames;

    if (sub) hipLaunchKernelGGL(column_min_kernel<T>, dim3(L.nframes), dim3(kThreads), 0, st, raw, L.nr, L.nc,
                                L.offset, L.conv, sub);

    DetectArgs A;
    A.raw = L.frames;
    A.nr = L.nr;
    A.nc = L.nc;
    A.offset = L.offset;
    A.conv = L.conv;
    A.h = L.h;
    A.mr = L.mr;
    A.mc = L.mc;
    A.sub = sub;
    // the largest float <= cutoff: m > it  <=>  (double)m > cutoff
    float cf = (float)L.cutoff;
    if ((double)cf > L.cutoff) cf = std::nextafter(cf, -HUGE_VALF);
    A.cutoff_f = cf;
    A.dx = L.S / 2;
    // with aggregation the PSF border test follows the merge (aggregate_kernel)
    const double psf_half = L.psf_size > 0.0 ? L.psf_size / 2.0 : 0.0;
    A.psf_half = agg ? 0.0 : psf_half;
    A.mask = mask;
    A.wpc = w.wpc;
    const size_t lds = front_detect_lds_bytes(L.mr, L.mc);
    const DetectFn fn = detect_fn<T>(L.mr, std::make_index_sequence<kMaxTemplR + 1>{});
    A.tiles_x = (L.nc + kTC - 1) / kTC;
    const int per_frame = A.tiles_x * w.wpc;
    const int fchunk = (int)((1L << 30) / per_frame) > 0 ? (int)((1L << 30) / per_frame) : 1;   // grid < 2^30
    for (int f0 = 0; f0 < L.nframes; f0 += fchunk) {
        A.f0 = f0;
        const int nf = L.nframes - f0 < fchunk ? L.nframes - f0 : fchunk;
        const dim3 grid((unsigned)(per_frame * nf));
        void *args[] = {&A};
        const hipError_t e = hipLaunchKernel((const void *)fn, grid, dim3(kThreads), args, lds, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(count_kernel, dim3((unsigned)w.nb), dim3(kThreads), 0, st, mask, w.W, bsum);
    if (!agg) {
        hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(kThreads), 0, st, bsum, w.nb, L.count);
        hipLaunchKernelGGL(emit_kernel, dim3((unsigned)w.nb), dim3(kThreads), 0, st, mask, w.W, bsum, w.wpc, L.nc,
                           L.first_frame, L.cap, L.peaks);
    } else {
        // every maximum into the candidate list, merged per frame, packed
        int *cand = (int *)(work + w.cand), *ccount = (int *)(work + w.ccount);
        int *fkept = (int *)(work + w.fkept), *foff = (int *)(work + w.foff), *fseg = (int *)(work + w.fseg);
        double *cval = L.value ? (double *)(work + w.cval) : nullptr;
        hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(kThreads), 0, st, bsum, w.nb, ccount);
        hipLaunchKernelGGL(emit_kernel, dim3((unsigned)w.nb), dim3(kThreads), 0, st, mask, w.W, bsum, w.wpc, L.nc,
                           L.first_frame, w.ccap, cand);
        AggArgs G;
        G.raw = L.frames;
        G.nr = L.nr;
        G.nc = L.nc;
        G.offset = L.offset;
        G.conv = L.conv;
        G.h = L.h;
        G.mr = L.mr;
        G.mc = L.mc;
        G.sub = sub;
        G.cand = cand;
        G.cval = cval;
        G.ccount = ccount;
        G.ccap = w.ccap;
        G.first_frame = L.first_frame;
        G.d2max = w.grid.d2max;
        G.cnext = (int *)(work + w.cnext);
        G.chead = (int *)(work + w.chead);
        G.cs = w.grid.cs;
        G.gx = w.grid.gx;
        G.gy = w.grid.gy;
        G.psf_half = psf_half;
        G.fkept = fkept;
        G.foff = foff;
        G.fseg = fseg;
        hipLaunchKernelGGL(aggregate_kernel<T>, dim3((unsigned)L.nframes), dim3(kThreads), 0, st, G);
        hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(kThreads), 0, st, foff, L.nframes, L.count);
        hipLaunchKernelGGL(gather_kernel, dim3((unsigned)L.nframes), dim3(kThreads), 0, st, cand, cval, w.ccap, fkept,
                           foff, fseg, L.cap, L.peaks, L.value);
    }
    if (L.cap > 0) {
        CutArgs C;
        C.raw = L.frames;
        C.nr = L.nr;
        C.nc = L.nc;
        C.offset = L.offset;
        C.conv = L.conv;
        C.peaks = L.peaks;
        C.count = L.count;
        C.cap = L.cap;
        C.first_frame = L.first_frame;
        C.S = L.S;
        C.dx = L.S / 2;
        C.mirror = L.mirror;
        C.stack = L.stack;
        C.value = agg ? nullptr : L.value;   // with aggregation the gather wrote the values
        C.h = L.h;
        C.mr = L.mr;
        C.mc = L.mc;
        C.sub = sub;
        const int blocks = L.cap < 8192 ? L.cap : 8192;
        hipLaunchKernelGGL(cut_kernel<T>, dim3((unsigned)blocks), dim3(64), 0, st, C);
    }
    return hipGetLastError();
}

}  // namespace

size_t front_detect_lds_bytes(int mr, int mc)
{
    const int XR = kFR + mr - 1, XC = kFC + mc - 1;
    return sizeof(float) * ((size_t)XC * (XR | 1) + (size_t)kFC * kFLD);
}

size_t front_work_bytes(int nr, int nc, int nframes, double agg_dist)
{
    return work_layout(nr, nc, nframes, agg_dist).total;
}

// Workspace of the allocation-free entry point (bsf_find_rois_device_noalloc):
// one array per GPU for the life of the process.  Nothing in it needs to
// start at zero (the column minima, every mask word and every block sum are
// written before they are read), and calls that overlap on one GPU share it
// and race.  In host code `scratch` only stands in for the symbol: its device
// address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[kFrontScratchBytes];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kFrontScratchDevices];
}  // namespace

hipError_t front_scratch(void **ptr)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kFrontScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return hipSuccess;
}

hipError_t launch_find_rois(const FrontLaunch &L, hipStream_t st)
{
    if (L.nframes <= 0) {
        hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, st, L.count);
        return hipGetLastError();
    }
    return L.pixel_type == 0 ? launch_typed<uint16_t>(L, st) : launch_typed<float>(L, st);
}

}  // namespace bsf
