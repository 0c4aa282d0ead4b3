/*
 * Oracle (test infrastructure, not shipped): CPU restatement of the
 * reference's frame -> ROI front-end, the part of the GUI fitter that turns
 * camera frames into the ROI stack the MEX fits.
 *
 *   MultiD_fitter/simplefitter_spline.m:259   imphot = (single(image) - p.offset) * p.conversion
 *   MultiD_fitter/simplefitter_spline.m:269-276  impf = filter2(h, imphot) (Gaussian mode) or
 *                                            filter2(h, imphot - min(imphot(:,1))) (DoG / PSF modes)
 *   MultiD_fitter/shared/maximumfindcall.m:4-21, maximumfind2.c:18-65
 *        strict 8-neighbour maxima of the filtered frame, interior pixels only,
 *        scanned column by column (x = column + 1, y = row + 1, value)
 *   MultiD_fitter/simplefitter_spline.m:277-280  value > peakcutoff, dx < x <= ncols - dx,
 *                                            dx < y <= nrows - dx
 *   MultiD_fitter/simplefitter_spline.m:281-283, 653-673  with p.agg: aggre merges maxima
 *                                            closer than p.roifit * p.aggsz (orc_aggregate)
 *   MultiD_fitter/simplefitter_spline.m:284-295  PSF-convolution mode: drop maxima within
 *                                            size(PSF,1)/2 of the border (nrows used for
 *                                            both axes, as the reference does)
 *   MultiD_fitter/simplefitter_spline.m:296-307  ROI = imphot(y-dx:y+dx, x-dx:x+dx), columns
 *                                            reversed when p.mirror; peakcoordinates = [x, y, F]
 *
 * Arithmetic (single precision, as MATLAB computes on singles): the photon
 * conversion rounds after the subtraction and after the product; filter2 is
 * the 'same'-size correlation with zero padding, tap (u, v) of an mr x mc
 * kernel reading X(i + u - floor((mr-1)/2), j + v - floor((mc-1)/2));
 * its sum is accumulated here as acc = fmaf(h[v*mr+u], X, acc) from 0, v
 * outer, u inner, zero-padded taps included.  MATLAB's own conv2 summation
 * order is not published, so this order is a definition, shared with the
 * GPU kernel; the tests pin it against an FP64 evaluation of filter2's
 * documented 'same' definition.  Arrays are column-major: element (row i,
 * column j) of frame f at f*nr*nc + j*nr + i.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static float photon_u16(uint16_t v, float offset, float conv)
{
    const float d = (float)v - offset;
    return d * conv;
}

static float photon_f32(float v, float offset, float conv)
{
    const float d = v - offset;
    return d * conv;
}

/* pixel_type 0: uint16, 1: float32 */
void orc_photons(const void *raw, int pixel_type, long n, float offset, float conv, float *out)
{
    for (long k = 0; k < n; k++)
        out[k] = pixel_type == 0 ? photon_u16(((const uint16_t *)raw)[k], offset, conv)
                                 : photon_f32(((const float *)raw)[k], offset, conv);
}

/* filter2(h, x, 'same') in the order stated above. */
void orc_filter2(const float *x, int nr, int nc, const float *h, int mr, int mc, float *out)
{
    const int ar = (mr - 1) / 2, ac = (mc - 1) / 2;
    for (int j = 0; j < nc; j++)
        for (int i = 0; i < nr; i++) {
            float acc = 0.0f;
            for (int v = 0; v < mc; v++) {
                const int jj = j - ac + v;
                for (int u = 0; u < mr; u++) {
                    const int ii = i - ar + u;
                    const float xv = (ii >= 0 && ii < nr && jj >= 0 && jj < nc) ? x[(long)jj * nr + ii] : 0.0f;
                    acc = fmaf(h[v * mr + u], xv, acc);
                }
            }
            out[(long)j * nr + i] = acc;
        }
}

/* maximumfind2's strict maxima plus the caller's filters, in scan order.
 * Writes up to cap entries (x, y 1-based; value) and returns the total. */
long orc_find_maxima(const float *F, int nr, int nc, double cutoff, int dx, double psf_size, int *xs, int *ys,
                     float *vals, long cap)
{
    long n = 0;
    for (int j = 1; j < nc - 1; j++)
        for (int i = 1; i < nr - 1; i++) {
            const float m = F[(long)j * nr + i];
            int is_max = 1;
            for (int dj = -1; dj <= 1 && is_max; dj++)
                for (int di = -1; di <= 1; di++) {
                    if (!di && !dj) continue;
                    if (!(m > F[(long)(j + dj) * nr + i + di])) {
                        is_max = 0;
                        break;
                    }
                }
            if (!is_max) continue;
            const int x = j + 1, y = i + 1;
            if (!((double)m > cutoff)) continue;
            if (!(x > dx && x <= nc - dx && y > dx && y <= nr - dx)) continue;
            if (psf_size > 0) {
                const double hp = psf_size / 2.0;
                if ((double)x < hp || (double)x > nr - hp || (double)y < hp || (double)y > nr - hp) continue;
            }
            if (n < cap) {
                xs[n] = x;
                ys[n] = y;
                if (vals) vals[n] = m;
            }
            n++;
        }
    return n;
}

/* aggre (MultiD_fitter/simplefitter_spline.m:653-673) on one frame's maxima
 * [x, y, value] (the reference's maxgood, double): D(ii, jj) is the distance
 * between maxima ii < jj, D == 0 counts as NaN (never merged); while some
 * D < thr, the first such pair in MATLAB's column-major find order (smallest
 * jj, then smallest ii) becomes round(mean) of the two rows in row ii (x, y
 * and value alike; MATLAB's round is half away from zero, as C's) and row jj
 * is deleted.  Returns the new count.  Literal O(n^2) search per merge. */
long orc_aggregate(int *xs, int *ys, double *vs, long n, double thr)
{
    for (;;) {
        long fi = -1, fj = -1;
        for (long jj = 1; jj < n && fj < 0; jj++)
            for (long ii = 0; ii < jj; ii++) {
                const double dxv = (double)xs[ii] - (double)xs[jj], dyv = (double)ys[ii] - (double)ys[jj];
                const double D = sqrt(dxv * dxv + dyv * dyv);
                if (D != 0.0 && D < thr) {
                    fi = ii;
                    fj = jj;
                    break;
                }
            }
        if (fj < 0) return n;
        xs[fi] = (int)round(((double)xs[fi] + (double)xs[fj]) / 2.0);
        ys[fi] = (int)round(((double)ys[fi] + (double)ys[fj]) / 2.0);
        vs[fi] = round((vs[fi] + vs[fj]) / 2.0);
        for (long k = fj; k + 1 < n; k++) {
            xs[k] = xs[k + 1];
            ys[k] = ys[k + 1];
            vs[k] = vs[k + 1];
        }
        n--;
    }
}

/* The whole front-end over nframes frames.  Outputs: stack [S, S, cap],
 * peaks [cap, 3] column-major (x, y, frame number), vals [cap] or NULL.
 * With agg_dist > 0 (p.agg: p.roifit * p.aggsz) each frame's maxima go
 * through orc_aggregate before the PSF-convolution mode's border test, in
 * the reference's order (simplefitter_spline.m:281-295); vals then holds the
 * aggregated value column, round(mean) for merged maxima.
 * Returns the number of ROIs found (which may exceed cap). */
long orc_find_rois(const void *raw, int pixel_type, int nr, int nc, int nframes, int first_frame, float offset,
                   float conv, const float *h, int mr, int mc, int subtract_min, double cutoff, int S, int mirror,
                   double psf_size, long cap, float *stack, int *peaks, float *vals, double agg_dist)
{
    const long np = (long)nr * nc;
    float *ph = (float *)malloc(sizeof(float) * np);
    float *X = (float *)malloc(sizeof(float) * np);
    float *F = (float *)malloc(sizeof(float) * np);
    int *xs = (int *)malloc(sizeof(int) * (np / 4 + 8));
    int *ys = (int *)malloc(sizeof(int) * (np / 4 + 8));
    float *vs = (float *)malloc(sizeof(float) * (np / 4 + 8));
    double *vd = (double *)malloc(sizeof(double) * (np / 4 + 8));
    const int dx = S / 2;
    const int agg = agg_dist > 0.0;
    long total = 0;
    for (int f = 0; f < nframes; f++) {
        const char *fr = (const char *)raw + (size_t)f * np * (pixel_type == 0 ? 2 : 4);
        orc_photons(fr, pixel_type, np, offset, conv, ph);
        if (subtract_min) {
            float mn = ph[0];
            for (int i = 1; i < nr; i++) mn = fminf(mn, ph[i]);
            for (long k = 0; k < np; k++) X[k] = ph[k] - mn;
        } else {
            memcpy(X, ph, sizeof(float) * np);
        }
        orc_filter2(X, nr, nc, h, mr, mc, F);
        long nf = orc_find_maxima(F, nr, nc, cutoff, dx, agg ? 0.0 : psf_size, xs, ys, vs, np / 4 + 8);
        for (long q = 0; q < nf; q++) vd[q] = (double)vs[q];
        if (agg) {
            nf = orc_aggregate(xs, ys, vd, nf, agg_dist);
            if (psf_size > 0) {   /* the border test after the merge, in place */
                const double hp = psf_size / 2.0;
                long m = 0;
                for (long q = 0; q < nf; q++) {
                    if ((double)xs[q] < hp || (double)xs[q] > nr - hp || (double)ys[q] < hp || (double)ys[q] > nr - hp)
                        continue;
                    xs[m] = xs[q];
                    ys[m] = ys[q];
                    vd[m] = vd[q];
                    m++;
                }
                nf = m;
            }
        }
        for (long q = 0; q < nf; q++, total++) {
            if (total >= cap) continue;
            const int x = xs[q] - 1, y = ys[q] - 1;
            for (int c = 0; c < S; c++) {
                const int gc = mirror ? x + dx - c : x - dx + c;
                for (int r = 0; r < S; r++)
                    stack[(size_t)total * S * S + (size_t)c * S + r] = ph[(long)gc * nr + (y - dx + r)];
            }
            peaks[total] = xs[q];
            peaks[cap + total] = ys[q];
            peaks[2 * cap + total] = first_frame + f;
            if (vals) vals[total] = (float)vd[q];
        }
    }
    free(ph);
    free(X);
    free(F);
    free(xs);
    free(ys);
    free(vs);
    free(vd);
    return total;
}
