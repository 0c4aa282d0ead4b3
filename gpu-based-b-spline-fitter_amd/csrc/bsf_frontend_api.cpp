// bsf_frontend_api.cpp -- C-ABI of the frame -> ROI front-end
// (include/bsfit.h): argument checks, the peak-finding filter kernels of the
// GUI fitter, device buffers for the host entry point, launches.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/bsfit.h"
#include "bsf_kernels.h"

namespace {

using bsf::fail;

int check_gfx950()
{
    bsf::DeviceInfo d;
    int rc = bsf::device_info(&d);
    if (rc) return rc;
    if (!d.gfx950)
        return fail(BSF_ERR_DEVICE, "bsf: device %d is %s; this build targets gfx950 (MI355X)", d.dev, d.arch);
    return BSF_OK;
}

int validate_rois(const bsf_roi_request *r, const bsf_roi_result *res)
{
    if (!r || !res) return fail(BSF_ERR_ARG, "bsf: NULL request or result");
    if (r->pixel_type != BSF_PIXEL_U16 && r->pixel_type != BSF_PIXEL_F32)
        return fail(BSF_ERR_ARG, "bsf: unknown pixel type %d", r->pixel_type);
    if (r->nframes < 0) return fail(BSF_ERR_ARG, "bsf: nframes must be >= 0");
    if (r->nrows < 3 || r->ncols < 3)
        return fail(BSF_ERR_SHAPE, "bsf: frames must be at least 3 x 3 (got %d x %d)", r->nrows, r->ncols);
    if (r->nframes > 0 && !r->frames) return fail(BSF_ERR_ARG, "bsf: NULL frames");
    if (r->roi < 3 || r->roi > BSF_MAX_ROI || r->roi % 2 == 0)
        return fail(BSF_ERR_ROI_SIZE, "bsf: the ROI size must be odd, 3..%d (got %d)", BSF_MAX_ROI, r->roi);
    if (!r->filter) return fail(BSF_ERR_ARG, "bsf: NULL filter kernel");
    if (r->filter_rows < 1 || r->filter_cols < 1 || r->filter_rows > BSF_MAX_PEAK_FILTER ||
        r->filter_cols > BSF_MAX_PEAK_FILTER)
        return fail(BSF_ERR_SHAPE, "bsf: filter kernel must be 1..%d on each side (got %d x %d)", BSF_MAX_PEAK_FILTER,
                    r->filter_rows, r->filter_cols);
    if (r->capacity < 0) return fail(BSF_ERR_ARG, "bsf: capacity must be >= 0");
    if (r->capacity > 0 && (!res->imstack || !res->peaks))
        return fail(BSF_ERR_ARG, "bsf: NULL imstack or peaks output");
    if (!std::isfinite(r->offset) || !std::isfinite(r->conversion))
        return fail(BSF_ERR_ARG, "bsf: offset and conversion must be finite");
    if (!(r->agg_dist >= 0.0) || !std::isfinite(r->agg_dist))
        return fail(BSF_ERR_ARG, "bsf: agg_dist must be finite and >= 0 (0: no aggregation)");
    // strict maxima are never adjacent: at most ceil(nr/2) * ceil(nc/2) per frame
    const double most = (double)r->nframes * ((r->nrows + 1) / 2) * ((r->ncols + 1) / 2);
    if (most > (double)INT_MAX)
        return fail(BSF_ERR_ARG, "bsf: too many frames in one call (maxima could exceed 2^31)");
    return BSF_OK;
}

// workspace of a request (aggregation holds every maximum of the frames)
size_t work_bytes_of(const bsf_roi_request *r)
{
    return bsf::front_work_bytes(r->nrows, r->ncols, r->nframes, r->agg_dist);
}

bsf::FrontLaunch make_front(const bsf_roi_request *r, const void *frames, const float *h, float *stack, int *peaks,
                            float *value, int *count, void *work)
{
    bsf::FrontLaunch L;
    std::memset(&L, 0, sizeof L);
    L.frames = frames;
    L.pixel_type = r->pixel_type;
    L.nr = r->nrows;
    L.nc = r->ncols;
    L.nframes = r->nframes;
    L.first_frame = r->first_frame;
    L.offset = r->offset;
    L.conv = r->conversion;
    L.h = h;
    L.mr = r->filter_rows;
    L.mc = r->filter_cols;
    L.subtract_min = r->subtract_min != 0;
    L.cutoff = r->cutoff;
    L.S = r->roi;
    L.mirror = r->mirror != 0;
    L.psf_size = r->psf_size;
    L.agg_dist = r->agg_dist;
    L.cap = r->capacity;
    L.stack = stack;
    L.peaks = peaks;
    L.value = value;
    L.count = count;
    L.work = work;
    return L;
}

// fspecial('gaussian', n, sigma): exp(-(x^2 + y^2) / (2 sigma^2)) on the n x n
// grid centred at (n-1)/2, values below eps * max set to 0, normalised to sum 1.
void gaussian(int n, double sigma, std::vector<double> &h)
{
    h.assign((size_t)n * n, 0.0);
    const double c = (n - 1) / 2.0;
    double mx = 0.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            const double x = j - c, y = i - c;
            const double v = std::exp(-(x * x + y * y) / (2.0 * sigma * sigma));
            h[(size_t)j * n + i] = v;
            mx = v > mx ? v : mx;
        }
    double s = 0.0;
    for (double &v : h) {
        if (v < 2.220446049250313e-16 * mx) v = 0.0;
        s += v;
    }
    if (s != 0.0)
        for (double &v : h) v /= s;
}

}  // namespace

extern "C" {

int bsf_peak_filter(int kind, double sigma, float *h, int cap, int *size)
{
    bsf::clear_error();
    if (!h || !size) return fail(BSF_ERR_ARG, "bsf: NULL output");
    if (!(sigma > 0.0) || !std::isfinite(sigma)) return fail(BSF_ERR_ARG, "bsf: sigma must be > 0");
    int n;
    std::vector<double> a, b;
    if (kind == BSF_PEAK_GAUSS) {
        n = (int)std::ceil(3.0 * sigma + 1.0);
        n = n > 3 ? n : 3;
    } else if (kind == BSF_PEAK_DOG) {
        n = (int)std::ceil(6.0 * sigma + 1.0);
        n = n > 3 ? n : 3;
    } else {
        return fail(BSF_ERR_ARG, "bsf: unknown peak filter kind %d", kind);
    }
    if (n > BSF_MAX_PEAK_FILTER) return fail(BSF_ERR_SHAPE, "bsf: sigma %g gives a %d x %d kernel (max %d)", sigma, n, n,
                                             BSF_MAX_PEAK_FILTER);
    *size = n;
    if (cap < n * n) return fail(BSF_ERR_ARG, "bsf: the kernel needs %d floats (cap %d)", n * n, cap);
    gaussian(n, sigma, a);
    if (kind == BSF_PEAK_DOG) {
        gaussian(n, sigma * 2.5 > 1.0 ? sigma * 2.5 : 1.0, b);
        for (size_t k = 0; k < a.size(); k++) a[k] -= b[k];
    }
    for (size_t k = 0; k < a.size(); k++) h[k] = (float)a[k];
    return BSF_OK;
}

int bsf_peak_filter_psf(int kind, const float *psf, const int *psf_dims, int roi, float *h, int cap, int *rows,
                        int *cols)
{
    bsf::clear_error();
    if (!psf || !psf_dims || !h || !rows || !cols) return fail(BSF_ERR_ARG, "bsf: NULL argument");
    const int n = psf_dims[0], n2 = psf_dims[1], nz = psf_dims[2];
    if (n < 1 || n2 != n || nz < 1) return fail(BSF_ERR_SHAPE, "bsf: the PSF must be n x n x nz (got %d x %d x %d)",
                                                n, n2, nz);
    const size_t plane = (size_t)n * n;
    if (kind == BSF_PEAK_PSF_WINDOW) {
        // rows/columns floor((n - roi)/2) .. floor((n + roi)/2) (1-based) of the
        // z-mean, minus its mean, scaled to 0.05 at its maximum
        const int lo = (n - roi) / 2, hi = (n + roi) / 2;      // n, roi >= 1
        if (roi < 1 || n - roi < 2) return fail(BSF_ERR_SHAPE, "bsf: PSF of size %d too small for ROI %d", n, roi);
        const int m = hi - lo + 1;
        if (m > BSF_MAX_PEAK_FILTER) return fail(BSF_ERR_SHAPE, "bsf: kernel of %d x %d is too large", m, m);
        *rows = *cols = m;
        if (cap < m * m) return fail(BSF_ERR_ARG, "bsf: the kernel needs %d floats (cap %d)", m * m, cap);
        std::vector<double> k((size_t)m * m, 0.0);
        for (int j = 0; j < m; j++)
            for (int i = 0; i < m; i++) {
                double s = 0.0;
                for (int z = 0; z < nz; z++) s += psf[z * plane + (size_t)(lo - 1 + j) * n + (lo - 1 + i)];
                k[(size_t)j * m + i] = s / nz;
            }
        double mean = 0.0;
        for (double v : k) mean += v;
        mean /= (double)k.size();
        double mx = -HUGE_VAL;
        for (double &v : k) {
            v -= mean;
            mx = v > mx ? v : mx;
        }
        for (size_t q = 0; q < k.size(); q++) h[q] = (float)(0.05 * k[q] / mx);
        return BSF_OK;
    }
    if (kind == BSF_PEAK_PSF_MEAN) {
        // mean over z of 5 * PSF(:, :, z) minus its column means: by linearity
        // the z-mean of the filtered frames is this one kernel's filter
        if (n > BSF_MAX_PEAK_FILTER) return fail(BSF_ERR_SHAPE, "bsf: PSF of %d x %d is too large", n, n);
        *rows = *cols = n;
        if (cap < n * n) return fail(BSF_ERR_ARG, "bsf: the kernel needs %d floats (cap %d)", n * n, cap);
        std::vector<double> k(plane, 0.0);
        for (int z = 0; z < nz; z++)
            for (int j = 0; j < n; j++) {
                double cm = 0.0;
                for (int i = 0; i < n; i++) cm += 5.0 * psf[z * plane + (size_t)j * n + i];
                cm /= n;
                for (int i = 0; i < n; i++) k[(size_t)j * n + i] += 5.0 * psf[z * plane + (size_t)j * n + i] - cm;
            }
        for (size_t q = 0; q < plane; q++) h[q] = (float)(k[q] / nz);
        return BSF_OK;
    }
    return fail(BSF_ERR_ARG, "bsf: unknown PSF filter kind %d", kind);
}

size_t bsf_find_rois_work_bytes(const bsf_roi_request *r)
{
    if (!r || r->nrows < 1 || r->ncols < 1 || r->nframes < 0) return 0;
    return work_bytes_of(r);
}

int bsf_find_rois_device(const bsf_roi_request *r, bsf_roi_result *res, int *count, void *work, size_t work_bytes,
                         void *stream)
{
    bsf::clear_error();
    int rc = validate_rois(r, res);
    if (rc) return rc;
    if (!count) return fail(BSF_ERR_ARG, "bsf: NULL count");
    if (r->nframes > 0 && (!work || work_bytes < work_bytes_of(r)))
        return fail(BSF_ERR_ARG, "bsf: work buffer of %zu B is smaller than bsf_find_rois_work_bytes (%zu B)",
                    work_bytes, work_bytes_of(r));
    if ((rc = check_gfx950())) return rc;
    bsf::FrontLaunch L = make_front(r, r->frames, r->filter, res->imstack, res->peaks, res->value, count, work);
    hipError_t e = bsf::launch_find_rois(L, (hipStream_t)stream);
    if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: launch failed: %s", hipGetErrorString(e));
    return BSF_OK;
}

int bsf_find_rois_device_noalloc(const bsf_roi_request *r, bsf_roi_result *res, int *count, void *stream)
{
    bsf::clear_error();
    int rc = validate_rois(r, res);
    if (rc) return rc;
    if (!count) return fail(BSF_ERR_ARG, "bsf: NULL count");
    const size_t need = work_bytes_of(r);
    if (need > bsf::kFrontScratchBytes)
        return fail(BSF_ERR_NOMEM,
                    "bsf: %d frames of %d x %d need %zu B of workspace; the allocation-free entry point has %zu B "
                    "(use bsf_find_rois_device with a work buffer)",
                    r->nframes, r->nrows, r->ncols, need, bsf::kFrontScratchBytes);
    if ((rc = check_gfx950())) return rc;
    void *work = nullptr;
    hipError_t e = bsf::front_scratch(&work);
    if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: workspace lookup failed: %s", hipGetErrorString(e));
    bsf::FrontLaunch L = make_front(r, r->frames, r->filter, res->imstack, res->peaks, res->value, count, work);
    e = bsf::launch_find_rois(L, (hipStream_t)stream);
    if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: launch failed: %s", hipGetErrorString(e));
    return BSF_OK;
}

int bsf_find_rois_host(const bsf_roi_request *r, bsf_roi_result *res)
{
    bsf::clear_error();
    int rc = validate_rois(r, res);
    if (rc) return rc;
    if ((rc = check_gfx950())) return rc;
    const size_t px = r->pixel_type == BSF_PIXEL_U16 ? 2 : 4;
    const size_t fbytes = px * (size_t)r->nrows * r->ncols * r->nframes;
    const size_t hbytes = sizeof(float) * (size_t)r->filter_rows * r->filter_cols;
    const size_t S2 = (size_t)r->roi * r->roi, cap = (size_t)r->capacity;
    const size_t wbytes = work_bytes_of(r);
    void *d_frames = nullptr, *d_work = nullptr;
    float *d_h = nullptr, *d_stack = nullptr, *d_value = nullptr;
    int *d_peaks = nullptr, *d_count = nullptr;
    hipError_t e = hipSuccess;
    auto alloc = [&](void **p, size_t b) {
        if (e == hipSuccess) e = hipMalloc(p, b ? b : 4);
    };
    alloc(&d_frames, fbytes);
    alloc((void **)&d_h, hbytes);
    alloc(&d_work, wbytes);
    alloc((void **)&d_stack, sizeof(float) * S2 * cap);
    alloc((void **)&d_peaks, sizeof(int) * 3 * cap);
    if (res->value) alloc((void **)&d_value, sizeof(float) * cap);
    alloc((void **)&d_count, sizeof(int));
    int count = 0;
    if (e != hipSuccess) {
        rc = fail(BSF_ERR_NOMEM, "bsf: device allocation failed: %s", hipGetErrorString(e));
    } else {
        if (fbytes) e = hipMemcpy(d_frames, r->frames, fbytes, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_h, r->filter, hbytes, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            bsf::FrontLaunch L = make_front(r, d_frames, d_h, d_stack, d_peaks, d_value, d_count, d_work);
            e = bsf::launch_find_rois(L, nullptr);
        }
        if (e == hipSuccess) e = hipMemcpy(&count, d_count, sizeof(int), hipMemcpyDeviceToHost);
        const size_t n = (size_t)count < cap ? (size_t)count : cap;
        if (e == hipSuccess && n) e = hipMemcpy(res->imstack, d_stack, sizeof(float) * S2 * n, hipMemcpyDeviceToHost);
        for (int col = 0; col < 3 && e == hipSuccess && n; col++)
            e = hipMemcpy(res->peaks + col * cap, d_peaks + col * cap, sizeof(int) * n, hipMemcpyDeviceToHost);
        if (e == hipSuccess && n && res->value)
            e = hipMemcpy(res->value, d_value, sizeof(float) * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(BSF_ERR_DEVICE, "bsf: HIP error: %s", hipGetErrorString(e));
    }
    for (void *p : {d_frames, (void *)d_h, d_work, (void *)d_stack, (void *)d_peaks, (void *)d_value, (void *)d_count})
        (void)hipFree(p);
    res->count = rc ? 0 : count;
    return rc;
}

int bsf_fit_frames_device(const bsf_roi_request *rq, const bsf_request *fq, bsf_roi_result *rois, int *count,
                          bsf_result *res, void *work, size_t work_bytes, void *stream)
{
    bsf::clear_error();
    int rc = validate_rois(rq, rois);
    if (rc) return rc;
    if (!fq || !count) return fail(BSF_ERR_ARG, "bsf: NULL fit request or count");
    if (rq->nframes > 0 && (!work || work_bytes < work_bytes_of(rq)))
        return fail(BSF_ERR_ARG, "bsf: work buffer of %zu B is smaller than bsf_find_rois_work_bytes (%zu B)",
                    work_bytes, work_bytes_of(rq));
    // the fit reads the ROI stack in place: sz = roi, nfits = capacity (the
    // outputs' stride); only the first *count ROIs are fitted
    bsf_request q = *fq;
    q.imstack = rois->imstack;
    q.sz = rq->roi;
    q.nfits = rq->capacity;
    bsf::FitLaunch L;
    if ((rc = bsf::prepare_fit_launch(&q, res, &L))) return rc;
    L.ncount = count;
    const hipStream_t st = (hipStream_t)stream;
    bsf::FrontLaunch F = make_front(rq, rq->frames, rq->filter, rois->imstack, rois->peaks, rois->value, count, work);
    hipError_t e = bsf::launch_find_rois(F, st);
    if (e == hipSuccess && rq->capacity > 0) e = bsf::launch_fit_all(L, st);
    if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: launch failed: %s", hipGetErrorString(e));
    return BSF_OK;
}

int bsf_fit_frames_host(const bsf_roi_request *rq, const bsf_request *fq, bsf_roi_result *rois, bsf_result *res)
{
    bsf::clear_error();
    int rc = validate_rois(rq, rois);
    if (rc) return rc;
    if (!fq || !res || !fq->coeff) return fail(BSF_ERR_ARG, "bsf: NULL fit request, result or coeff");
    const int NV = bsf_num_params(fq->fittype);
    if (NV < 0) return fail(BSF_ERR_FITTYPE, "bsf: fittype %d is not a B-spline fittype (6, 7, 8)", fq->fittype);
    if (fq->coeff_ndim < 1 || fq->coeff_ndim > 5) return fail(BSF_ERR_SHAPE, "bsf: bad coeff_ndim %d", fq->coeff_ndim);
    if (rq->capacity > 0 && (!res->P || !res->CRLB || !res->LL)) return fail(BSF_ERR_ARG, "bsf: NULL fit output");
    if ((rc = check_gfx950())) return rc;
    const size_t px = rq->pixel_type == BSF_PIXEL_U16 ? 2 : 4;
    const size_t fbytes = px * (size_t)rq->nrows * rq->ncols * rq->nframes;
    const size_t hbytes = sizeof(float) * (size_t)rq->filter_rows * rq->filter_cols;
    const size_t S2 = (size_t)rq->roi * rq->roi, cap = (size_t)rq->capacity;
    const size_t wbytes = work_bytes_of(rq);
    size_t cn = 1;
    for (int a = 0; a < fq->coeff_ndim; a++) cn *= (size_t)(fq->coeff_dims[a] > 0 ? fq->coeff_dims[a] : 0);
    const size_t c3n = fq->fittype == 8 && fq->coeff_ndim >= 3
                           ? (size_t)fq->coeff_dims[0] * fq->coeff_dims[1] * fq->coeff_dims[2] : 0;
    void *d_frames = nullptr, *d_work = nullptr;
    float *d_h = nullptr, *d_stack = nullptr, *d_value = nullptr, *d_c = nullptr, *d_c3 = nullptr;
    float *d_P = nullptr, *d_CR = nullptr, *d_LL = nullptr;
    int *d_peaks = nullptr, *d_count = nullptr;
    hipError_t e = hipSuccess;
    auto alloc = [&](void **p, size_t b) {
        if (e == hipSuccess) e = hipMalloc(p, b ? b : 4);
    };
    alloc(&d_frames, fbytes);
    alloc((void **)&d_h, hbytes);
    alloc(&d_work, wbytes);
    alloc((void **)&d_stack, sizeof(float) * S2 * cap);
    alloc((void **)&d_peaks, sizeof(int) * 3 * cap);
    if (rois->value) alloc((void **)&d_value, sizeof(float) * cap);
    alloc((void **)&d_count, sizeof(int));
    alloc((void **)&d_c, sizeof(float) * cn);
    if (c3n && fq->coeff3) alloc((void **)&d_c3, sizeof(float) * c3n);
    alloc((void **)&d_P, sizeof(float) * cap * (NV + 1));
    alloc((void **)&d_CR, sizeof(float) * cap * NV);
    alloc((void **)&d_LL, sizeof(float) * cap);
    int count = 0;
    if (e != hipSuccess) {
        rc = fail(BSF_ERR_NOMEM, "bsf: device allocation failed: %s", hipGetErrorString(e));
    } else {
        if (fbytes) e = hipMemcpy(d_frames, rq->frames, fbytes, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_h, rq->filter, hbytes, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_c, fq->coeff, sizeof(float) * cn, hipMemcpyHostToDevice);
        if (e == hipSuccess && d_c3) e = hipMemcpy(d_c3, fq->coeff3, sizeof(float) * c3n, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            rc = fail(BSF_ERR_DEVICE, "bsf: HIP error: %s", hipGetErrorString(e));
        } else {
            bsf_roi_request q = *rq;
            q.frames = d_frames;
            q.filter = d_h;
            bsf_request f = *fq;
            f.coeff = d_c;
            f.coeff3 = d_c3;
            bsf_roi_result dr = {d_stack, d_peaks, d_value, 0};
            bsf_result dres = {d_P, d_CR, d_LL};
            rc = bsf_fit_frames_device(&q, &f, &dr, d_count, &dres, d_work, wbytes, nullptr);
            if (!rc) {
                e = hipMemcpy(&count, d_count, sizeof(int), hipMemcpyDeviceToHost);
                const size_t n = (size_t)count < cap ? (size_t)count : cap;
                if (e == hipSuccess && n) e = hipMemcpy(rois->imstack, d_stack, sizeof(float) * S2 * n,
                                                        hipMemcpyDeviceToHost);
                for (int c = 0; c < 3 && e == hipSuccess && n; c++)
                    e = hipMemcpy(rois->peaks + c * cap, d_peaks + c * cap, sizeof(int) * n, hipMemcpyDeviceToHost);
                if (e == hipSuccess && n && rois->value)
                    e = hipMemcpy(rois->value, d_value, sizeof(float) * n, hipMemcpyDeviceToHost);
                for (int c = 0; c <= NV && e == hipSuccess && n; c++)
                    e = hipMemcpy(res->P + c * cap, d_P + c * cap, sizeof(float) * n, hipMemcpyDeviceToHost);
                for (int c = 0; c < NV && e == hipSuccess && n; c++)
                    e = hipMemcpy(res->CRLB + c * cap, d_CR + c * cap, sizeof(float) * n, hipMemcpyDeviceToHost);
                if (e == hipSuccess && n) e = hipMemcpy(res->LL, d_LL, sizeof(float) * n, hipMemcpyDeviceToHost);
                if (e != hipSuccess) rc = fail(BSF_ERR_DEVICE, "bsf: HIP error: %s", hipGetErrorString(e));
            }
        }
    }
    for (void *p : {d_frames, (void *)d_h, d_work, (void *)d_stack, (void *)d_peaks, (void *)d_value, (void *)d_count,
                    (void *)d_c, (void *)d_c3, (void *)d_P, (void *)d_CR, (void *)d_LL})
        (void)hipFree(p);
    rois->count = rc ? 0 : count;
    return rc;
}

}  // extern "C"
