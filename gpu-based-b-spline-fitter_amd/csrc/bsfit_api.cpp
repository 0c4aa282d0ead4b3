// bsfit_api.cpp -- the C-ABI (include/bsfit.h): argument checks with the
// reference MEX's semantics, device buffers, launches, error reporting.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <cmath>
#include <cstdlib>
#include <cstddef>

#include "../../include/bsfit.h"
#include "bsf_kernels.h"

namespace bsf {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

void clear_error() { g_err.clear(); }

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_dev_mu;
DeviceInfo g_dev[kMaxDevices];
bool g_dev_known[kMaxDevices];
}  // namespace

int device_info(DeviceInfo *out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: no HIP device (%s)", hipGetErrorString(e));
    const bool cacheable = dev >= 0 && dev < kMaxDevices;
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (cacheable && g_dev_known[dev]) {
        *out = g_dev[dev];
        return BSF_OK;
    }
    hipDeviceProp_t p;
    e = hipGetDeviceProperties(&p, dev);
    if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: hipGetDeviceProperties: %s", hipGetErrorString(e));
    DeviceInfo d;
    d.dev = dev;
    d.gfx950 = std::strncmp(p.gcnArchName, "gfx950", 6) == 0;
    d.lds_limit = p.sharedMemPerBlock;
    std::snprintf(d.arch, sizeof d.arch, "%s", p.gcnArchName);
    if (cacheable) {
        g_dev[dev] = d;
        g_dev_known[dev] = true;
    }
    *out = d;
    return BSF_OK;
}

}  // namespace bsf

namespace {

using bsf::fail;
using bsf::g_err;

int nv_of(int ft) { return ft == 6 ? 5 : ft == 7 ? 6 : ft == 8 ? 8 : -1; }
int dim_of(int ft) { return ft == 6 ? 3 : ft == 7 ? 4 : ft == 8 ? 5 : -1; }

size_t coeff_numel(const bsf_request *r, int nd)
{
    size_t n = 1;
    for (int a = 0; a < nd; a++) n *= (size_t)r->coeff_dims[a];
    return n;
}

// Argument checks mirroring GPUmleFit_LM/mexFunction.cpp:247-385 for
// fittypes 6-8, with explicit errors where the MEX would crash.
int validate(const bsf_request *r, const bsf_result *res)
{
    if (!r || !res) return fail(BSF_ERR_ARG, "bsf: NULL request or result");
    const int D = dim_of(r->fittype);
    if (D < 0) return fail(BSF_ERR_FITTYPE, "bsf: fittype %d is not a B-spline fittype (6, 7, 8)", r->fittype);
    if (r->nfits < 0) return fail(BSF_ERR_ARG, "bsf: nfits must be >= 0");
    if (r->sz < 1 || r->sz > BSF_MAX_ROI)
        return fail(BSF_ERR_ROI_SIZE, "bsf: X,Y dimension of data must be between 1 and %d (got %d)", BSF_MAX_ROI,
                    r->sz);
    if (r->iterations < 0) return fail(BSF_ERR_ARG, "bsf: iterations must be >= 0");
    if (r->nfits > 0 && (!r->imstack || !res->P || !res->CRLB || !res->LL))
        return fail(BSF_ERR_ARG, "bsf: NULL data or output pointer");
    if (!r->coeff) return fail(BSF_ERR_ARG, "bsf: coeff must not be NULL");
    if (r->coeff_ndim != D)
        return fail(BSF_ERR_SHAPE, "bsf: fittype %d needs a %d-D coefficient array (got %d-D)", r->fittype, D,
                    r->coeff_ndim);
    for (int a = 0; a < D; a++)
        if (r->coeff_dims[a] < 5)
            return fail(BSF_ERR_SHAPE, "bsf: coefficient dim %d is %d; each axis needs >= 2 data samples (+3)", a,
                        r->coeff_dims[a]);
    // the kernels address the arrays with 32-bit byte offsets (buffer records)
    if (coeff_numel(r, D) * 4 >= ((size_t)1 << 31))
        return fail(BSF_ERR_SHAPE, "bsf: the coefficient array has %zu elements; at most 2^29 - 1 are supported",
                    coeff_numel(r, D));
    if (r->fittype == 8 && !r->coeff3)
        return fail(BSF_ERR_ARG, "bsf: fittype 8 needs coeff3 (the 3D coefficient array)");
    if (r->initP) {
        int need = r->fittype == 7 ? 2 : r->fittype == 8 ? 4 : 0;
        if (r->n_initP < need)
            return fail(BSF_ERR_ARG, "bsf: initP for fittype %d needs %d values (got %d)", r->fittype, need,
                        r->n_initP);
    }
    if (r->precision != BSF_PREC_MIX && r->precision != BSF_PREC_F64)
        return fail(BSF_ERR_ARG, "bsf: unknown precision mode %d", r->precision);
    return BSF_OK;
}

// The device is a gfx950 and one workgroup's LDS fits.
int check_device(int S, int ft, int f64)
{
    bsf::DeviceInfo d;
    int rc = bsf::device_info(&d);
    if (rc) return rc;
    if (!d.gfx950)
        return fail(BSF_ERR_DEVICE, "bsf: device %d is %s; this build targets gfx950 (MI355X)", d.dev, d.arch);
    size_t lds = bsf::lds_bytes(ft, f64, S);
    if (lds > d.lds_limit)
        return fail(BSF_ERR_ROI_SIZE, "bsf: ROI size %d with fittype %d needs %zu B of LDS (device limit %zu B)", S,
                    ft, lds, d.lds_limit);
    return BSF_OK;
}

bsf::FitLaunch make_launch(const bsf_request *r, const float *data, const float *coeff, const float *coeff3,
                           float *P, float *CRLB, float *LL)
{
    bsf::FitLaunch L;
    std::memset(&L, 0, sizeof L);
    L.fittype = r->fittype;
    L.f64 = r->precision == BSF_PREC_F64;
    L.data = data;
    L.coeff = coeff;
    L.coeff3 = coeff3;
    L.P = P;
    L.CRLB = CRLB;
    L.LL = LL;
    L.N = r->nfits;
    L.S = r->sz;
    L.iterations = r->iterations;
    L.ndim = r->coeff_ndim;
    for (int a = 0; a < 5; a++) L.cdims[a] = a < r->coeff_ndim ? r->coeff_dims[a] : 1;
    L.has_initP = r->initP != nullptr;
    if (r->initP) {
        int n = r->fittype == 7 ? 2 : 4;
        for (int i = 0; i < n; i++) L.initP[i] = r->initP[i];
    }
    return L;
}

// Launch in grid chunks (one workgroup per ROI).
hipError_t launch_all(bsf::FitLaunch L, hipStream_t st)
{
    const int chunk = 1 << 24;
    for (int off = 0; off < L.N; off += chunk) {
        L.roi_offset = off;
        L.count = (L.N - off) < chunk ? (L.N - off) : chunk;
        hipError_t e = bsf::launch_fit(L, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

namespace bsf {

int prepare_fit_launch(const bsf_request *q, const bsf_result *res, FitLaunch *L)
{
    int rc = validate(q, res);
    if (rc) return rc;
    if ((rc = check_device(q->sz, q->fittype, q->precision == BSF_PREC_F64))) return rc;
    *L = make_launch(q, q->imstack, q->coeff, q->coeff3, res->P, res->CRLB, res->LL);
    return BSF_OK;
}

hipError_t launch_fit_all(const FitLaunch &L, hipStream_t st) { return launch_all(L, st); }

}  // namespace bsf

extern "C" {

int bsf_num_params(int fittype) { return nv_of(fittype); }

const char *bsf_last_error(void) { return g_err.c_str(); }

static int fit_device(const bsf_request *r, bsf_result *res, unsigned char *status, void *stream,
                      const bsf_roi *rois = nullptr, int *evals = nullptr, int tile = 0)
{
    int rc = validate(r, res);
    if (rc) return rc;
    if (tile < 0 || tile > 2) return fail(BSF_ERR_ARG, "bsf: tile %d (0, 1 or 2 ROIs per wavefront)", tile);
    if ((rc = check_device(r->sz, r->fittype, r->precision == BSF_PREC_F64))) return rc;
    if (r->nfits == 0) return BSF_OK;
    bsf::FitLaunch L = make_launch(r, r->imstack, r->coeff, r->coeff3, res->P, res->CRLB, res->LL);
    L.status = status;
    L.evals = evals;
    L.rois = (const bsf::RoiDesc *)rois;
    L.tile = tile;
    hipError_t e = launch_all(L, (hipStream_t)stream);
    if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: kernel launch failed: %s", hipGetErrorString(e));
    return BSF_OK;
}

int bsf_fit_device(const bsf_request *r, bsf_result *res, void *stream)
{
    g_err.clear();
    return fit_device(r, res, nullptr, stream);
}

int bsf_fit_device_status(const bsf_request *r, bsf_result *res, unsigned char *status, void *stream)
{
    g_err.clear();
    if (!status) return fail(BSF_ERR_ARG, "bsf: NULL status array");
    return fit_device(r, res, status, stream);
}

static_assert(sizeof(bsf_roi) == sizeof(bsf::RoiDesc) && offsetof(bsf_roi, off) == offsetof(bsf::RoiDesc, off) &&
                  offsetof(bsf_roi, cstride) == offsetof(bsf::RoiDesc, cstride) &&
                  offsetof(bsf_roi, sz) == offsetof(bsf::RoiDesc, S),
              "bsf_roi and the kernels' ROI descriptor share one layout");

int bsf_fit_device_ex(const bsf_request *r, bsf_result *res, const bsf_fit_extras *x, void *stream)
{
    g_err.clear();
    if (!x) return fit_device(r, res, nullptr, stream);
    return fit_device(r, res, x->status, stream, x->rois, x->evals, x->tile);
}

int bsf_tune(const bsf_request *r, bsf_result *res, int reps, int *tile, float *ms, void *stream)
{
    g_err.clear();
    int rc = validate(r, res);
    if (rc) return rc;
    if (!tile) return fail(BSF_ERR_ARG, "bsf: NULL tile");
    if (reps < 1 || reps > 1000) return fail(BSF_ERR_ARG, "bsf: reps %d (1..1000)", reps);
    const int f64 = r->precision == BSF_PREC_F64;
    if ((rc = check_device(r->sz, r->fittype, f64))) return rc;
    if (ms) ms[0] = ms[1] = 0.f;
    *tile = 1;
    if (r->nfits == 0) return BSF_OK;
    const hipStream_t st = (hipStream_t)stream;
    bsf::FitLaunch L = make_launch(r, r->imstack, r->coeff, r->coeff3, res->P, res->CRLB, res->LL);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float best = 0.f;
    int last = 1;
    for (int t = 1; t <= 2 && e == hipSuccess; t++) {
        if (!bsf::tile_possible(r->fittype, f64, r->sz, t)) continue;
        L.tile = last = t;
        e = launch_all(L, st);   // warm-up: code object load, caches
        if (e == hipSuccess) e = hipEventRecord(e0, st);
        for (int k = 0; k < reps && e == hipSuccess; k++) e = launch_all(L, st);
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t_ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t_ms, e0, e1);
        if (e != hipSuccess) break;
        t_ms /= reps;
        if (ms) ms[t - 1] = t_ms;
        if (t == 1 || t_ms < best) best = t_ms, *tile = t;
    }
    // res holds the fits of the chosen tile (fittype 8's two tiles sum the
    // pixels in different orders)
    if (e == hipSuccess && *tile != last) {
        L.tile = *tile;
        e = launch_all(L, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: tuning launch failed: %s", hipGetErrorString(e));
    bsf::set_fit_tile(r->fittype, f64, r->sz, *tile);
    return BSF_OK;
}

int bsf_fit_tile(int fittype, int precision, int sz) { return bsf::fit_tile(fittype, precision == BSF_PREC_F64, sz); }

int bsf_set_fit_tile(int fittype, int precision, int sz, int tile)
{
    g_err.clear();
    if (precision != BSF_PREC_MIX && precision != BSF_PREC_F64)
        return fail(BSF_ERR_ARG, "bsf: unknown precision mode %d", precision);
    if (bsf::set_fit_tile(fittype, precision == BSF_PREC_F64, sz, tile))
        return fail(BSF_ERR_ARG, "bsf: tile %d is not possible for fittype %d, ROI size %d", tile, fittype, sz);
    return BSF_OK;
}

// Host-array fit from nstarts starts (rows of `starts`, n_initP values each;
// NULL starts: the request's initP, one start): one upload, one launch per
// start (later starts keep a result only where it beats the kept
// log-likelihood), one download.
static int fit_host_starts(const bsf_request *r, const float *starts, int nstarts, int strict, bsf_result *res,
                           unsigned char *status = nullptr)
{
    int rc = validate(r, res);
    if (rc) return rc;
    if ((rc = check_device(r->sz, r->fittype, r->precision == BSF_PREC_F64))) return rc;
    if (r->nfits == 0) return BSF_OK;
    const int D = dim_of(r->fittype), NV = nv_of(r->fittype);
    const size_t N = (size_t)r->nfits, S2 = (size_t)r->sz * r->sz;
    const size_t nc = coeff_numel(r, D);
    const size_t nc3 = r->fittype == 8 ? (size_t)r->coeff_dims[0] * r->coeff_dims[1] * r->coeff_dims[2] : 0;
    const size_t bytes_data = N * S2 * 4, bytes_c = nc * 4, bytes_c3 = nc3 * 4;
    const size_t bytes_out = N * (size_t)(NV + 1 + NV + 1) * 4;
    char *buf = nullptr;
    const size_t a = 256;
    auto up = [&](size_t x) { return (x + a - 1) / a * a; };
    const size_t bytes_st = status ? N : 0;
    const size_t total = up(bytes_data) + up(bytes_c) + up(bytes_c3) + up(bytes_out) + up(bytes_st);
    hipError_t e = hipMalloc((void **)&buf, total);
    if (e != hipSuccess) return fail(BSF_ERR_NOMEM, "bsf: could not allocate %zu bytes on the device", total);
    float *d_data = (float *)buf;
    float *d_c = (float *)(buf + up(bytes_data));
    float *d_c3 = (float *)(buf + up(bytes_data) + up(bytes_c));
    float *d_out = (float *)(buf + up(bytes_data) + up(bytes_c) + up(bytes_c3));
    float *d_P = d_out, *d_CR = d_out + N * (NV + 1), *d_LL = d_CR + N * NV;
    unsigned char *d_st = status ? (unsigned char *)(buf + up(bytes_data) + up(bytes_c) + up(bytes_c3) + up(bytes_out))
                                 : nullptr;
    hipStream_t st = hipStreamPerThread;
    rc = BSF_OK;
    e = hipMemcpyAsync(d_data, r->imstack, bytes_data, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_c, r->coeff, bytes_c, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nc3) e = hipMemcpyAsync(d_c3, r->coeff3, bytes_c3, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        bsf::FitLaunch L = make_launch(r, d_data, d_c, nc3 ? d_c3 : nullptr, d_P, d_CR, d_LL);
        L.status = d_st;
        for (int k = 0; k < nstarts && e == hipSuccess; k++) {
            if (starts) {
                for (int q = 0; q < r->n_initP && q < 4; q++) L.initP[q] = starts[(size_t)k * r->n_initP + q];
                L.has_initP = 1;
            }
            L.keep = k == 0 ? 0 : (strict ? 2 : 1);
            e = launch_all(L, st);
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(res->P, d_P, N * (NV + 1) * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(res->CRLB, d_CR, N * NV * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(res->LL, d_LL, N * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && status) e = hipMemcpyAsync(status, d_st, N, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = fail(BSF_ERR_DEVICE, "bsf: %s", hipGetErrorString(e));
    (void)hipFree(buf);
    return rc;
}

int bsf_fit_host(const bsf_request *r, bsf_result *res)
{
    g_err.clear();
    return fit_host_starts(r, nullptr, 1, 0, res);
}

int bsf_fit_host_status(const bsf_request *r, bsf_result *res, unsigned char *status)
{
    g_err.clear();
    if (!status) return fail(BSF_ERR_ARG, "bsf: NULL status array");
    return fit_host_starts(r, nullptr, 1, 0, res, status);
}

// The starts of a multi-start request: fittype 7 or 8, 1..64 starts of
// n_initP (>= 2 / 4) values.
static int validate_starts(const bsf_request *r, const float *starts, int nstarts)
{
    if (!r) return fail(BSF_ERR_ARG, "bsf: NULL request");
    if (r->fittype != 7 && r->fittype != 8)
        return fail(BSF_ERR_FITTYPE, "bsf: multi-start fits take fittype 7 or 8 (got %d)", r->fittype);
    if (!starts || nstarts < 1 || nstarts > 64)
        return fail(BSF_ERR_ARG, "bsf: multi-start fits need 1..64 starts (got %d)", nstarts);
    const int need = r->fittype == 7 ? 2 : 4;
    if (r->n_initP < need)
        return fail(BSF_ERR_ARG, "bsf: each start of fittype %d needs %d values (n_initP = %d)", r->fittype, need,
                    r->n_initP);
    return BSF_OK;
}

static int multistart_host(const bsf_request *r, const float *starts, int nstarts, int strict, bsf_result *res,
                           unsigned char *status)
{
    int rc = validate_starts(r, starts, nstarts);
    if (rc) return rc;
    bsf_request q = *r;
    q.initP = starts;   // row 0 passes the single-start checks
    return fit_host_starts(&q, starts, nstarts, strict, res, status);
}

int bsf_fit_multistart_host(const bsf_request *r, const float *starts, int nstarts, int strict, bsf_result *res)
{
    g_err.clear();
    return multistart_host(r, starts, nstarts, strict, res, nullptr);
}

int bsf_fit_multistart_host_status(const bsf_request *r, const float *starts, int nstarts, int strict,
                                   bsf_result *res, unsigned char *status)
{
    g_err.clear();
    if (!status) return fail(BSF_ERR_ARG, "bsf: NULL status array");
    return multistart_host(r, starts, nstarts, strict, res, status);
}

// The launches of a device multi-start: the kernel writes a ROI's outputs
// and status flags only where a later start's result is kept.
static int multistart_device(const bsf_request *r, const float *starts, int nstarts, int strict, bsf_result *res,
                             unsigned char *status, void *stream)
{
    int rc = validate_starts(r, starts, nstarts);
    if (rc) return rc;
    bsf_request q = *r;
    q.initP = starts;
    if ((rc = validate(&q, res))) return rc;
    if ((rc = check_device(q.sz, q.fittype, q.precision == BSF_PREC_F64))) return rc;
    if (q.nfits == 0) return BSF_OK;
    bsf::FitLaunch L = make_launch(&q, q.imstack, q.coeff, q.coeff3, res->P, res->CRLB, res->LL);
    L.status = status;
    hipError_t e = hipSuccess;
    for (int k = 0; k < nstarts && e == hipSuccess; k++) {
        for (int i = 0; i < q.n_initP && i < 4; i++) L.initP[i] = starts[(size_t)k * q.n_initP + i];
        L.has_initP = 1;
        L.keep = k == 0 ? 0 : (strict ? 2 : 1);
        e = launch_all(L, (hipStream_t)stream);
    }
    if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: kernel launch failed: %s", hipGetErrorString(e));
    return BSF_OK;
}

int bsf_fit_multistart_device(const bsf_request *r, const float *starts, int nstarts, int strict, bsf_result *res,
                              void *stream)
{
    g_err.clear();
    return multistart_device(r, starts, nstarts, strict, res, nullptr, stream);
}

int bsf_fit_multistart_device_status(const bsf_request *r, const float *starts, int nstarts, int strict,
                                     bsf_result *res, unsigned char *status, void *stream)
{
    g_err.clear();
    if (!status) return fail(BSF_ERR_ARG, "bsf: NULL status array");
    return multistart_device(r, starts, nstarts, strict, res, status, stream);
}

static int mex_fit(const float *imstack, const size_t *imdims, int im_ndim, int fittype, int iterations,
                   const float *coeff, const size_t *coeff_dims, int coeff_ndim, int silent, const float *initP,
                   size_t initP_numel, const float *coeff3, float *P, float *CRLB, float *LL, unsigned char *status)
{
    if (!imdims || im_ndim < 2 || im_ndim > 3) return fail(BSF_ERR_ARG, "bsf: imstack must be 2-D or 3-D");
    if (imdims[0] > BSF_MAX_ROI)
        return fail(BSF_ERR_ROI_SIZE, "bsf: X,Y dimension of data must be smaller than %d", BSF_MAX_ROI);
    if (imdims[1] != imdims[0]) return fail(BSF_ERR_SHAPE, "bsf: Fit Box must be square");
    if (!coeff_dims || coeff_ndim < 1 || coeff_ndim > 5) return fail(BSF_ERR_SHAPE, "bsf: bad coefficient dims");
    bsf_request r;
    std::memset(&r, 0, sizeof r);
    r.fittype = fittype;
    r.imstack = imstack;
    r.sz = (int)imdims[0];
    r.nfits = im_ndim == 2 ? 1 : (int)imdims[2];
    r.iterations = iterations;
    r.coeff = coeff;
    r.coeff_ndim = coeff_ndim;
    for (int a = 0; a < coeff_ndim; a++) r.coeff_dims[a] = (int)coeff_dims[a];
    r.coeff3 = coeff3;
    r.initP = (fittype > 6) ? initP : nullptr;
    r.n_initP = (int)initP_numel;
    r.precision = BSF_PREC_MIX;
    bsf_result res = {P, CRLB, LL};
    auto t0 = std::chrono::steady_clock::now();
    int rc = fit_host_starts(&r, nullptr, 1, 0, &res, status);
    double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc == BSF_OK && !silent)
        std::printf("subregion size: %d fittype: %d\nNumber of fits: %d\nfit time %f s, fits per second %f\n", r.sz,
                    fittype, r.nfits, dt, dt > 0 ? r.nfits / dt : 0.0);
    return rc;
}

int bsf_GPUmleFit_LM_multiD(const float *imstack, const size_t *imdims, int im_ndim, int fittype, int iterations,
                            const float *coeff, const size_t *coeff_dims, int coeff_ndim, const float *varmap,
                            size_t varmap_numel, int silent, const float *initP, size_t initP_numel,
                            const float *coeff3, float *P, float *CRLB, float *LL)
{
    (void)varmap;
    (void)varmap_numel;
    g_err.clear();
    return mex_fit(imstack, imdims, im_ndim, fittype, iterations, coeff, coeff_dims, coeff_ndim, silent, initP,
                   initP_numel, coeff3, P, CRLB, LL, nullptr);
}

int bsf_GPUmleFit_LM_multiD_status(const float *imstack, const size_t *imdims, int im_ndim, int fittype,
                                   int iterations, const float *coeff, const size_t *coeff_dims, int coeff_ndim,
                                   const float *varmap, size_t varmap_numel, int silent, const float *initP,
                                   size_t initP_numel, const float *coeff3, float *P, float *CRLB, float *LL,
                                   unsigned char *status)
{
    (void)varmap;
    (void)varmap_numel;
    g_err.clear();
    if (!status) return fail(BSF_ERR_ARG, "bsf: NULL status array");
    return mex_fit(imstack, imdims, im_ndim, fittype, iterations, coeff, coeff_dims, coeff_ndim, silent, initP,
                   initP_numel, coeff3, P, CRLB, LL, status);
}

// ------------------------------------------------------------------ 4Pi
static int nv_4pi(const unsigned int *shared, int nch)
{
    if (!shared || nch < 1 || nch > 8) return -1;
    int n = 0;
    for (int p = 0; p < 6; p++) n += shared[p] ? 1 : nch;
    return n;
}

int bsf_num_params_4pi(const unsigned int *shared, int nch) { return nv_4pi(shared, nch); }

static int validate_4pi(const bsf_request_4pi *r, const bsf_result *res)
{
    if (!r || !res) return fail(BSF_ERR_ARG, "bsf: NULL request or result");
    const int nv = nv_4pi(r->shared, r->nch);
    if (nv < 0) return fail(BSF_ERR_ARG, "bsf: 4Pi needs a shared[6] mask and 1..8 channels (got %d)", r->nch);
    if (nv > 24) return fail(BSF_ERR_ARG, "bsf: 4Pi parameter count %d exceeds 24", nv);
    if (r->nfits < 0 || r->iterations < 0) return fail(BSF_ERR_ARG, "bsf: nfits and iterations must be >= 0");
    if (r->sz < 1 || r->sz > BSF_MAX_ROI)
        return fail(BSF_ERR_ROI_SIZE, "bsf: X,Y dimension of data must be between 1 and %d (got %d)", BSF_MAX_ROI,
                    r->sz);
    if (r->nfits > 0 && (!r->imstack || !res->P || !res->CRLB || !res->LL))
        return fail(BSF_ERR_ARG, "bsf: NULL data or output pointer");
    if (!r->coeff || !r->phaseShift) return fail(BSF_ERR_ARG, "bsf: coeff and phaseShift must not be NULL");
    for (int a = 0; a < 4; a++)
        if (r->coeff_dims[a] < 5) return fail(BSF_ERR_SHAPE, "bsf: coefficient dim %d is %d (needs >= 5)", a, r->coeff_dims[a]);
    const size_t nc = (size_t)r->coeff_dims[0] * r->coeff_dims[1] * r->coeff_dims[2] * r->coeff_dims[3];
    if (nc * 4 >= ((size_t)1 << 31))
        return fail(BSF_ERR_SHAPE, "bsf: the coefficient array has %zu elements; at most 2^29 - 1 are supported", nc);
    if (r->precision != BSF_PREC_MIX && r->precision != BSF_PREC_F64)
        return fail(BSF_ERR_ARG, "bsf: unknown precision mode %d", r->precision);
    return BSF_OK;
}

static int check_device_4pi(const bsf_request_4pi *r)
{
    int rc = check_device(1, 6, 0);
    if (rc) return rc;
    bsf::DeviceInfo d;
    if ((rc = bsf::device_info(&d))) return rc;
    const int f64 = r->precision == BSF_PREC_F64, nv = nv_4pi(r->shared, r->nch);
    size_t lds = bsf::lds_bytes_4pi(f64, r->sz, r->nch, nv);
    if (lds > d.lds_limit)
        return fail(BSF_ERR_ROI_SIZE, "bsf: 4Pi ROI size %d x %d channels needs %zu B of LDS (limit %zu B)", r->sz,
                    r->nch, lds, d.lds_limit);
    return BSF_OK;
}

// The wave-per-channel 4Pi kernel finishes a ROI sooner but moves fewer
// ROIs per second once the GPU is full: launches of fewer ROIs than this use
// it (measured crossover between 400 and 1,600 ROIs of 4 x 13 x 13).
constexpr int kWpcMaxRois = 1024;

static bsf::Fit4PiLaunch make_launch_4pi(const bsf_request_4pi *r, const float *data, const float *coeff,
                                         const float *dT, const float *iz, const float *ip, float *P, float *CR,
                                         float *LL)
{
    bsf::Fit4PiLaunch L;
    std::memset(&L, 0, sizeof L);
    L.f64 = r->precision == BSF_PREC_F64;
    L.data = data;
    L.coeff = coeff;
    L.P = P;
    L.CRLB = CR;
    L.LL = LL;
    L.N = r->nfits;
    L.S = r->sz;
    L.iterations = r->iterations;
    L.nch = r->nch;
    for (int a = 0; a < 4; a++) L.cdims[a] = r->coeff_dims[a];
    for (int p = 0; p < 6; p++) L.shared[p] = r->shared[p] ? 1 : 0;
    for (int k = 0; k < r->nch; k++) L.phaseShift[k] = r->phaseShift[k];
    L.dT = dT;
    L.initZ = iz;
    L.initPhi = ip;
    bsf::DeviceInfo d;
    const size_t limit = bsf::device_info(&d) == BSF_OK ? d.lds_limit : 0;
    L.wpc = bsf::use_wpc_4pi(r->nfits < kWpcMaxRois, L.f64, L.S, L.nch, nv_4pi(r->shared, r->nch), limit);
    return L;
}

static hipError_t launch_all_4pi(bsf::Fit4PiLaunch L, hipStream_t st)
{
    const int chunk = 1 << 24;
    for (int off = 0; off < L.N; off += chunk) {
        L.roi_offset = off;
        L.count = (L.N - off) < chunk ? (L.N - off) : chunk;
        hipError_t e = bsf::launch_fit_4pi(L, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Test hook (not in the header): force the 4Pi kernel (-1 automatic, 0 one
// wave per ROI, 1 one wave per channel) for every later launch.
int bsf_test_set_4pi_kernel(int mode)
{
    bsf::g_4pi_kernel = mode < 0 ? -1 : mode ? 1 : 0;
    return BSF_OK;
}

static int fit_4pi_device(const bsf_request_4pi *r, bsf_result *res, unsigned char *status, void *stream,
                          const bsf_roi *rois = nullptr, int *evals = nullptr)
{
    int rc = validate_4pi(r, res);
    if (rc) return rc;
    if ((rc = check_device_4pi(r))) return rc;
    if (r->nfits == 0) return BSF_OK;
    bsf::Fit4PiLaunch L = make_launch_4pi(r, r->imstack, r->coeff, r->dT, r->initZ, r->initPhi, res->P, res->CRLB,
                                          res->LL);
    L.status = status;
    L.evals = evals;
    L.rois = (const bsf::RoiDesc *)rois;
    hipError_t e = launch_all_4pi(L, (hipStream_t)stream);
    if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: kernel launch failed: %s", hipGetErrorString(e));
    return BSF_OK;
}

int bsf_fit_4pi_device(const bsf_request_4pi *r, bsf_result *res, void *stream)
{
    g_err.clear();
    return fit_4pi_device(r, res, nullptr, stream);
}

int bsf_fit_4pi_device_status(const bsf_request_4pi *r, bsf_result *res, unsigned char *status, void *stream)
{
    g_err.clear();
    if (!status) return fail(BSF_ERR_ARG, "bsf: NULL status array");
    return fit_4pi_device(r, res, status, stream);
}

int bsf_fit_4pi_device_ex(const bsf_request_4pi *r, bsf_result *res, const bsf_fit_extras *x, void *stream)
{
    g_err.clear();
    if (!x) return fit_4pi_device(r, res, nullptr, stream);
    return fit_4pi_device(r, res, x->status, stream, x->rois, x->evals);
}

static int fit_4pi_host(const bsf_request_4pi *r, bsf_result *res, unsigned char *status)
{
    int rc = validate_4pi(r, res);
    if (rc) return rc;
    if ((rc = check_device_4pi(r))) return rc;
    if (r->nfits == 0) return BSF_OK;
    const int nv = nv_4pi(r->shared, r->nch);
    const size_t N = (size_t)r->nfits, S2 = (size_t)r->sz * r->sz;
    size_t nc = 1;
    for (int a = 0; a < 4; a++) nc *= (size_t)r->coeff_dims[a];
    const size_t bd = N * S2 * r->nch * 4, bc = nc * 4, bt = r->dT ? N * 6 * r->nch * 4 : 0;
    const size_t bz = r->initZ ? N * 4 : 0, bp = r->initPhi ? N * 4 : 0, bo = N * (2 * nv + 2) * 4;
    const size_t bs = status ? N : 0;
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    char *buf = nullptr;
    hipError_t e = hipMalloc((void **)&buf, up(bd) + up(bc) + up(bt) + up(bz) + up(bp) + up(bo) + up(bs));
    if (e != hipSuccess) return fail(BSF_ERR_NOMEM, "bsf: device allocation failed");
    char *q = buf;
    float *d_data = (float *)q; q += up(bd);
    float *d_c = (float *)q; q += up(bc);
    float *d_dT = bt ? (float *)q : nullptr; q += up(bt);
    float *d_iz = bz ? (float *)q : nullptr; q += up(bz);
    float *d_ip = bp ? (float *)q : nullptr; q += up(bp);
    float *d_P = (float *)q, *d_CR = d_P + N * (nv + 1), *d_LL = d_CR + N * nv;
    q += up(bo);
    unsigned char *d_st = bs ? (unsigned char *)q : nullptr;
    hipStream_t st = hipStreamPerThread;
    e = hipMemcpyAsync(d_data, r->imstack, bd, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_c, r->coeff, bc, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && bt) e = hipMemcpyAsync(d_dT, r->dT, bt, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && bz) e = hipMemcpyAsync(d_iz, r->initZ, bz, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && bp) e = hipMemcpyAsync(d_ip, r->initPhi, bp, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        bsf::Fit4PiLaunch L = make_launch_4pi(r, d_data, d_c, d_dT, d_iz, d_ip, d_P, d_CR, d_LL);
        L.status = d_st;
        e = launch_all_4pi(L, st);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(res->P, d_P, N * (nv + 1) * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(res->CRLB, d_CR, N * nv * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(res->LL, d_LL, N * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && bs) e = hipMemcpyAsync(status, d_st, N, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    rc = BSF_OK;
    if (e != hipSuccess) rc = fail(BSF_ERR_DEVICE, "bsf: %s", hipGetErrorString(e));
    (void)hipFree(buf);
    return rc;
}

int bsf_fit_4pi_host(const bsf_request_4pi *r, bsf_result *res)
{
    g_err.clear();
    return fit_4pi_host(r, res, nullptr);
}

int bsf_fit_4pi_host_status(const bsf_request_4pi *r, bsf_result *res, unsigned char *status)
{
    g_err.clear();
    if (!status) return fail(BSF_ERR_ARG, "bsf: NULL status array");
    return fit_4pi_host(r, res, status);
}

// ------------------------------------------------------ coefficient builder
// Argument checks of bsarray.m:58-173 for degree 3 (and the builder's limits).
static int validate_coeff(const int *dims, int ndim, const double *lambda)
{
    if (!dims || ndim < 1 || ndim > bsf::kMaxCoeffDims)
        return fail(BSF_ERR_SHAPE, "bsf: coefficient builder takes 1..%d dimensions (got %d)", bsf::kMaxCoeffDims,
                    ndim);
    size_t n = 1;
    for (int a = 0; a < ndim; a++) {
        if (dims[a] < 2) return fail(BSF_ERR_SHAPE, "bsf: dimension %d has %d samples; each needs >= 2", a, dims[a]);
        n *= (size_t)dims[a];
    }
    if (n >= ((size_t)1 << 31)) return fail(BSF_ERR_SHAPE, "bsf: sample arrays must have fewer than 2^31 elements");
    for (int a = 0; lambda && a < ndim; a++)
        if (!(lambda[a] >= 0.0) || std::isinf(lambda[a]))
            return fail(BSF_ERR_ARG, "bsf: lambda must be finite and >= 0 (axis %d: %g)", a, lambda[a]);
    return BSF_OK;
}

static size_t numel_pad(const int *dims, int ndim, int pad)
{
    size_t n = 1;
    for (int a = 0; a < ndim; a++) n *= (size_t)(dims[a] + pad);
    return n;
}

size_t bsf_coeff_work_bytes(const int *dims, int ndim)
{
    if (validate_coeff(dims, ndim, nullptr)) return 0;
    return bsf::coeff_work_bytes(dims, ndim);
}

int bsf_coeff_build_device(const float *samples, const int *dims, int ndim, const double *lambda, float *coeff,
                           void *work, void *stream)
{
    g_err.clear();
    int rc = validate_coeff(dims, ndim, lambda);
    if (rc) return rc;
    if (!samples || !coeff || !work) return fail(BSF_ERR_ARG, "bsf: NULL samples, coeff or work pointer");
    if ((rc = check_device(1, 6, 0))) return rc;
    hipError_t e = bsf::coeff_build(samples, dims, ndim, lambda, coeff, work, (hipStream_t)stream);
    if (e != hipSuccess) return fail(BSF_ERR_DEVICE, "bsf: coefficient build failed: %s", hipGetErrorString(e));
    return BSF_OK;
}

int bsf_coeff_build_host(const float *samples, const int *dims, int ndim, const double *lambda, float *coeff)
{
    g_err.clear();
    int rc = validate_coeff(dims, ndim, lambda);
    if (rc) return rc;
    if (!samples || !coeff) return fail(BSF_ERR_ARG, "bsf: NULL samples or coeff pointer");
    if ((rc = check_device(1, 6, 0))) return rc;
    const size_t n = numel_pad(dims, ndim, 0), nout = numel_pad(dims, ndim, 3);
    const size_t bw = bsf::coeff_work_bytes(dims, ndim);
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    char *buf = nullptr;
    hipError_t e = hipMalloc((void **)&buf, up(n * 4) + up(nout * 4) + bw);
    if (e != hipSuccess) return fail(BSF_ERR_NOMEM, "bsf: device allocation failed");
    float *d_x = (float *)buf, *d_c = (float *)(buf + up(n * 4));
    void *d_w = buf + up(n * 4) + up(nout * 4);
    hipStream_t st = hipStreamPerThread;
    e = hipMemcpyAsync(d_x, samples, n * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = bsf::coeff_build(d_x, dims, ndim, lambda, d_c, d_w, st);
    if (e == hipSuccess) e = hipMemcpyAsync(coeff, d_c, nout * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    rc = e == hipSuccess ? BSF_OK : fail(BSF_ERR_DEVICE, "bsf: %s", hipGetErrorString(e));
    (void)hipFree(buf);
    return rc;
}

// Test hook, not part of include/bsfit.h: one deviance / gradient / Hessian
// evaluation per ROI at the given parameters (theta: host [nfits][NV]);
// out (host): [nfits][1 + NV + NV*NV] = err, J, H.  Lets tests compare the
// GPU's model, derivatives and pixel sums with the oracle's directly.
int bsf_test_eval_pass(const bsf_request *r, const float *theta, float *out)
{
    g_err.clear();
    float dummy = 0.f;
    bsf_result res = {&dummy, &dummy, &dummy};
    int rc = validate(r, &res);
    if (rc) return rc;
    if ((rc = check_device(r->sz, r->fittype, r->precision == BSF_PREC_F64))) return rc;
    if (r->nfits == 0) return BSF_OK;
    const int D = dim_of(r->fittype), NV = nv_of(r->fittype);
    const size_t N = (size_t)r->nfits, S2 = (size_t)r->sz * r->sz;
    const size_t nc = coeff_numel(r, D);
    const size_t nc3 = r->fittype == 8 ? (size_t)r->coeff_dims[0] * r->coeff_dims[1] * r->coeff_dims[2] : 0;
    const size_t no = N * (1 + NV + NV * NV);
    float *d_data = nullptr, *d_c = nullptr, *d_c3 = nullptr, *d_th = nullptr, *d_o = nullptr, *d_out = nullptr;
    hipError_t e = hipMalloc((void **)&d_data, N * S2 * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&d_c, nc * 4);
    if (e == hipSuccess && nc3) e = hipMalloc((void **)&d_c3, nc3 * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&d_th, N * NV * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&d_o, no * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&d_out, N * (2 * NV + 2) * 4);
    if (e == hipSuccess) e = hipMemcpy(d_data, r->imstack, N * S2 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_c, r->coeff, nc * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && nc3) e = hipMemcpy(d_c3, r->coeff3, nc3 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_th, theta, N * NV * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        bsf::FitLaunch L = make_launch(r, d_data, d_c, d_c3, d_out, d_out + N * (NV + 1), d_out + N * (2 * NV + 1));
        L.theta_in = d_th;
        L.dbg = d_o;
            e = launch_all(L, nullptr);
    }
    if (e == hipSuccess) e = hipMemcpy(out, d_o, no * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(BSF_ERR_DEVICE, "bsf: %s", hipGetErrorString(e));
    (void)hipFree(d_data);
    (void)hipFree(d_c);
    (void)hipFree(d_c3);
    (void)hipFree(d_th);
    (void)hipFree(d_o);
    (void)hipFree(d_out);
    return rc;
}

}  // extern "C"
