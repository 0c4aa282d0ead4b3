/*
 * bsfit.h -- C-ABI of the MI355X (gfx950) multi-D B-spline PSF Poisson-MLE
 * Levenberg-Marquardt fitter.  Plain pointers and sizes only.
 *
 * The entry points replace the reference's MATLAB MEX boundary:
 *   MultiD_bspline/visual code/GPUmleFit_LM/mexFunction.cpp:208
 *       [P, CRLB, LogL] = GPUmleFit_LM_multiD(imstack, fittype, iterations,
 *                                             coeff, varmap, silent, initP, coeff3)
 *   and its per-ROI kernels / launch wrappers
 *   MultiD_bspline/visual code/GPUmleFit_LM/wrapper.cu:256-279 (fittypes 6-8).
 *
 * Conventions (MATLAB column-major, as the MEX):
 *   imstack  float [sz, sz, nfits]; element (row ii, col jj) of ROI n at
 *            n*sz*sz + jj*sz + ii.
 *   coeff    float bsarray coefficients, dims = data grid + 3 per axis,
 *            3 axes (fittype 6), 4 (fittype 7) or 5 (fittype 8).
 *   coeff3   fittype 8 only: 3D coefficients with coeff's first three dims.
 *   initP    fittype 7: {z, w}; fittype 8: {z, w, u, g2}; one start shared
 *            by all ROIs (the MEX passes one vector).  NULL selects
 *            nz/2, nw/2 (, nu/2, g2 = 0.9).
 *   P        float [nfits, NV+1]: parameters, last column = iterations used.
 *            fittype 6: x, y, N, bg, z;  7: + w;  8: x, y, N, bg, z, w, u, g2
 *            (x is the column position theta0 as the reference reports it).
 *   CRLB     float [nfits, NV]: variances (diagonal of the inverse Fisher).
 *   LL       float [nfits]: log-likelihood (data-normalised, as the MEX).
 *
 * Every function returns a bsf_status; bsf_last_error() describes the last
 * failure on the calling thread.  Nothing here exits or resets the device.
 */
#ifndef BSFIT_H
#define BSFIT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum bsf_status {
    BSF_OK = 0,
    BSF_ERR_ARG = 1,        /* NULL pointer, bad count, bad dims        */
    BSF_ERR_FITTYPE = 2,    /* fittype not in {6, 7, 8}                 */
    BSF_ERR_ROI_SIZE = 3,   /* sz < 1, sz > 51, or window beyond LDS    */
    BSF_ERR_SHAPE = 4,      /* non-square ROIs / coeff rank mismatch    */
    BSF_ERR_NOMEM = 5,      /* device allocation failed                 */
    BSF_ERR_DEVICE = 6      /* no gfx950 device / HIP runtime error     */
};

enum bsf_precision {
    BSF_PREC_MIX = 0,       /* FP32 spline evaluation and pixel sums, FP64 solve and Fisher inverse */
    BSF_PREC_F64 = 1        /* everything in FP64 */
};

#define BSF_MAX_ROI 51      /* IMSZBIG, MultiD_bspline/visual code/GPUmleFit_LM/definitions.h:13 */

typedef struct bsf_request {
    int fittype;            /* 6, 7 or 8 */
    const float *imstack;   /* [sz, sz, nfits] */
    int sz;
    int nfits;
    int iterations;         /* the MEX default is 50 */
    const float *coeff;
    int coeff_ndim;         /* 3, 4 or 5 */
    int coeff_dims[5];      /* stored dims (data grid + 3) */
    const float *coeff3;    /* fittype 8 */
    const float *initP;     /* may be NULL */
    int n_initP;
    int precision;          /* enum bsf_precision */
} bsf_request;

typedef struct bsf_result {
    float *P;
    float *CRLB;
    float *LL;
} bsf_result;

/* Number of fitted parameters NV for a fittype (5, 6, 8), or -1.
 * Replaces the MEX's output sizing, mexFunction.cpp:528-586. */
int bsf_num_params(int fittype);

/* Synchronous fit, host pointers in and out (copies to and from the device).
 * Replaces mexFunction (GPUmleFit_LM/mexFunction.cpp:208) for fittypes 6-8:
 * H2D copies (:466-518), the kernel launch (:661-674), D2H copies (:725-734). */
int bsf_fit_host(const bsf_request *req, bsf_result *res);

/* Asynchronous fit on device-resident buffers: imstack, coeff, coeff3 and the
 * outputs are device pointers (initP stays a host pointer: a few scalars
 * read at launch); the work is queued on `stream` (a hipStream_t, NULL = the
 * null stream) and nothing is copied or allocated.
 * Replaces the launch wrappers kernel_bsplineMLEFit_z_EMCCD_wrapper,
 * kernel_bsplineMLEFit_4D_EMCCD_wrapper, kernel_splineMLEFit_5n3D_EMCCD_wrapper
 * (GPUmleFit_LM/wrapper.cu:256-279). */
int bsf_fit_device(const bsf_request *req, bsf_result *res, void *stream);

/* Per-ROI status flags, one byte per ROI (bits of enum bsf_roi_flag; 0: none
 * of them).  The reference reports failures only for the whole call
 * (mexErrMsgIdAndTxt + exit, GPUmleFit_LM/mexFunction.cpp:140-176) and
 * leaves per-ROI trouble silent in its outputs; these flags name it. */
enum bsf_roi_flag {
    BSF_ROI_MAX_ITER = 1,        /* the iteration limit ended the fit (the relative
                                    deviance change, and for fittype 8 the summed
                                    step, stayed above the tolerance; P[:, NV] ==
                                    iterations, CPUmleFit_LM.cpp:2555-2557)       */
    BSF_ROI_SOLVE_FAILED = 2,    /* the damped normal equations of the last
                                    iteration had no solution (a negative Cholesky
                                    pivot, or a non-finite step): that step was not
                                    taken, as the reference's kernel_cholesky
                                    failure path (CPUmleFit_LM.cpp:2592, 2701-2704) */
    BSF_ROI_FISHER_SINGULAR = 4, /* a CRLB entry is not a finite positive number:
                                    the Fisher information at the result is
                                    singular or indefinite (e.g. a fit ending where
                                    the spline's z derivative vanishes)           */
    BSF_ROI_NONFINITE = 8,       /* a parameter or the log-likelihood is not finite */
    BSF_ROI_FIRST_SOLVE_FAILED = 16 /* the damped normal equations of the first
                                    iteration had no solution (fittype 7's first
                                    Hessian is singular by construction,
                                    CPUmleFit_LM.cpp:2896, so rounding decides
                                    whether its first step is taken)             */
};

/* bsf_fit_host / bsf_fit_device plus the per-ROI status flags: `status` holds
 * nfits bytes (host memory for the host entry point, device memory for the
 * device one).  The fits are the same, bit for bit. */
int bsf_fit_host_status(const bsf_request *req, bsf_result *res, unsigned char *status);
int bsf_fit_device_status(const bsf_request *req, bsf_result *res, unsigned char *status, void *stream);

/* Placement and size of one ROI in a launch that mixes ROI sizes: its
 * column-major sz x sz image is at imstack + off (floats), 4Pi channel k at
 * imstack + off + k * cstride.  No counterpart in the reference, whose MEX
 * takes one ROI size per call (GPUmleFit_LM/mexFunction.cpp:251-263); for
 * callers that fit ROIs of several sizes (a frame's ROIs cut at the size each
 * emitter needs) in one launch. */
typedef struct bsf_roi {
    long long off;
    int cstride;
    int sz;
} bsf_roi;

/* Optional device-side inputs and outputs of the extended entry points
 * (each may be NULL):
 *   rois    [nfits] per-ROI placement and size; req->sz is then the largest
 *           size of the launch (every rois[n].sz in 1..req->sz: the LDS of a
 *           workgroup is sized for it), imstack the base `off` is counted from
 *   status  [nfits] per-ROI flags (enum bsf_roi_flag)
 *   evals   [nfits] ROI evaluations each fit ran (model, derivatives and
 *           pixel sums over the whole ROI; the initial one, one per accepted or
 *           rejected step, and the final Fisher pass)
 * and the launch's tile (ROIs per wavefront, see bsf_tune): 0 takes the
 * recorded choice for (fittype, precision, sz), 1 or 2 forces it (2 is
 * taken by fittypes 6, 7 and FP32 fittype 8 up to sz 32 where two ROIs'
 * LDS fits 64 KB; elsewhere the launch runs with 1). */
typedef struct bsf_fit_extras {
    const bsf_roi *rois;
    unsigned char *status;
    int *evals;
    int tile;
} bsf_fit_extras;

/* bsf_fit_device (the launch wrappers, GPUmleFit_LM/wrapper.cu:256-279) with
 * the extras above; x == NULL is bsf_fit_device.  With rois every ROI is
 * fitted by the same kernel and code as a uniform launch of its size; a
 * launch of mixed sizes gives each ROI the results it has alone, bit for bit. */
int bsf_fit_device_ex(const bsf_request *req, bsf_result *res, const bsf_fit_extras *x, void *stream);

/* Autotuner of the fit kernel's tile: one ROI per 64-lane wavefront, or two
 * (each half-wave fits its own ROI, and the work of a fit that does not
 * spread over pixels -- spline geometry, the damped solve, the LM control --
 * runs once for both).  The reference has one shape, a thread per ROI in
 * blocks of 64 (GPUmleFit_LM/mexFunction.cpp:596-601), so this has no
 * counterpart there; it is part of the session API's tuning surface.
 * Fits `req` (device buffers, as bsf_fit_device) with each possible tile:
 * one untimed launch, then `reps` launches timed with HIP events on
 * `stream` (waited for); records the fastest for (fittype, precision, sz)
 * in the process-wide table that later fits of that kind use, and returns
 * it in *tile; ms (2 floats, may be NULL) receives each tile's mean launch
 * time in milliseconds (0 where the tile is not possible); res holds the
 * fits of the chosen tile afterwards.  For fittypes 6 and 7 both tiles give
 * the same results bit for bit.  A two-ROI fittype-8 wave (FP32 only) keeps
 * one set of pixel sums per lane, added in another order than the one-ROI
 * kernel's, so its fits differ in the last bits (same accuracy against the
 * FP64 oracle); a launch takes the tile recorded for its largest ROI size,
 * so with tile 2 recorded for some fittype-8 sizes and not others a ROI's
 * bits depend on the largest size it is launched with (sessions: on its
 * frame).  Measured on MI355X one ROI per wave is faster for fittype 8 at
 * every size, the default.  FP64 fittype 8 and the 4Pi fitter have one
 * tile. */
int bsf_tune(const bsf_request *req, bsf_result *res, int reps, int *tile, float *ms, void *stream);

/* The recorded tile for (fittype, precision, sz) (1 or 2), and setting it
 * by hand (tile 0 restores the measured default; BSF_ERR_ARG where the tile
 * is not possible).  No reference counterpart (see bsf_tune). */
int bsf_fit_tile(int fittype, int precision, int sz);
int bsf_set_fit_tile(int fittype, int precision, int sz, int tile);

/* Multi-start fit: every ROI is fitted from nstarts starts (rows of
 * `starts`, req->n_initP values each: fittype 7 {z, w}, fittype 8
 * {z, w, u, g2}) and the outputs keep, start by start, a result whose
 * log-likelihood beats the kept one by at least 1e-4 (strict != 0: by more
 * than 1e-4), compared in FP32.  Replaces the reference callers' loop over
 * GPUmleFit_LM_multiD calls with the same rule
 * (demo_orientation/store/fitb5n3.m:32-46 (>=), MultiD_fitter/
 * simplefitter_spline.m:482-494 (>) and 512-520 (>=)): one upload, one
 * launch per start on the device buffers, one download.  req->initP is
 * ignored; fittypes 7 and 8; 1..64 starts. */
int bsf_fit_multistart_host(const bsf_request *req, const float *starts, int nstarts, int strict, bsf_result *res);

/* The same on device buffers (as bsf_fit_device; starts is a host array). */
int bsf_fit_multistart_device(const bsf_request *req, const float *starts, int nstarts, int strict, bsf_result *res,
                              void *stream);

/* Multi-start fits plus the per-ROI status flags (enum bsf_roi_flag) of the
 * kept result of every ROI (host / device status array of nfits bytes). */
int bsf_fit_multistart_host_status(const bsf_request *req, const float *starts, int nstarts, int strict,
                                   bsf_result *res, unsigned char *status);
int bsf_fit_multistart_device_status(const bsf_request *req, const float *starts, int nstarts, int strict,
                                     bsf_result *res, unsigned char *status, void *stream);

/* The MEX argument list as plain C (GPUmleFit_LM/mexFunction.cpp:14-55):
 * imdims = size(imstack) (2 dims: one ROI), coeff_dims = size(coeff).
 * varmap is accepted and ignored as the MEX ignores it for fittypes 6-8
 * (:363-373, 661-674); silent = 0 prints the fit rate to stdout (:753-760).
 * P, CRLB and LL are caller-owned host arrays sized as bsf_fit_host's. */
int bsf_GPUmleFit_LM_multiD(const float *imstack, const size_t *imdims, int im_ndim, int fittype,
                            int iterations, const float *coeff, const size_t *coeff_dims, int coeff_ndim,
                            const float *varmap, size_t varmap_numel, int silent, const float *initP,
                            size_t initP_numel, const float *coeff3, float *P, float *CRLB, float *LL);

/* The same plus the per-ROI status flags (enum bsf_roi_flag, one byte per
 * ROI): the gateway's optional fourth output.  The reference MEX has three
 * outputs and no such flags. */
int bsf_GPUmleFit_LM_multiD_status(const float *imstack, const size_t *imdims, int im_ndim, int fittype,
                                   int iterations, const float *coeff, const size_t *coeff_dims, int coeff_ndim,
                                   const float *varmap, size_t varmap_numel, int silent, const float *initP,
                                   size_t initP_numel, const float *coeff3, float *P, float *CRLB, float *LL,
                                   unsigned char *status);

/* ------------------------------------------------------------------ 4Pi
 * Multi-channel 4Pi fit.  Replaces the MEX the reference's 4Pi demo calls,
 *   [P, CRLB, LogL] = GPUmleFit_LM_4Pi_bspline(imstack, uint32(shared), iterations,
 *                         coeffs, dT, initZ, initPhi, phaseShift)
 * (demo_4pi/bspline_based_4piloc_v4.m:196-206), whose source the reference
 * does not ship.
 *   imstack    float [sz, sz, nfits, nch], nch <= 8 channels
 *   shared     6 flags for x, y, N, bg, z, phi; NV = sum(shared ? 1 : nch) <= 24
 *              (calculate_CRLB_YL_shared.m:11-13 layout: an unshared
 *              parameter takes nch consecutive slots)
 *   coeff      4D bsarray coefficients (x, y, z, phase); the phase axis has
 *              period nw-2 samples (nw = coeff_dims[3]-3)
 *   dT         float [6, nch, nfits] per-channel offsets (x, y used), or NULL
 *   initZ, initPhi   float [nfits] starts (spline units), or NULL (nz/2, nw/2)
 *   phaseShift float [nch] channel phase offsets in phase-axis (w) units
 *   P [nfits, NV+1] (last column: iterations), CRLB [nfits, NV], LL [nfits]. */
typedef struct bsf_request_4pi {
    const float *imstack;
    int sz;
    int nfits;
    int nch;
    const unsigned int *shared;
    int iterations;
    const float *coeff;
    int coeff_dims[4];
    const float *dT;
    const float *initZ;
    const float *initPhi;
    const float *phaseShift;
    int precision;
} bsf_request_4pi;

/* NV for a shared mask and channel count, or -1. */
int bsf_num_params_4pi(const unsigned int *shared, int nch);

/* Synchronous 4Pi fit on host arrays (copies to and from the device). */
int bsf_fit_4pi_host(const bsf_request_4pi *req, bsf_result *res);

/* Asynchronous 4Pi fit on device buffers (imstack, coeff, dT, initZ, initPhi
 * and the outputs are device pointers; shared and phaseShift are host). */
int bsf_fit_4pi_device(const bsf_request_4pi *req, bsf_result *res, void *stream);

/* 4Pi fits plus the per-ROI status flags (enum bsf_roi_flag): host / device
 * status array of nfits bytes. */
int bsf_fit_4pi_host_status(const bsf_request_4pi *req, bsf_result *res, unsigned char *status);
int bsf_fit_4pi_device_status(const bsf_request_4pi *req, bsf_result *res, unsigned char *status, void *stream);
/* bsf_fit_4pi_device with the extras of bsf_fit_device_ex (rois: per-ROI
 * placement of channel 0 and the channel stride). */
int bsf_fit_4pi_device_ex(const bsf_request_4pi *req, bsf_result *res, const bsf_fit_extras *x, void *stream);

/* Text of the last error on this thread ("" if none).
 * Replaces mexErrMsgIdAndTxt (GPUmleFit_LM/mexFunction.cpp:140-176). */
const char *bsf_last_error(void);

/* ------------------------------------------------ coefficient builder
 * Cubic B-spline coefficients of a sampled PSF model, the array the fit entry
 * points take as coeff / coeff3.  Replaces the reference's MATLAB
 *   b = bsarray(a, 'lambda', lambda); coeff = single(b.coeffs)
 * with degree 3 on every axis (MultiD_bspline/bsarray/bsarray.m:58-173,
 * directFilter.m:1-297, getBSplineFiltCoeffs.m, symExpFilt.m:26-45), computed
 * in FP64 on the GPU:
 *   samples  float [dims[0], ..., dims[ndim-1]] column-major, ndim 1..8,
 *            every dim >= 2
 *   lambda   one smoothing weight >= 0 per axis (0 = interpolation), or NULL
 *   coeff    float [dims[0]+3, ..., dims[ndim-1]+3] column-major */
int bsf_coeff_build_host(const float *samples, const int *dims, int ndim, const double *lambda, float *coeff);

/* Device scratch bytes bsf_coeff_build_device needs for these dims (0 if the
 * dims are invalid). */
size_t bsf_coeff_work_bytes(const int *dims, int ndim);

/* The same on device buffers (samples, coeff, work), queued on `stream`;
 * lambda stays a host array; nothing is allocated. */
int bsf_coeff_build_device(const float *samples, const int *dims, int ndim, const double *lambda, float *coeff,
                           void *work, void *stream);

/* -------------------------------------------------------------- sessions
 * A session keeps models (coefficient arrays) resident on one device and
 * fits frames asynchronously.  A frame is a list of ROI groups; a group is
 * one MEX call's worth of ROIs (one model, one ROI size, one initP), as the
 * reference's callers batch them (GPUmleFit_LM/mexFunction.cpp:259-263: one
 * size per call).  The session packs all groups of a frame that share a
 * model, iteration count and start into one kernel launch (mixed ROI sizes),
 * runs the launches of a frame concurrently, and copies the results back.
 *
 * Device memory: the models plus a frame arena of fixed size, allocated at
 * creation; both count against hbm_budget.  A frame is admitted when its
 * inputs and outputs fit in the arena next to the frames still in flight
 * (HBM-budget admission); a blocking submit waits for the oldest frames,
 * a non-blocking one returns BSF_BUSY.  Submits and polls allocate nothing.
 * All calls on one session are serialised by the session (thread-safe).
 *
 * No counterpart in the reference, which allocates, copies and frees per
 * call (mexFunction.cpp:466-586, 741-748); this is the serving path for
 * callers that fit a stream of frames. */
#define BSF_FT_4PI 104      /* model fittype of the 4Pi multi-channel fitter */

enum {
    BSF_BUSY = 7,           /* non-blocking submit: the arena has no room now */
    BSF_PENDING = 8         /* poll: the frame has not finished */
};

typedef struct bsf_session bsf_session;

typedef struct bsf_session_config {
    int device;             /* HIP device ordinal */
    size_t hbm_budget;      /* device bytes for models + arena (0: 2 GiB) */
    size_t arena_bytes;     /* frame arena, part of the budget (0: 256 MiB) */
    int kernel_streams;     /* streams that run fit launches, 1..16 (0: 2); with the two
                               copy streams, keep within the HIP runtime's hardware
                               queues (GPU_MAX_HW_QUEUES, default 4) */
    int precision;          /* enum bsf_precision for submitted frames */
} bsf_session_config;

typedef struct bsf_model_desc {
    int fittype;                /* 6, 7, 8 or BSF_FT_4PI */
    const float *coeff;         /* host bsarray coefficients (as bsf_request) */
    int coeff_ndim;             /* 3, 4, 5 (4 for BSF_FT_4PI) */
    int coeff_dims[5];
    const float *coeff3;        /* fittype 8: 3D coefficients */
    int nch;                    /* BSF_FT_4PI: channels (1..8) */
    unsigned int shared[6];     /* BSF_FT_4PI: shared-parameter mask */
    float phaseShift[8];        /* BSF_FT_4PI: channel phase offsets (w units) */
} bsf_model_desc;

typedef struct bsf_roi_group {
    int model;                  /* id from bsf_model_register */
    int sz;                     /* ROI size, 1..51 */
    int count;                  /* ROIs in the group (0 allowed) */
    int iterations;
    const float *data;          /* host [sz, sz, count] (4Pi: per channel) */
    size_t chan_stride;         /* 4Pi: floats between channels, 0: count*sz*sz */
    const float *initP;         /* fittype 7: 2, 8: 4 values, or NULL */
    int n_initP;
    const float *dT;            /* 4Pi: host [6, nch, count] or NULL */
    const float *initZ;         /* 4Pi: host [count] or NULL */
    const float *initPhi;       /* 4Pi: host [count] or NULL */
    float *P, *CRLB, *LL;       /* host outputs as bsf_result (stride count);
                                   written by bsf_poll / a blocking submit,
                                   must stay valid until the frame completes */
    int nstarts;                /* fittypes 7, 8: starts in initP (nstarts rows of
                                   n_initP values), 0 or 1: one; up to 8.  Kept
                                   as bsf_fit_multistart_* keeps them */
    int strict;                 /* multi-start rule: 0 LL gain >= 1e-4, 1 > 1e-4 */
} bsf_roi_group;

typedef struct bsf_frame {
    const bsf_roi_group *groups;
    int ngroups;
} bsf_frame;

#define BSF_SUBMIT_NONBLOCK 1

/* Creates a session on cfg->device: streams, events, the frame arena and
 * its pinned host mirror. */
int bsf_session_create(const bsf_session_config *cfg, bsf_session **out);

/* Waits for the frames in flight, delivers their results, frees everything. */
int bsf_session_destroy(bsf_session *s);

/* Copies a model's coefficients to the device; *model_id names it in
 * groups.  Fails with BSF_ERR_NOMEM when models + arena exceed hbm_budget. */
int bsf_model_register(bsf_session *s, const bsf_model_desc *m, int *model_id);

/* Registers a model from its samples: the coefficients (and, fittype 8, the
 * 3D coefficients) are built on the device (bsf_coeff_build) straight into
 * the model's buffers.  Dims are the sample dims (coefficients get +3);
 * fittype 8's 3D samples share dims[0..2].  Same result as building with
 * bsf_coeff_build_host and registering the coefficients. */
typedef struct bsf_model_samples {
    int fittype;                /* 6, 7, 8 or BSF_FT_4PI */
    const float *samples;       /* host, [dims] column-major */
    int ndim;                   /* 3, 4, 5 (4 for BSF_FT_4PI) */
    int dims[5];
    const double *lambda;       /* per-axis smoothing, or NULL */
    const float *samples3;      /* fittype 8: host [dims[0], dims[1], dims[2]] */
    const double *lambda3;      /* fittype 8: 3 values, or NULL */
    int nch;                    /* BSF_FT_4PI, as bsf_model_desc */
    unsigned int shared[6];
    float phaseShift[8];
} bsf_model_samples;

int bsf_model_register_samples(bsf_session *s, const bsf_model_samples *m, int *model_id);

/* Frees a model's device memory (waits for frames in flight first). */
int bsf_model_release(bsf_session *s, int model_id);

/* Precision of frames submitted from now on (enum bsf_precision). */
int bsf_session_set_precision(bsf_session *s, int precision);

/* Queues a frame.  Host inputs are copied before return (the caller may
 * reuse them); outputs are delivered by bsf_poll.  flags: 0 = wait for
 * arena room, BSF_SUBMIT_NONBLOCK = return BSF_BUSY instead.  A frame larger
 * than the whole arena fails with BSF_ERR_NOMEM. */
int bsf_submit(bsf_session *s, const bsf_frame *f, int flags, unsigned long long *ticket);

/* BSF_OK once the frame's results are in its groups' host outputs,
 * BSF_PENDING while it runs (wait = 0), BSF_ERR_ARG for an unknown ticket.
 * wait != 0 blocks until the frame completes. */
int bsf_poll(bsf_session *s, unsigned long long ticket, int wait);

/* Occupancy of the session: frames in flight and arena bytes they hold. */
int bsf_session_inflight(bsf_session *s, int *frames, size_t *arena_used);

/* Graph mode: slots = 1..64 cuts the arena into that many equal slots, each
 * holding one frame in flight (16 measured best for the streamed mix);
 * 0 returns to per-call launches, the default.  Waits for the frames in
 * flight.  A frame whose launches match a captured plan -- the same models,
 * iterations, starts and optional inputs, with ROI counts and sizes within
 * the plan's capacities, which are twice the counts of the frame that made
 * it -- is submitted as one host-to-device copy, one hipGraphLaunch of the
 * plan's graph for its slot (its launches read their ROI counts from device
 * memory) and one device-to-host copy.  A frame no plan fits is captured
 * anew; one that a slot cannot hold is launched call by call in as many
 * slots as it needs (counted in `eager`).  Results are the same, bit for
 * bit, in both modes.  Measured on the streamed mix, per-call launches
 * sustain more frames and have the lower latency, hence the default.  No
 * reference counterpart (the reference launches per MEX call,
 * GPUmleFit_LM/mexFunction.cpp:596-601). */
int bsf_session_set_graphs(bsf_session *s, int slots);
/* Heterogeneous launches (on = 1; 0, the default, launches per model):
 * a frame's single-start groups of fittypes 6, 7 and 8 -- up to four
 * launches' worth, across models -- run as one launch in which every
 * workgroup runs its ROI's fittype (4Pi and multi-start groups launch as
 * before).  Results are the same, bit for bit.  That kernel holds every
 * fittype's code and takes fittype 8's register footprint for all of them,
 * so 3D and colour-4D ROIs run at half their own kernels' occupancy; measured
 * on the streamed mix it is slower, hence the default.  Applies to frames
 * launched call by call (not to graph replays).  No reference counterpart. */
int bsf_session_set_hetero(bsf_session *s, int on);
/* Graph mode counters: graphs captured, frames replayed, frames launched
 * call by call in graph mode (each may be NULL). */
int bsf_session_graph_stats(bsf_session *s, unsigned long long *captures, unsigned long long *replays,
                            unsigned long long *eager);

/* ------------------------------------------------ frames -> ROI stack
 * The GUI fitter's per-frame front-end (MultiD_fitter/simplefitter_spline.m:
 * 197-307 with MultiD_fitter/shared/maximumfindcall.m:4-21 and
 * maximumfind2.c:18-65), which cuts the ROIs that GPUmleFit_LM_multiD fits:
 *   imphot = (single(frame) - offset) * conversion          (:259)
 *   impf   = filter2(h, imphot)                  Gaussian mode (:275-276)
 *          = filter2(h, imphot - min(imphot(:,1)))  DoG / PSF modes (:266-273)
 *   maxima = pixels of impf above all 8 neighbours, interior pixels, in
 *            column-major scan order, x = column + 1, y = row + 1
 *   kept   where impf > cutoff, dx < x <= ncols - dx, dx < y <= nrows - dx
 *            (:277-280; dx = floor(roi/2))
 *   merged with agg_dist > 0 (p.agg), per frame, as aggre (:281-283,
 *            653-673): while two maxima ii < jj (scan order) lie at a
 *            distance 0 < D < agg_dist, the first such pair in column-major
 *            order over (ii, jj) (smallest jj, then smallest ii) becomes
 *            round(mean) of the two in row ii (x, y and value; half away
 *            from zero) and row jj is deleted
 *   kept   with psf_size > 0 where psf_size/2 <= x, y <= nrows - psf_size/2
 *            (:284-295)
 *   ROI k  = imphot(y-dx:y+dx, x-dx:x+dx), columns reversed if mirror
 *            (:296-307); peaks(k, :) = [x, y, frame] as peakcoordinates.
 * Frames are [nrows, ncols, nframes] column-major; frame f is numbered
 * first_frame + f (the GUI counts from 1).  The filter2 sum is accumulated
 * in FP32 with fused multiply-adds, kernel column by column (v outer, row u
 * inner). */
#define BSF_MAX_PEAK_FILTER 51

enum bsf_pixel_type { BSF_PIXEL_U16 = 0, BSF_PIXEL_F32 = 1 };

typedef struct bsf_roi_request {
    const void *frames;     /* host (bsf_find_rois_host) or device pointer */
    int pixel_type;         /* enum bsf_pixel_type */
    int nrows, ncols, nframes;
    int first_frame;
    float offset;           /* p.offset, ADU */
    float conversion;       /* p.conversion, photons per ADU */
    const float *filter;    /* [filter_rows, filter_cols] column-major; host or device as frames */
    int filter_rows, filter_cols;
    int subtract_min;       /* filter imphot - min(imphot(:,1)) (DoG and PSF modes) */
    double cutoff;          /* p.peakcutoff */
    int roi;                /* p.roifit, odd, 3..51 */
    int mirror;             /* p.mirror */
    double psf_size;        /* PSF-convolution mode: size(cal.PSF, 1); 0 = off */
    int capacity;           /* ROIs the outputs hold */
    double agg_dist;        /* p.agg: p.roifit * p.aggsz, maxima closer than this are
                               merged (aggre); 0 = off.  Aggregation enlarges the
                               workspace (bsf_find_rois_work_bytes: ~24 B per 4 pixels
                               of all frames, plus at most 1/4 B per pixel of cell
                               heads); the search for the next pair looks only at the
                               maxima of the neighbouring cells, so dense frames (a
                               low cutoff) stay linear in their maxima per merge */
} bsf_roi_request;

typedef struct bsf_roi_result {
    float *imstack;         /* [roi, roi, capacity] photons */
    int *peaks;             /* [capacity, 3] column-major: x, y, frame number */
    float *value;           /* [capacity] filtered value at the maximum (with
                               aggregation, round(mean) where maxima merged, as the
                               reference's maxima column 3), or NULL */
    int count;              /* bsf_find_rois_host: ROIs found; when above capacity
                               only the first capacity are written */
} bsf_roi_result;

/* Host frames in, host ROI stack out (copies to and from the device). */
int bsf_find_rois_host(const bsf_roi_request *req, bsf_roi_result *res);

/* Device frames, filter and outputs; the number of ROIs found goes to the
 * device int *count; work is device scratch of bsf_find_rois_work_bytes
 * bytes.  Queued on stream; allocates nothing. */
size_t bsf_find_rois_work_bytes(const bsf_roi_request *req);
int bsf_find_rois_device(const bsf_roi_request *req, bsf_roi_result *res, int *count, void *work, size_t work_bytes,
                         void *stream);

/* bsf_find_rois_device without the work buffer, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * forbid allocation).  No counterpart in the reference.  The workspace is a
 * device array inside the library, 16 MiB per GPU, held until the process
 * ends: requests needing more than that (bsf_find_rois_work_bytes; about
 * 8 B per 64 pixels of all frames, e.g. up to 120 frames of 1024 x 1024;
 * with aggregation about 6.3 B per pixel, e.g. 2 such frames) return
 * BSF_ERR_NOMEM.  The call allocates nothing.  Calls that overlap on
 * one GPU share that array and race: serialise them (one stream, or events). */
int bsf_find_rois_device_noalloc(const bsf_roi_request *req, bsf_roi_result *res, int *count, void *stream);

/* Frames -> ROIs -> fits in one call: the GUI's per-frame loop followed by
 * its fitspline call on the ROI stack (simplefitter_spline.m:197-316, 479-546)
 * without a round trip to the host.  fit gives fittype, iterations, coeff
 * (and coeff3, initP, precision) as for bsf_fit_device; its imstack, sz and
 * nfits are taken from the ROI stack (sz = roi, nfits = capacity).  The fit
 * outputs res->P [capacity, NV+1], CRLB [capacity, NV], LL [capacity] keep
 * stride capacity; the first *count rows (at most capacity) are the fits of
 * the ROIs in rois, in the same order.  Device pointers, queued on stream,
 * the count read by the fit on the device; allocates nothing. */
int bsf_fit_frames_device(const bsf_roi_request *rq, const bsf_request *fit, bsf_roi_result *rois, int *count,
                          bsf_result *res, void *work, size_t work_bytes, void *stream);

/* The same from host frames and host coefficients into host outputs (sized
 * for capacity ROIs; rois->count receives the number found). */
int bsf_fit_frames_host(const bsf_roi_request *rq, const bsf_request *fit, bsf_roi_result *rois, bsf_result *res);

/* The GUI's peak-finding kernels (simplefitter_spline.m:195-212), computed
 * in FP64 and rounded to FP32:
 *   BSF_PEAK_GAUSS  fspecial('gaussian', max(3, ceil(3 sigma + 1)), sigma)
 *   BSF_PEAK_DOG    fspecial('gaussian', n, sigma) - fspecial('gaussian', n,
 *                   max(1, 2.5 sigma)), n = max(ceil(6 sigma + 1), 3)
 * *size receives n; h must hold n * n floats (cap). */
enum bsf_peak_filter_kind {
    BSF_PEAK_GAUSS = 0, BSF_PEAK_DOG = 1, BSF_PEAK_PSF_WINDOW = 2, BSF_PEAK_PSF_MEAN = 3
};
int bsf_peak_filter(int kind, double sigma, float *h, int cap, int *size);

/* Kernels from the calibration PSF [n, n, nz] (column-major):
 *   BSF_PEAK_PSF_WINDOW  p.conv (:199-207): the z-mean of rows/columns
 *                        floor((n-roi)/2)..floor((n+roi)/2) (1-based), minus
 *                        its mean, scaled to 0.05 at its maximum
 *   BSF_PEAK_PSF_MEAN    'Conv PSF' background mode (:262-268): the z-mean of
 *                        5 PSF(:,:,z) minus its column means (one filter with
 *                        the mean kernel instead of the mean of nz filters). */
int bsf_peak_filter_psf(int kind, const float *psf, const int *psf_dims, int roi, float *h, int cap, int *rows,
                        int *cols);

/* ---------------------------------------------------- GUI result table
 * The GUI fitter's localisation table from one fit block
 * (MultiD_fitter/simplefitter_spline.m:550-630): results [N, ncols]
 * column-major double, ncols = bsf_gui_results_cols(fittype) (12, 17, 19),
 * from P [N, NV+1], CRLB [N, NV], LL [N] (the fit's outputs) and peaks
 * [N, 3] double (the GUI's peakcoordinates: x, y, frame).  dx = p.dx, z0 =
 * p.z0, dz = p.dz, du = p.du, dv = p.dv (fittype 8), mirror = p.mirror.
 *   column 1 frame; 2 x = P(:,1) (fittypes 7, 8) or P(:,2) (fittype 6) - dx +
 *   peak x, or dx - P(:,2) + peak x when mirrored; 3 y; then z (fittypes 6,
 *   7: -(z - z0) dz, fittype 8: (z - z0) dz), photons, background, errors
 *   as real(sqrt(CRLB)), LL, ... exactly as the GUI lays them out, each value
 *   computed in single precision as MATLAB does with single P and CRLB.
 * Host arithmetic only (no device needed). */
int bsf_gui_results_cols(int fittype);
int bsf_gui_results(int fittype, int mirror, double dx, double z0, double dz, double du, double dv, const float *P,
                    const float *CRLB, const float *LL, const double *peaks, int N, double *results);

/* The colour demo's class from the fitted colour coordinate w = P(:,6)
 * (demo_color_classification/bspline_based_color_classification_v3.m:
 * 115-117): (min(3, max(2, w)) - 2.5) * 2 in single precision. */
int bsf_color_class(const float *w, int N, float *cls);

#ifdef __cplusplus
}
#endif

#endif /* BSFIT_H */
