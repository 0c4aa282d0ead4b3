// Internal launch interface between the C-ABI (bsfit_api.cpp) and the
// gfx950 kernels (bsf_fit_kernels.hip).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include <complex>

namespace bsf {

// Placement of one ROI in a launch that mixes ROI sizes (the serving
// session): size S, channel-0 image at data + off (column-major S x S),
// channel k at data + off + k * cstride.
struct RoiDesc {
    long long off;
    int cstride;
    int S;
};

struct FitLaunch {
    int fittype;            // 6, 7, 8
    int f64;                // 0: FP32 evaluation + FP64 solve, 1: all FP64
    const float *data;      // device, column-major [S, S, N]
    const float *coeff;     // device, bsarray layout, dims cdims[0..ndim)
    const float *coeff3;    // device, fittype 8 only
    float *P, *CRLB, *LL;   // device outputs, column-major with stride N
    int N, S, iterations;   // S: the ROI size, or the largest one when rois is set
    int roi_offset, count;  // ROIs [roi_offset, roi_offset + count)
    const int *ncount;      // device: only ROIs below this count are fitted (outputs keep stride N), or NULL
    const RoiDesc *rois;    // device, per-ROI placement, or NULL
    int uniform;            // with rois: every ROI has size S
    int cdims[5], ndim;
    float initP[4];
    int has_initP;
    int keep;               // multi-start: 0 write, 1 keep if LL gains >= 1e-4, 2 if > 1e-4
    const float *theta_in;  // test hook (device): NV parameters per ROI, or NULL
    float *dbg;             // test hook (device): err, J, H per ROI
    unsigned char *status;  // device: per-ROI status flags (bsf_roi_flag), or NULL
    int *evals;             // device: per-ROI evaluations run, or NULL
    int tile;               // ROIs per wave: 1, or 2 (fittypes 6 / 7)
};

hipError_t launch_fit(const FitLaunch &L, hipStream_t st);

// One launch over the ROIs of up to kHeteroMax fit launches of fittypes 6-8
// (one precision, one ROI per wave): workgroups are dealt to the entries in
// order, count blocks each, and every workgroup runs its entry's fittype.
// The kernel holds every fittype's code, so it takes the largest register
// footprint (fittype 8's) for all of them.
constexpr int kHeteroMax = 4;
hipError_t launch_fit_hetero(const FitLaunch *L, int n, hipStream_t st);

// ROIs per wave of the fit kernel, per (fittype, precision, ROI size) of a
// launch (its largest size with a ROI table): the measured default, or the
// tile set_fit_tile / the autotuner recorded.  A launch whose FitLaunch.tile
// is 0 takes it from here.  FP64 fittype 8, device-side counts and tiles
// whose LDS (two regions) exceed kTileLdsMax run one ROI per wave.
constexpr int kTileMaxS = 64;
constexpr size_t kTileLdsMax = 64 * 1024;
int fit_tile(int fittype, int f64, int S);
int set_fit_tile(int fittype, int f64, int S, int tile);   // tile 0: back to the default; -1 if out of range
int tile_possible(int fittype, int f64, int S, int tile);  // the launch would run with this tile
// LDS bytes of one fit workgroup (one wave) for ROI size S
size_t lds_bytes(int fittype, int f64, int S);

struct Fit4PiLaunch {
    int f64;
    const float *data;        // device, [S, S, N, nch]
    const float *coeff;       // device, 4D (x, y, z, phase) coefficients
    float *P, *CRLB, *LL;     // device outputs
    int N, S, iterations, nch;   // S: the ROI size, or the largest one when rois is set
    int roi_offset, count;
    const RoiDesc *rois;      // device, per-ROI placement, or NULL
    int cdims[4];
    int shared[6];
    float phaseShift[8];      // w units
    const float *dT;          // device [6, nch, N] or NULL
    const float *initZ, *initPhi;  // device [N] or NULL
    int wpc;                  // 1: a workgroup of nch waves per ROI, one per channel; 0: one wave per ROI
    unsigned char *status;    // device: per-ROI status flags (bsf_roi_flag), or NULL
    int *evals;               // device: per-ROI evaluations run, or NULL
    const int *ncount;        // device: ROIs to fit (the rest of the grid exits), or NULL; needs rois
};

hipError_t launch_fit_4pi(const Fit4PiLaunch &L, hipStream_t st);
size_t lds_bytes_4pi(int f64, int S, int nch, int nv);
size_t lds_bytes_4pi_wpc(int f64, int S, int nch, int nv);
// 4Pi kernel choice: -1 automatic, 0 one wave per ROI, 1 one wave per channel
extern int g_4pi_kernel;
// the wave-per-channel kernel when asked for (prefer = 1) or forced, and its LDS fits
int use_wpc_4pi(int prefer, int f64, int S, int nch, int nv, size_t lds_limit);

// Cubic B-spline coefficients (bsarray, degree 3): samples [dims] column-major
// (device, FP32), lambda one per axis or NULL, coeff [dims + 3] (device,
// FP32); work is device scratch of coeff_work_bytes(dims, ndim) bytes.
constexpr int kMaxCoeffDims = 8;
hipError_t coeff_build(const float *samples, const int *dims, int ndim, const double *lambda, float *coeff,
                       void *work, hipStream_t st);
size_t coeff_work_bytes(const int *dims, int ndim);

// Error text of the calling thread (bsf_last_error): set it and return code.
// Frame -> ROI front-end.  frames [nr, nc, nframes] column-major (device;
// pixel_type 0 = uint16, 1 = float32); h [mr, mc] the correlation kernel
// (device); outputs on the device: stack [S, S, cap], peaks [cap, 3] (x, y,
// frame), value [cap] or NULL, count (one int: the ROIs found, possibly more
// than cap); work is device scratch of front_work_bytes(nr, nc, nframes, agg_dist).
struct FrontLaunch {
    const void *frames;
    int pixel_type;
    int nr, nc, nframes, first_frame;
    float offset, conv;
    const float *h;
    int mr, mc;
    int subtract_min;
    double cutoff;
    int S, mirror;
    double psf_size;
    double agg_dist;      // > 0: merge each frame's maxima closer than this (the GUI's aggre)
    int cap;
    float *stack;
    int *peaks;
    float *value;
    int *count;
    void *work;
};
hipError_t launch_find_rois(const FrontLaunch &L, hipStream_t st);
size_t front_work_bytes(int nr, int nc, int nframes, double agg_dist);   // agg_dist <= 0: no merge
// The allocation-free front-end's workspace: a file-scope device array of
// kFrontScratchBytes per GPU; front_scratch returns the current device's
// address of it (looked up once per device).
constexpr size_t kFrontScratchBytes = (size_t)16 << 20;
constexpr int kFrontScratchDevices = 64;
hipError_t front_scratch(void **ptr);
size_t front_detect_lds_bytes(int mr, int mc);

// Checks a fit request as bsf_fit_device does and builds its launch (the
// request's imstack and outputs are device pointers).
int prepare_fit_launch(const struct bsf_request *q, const struct bsf_result *res, FitLaunch *L);
hipError_t launch_fit_all(const FitLaunch &L, hipStream_t st);

int fail(int code, const char *fmt, ...);
void clear_error();

// The current device: its index, whether it is a gfx950 and its LDS limit per
// workgroup, queried from HIP once per device (hipGetDeviceProperties takes
// tens of microseconds, a small launch's worth).  BSF_OK or BSF_ERR_DEVICE.
struct DeviceInfo {
    int dev;
    int gfx950;
    size_t lds_limit;
    char arch[64];
};
int device_info(DeviceInfo *out);

}  // namespace bsf
