// bsf_fit_kernels.hip -- gfx950 fit kernels of the multi-D B-spline PSF
// Poisson-MLE Levenberg-Marquardt fitter (fittypes 6: 3D, 7: 4D, 8: 5D+3D):
// LDS sizing, the generic instantiations of the kernel template in
// bsf_fit_kernel.h (any ROI size, FP32 and FP64) and the launch dispatch,
// which sends uniform-size fittype-6/7 launches to the size-specialised
// kernels (bsf_fit_sized6.hip, bsf_fit_sized7.hip).
#include "bsf_fit_kernel.h"

namespace bsf {

// ------------------------------------------------------------ launchers
// planes, the column sums of the two-pass stage B after them, and the
// reduction scratch over both
static size_t plane_elems(int fittype, int f64, int S)
{
    const int NV = fittype == 6 ? 5 : fittype == 7 ? 6 : 8;
    const int NPL = fittype == 6 ? 2 : fittype == 7 ? 3 : 8;
    const size_t W = S + 4;
    const size_t pl = (size_t)NPL * W * W;
    const size_t h = col_sum_elems(fittype, S) + 3;   // after the planes, 16-byte aligned
    const size_t red = (size_t)red_elems(1 + NV + NV * (NV + 1) / 2);
    const size_t e = (fittype != 8 && !f64) || stage_b_split(fittype, f64, S) ? pl + h : pl;
    return e > red ? e : red;
}

size_t lds_bytes(int fittype, int f64, int S)
{
    const int NV = fittype == 6 ? 5 : fittype == 7 ? 6 : 8;
    return fit_lds_layout(S, f64 ? 8 : 4, plane_elems(fittype, f64, S), (NV + NV * (NV + 1) / 2) * (f64 ? 1 : 2),
                          fit_data_in_lds(fittype))
        .total;
}

hipError_t launch_fit(const FitLaunch &L, hipStream_t st)
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
    const int nb = L.count;
    const int sized = !L.rois || L.uniform ? L.S : 0;
    switch (L.fittype * 2 + (L.f64 ? 1 : 0)) {
    case 12: return sized_kernel(6, sized) ? launch_sized6(A, nb, st, sized) : launch_t<6, float>(A, nb, st);
    case 13: return sized_kernel_f64(6, sized) ? launch_sized6d(A, nb, st, sized) : launch_t<6, double>(A, nb, st);
    case 14: return sized_kernel(7, sized) ? launch_sized7(A, nb, st, sized) : launch_t<7, float>(A, nb, st);
    case 15: return sized_kernel_f64(7, sized) ? launch_sized7d(A, nb, st, sized) : launch_t<7, double>(A, nb, st);
    case 16: return launch_t<8, float>(A, nb, st);
    case 17: return launch_t<8, double>(A, nb, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace bsf
