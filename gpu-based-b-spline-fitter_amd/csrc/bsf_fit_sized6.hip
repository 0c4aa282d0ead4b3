// bsf_fit_sized6.hip -- fittype-6 FP32 fit kernels specialised on the ROI
// size (see sized_kernel in bsf_fit_kernel.h); a translation unit of their
// own so that they compile in parallel with the generic kernels.
#include "bsf_fit_kernel.h"

namespace bsf {

hipError_t launch_sized6(const Dev &A, int nrois, hipStream_t st, int S, int R)
{
    // R = 2: two ROIs per wave (fit_kernel's tile of two half-waves)
    switch (S * 4 + R) {
    case 11 * 4 + 1: return launch_t<6, float, 11>(A, nrois, st);
    case 13 * 4 + 1: return launch_t<6, float, 13>(A, nrois, st);
    case 15 * 4 + 1: return launch_t<6, float, 15>(A, nrois, st);
    case 17 * 4 + 1: return launch_t<6, float, 17>(A, nrois, st);
    case 11 * 4 + 2: return launch_t<6, float, 11, 2>(A, nrois, st);
    case 13 * 4 + 2: return launch_t<6, float, 13, 2>(A, nrois, st);
    case 15 * 4 + 2: return launch_t<6, float, 15, 2>(A, nrois, st);
    case 17 * 4 + 2: return launch_t<6, float, 17, 2>(A, nrois, st);
    default: return hipErrorInvalidValue;
    }
}

// FP64: the specialised kernel needs 164 VGPRs at S <= 15 (the generic one
// 205), 3 waves per SIMD instead of 2: 16-28 % faster there, 4 % at S = 17
hipError_t launch_sized6d(const Dev &A, int nblocks, hipStream_t st, int S)
{
    switch (S) {
    case 11: return launch_t<6, double, 11>(A, nblocks, st);
    case 13: return launch_t<6, double, 13>(A, nblocks, st);
    case 15: return launch_t<6, double, 15>(A, nblocks, st);
    case 17: return launch_t<6, double, 17>(A, nblocks, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace bsf
