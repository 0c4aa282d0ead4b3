// bsf_fit_kernels.hip -- gfx950 fit kernels of the multi-D B-spline PSF
// Poisson-MLE Levenberg-Marquardt fitter (fittypes 6: 3D, 7: 4D, 8: 5D+3D):
// LDS sizing, the generic instantiations of the kernel template in
// bsf_fit_kernel.h (any ROI size, FP32 and FP64) and the launch dispatch,
// which sends uniform-size fittype-6/7 launches to the size-specialised
// kernels (bsf_fit_sized6.hip, bsf_fit_sized7.hip).
#include "bsf_fit_kernel.h"

#include <atomic>

namespace bsf {

// ------------------------------------------------------------ launchers
size_t lds_bytes(int fittype, int f64, int S) { return fit_lds_bytes(fittype, f64, S); }

// --------------------------------------------------------------- tiles
namespace {
// recorded tiles, [fittype 6 / 7 / 8][FP64][ROI size]; 0: the default
std::atomic<int> g_tile[3][2][kTileMaxS + 1];

// measured on MI355X (131,072 ROIs per launch, both tiles' bsf_tune
// timings): two ROIs per wave win for FP32 3D fits of up to 9 x 9 pixels
// (2-9 %), where the per-ROI work that does not spread over pixels weighs
// most; from 11 x 11 up, and for fittype 7 and FP64 throughout, one ROI per
// wave is faster (a two-ROI wave holds twice the LDS and registers, so fewer
// waves hide each other's latency)
int default_tile(int fittype, int f64, int S) { return fittype == 6 && !f64 && S <= 9 ? 2 : 1; }
}  // namespace

// two ROIs per wave: fittypes 6 and 7, and fittype 8 in FP32 (in FP64 its
// registers and LDS hold one or two waves of one ROI per SIMD, no room for
// a second ROI per wave)
int tile_possible(int fittype, int f64, int S, int tile)
{
    if (tile == 1) return 1;
    // (a half-wave's 32 lanes hold the tap rows of at most 32 ROI rows)
    return tile == 2 && (fittype == 6 || fittype == 7 || (fittype == 8 && !f64)) && S <= 32 &&
           2 * lds_bytes(fittype, f64, S) <= kTileLdsMax;
}

int fit_tile(int fittype, int f64, int S)
{
    if (fittype < 6 || fittype > 8 || S < 1 || S > kTileMaxS) return 1;
    const int t = g_tile[fittype - 6][f64 ? 1 : 0][S].load(std::memory_order_relaxed);
    return t ? t : default_tile(fittype, f64, S);
}

int set_fit_tile(int fittype, int f64, int S, int tile)
{
    if (fittype < 6 || fittype > 8 || S < 1 || S > kTileMaxS || tile < 0 || tile > 2) return -1;
    if (tile && !tile_possible(fittype, f64, S, tile)) return -1;
    g_tile[fittype - 6][f64 ? 1 : 0][S].store(tile, std::memory_order_relaxed);
    return 0;
}

// FP64 fittype 8 runs with registers for two waves per SIMD where its LDS
// lets a CU (160 KB) hold enough of its waves for the spills to pay: five
// with the one-pass stage B (up to S = 17), six with the two-pass one
// (which fittype 8 no longer takes, kSplit8MaxF64).  Measured on MI355X
// (40,000 ROIs per size, ms per launch, two-wave against one-wave
// registers, same bits): S = 9: 20.0 / 27.0, 11: 29.3 / 33.3, 15: 38.8 /
// 47.4, 16: 50.2 / 50.1, 17: 50.4 / 53.5 (S = 9 and 11 with the two-pass
// form); 100,000 ROIs of sizes 11-17 in one launch: 100.3 / 103.3.
constexpr size_t kCuLds = 160 * 1024;
inline bool f64_two_waves(int S)
{
    const size_t waves = stage_b_split(8, true, S) ? 6 : 5;
    return waves * lds_bytes(8, 1, S) <= kCuLds;
}

hipError_t launch_fit(const FitLaunch &L, hipStream_t st)
{
    const Dev A = fit_dev(L);
    const int nb = L.count;
    const int sized = !L.rois || L.uniform ? L.S : 0;
    // two ROIs per wave where the launch's tile asks for it (fittypes 6 / 7,
    // no device-side count and no test hook)
    const int tile = L.tile ? L.tile : fit_tile(L.fittype, L.f64, L.S);
    if (tile == 2 && tile_possible(L.fittype, L.f64, L.S, 2) && !L.ncount && !L.theta_in) {
        switch (L.fittype * 2 + (L.f64 ? 1 : 0)) {
        case 12: return sized_kernel(6, sized) ? launch_sized6(A, nb, st, sized, 2) : launch_t<6, float, 0, 2>(A, nb, st);
        case 13: return launch_t<6, double, 0, 2>(A, nb, st);
        case 14: return sized_kernel(7, sized) ? launch_sized7(A, nb, st, sized, 2) : launch_t<7, float, 0, 2>(A, nb, st);
        case 15: return launch_t<7, double, 0, 2>(A, nb, st);
        case 16: return launch_t<8, float, 0, 2>(A, nb, st);
        default: return hipErrorInvalidValue;
        }
    }
    switch (L.fittype * 2 + (L.f64 ? 1 : 0)) {
    case 12: return sized_kernel(6, sized) ? launch_sized6(A, nb, st, sized) : launch_t<6, float>(A, nb, st);
    case 13: return sized_kernel_f64(6, sized) ? launch_sized6d(A, nb, st, sized) : launch_t<6, double>(A, nb, st);
    case 14: return sized_kernel(7, sized) ? launch_sized7(A, nb, st, sized) : launch_t<7, float>(A, nb, st);
    case 15: return sized_kernel_f64(7, sized) ? launch_sized7d(A, nb, st, sized) : launch_t<7, double>(A, nb, st);
    case 16: return launch_t<8, float>(A, nb, st);
    case 17: return f64_two_waves(L.S) ? launch_t<8, double, 0, 1, 1>(A, nb, st) : launch_t<8, double>(A, nb, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace bsf
