// bsf_gui.cpp -- the GUI fitter's localisation table (include/bsfit.h:
// bsf_gui_results), MultiD_fitter/simplefitter_spline.m:550-630, and the
// colour demo's class mapping,
// demo_color_classification/bspline_based_color_classification_v3.m:115-117.
//
// MATLAB evaluates these expressions with single-precision P and CRLB and
// double parameters; an operation with a single operand is carried out in
// single precision (the double converted first), so every expression below
// is FP32 arithmetic on FP32-rounded parameters, in the GUI's order, and the
// result is stored into the GUI's double table.  real(sqrt(v)) is 0 for
// v < 0 (the real part of an imaginary root).
#include <cmath>

#include "../../include/bsfit.h"
#include "bsf_kernels.h"

// keep each single-precision operation rounded on its own
#pragma clang fp contract(off)

namespace {

float real_sqrt(float v) { return v < 0.0f ? 0.0f : std::sqrt(v); }

}  // namespace

extern "C" {

int bsf_gui_results_cols(int fittype) { return fittype == 6 ? 12 : fittype == 7 ? 17 : fittype == 8 ? 19 : -1; }

int bsf_gui_results(int fittype, int mirror, double dx, double z0, double dz, double du, double dv, const float *P,
                    const float *CRLB, const float *LL, const double *peaks, int N, double *results)
{
    bsf::clear_error();
    const int NV = bsf_num_params(fittype);
    if (NV < 0) return bsf::fail(BSF_ERR_FITTYPE, "bsf: fittype %d has no GUI table (6, 7, 8)", fittype);
    if (N < 0) return bsf::fail(BSF_ERR_ARG, "bsf: N must be >= 0");
    if (N > 0 && (!P || !CRLB || !LL || !peaks || !results)) return bsf::fail(BSF_ERR_ARG, "bsf: NULL argument");
    const float fdx = (float)dx, fz0 = (float)z0, fdz = (float)dz, fdu = (float)du, fdv = (float)dv;
    const size_t n = (size_t)N;
    auto p = [&](int col, size_t k) { return P[(size_t)(col - 1) * n + k]; };        // P(k, col), 1-based col
    auto c = [&](int col, size_t k) { return CRLB[(size_t)(col - 1) * n + k]; };
    double *R = results;
    auto out = [&](int col, size_t k, float v) { R[(size_t)(col - 1) * n + k] = (double)v; };
    for (size_t k = 0; k < n; k++) {
        const float px = (float)peaks[k], py = (float)peaks[n + k];
        R[k] = peaks[2 * n + k];                                                       // frame
        // x: mirrored data, else the 4D/5D calibrations' x/y exchange (:550-560)
        if (mirror)
            out(2, k, (fdx - p(2, k)) + px);
        else if (fittype == 7 || fittype == 8)
            out(2, k, (p(1, k) - fdx) + px);
        else
            out(2, k, (p(2, k) - fdx) + px);
        if (fittype == 6) {                                                            // :573-582 (spline)
            out(3, k, (p(1, k) - fdx) + py);
            out(4, k, -((p(5, k) - fz0) * fdz));
            out(5, k, p(3, k));
            out(6, k, p(4, k));
            out(7, k, real_sqrt(c(2, k)));
            out(8, k, real_sqrt(c(1, k)));
            out(9, k, real_sqrt(c(5, k) * fdz));
            out(10, k, real_sqrt(c(3, k)));
            out(11, k, real_sqrt(c(4, k)));
            out(12, k, LL[k]);
        } else if (fittype == 7) {                                                     // :598-612
            out(3, k, (p(2, k) - fdx) + py);
            out(4, k, -((p(5, k) - fz0) * fdz));
            out(5, k, p(6, k));
            out(6, k, p(3, k));
            out(7, k, p(4, k));
            out(8, k, real_sqrt(c(2, k)));
            out(9, k, real_sqrt(c(1, k)));
            out(10, k, real_sqrt(c(5, k) * fdz));
            out(11, k, real_sqrt(c(6, k)));
            out(12, k, real_sqrt(c(3, k)));
            out(13, k, real_sqrt(c(4, k)));
            out(14, k, LL[k]);
            out(15, k, p(NV + 1, k));
            out(16, k, p(2, k));
            out(17, k, p(1, k));
        } else {                                                                       // :614-630
            out(3, k, (p(2, k) - fdx) + py);
            out(4, k, (p(5, k) - fz0) * fdz);
            out(5, k, (p(6, k) - 1.0f) * fdu);
            out(6, k, (p(7, k) - 1.0f) * fdv);
            out(7, k, p(8, k));
            out(8, k, p(3, k));
            out(9, k, p(4, k));
            out(10, k, real_sqrt(c(2, k)));
            out(11, k, real_sqrt(c(1, k)));
            out(12, k, real_sqrt(c(5, k) * fdz));
            out(13, k, real_sqrt(c(6, k)) * fdu);
            out(14, k, real_sqrt(c(7, k)) * fdv);
            out(15, k, real_sqrt(c(7, k)));                                           // as the GUI: CRLB column 7
            out(16, k, real_sqrt(c(3, k)));
            out(17, k, real_sqrt(c(4, k)));
            out(18, k, LL[k]);
            out(19, k, p(NV + 1, k));
        }
    }
    return BSF_OK;
}

int bsf_color_class(const float *w, int N, float *cls)
{
    bsf::clear_error();
    if (N < 0 || (N > 0 && (!w || !cls))) return bsf::fail(BSF_ERR_ARG, "bsf: bad arguments");
    for (int k = 0; k < N; k++) {
        float v = w[k];
        v = std::fmax(2.0f, v);                   // max(2, P(:,6)): NaN gives 2, as MATLAB's max
        v = std::fmin(3.0f, v);
        cls[k] = (v - 2.5f) * 2.0f;
    }
    return BSF_OK;
}

}  // extern "C"
