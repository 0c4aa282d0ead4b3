// bsf_fit_hetero.hip -- one launch over the ROIs of several fit launches of
// fittypes 6-8 (a frame's models): each workgroup looks up its entry and
// runs that entry's fittype of the fit kernel (fit_roi in bsf_fit_kernel.h)
// on its place within the entry, so every ROI is fitted by the same code as
// in a launch of its own.
#include "bsf_fit_kernel.h"

#include <algorithm>
#include <cstring>

namespace bsf {

struct DevHet {
    Dev d[kHeteroMax];
    int ft[kHeteroMax];
    int first[kHeteroMax + 1];   // entry k: workgroups [first[k], first[k + 1])
    int n;
};

template <typename E>
__global__ __launch_bounds__(64) void fit_hetero_kernel(DevHet H)
{
    const int b = (int)blockIdx.x;
    int k = 0;
    while (k + 1 < H.n && b >= H.first[k + 1]) k++;
    const int lb = b - H.first[k], G = H.first[k + 1] - H.first[k];
    if (H.ft[k] == 6) fit_roi<6, E, 0, 1>(H.d[k], lb, G);
    else if (H.ft[k] == 7) fit_roi<7, E, 0, 1>(H.d[k], lb, G);
    else fit_roi<8, E, 0, 1>(H.d[k], lb, G);
}

hipError_t launch_fit_hetero(const FitLaunch *L, int n, hipStream_t st)
{
    if (n < 1 || n > kHeteroMax) return hipErrorInvalidValue;
    DevHet H;
    std::memset(&H, 0, sizeof H);
    H.n = n;
    size_t lds = 0;
    int blocks = 0;
    for (int k = 0; k < n; k++) {
        if (L[k].fittype < 6 || L[k].fittype > 8 || L[k].f64 != L[0].f64 || L[k].theta_in || L[k].count < 0)
            return hipErrorInvalidValue;
        H.d[k] = fit_dev(L[k]);
        H.ft[k] = L[k].fittype;
        H.first[k] = blocks;
        blocks += L[k].count;
        lds = std::max(lds, lds_bytes(L[k].fittype, L[k].f64, L[k].S));
    }
    H.first[n] = blocks;
    if (blocks == 0) return hipSuccess;
    if (L[0].f64)
        hipLaunchKernelGGL((fit_hetero_kernel<double>), dim3(blocks), dim3(64), lds, st, H);
    else
        hipLaunchKernelGGL((fit_hetero_kernel<float>), dim3(blocks), dim3(64), lds, st, H);
    return hipGetLastError();
}

}  // namespace bsf
