// hydro_fp32.hip -- the FP32 instantiations of the MFV physics kernels
// (pair arithmetic in FP32; FP64 for E_i and its inverse, the limiter
// extrema, the Newton residual, the flux sums and the state), compiled with
// FMA contraction within expressions.
#define MLH_FP32_TU
// before the includes: the Riemann solver and the helpers of the headers
// are compiled with contraction here too (the FP64 units include them
// without it)
#pragma clang fp contract(on)
#include "hydro_impl.h"

namespace mlh {

void launch_density_fp32(const HydroArgs &a, int D, hipStream_t s) {
    dim3 grid((a.N + 255) / 256);
    if (D == 2) hipLaunchKernelGGL((k_density<2, float>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_density<3, float>), grid, dim3(256), 0, s, a);
}

void launch_gradients_fp32(const HydroArgs &a, int D, hipStream_t s) {
    dim3 grid((a.N + 255) / 256);
    if (D == 2) hipLaunchKernelGGL((k_gradients<2, float>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_gradients<3, float>), grid, dim3(256), 0, s, a);
}

void launch_faces_fp32(const HydroArgs &a, int D, hipStream_t s) {
    dim3 grid((a.N + 255) / 256);
    if (a.learned) {
        // the face loop defers the faces whose standard guess was far off to
        // k_faces_deferred (a small grid-stride launch over the records)
        hipLaunchKernelGGL(k_defer_reset, dim3(1), dim3(1), 0, s, a.defer_n);
        if (D == 2) hipLaunchKernelGGL((k_faces<2, float, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_faces<3, float, true>), grid, dim3(256), 0, s, a);
        dim3 dgrid(grid.x < 1024 ? grid.x : 1024);
        if (D == 2) hipLaunchKernelGGL((k_faces_deferred<2>), dgrid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_faces_deferred<3>), dgrid, dim3(256), 0, s, a);
    } else {
        if (D == 2) hipLaunchKernelGGL((k_faces<2, float, false>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_faces<3, float, false>), grid, dim3(256), 0, s, a);
    }
}

}  // namespace mlh
