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
#include "stages_impl.h"

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
        // faces whose standard guess was far off are solved after the row
        // from the network's seed, inside the same kernel
        if (D == 2) hipLaunchKernelGGL((k_faces<2, float, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_faces<3, float, true>), grid, dim3(256), 0, s, a);
    } else {
        if (D == 2) hipLaunchKernelGGL((k_faces<2, float, false>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_faces<3, float, false>), grid, dim3(256), 0, s, a);
    }
}

void launch_stage_fp32(const HydroArgs &a, int D, int op, int what, const FaceArrays &fa,
                       const double *f, double *grad, hipStream_t s) {
    dim3 grid((a.N + 255) / 256), blk(256);
    switch (op) {
    case kOpLimit:
        if (D == 2) hipLaunchKernelGGL((k_limit<2, float>), grid, blk, 0, s, a);
        else hipLaunchKernelGGL((k_limit<3, float>), grid, blk, 0, s, a);
        break;
    case kOpGradField:
        if (D == 2) hipLaunchKernelGGL((k_gradient_field<2, float>), grid, blk, 0, s, a, f, grad);
        else hipLaunchKernelGGL((k_gradient_field<3, float>), grid, blk, 0, s, a, f, grad);
        break;
    case kOpFaceStates:
        if (D == 2) hipLaunchKernelGGL((k_face_states<2, float>), grid, blk, 0, s, a, fa, what);
        else hipLaunchKernelGGL((k_face_states<3, float>), grid, blk, 0, s, a, fa, what);
        break;
    case kOpSolveOwn:
    case kOpSolveMirror: {
        const int pass = op == kOpSolveOwn ? 0 : 1;
        if (D == 2) hipLaunchKernelGGL((k_face_solve<2, float>), grid, blk, 0, s, a, fa, pass);
        else hipLaunchKernelGGL((k_face_solve<3, float>), grid, blk, 0, s, a, fa, pass);
        break;
    }
    case kOpCollect:
        if (D == 2) hipLaunchKernelGGL((k_collect<2, float>), grid, blk, 0, s, a, fa);
        else hipLaunchKernelGGL((k_collect<3, float>), grid, blk, 0, s, a, fa);
        break;
    }
}

}  // namespace mlh
