// hydro.hip -- the FP64 instantiations of the MFV physics kernels, the
// kernels common to both precisions and the launchers.  The FP32
// instantiations are in hydro_fp32.hip (compiled with FMA contraction).
#include "hydro_impl.h"
#include "stages_impl.h"

namespace mlh {


// ------------------------------------------------------------------ launchers
#define MLH_DISPATCH(KERNEL, D, P, GRID, ...)                                             \
    do {                                                                                  \
        if (D == 2 && P == 0) hipLaunchKernelGGL((KERNEL<2, double>), GRID, dim3(256), 0, s, __VA_ARGS__); \
        else if (D == 2) hipLaunchKernelGGL((KERNEL<2, float>), GRID, dim3(256), 0, s, __VA_ARGS__);  \
        else if (P == 0) hipLaunchKernelGGL((KERNEL<3, double>), GRID, dim3(256), 0, s, __VA_ARGS__); \
        else hipLaunchKernelGGL((KERNEL<3, float>), GRID, dim3(256), 0, s, __VA_ARGS__);              \
    } while (0)

int dt_blocks(int N) { return (N + 255) / 256; }

void launch_density(const HydroArgs &a, int D, int precision, hipStream_t s) {
    dim3 grid((a.N + 255) / 256);
    if (precision != 0) launch_density_fp32(a, D, s);
    else if (D == 2) hipLaunchKernelGGL((k_density<2, double>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_density<3, double>), grid, dim3(256), 0, s, a);
}
void launch_gradients(const HydroArgs &a, int D, int precision, hipStream_t s) {
    dim3 grid((a.N + 255) / 256);
    if (precision != 0) launch_gradients_fp32(a, D, s);
    else if (D == 2) hipLaunchKernelGGL((k_gradients<2, double>), grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_gradients<3, double>), grid, dim3(256), 0, s, a);
}
void launch_dt_finalize(const HydroArgs &a, double dt_fixed, double t_now, double t_clip,
                        const double *dt_in, double *dt_hist, int step, hipStream_t s) {
    hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(256), 0, s, a.dt_block, dt_blocks(a.N),
                       dt_fixed, t_now, t_clip, dt_in, a.dt, dt_hist, step, a.status);
}
void launch_faces(const HydroArgs &a, int D, int precision, hipStream_t s) {
    dim3 grid((a.N + 255) / 256);
    if (precision != 0) {
        launch_faces_fp32(a, D, s);
    } else if (a.learned) {
        if (D == 2) hipLaunchKernelGGL((k_faces<2, double, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_faces<3, double, true>), grid, dim3(256), 0, s, a);
    } else {
        if (D == 2) hipLaunchKernelGGL((k_faces<2, double, false>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_faces<3, double, false>), grid, dim3(256), 0, s, a);
    }
}
void launch_riemann_flux(const HydroArgs &a, int D, long long n, const double *wl,
                         const double *wr, const double *vf, const double *aij, double *f,
                         int *flag, hipStream_t s) {
    dim3 grid((unsigned)((n + 255) / 256));
    if (D == 2)
        hipLaunchKernelGGL(k_riemann_flux<2>, grid, dim3(256), 0, s, a, n, wl, wr, vf, aij, f, flag);
    else
        hipLaunchKernelGGL(k_riemann_flux<3>, grid, dim3(256), 0, s, a, n, wl, wr, vf, aij, f, flag);
}

void launch_stage(const HydroArgs &a, int D, int precision, int op, int what,
                  const FaceArrays &fa, const double *f, double *grad, hipStream_t s) {
    if (precision != 0) {
        launch_stage_fp32(a, D, op, what, fa, f, grad, s);
        return;
    }
    dim3 grid((a.N + 255) / 256), blk(256);
    switch (op) {
    case kOpLimit:
        if (D == 2) hipLaunchKernelGGL((k_limit<2, double>), grid, blk, 0, s, a);
        else hipLaunchKernelGGL((k_limit<3, double>), grid, blk, 0, s, a);
        break;
    case kOpGradField:
        if (D == 2) hipLaunchKernelGGL((k_gradient_field<2, double>), grid, blk, 0, s, a, f, grad);
        else hipLaunchKernelGGL((k_gradient_field<3, double>), grid, blk, 0, s, a, f, grad);
        break;
    case kOpFaceStates:
        if (D == 2) hipLaunchKernelGGL((k_face_states<2, double>), grid, blk, 0, s, a, fa, what);
        else hipLaunchKernelGGL((k_face_states<3, double>), grid, blk, 0, s, a, fa, what);
        break;
    case kOpSolveOwn:
    case kOpSolveMirror: {
        const int pass = op == kOpSolveOwn ? 0 : 1;
        if (D == 2) hipLaunchKernelGGL((k_face_solve<2, double>), grid, blk, 0, s, a, fa, pass);
        else hipLaunchKernelGGL((k_face_solve<3, double>), grid, blk, 0, s, a, fa, pass);
        break;
    }
    case kOpCollect:
        if (D == 2) hipLaunchKernelGGL((k_collect<2, double>), grid, blk, 0, s, a, fa);
        else hipLaunchKernelGGL((k_collect<3, double>), grid, blk, 0, s, a, fa);
        break;
    }
}

void launch_flux_check(const HydroArgs &a, int D, const FaceArrays &fa, double tol,
                       unsigned long long *out, hipStream_t s) {
    dim3 grid((a.N + 255) / 256);
    if (D == 2) hipLaunchKernelGGL(k_flux_check<2>, grid, dim3(256), 0, s, a, fa, tol, out);
    else hipLaunchKernelGGL(k_flux_check<3>, grid, dim3(256), 0, s, a, fa, tol, out);
}

void launch_mirror_check(const HydroArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_mirror_check, dim3((a.N + 255) / 256), dim3(256), 0, s, a);
}

void launch_sums(const HydroArgs &a, int D, int precision, double *out, hipStream_t s) {
    MLH_DISPATCH(k_sums, D, precision, dim3((a.N + 255) / 256), a, out);
}

}  // namespace mlh
