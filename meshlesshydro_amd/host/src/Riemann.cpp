// Riemann.cpp -- host mirror of Riemann over the device flux batch.
#include "Riemann.h"

#include "Particles.h"

namespace mlh_host {

Riemann::Riemann(mlh_ctx *ctx, int dim, const double *WR, const double *WL, const double *vFrame,
                 const double *Aij, int i)
    : ctx_(ctx), dim_(dim), i_(i), wl_(WL, WL + dim + 2), wr_(WR, WR + dim + 2),
      vf_(vFrame, vFrame + dim), a_(Aij, Aij + dim) {}

int Riemann::exact(double *Fij, double gamma) {
    (void)gamma;  // the context's gamma is used (set at mlh_create)
    int32_t flag = 0;
    check(mlh_riemann_flux_batch(ctx_, 1, wl_.data(), wr_.data(), vf_.data(), a_.data(), Fij,
                                 &flag),
          "Riemann::exact");
    return flag;
}

void RiemannBatch::exact(mlh_ctx *ctx, long long n, const double *w_left, const double *w_right,
                         const double *v_frame, const double *a_ij, double *f_ij, int *flag) {
    check(mlh_riemann_flux_batch(ctx, n, w_left, w_right, v_frame, a_ij, f_ij, flag),
          "RiemannBatch::exact");
}

}  // namespace mlh_host
