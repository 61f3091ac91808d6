// Riemann.h -- host mirror of the reference's Riemann class.
//
// Reference: Riemann(double *WR, double *WL, double *vFrame, double *Aij,
// int i) and void exact(double *Fij, const double &gamma) (Riemann.h:19,26).
// The call site passes (WijL, WijR, ...) (Particles.cpp:1852), so the
// constructor's WR parameter receives particle j's state and WL particle
// i's; exact() solves with left = WL (particle i) and right = WR.  States
// are {rho, P, v[dim]} in the frame moving with vFrame.
//
// The face solve runs on the device (mlh_riemann_flux_batch, the code of
// the face kernel).  A single exact() call is one device round trip; use
// RiemannBatch for many faces.  The reference rotates WR / WL in place in
// the constructor; this mirror leaves its inputs unchanged.
#pragma once
#include <vector>

#include "../../../include/mlh.h"

namespace mlh_host {

class Riemann {
public:
    // ctx supplies dim, gamma (checked in exact), the solver tolerance and MFM
    Riemann(mlh_ctx *ctx, int dim, const double *WR, const double *WL, const double *vFrame,
            const double *Aij, int i);
    // Fij = {mass, energy, momentum[dim]}; returns the sampled side flag
    int exact(double *Fij, double gamma);

private:
    mlh_ctx *ctx_;
    int dim_, i_;
    std::vector<double> wl_, wr_, vf_, a_;
};

// n faces at once: left = particle i (the reference's WL), right = j
struct RiemannBatch {
    static void exact(mlh_ctx *ctx, long long n, const double *w_left, const double *w_right,
                      const double *v_frame, const double *a_ij, double *f_ij, int *flag);
};

}  // namespace mlh_host
