/*
 * mlh_oracle.h -- CPU FP64 oracle of the meshless finite-volume (MFV) step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it.  It is a restatement, in plain C, of the reference's
 * algorithm (jammartin/meshlessHydro, demonstrator/src/Particles.cpp,
 * Riemann.cpp, Helper.cpp, Domain.cpp, MeshlessScheme.cpp), with the
 * reference's data structures kept where they decide results: explicit
 * ghost particles in ghost-index order, a 3^D stencil grid search with
 * non-periodic cell indices, a reverse-index lookup per pair, each pair
 * solved once from the lower index and negated for the other side.
 */
#ifndef MLH_ORACLE_H
#define MLH_ORACLE_H

#include <stdint.h>
#include "../include/mlh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

orc *orc_create(const mlh_params *p, int32_t n);
void orc_destroy(orc *o);
int orc_set_state(orc *o, const double *x, const double *y, const double *z,
                  const double *vx, const double *vy, const double *vz,
                  const double *m, const double *u);
int orc_get_state(orc *o, double *x, double *y, double *z, double *vx,
                  double *vy, double *vz, double *m, double *u);

/* stages (same grouping as the C ABI) */
int orc_neighbours(orc *o);
int orc_density(orc *o);
int orc_timestep(orc *o, double *dt);
int orc_gradients(orc *o);
int orc_fluxes(orc *o, double dt);
int orc_update(orc *o, double dt);
/* full step: dt_fixed > 0 uses it, else CFL dt clipped to dt_max (> 0) */
int orc_step(orc *o, double dt_fixed, double dt_max, double *dt_used);

/* the gradient stage split as the reference's methods: compPsijTilde +
 * gradient (unlimited), slopeLimiter, and gradient(f) of any field */
int orc_psi_gradients(orc *o);
int orc_limit(orc *o);
int orc_gradient_field(orc *o, const double *f, double *grad);
/* per-pair arrays of the face stages in the C ABI's row layout
 * [N][maxr + maxg][components] (after orc_fluxes with the same dt) */
int orc_get_face_arrays(orc *o, double dt, double *A, double *WR, double *WL, double *vF,
                        double *F);
/* the ghost particles of the last neighbour search, in ghost-index order */
int orc_get_ghosts(orc *o, int32_t cap, int32_t *parent, int32_t *code, double *gx, double *gy,
                   double *gz, int32_t *n);
/* checkFluxSymmetry (Particles.cpp:2888-2976) on the stored F_ij */
int orc_check_flux_symmetry(orc *o, double tol, int64_t *asym, int64_t *missing);

int orc_get_neighbours(orc *o, int32_t *noi, int32_t *noi_ghost,
                       int32_t *nbr);
int orc_get_density(orc *o, double *rho, double *P, double *omega);
int orc_get_gradients(orc *o, double *grads, double *B);
int orc_get_fluxes(orc *o, double *mF, double *vF, double *eF);
int orc_conserved_sums(orc *o, double *out);
int orc_get_diag(orc *o, mlh_diag *d);
/* brute-force (O(N * N_ghost)) ghost search instead of the binned one */
void orc_set_bruteforce_ghosts(orc *o, int on);
void orc_riemann_flux_batch(orc *o, int64_t n, const double *w_left, const double *w_right,
                            const double *v_frame, const double *a_ij, double *f_ij,
                            int32_t *flag);
void orc_set_threads(int n);

/* exact Riemann solver (Toro, "Riemann Solvers and Numerical Methods for
 * Fluid Dynamics", 3rd ed., ch. 4): Newton-Raphson on
 * f(p) = f_L(p) + f_R(p) + (u_R - u_L) from the adaptive PVRS/TRRS/TSRS
 * guess, sampled at x/t = 0.  Returns the flag (-1 left, +1 right,
 * 0 vacuum) and the iteration count in *iters. */
int orc_riemann_solve(double gamma, double rhoL, double uL, double PL,
                      double rhoR, double uR, double PR, double tol,
                      int max_iter, double *rho_s, double *u_s, double *P_s,
                      int *iters);
/* batch form for ctypes */
void orc_riemann_batch(int64_t n, double gamma, const double *rhoL,
                       const double *uL, const double *PL, const double *rhoR,
                       const double *uR, const double *PR, double tol,
                       int max_iter, double *rho_s, double *u_s, double *P_s,
                       int32_t *flag, int32_t *iters);
/* the initial guess alone (for the learned-seed comparison) */
double orc_riemann_pstar(double gamma, double dl, double ul, double pl, double dr, double ur,
                         double pr, double tol, int max_iter, int *iters);
double orc_riemann_guess(double gamma, double rhoL, double uL, double PL,
                         double rhoR, double uR, double PR);
/* residual f(p) and f'(p) */
void orc_riemann_f(double gamma, double rhoL, double uL, double PL,
                   double rhoR, double uR, double PR, double p, double *f,
                   double *df);

/* Test access (the oracle's own pins): compRiemannStatesLR for one pair
 * (i, j) with given rho, P [N] and gradients [(D+2)][N][D]; WR / WL
 * [D+2] {rho, P, v...}, vFrame [D].  And pairwiseLimiter alone. */
int orc_test_face_states(orc *o, int i, int j, const double *xj, int is_ghost, double dt,
                         const double *rho, const double *P, const double *grads,
                         double *WR, double *WL, double *vFrame);
double orc_test_pairwise_limiter(const mlh_params *p, double phi0, double phi_i, double phi_j,
                                 double xijxi_abs, double xjxi_abs);

/* cubic spline kernel W(r, h) (Particles.cpp:7-24) */
double orc_kernel(int dim, double r, double h);

/* MLP 10-64-64-1, tanh hidden layers, FP64 forward/backward (loss = mean
 * squared error); params layout as mlh_mlp_set_params. */
double orc_mlp_loss_grad(int64_t n, const float *params, const float *feats,
                         const float *target, double *grad);
void orc_mlp_forward(int64_t n, const float *params, const float *feats,
                     double *out);
void orc_mlp_features(int64_t n, double gamma, const double *rhoL,
                      const double *uL, const double *PL, const double *rhoR,
                      const double *uR, const double *PR, float *feats,
                      double *p_ref);

#ifdef __cplusplus
}
#endif
#endif
