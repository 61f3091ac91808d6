/* sanitize_driver.c -- runs every entry point of the CPU oracle on small
 * cases for the AddressSanitizer / UndefinedBehaviorSanitizer build of the
 * oracle (make -C oracle sanitize).  Test infrastructure, like the oracle
 * itself.  The cases: 2D periodic and open lattices with and without the
 * pairwise limiter, MFM and the limiter's abs() modes; a 3D periodic
 * lattice with the 3D predictor quirk; the exact Riemann solver on random
 * and vacuum states; the MLP forward / backward and features.  Exits 1 on
 * a failed check; the sanitizers abort on the first error they find. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mlh_oracle.h"

static int g_fail = 0;  /* failed checks so far */
#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            ++g_fail;                                                        \
        }                                                                    \
    } while (0)

/* xorshift64*: uniform doubles in [0, 1) */
static uint64_t g_rng = 0x9e3779b97f4a7c15ull;
static double urand(void) {
    g_rng ^= g_rng >> 12;
    g_rng ^= g_rng << 25;
    g_rng ^= g_rng >> 27;
    return (double)((g_rng * 2685821657736338717ull) >> 11) * 0x1.0p-53;
}

static mlh_params base_params(int dim, int periodic, double h) {
    mlh_params p;
    memset(&p, 0, sizeof p);
    p.dim = dim;
    p.periodic = periodic;
    for (int a = 0; a < 3; ++a) {
        p.box_min[a] = 0.0;
        p.box_max[a] = 1.0;
    }
    p.kernel_size = h;
    p.gamma = 5.0 / 3.0;
    p.cfl = 0.4;
    p.max_interactions = 128;
    p.max_ghost_interactions = 128;
    p.slope_limiting = 1;
    p.beta = 4.0;
    p.pairwise_limiter = 0;
    p.psi1 = 0.5;
    p.psi2 = 0.25;
    p.enforce_flux_sym = 1;
    p.move_particles = 1;
    p.first_order_quad_point = 1;
    p.debug_level = 1;
    p.limiter_abs = MLH_ABS_INT_X86;
    p.precision = MLH_FP64;
    p.riemann_seed = MLH_SEED_STANDARD;
    p.riemann_tol = 1e-6;
    p.riemann_max_iter = 50;
    p.learned_iters = 3;
    p.work_order = MLH_ORDER_CELL;
    return p;
}

/* n^dim jittered lattice in the unit box with a shear flow and a pressure
 * bump */
static void lattice(int dim, int n, double *x[3], double *v[3], double *m, double *u) {
    const int N = dim == 2 ? n * n : n * n * n;
    const double dx = 1.0 / n;
    for (int k = 0; k < N; ++k) {
        int c[3] = {k % n, (k / n) % n, dim == 3 ? k / (n * n) : 0};
        for (int a = 0; a < 3; ++a) {
            x[a][k] = a < dim ? (c[a] + 0.5 + 0.2 * (urand() - 0.5)) * dx : 0.0;
            v[a][k] = 0.0;
        }
        v[0][k] = x[1][k] < 0.5 ? 0.5 : -0.5;
        v[1][k] = 0.05 * sin(4.0 * M_PI * x[0][k]);
        if (dim == 3) v[2][k] = 0.05 * cos(2.0 * M_PI * x[1][k]);
        const double r2 = (x[0][k] - 0.5) * (x[0][k] - 0.5) + (x[1][k] - 0.5) * (x[1][k] - 0.5);
        m[k] = (dim == 2 ? dx * dx : dx * dx * dx) * (1.0 + 0.5 * (x[1][k] > 0.25 && x[1][k] < 0.75));
        u[k] = 1.5 * (1.0 + 2.0 * exp(-r2 / 0.01));
    }
}

static void run_case(const char *name, mlh_params p, int n) {
    const int fail0 = g_fail;
    const int D = p.dim, N = D == 2 ? n * n : n * n * n;
    const int S = p.max_interactions + p.max_ghost_interactions;
    double *x[3], *v[3], *m = malloc(sizeof(double) * N), *u = malloc(sizeof(double) * N);
    for (int a = 0; a < 3; ++a) {
        x[a] = malloc(sizeof(double) * N);
        v[a] = malloc(sizeof(double) * N);
    }
    lattice(D, n, x, v, m, u);
    orc *o = orc_create(&p, N);
    CHECK(o != NULL);
    CHECK(orc_set_state(o, x[0], x[1], x[2], v[0], v[1], v[2], m, u) == 0);
    double c0[6], c1[6];
    CHECK(orc_conserved_sums(o, c0) == 0);
    for (int s = 0; s < 3; ++s) {
        double dt = 0.0;
        CHECK(orc_step(o, -1.0, 1e-3, &dt) == 0);
        CHECK(dt > 0.0 && dt <= 1e-3);
    }
    CHECK(orc_conserved_sums(o, c1) == 0);
    CHECK(fabs(c1[0] - c0[0]) <= 1e-12 * fabs(c0[0]));
    /* the stages one at a time, then every read-out */
    double dt = 0.0;
    CHECK(orc_neighbours(o) == 0);
    CHECK(orc_density(o) == 0);
    CHECK(orc_timestep(o, &dt) == 0 && dt > 0.0);
    CHECK(orc_psi_gradients(o) == 0);
    CHECK(orc_limit(o) == 0);
    CHECK(orc_gradients(o) == 0);
    CHECK(orc_fluxes(o, dt) == 0);
    int64_t asym = -1, missing = -1;
    CHECK(orc_check_flux_symmetry(o, 1e-20, &asym, &missing) == 0);
    CHECK(asym == 0 && missing == 0);
    double *f = malloc(sizeof(double) * N), *gf = malloc(sizeof(double) * N * D);
    for (int k = 0; k < N; ++k) f[k] = x[0][k] * x[0][k];
    CHECK(orc_gradient_field(o, f, gf) == 0);
    double *A = malloc(sizeof(double) * N * S * D), *WR = malloc(sizeof(double) * N * S * (D + 2));
    double *WL = malloc(sizeof(double) * N * S * (D + 2)), *vF = malloc(sizeof(double) * N * S * D);
    double *F = malloc(sizeof(double) * N * S * (D + 2));
    CHECK(orc_get_face_arrays(o, dt, A, WR, WL, vF, F) == 0);
    CHECK(orc_get_face_arrays(o, dt, NULL, NULL, NULL, NULL, F) == 0);
    int32_t *noi = malloc(sizeof(int32_t) * N), *nog = malloc(sizeof(int32_t) * N);
    int32_t *nbr = malloc(sizeof(int32_t) * N * S);
    CHECK(orc_get_neighbours(o, noi, nog, nbr) == 0);
    const int cap = 7 * N;
    int32_t *par = malloc(sizeof(int32_t) * cap), *code = malloc(sizeof(int32_t) * cap), ng = 0;
    double *gx = malloc(sizeof(double) * cap), *gy = malloc(sizeof(double) * cap);
    double *gz = malloc(sizeof(double) * cap);
    if (p.periodic) {
        CHECK(orc_get_ghosts(o, cap, par, code, gx, gy, gz, &ng) == 0);
        CHECK(ng > 0 && ng <= cap);
        for (int k = 0; k < ng; ++k) CHECK(par[k] >= 0 && par[k] < N && code[k] >= 0 && code[k] < 27);
        CHECK(orc_get_ghosts(o, 1, par, code, gx, gy, gz, &ng) == MLH_ERR_INVALID_ARGUMENT);
    } else {
        CHECK(orc_get_ghosts(o, cap, par, code, gx, gy, gz, &ng) == MLH_ERR_CALL_ORDER);
    }
    /* the brute-force ghost search, then the stages again up to the fluxes */
    orc_set_bruteforce_ghosts(o, 1);
    CHECK(orc_neighbours(o) == 0);
    orc_set_bruteforce_ghosts(o, 0);
    CHECK(orc_density(o) == 0);
    CHECK(orc_timestep(o, &dt) == 0 && dt > 0.0);
    CHECK(orc_gradients(o) == 0);
    CHECK(orc_fluxes(o, dt) == 0);
    double *rho = malloc(sizeof(double) * N), *P = malloc(sizeof(double) * N);
    double *om = malloc(sizeof(double) * N);
    CHECK(orc_get_density(o, rho, P, om) == 0);
    for (int k = 0; k < N; ++k) CHECK(rho[k] > 0.0 && P[k] > 0.0 && isfinite(om[k]));
    double *grads = malloc(sizeof(double) * (D + 2) * N * D), *B = malloc(sizeof(double) * N * D * D);
    CHECK(orc_get_gradients(o, grads, B) == 0);
    double *mF = malloc(sizeof(double) * N), *vFl = malloc(sizeof(double) * N * D);
    double *eF = malloc(sizeof(double) * N);
    CHECK(orc_get_fluxes(o, mF, vFl, eF) == 0);
    double msum = 0.0;
    for (int k = 0; k < N; ++k) msum += mF[k];
    CHECK(fabs(msum) < 1e-10);
    CHECK(orc_update(o, dt) == 0);
    mlh_diag d;
    CHECK(orc_get_diag(o, &d) == 0);
    CHECK(d.faces_solved > 0 && d.max_neighbours > 0);
    /* one pair through the predictor test hook, and the batch flux */
    const int j0 = (int)((uint32_t)nbr[0] & ((1u << 27) - 1u));
    CHECK(noi[0] > 0 && j0 < N);
    double wr[5], wl[5], vfr[3], xj[3] = {x[0][j0], x[1][j0], x[2][j0]};
    CHECK(orc_test_face_states(o, 0, j0, xj, 0, dt, rho, P, grads, wr, wl, vfr) == 0);
    double wl1[5] = {1.0, 1.0, 0.1, 0.0, 0.0}, wr1[5] = {0.125, 0.1, 0.0, 0.0, 0.0};
    double vf1[3] = {0.0, 0.0, 0.0}, a1[3] = {1e-3, 0.0, 0.0}, f1[5];
    int32_t fl1 = 0;
    orc_riemann_flux_batch(o, 1, wl1, wr1, vf1, a1, f1, &fl1);
    CHECK(isfinite(f1[0]));
    CHECK(orc_get_state(o, x[0], x[1], x[2], v[0], v[1], v[2], m, u) == 0);
    orc_destroy(o);
    void *bufs[] = {f, gf, A, WR, WL, vF, F, noi, nog, nbr, par, code, gx, gy, gz, rho, P, om,
                    grads, B, mF, vFl, eF, m, u, x[0], x[1], x[2], v[0], v[1], v[2]};
    for (size_t k = 0; k < sizeof bufs / sizeof bufs[0]; ++k) free(bufs[k]);
    fprintf(stderr, "[sanitize] %s: N=%d, %d ghosts, %s\n", name, N, ng,
            g_fail > fail0 ? "FAILED" : "ok");
}

static void riemann_cases(void) {
    const int fail0 = g_fail;
    enum { n = 2000 };
    static double rl[n], ul[n], pl[n], rr[n], ur[n], pr[n], rs[n], us[n], ps[n];
    static int32_t fl[n], it[n];
    for (int k = 0; k < n; ++k) {
        rl[k] = pow(10.0, 6.0 * urand() - 3.0);
        rr[k] = pow(10.0, 6.0 * urand() - 3.0);
        pl[k] = pow(10.0, 9.0 * urand() - 4.5);
        pr[k] = pow(10.0, 9.0 * urand() - 4.5);
        ul[k] = 4.0 * (urand() - 0.5);
        ur[k] = 4.0 * (urand() - 0.5);
    }
    rl[0] = rr[0] = pl[0] = pr[0] = 1.0; /* vacuum generation */
    ul[0] = -100.0;
    ur[0] = 100.0;
    orc_riemann_batch(n, 5.0 / 3.0, rl, ul, pl, rr, ur, pr, 1e-12, 50, rs, us, ps, fl, it);
    CHECK(fl[0] == 0);
    for (int k = 1; k < n; ++k) CHECK(isfinite(ps[k]) && ps[k] >= 0.0);
    int iters = 0;
    const double p1 = orc_riemann_pstar(1.4, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1e-12, 50, &iters);
    CHECK(fabs(p1 - 0.30313) < 1e-4); /* Toro's test 1 */
    CHECK(orc_riemann_guess(1.4, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1) > 0.0);
    double fv = 0.0, dfv = 0.0;
    orc_riemann_f(1.4, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, p1, &fv, &dfv);
    CHECK(fabs(fv) < 1e-9 && dfv > 0.0);
    double r1, u1, q1;
    int i1;
    orc_riemann_solve(1.4, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1e-12, 50, &r1, &u1, &q1, &i1);
    CHECK(isfinite(r1) && isfinite(u1) && isfinite(q1));
    fprintf(stderr, "[sanitize] exact Riemann solver: %s\n", g_fail > fail0 ? "FAILED" : "ok");
}

static void mlp_cases(void) {
    const int fail0 = g_fail;
    enum { n = 300, np = MLH_MLP_NPARAM };
    static float params[np], feats[n * MLH_MLP_IN], target[n];
    static double grad[np], out[n], pref[n];
    static double rl[n], ul[n], pl[n], rr[n], ur[n], pr[n];
    for (int k = 0; k < np; ++k) params[k] = (float)(0.3 * (urand() - 0.5));
    for (int k = 0; k < n * MLH_MLP_IN; ++k) feats[k] = (float)(urand() - 0.5);
    for (int k = 0; k < n; ++k) target[k] = (float)(urand() - 0.5);
    const double loss = orc_mlp_loss_grad(n, params, feats, target, grad);
    CHECK(loss > 0.0 && isfinite(loss));
    orc_mlp_forward(n, params, feats, out);
    for (int k = 0; k < n; ++k) {
        CHECK(isfinite(out[k]));
        rl[k] = 0.5 + urand();
        rr[k] = 0.5 + urand();
        pl[k] = 0.5 + urand();
        pr[k] = 0.5 + urand();
        ul[k] = urand() - 0.5;
        ur[k] = urand() - 0.5;
    }
    orc_mlp_features(n, 5.0 / 3.0, rl, ul, pl, rr, ur, pr, feats, pref);
    for (int k = 0; k < n; ++k) CHECK(pref[k] > 0.0);
    fprintf(stderr, "[sanitize] MLP: %s\n", g_fail > fail0 ? "FAILED" : "ok");
}

int main(void) {
    orc_set_threads(1);
    mlh_params p = base_params(2, 1, 0.12);
    run_case("2D periodic", p, 20);
    p.pairwise_limiter = 1;
    p.limiter_abs = MLH_ABS_FABS;
    p.mfm = 1;
    run_case("2D periodic, pairwise limiter, MFM", p, 20);
    p = base_params(2, 0, 0.12);
    p.move_particles = 0;
    run_case("2D open box, fixed particles", p, 16);
    p = base_params(3, 1, 0.3);
    p.quirk_3d_predictor = 1;
    p.pairwise_limiter = 1;
    run_case("3D periodic, pairwise limiter, quirk", p, 8);
    CHECK(orc_kernel(2, 0.01, 0.1) > 0.0 && orc_kernel(3, 0.2, 0.1) == 0.0);
    CHECK(isfinite(orc_test_pairwise_limiter(&p, 1.0, 1.2, 0.8, 0.5, 1.0)));
    riemann_cases();
    mlp_cases();
    return g_fail ? 1 : 0;
}
