/*
 * orc_riemann.c -- oracle exact Riemann solver for the ideal gas.
 *
 * TEST INFRASTRUCTURE ONLY (see mlh_oracle.h).
 *
 * The reference calls an un-vendored RiemannSolver::solve (fetched by
 * demonstrator/Makefile:94 from bwvdnbro/python_finite_volume_solver@master,
 * no commit pinned) with the contract of src/Riemann.cpp:84,93-94:
 *   solve(rhoL, uL, PL, rhoR, uR, PR, rho*, u*, P*) -> flag
 * where flag = -1 when the left side of the contact is sampled at x/t = 0,
 * +1 for the right side, 0 for vacuum (meaning fixed by the branches at
 * src/Riemann.cpp:104-127).  Its internals are not available here, so this
 * file restates the published exact solver of Toro, "Riemann Solvers and
 * Numerical Methods for Fluid Dynamics" (3rd ed., 2009), ch. 4:
 *   - pressure function f_K(p) and derivative, eqs. (4.6)-(4.7), (4.37);
 *   - adaptive initial guess (PVRS if p_max/p_min <= 2 and p_min <= p_pv <=
 *     p_max, else TRRS if p_pv < p_min, else TSRS), §4.3.2 / §9.5.1;
 *   - Newton-Raphson with the relative-change test
 *     2|p_k+1 - p_k| / (p_k+1 + p_k) < tol, negative iterates reset to tol;
 *   - sampling at S = x/t = 0, §4.5 (left of contact when S <= u*);
 *   - vacuum: §4.6 (vacuum right / left / generated).
 * Parity against the reference's own solver is unpinned (no golden data
 * exists for it); the checks are Toro's published tests 1-5 and the
 * invariants in tests/test_oracle_riemann.py.
 */
#include <math.h>
#include <stdint.h>
#include "mlh_oracle.h"

typedef struct {
    double g, g1, g2, g3, g4, g5, g6, g7, g8;
} gconst;

static gconst mk(double g) {
    gconst c;
    c.g = g;
    c.g1 = (g - 1.0) / (2.0 * g);
    c.g2 = (g + 1.0) / (2.0 * g);
    c.g3 = 2.0 * g / (g - 1.0);
    c.g4 = 2.0 / (g - 1.0);
    c.g5 = 2.0 / (g + 1.0);
    c.g6 = (g - 1.0) / (g + 1.0);
    c.g7 = (g - 1.0) / 2.0;
    c.g8 = g - 1.0;
    return c;
}

/* Toro (4.6)/(4.7) and derivatives (4.37) for one side K. */
static void prefun(const gconst *c, double p, double rk, double pk, double ck,
                   double *f, double *fd) {
    if (p <= pk) { /* rarefaction */
        double prat = p / pk;
        *f = c->g4 * ck * (pow(prat, c->g1) - 1.0);
        *fd = (1.0 / (rk * ck)) * pow(prat, -c->g2);
    } else { /* shock */
        double ak = c->g5 / rk;
        double bk = c->g6 * pk;
        double qrt = sqrt(ak / (bk + p));
        *f = (p - pk) * qrt;
        *fd = (1.0 - 0.5 * (p - pk) / (bk + p)) * qrt;
    }
}

static double guess(const gconst *c, double dl, double ul, double pl,
                    double cl, double dr, double ur, double pr, double cr,
                    double tol) {
    double cup = 0.25 * (dl + dr) * (cl + cr);
    double ppv = 0.5 * (pl + pr) + 0.5 * (ul - ur) * cup;
    if (ppv < 0.0) ppv = 0.0;
    double pmin = pl < pr ? pl : pr;
    double pmax = pl > pr ? pl : pr;
    double qmax = pmax / pmin;
    double pm;
    if (qmax <= 2.0 && pmin <= ppv && ppv <= pmax) {
        pm = ppv; /* PVRS */
    } else if (ppv < pmin) { /* two-rarefaction approximation (TRRS) */
        double pq = pow(pl / pr, c->g1);
        double um = (pq * ul / cl + ur / cr + c->g4 * (pq - 1.0)) /
                    (pq / cl + 1.0 / cr);
        double ptl = 1.0 + c->g7 * (ul - um) / cl;
        double ptr = 1.0 + c->g7 * (um - ur) / cr;
        pm = 0.5 * (pl * pow(ptl, c->g3) + pr * pow(ptr, c->g3));
    } else { /* two-shock approximation (TSRS) */
        double p0 = ppv > tol ? ppv : tol;
        double gel = sqrt((c->g5 / dl) / (c->g6 * pl + p0));
        double ger = sqrt((c->g5 / dr) / (c->g6 * pr + p0));
        pm = (gel * pl + ger * pr - (ur - ul)) / (gel + ger);
    }
    if (!(pm > tol)) pm = tol; /* also catches NaN from a failed TRRS base */
    return pm;
}

double orc_riemann_guess(double gamma, double rhoL, double uL, double PL,
                         double rhoR, double uR, double PR) {
    gconst c = mk(gamma);
    double cl = sqrt(gamma * PL / rhoL), cr = sqrt(gamma * PR / rhoR);
    return guess(&c, rhoL, uL, PL, cl, rhoR, uR, PR, cr, 1e-6 * 0.5 * (PL + PR));
}

void orc_riemann_f(double gamma, double rhoL, double uL, double PL,
                   double rhoR, double uR, double PR, double p, double *f,
                   double *df) {
    gconst c = mk(gamma);
    double cl = sqrt(gamma * PL / rhoL), cr = sqrt(gamma * PR / rhoR);
    double fl, fld, fr, frd;
    prefun(&c, p, rhoL, PL, cl, &fl, &fld);
    prefun(&c, p, rhoR, PR, cr, &fr, &frd);
    *f = fl + fr + (uR - uL);
    *df = fld + frd;
}

/* Sampling of a vacuum problem at S = 0 (Toro §4.6). */
static int sample_vacuum(const gconst *c, double dl, double ul, double pl,
                         double cl, double dr, double ur, double pr, double cr,
                         double *d, double *u, double *p) {
    const double s = 0.0;
    int left_vac = !(dl > 0.0);
    int right_vac = !(dr > 0.0);
    if (left_vac && right_vac) {
        *d = 0.0; *u = 0.0; *p = 0.0;
        return 0;
    }
    if (right_vac) { /* vacuum on the right, (4.76)-(4.77) */
        double shl = ul - cl;
        double stl = ul + c->g4 * cl;
        if (s <= shl) { *d = dl; *u = ul; *p = pl; return -1; }
        if (s < stl) {
            double cc = c->g5 * (cl + c->g7 * (ul - s));
            *u = c->g5 * (cl + c->g7 * ul + s);
            *d = dl * pow(cc / cl, c->g4);
            *p = pl * pow(cc / cl, c->g3);
            return -1;
        }
        *d = 0.0; *u = 0.0; *p = 0.0;
        return 0;
    }
    if (left_vac) { /* vacuum on the left, (4.79)-(4.80) */
        double shr = ur + cr;
        double str = ur - c->g4 * cr;
        if (s >= shr) { *d = dr; *u = ur; *p = pr; return 1; }
        if (s > str) {
            double cc = c->g5 * (cr - c->g7 * (ur - s));
            *u = c->g5 * (-cr + c->g7 * ur + s);
            *d = dr * pow(cc / cr, c->g4);
            *p = pr * pow(cc / cr, c->g3);
            return 1;
        }
        *d = 0.0; *u = 0.0; *p = 0.0;
        return 0;
    }
    /* vacuum generated between two rarefactions, (4.82)-(4.84) */
    double shl = ul - cl, sstl = ul + c->g4 * cl;
    double shr = ur + cr, sstr = ur - c->g4 * cr;
    if (s <= shl) { *d = dl; *u = ul; *p = pl; return -1; }
    if (s < sstl) {
        double cc = c->g5 * (cl + c->g7 * (ul - s));
        *u = c->g5 * (cl + c->g7 * ul + s);
        *d = dl * pow(cc / cl, c->g4);
        *p = pl * pow(cc / cl, c->g3);
        return -1;
    }
    if (s >= shr) { *d = dr; *u = ur; *p = pr; return 1; }
    if (s > sstr) {
        double cc = c->g5 * (cr - c->g7 * (ur - s));
        *u = c->g5 * (-cr + c->g7 * ur + s);
        *d = dr * pow(cc / cr, c->g4);
        *p = pr * pow(cc / cr, c->g3);
        return 1;
    }
    *d = 0.0; *u = 0.0; *p = 0.0;
    return 0;
}

/* Newton missed the change test within max_iter (near-vacuum states, where
 * f has an infinite slope at p = 0): f(p) = fL + fR + du increases with p
 * and f(0) = du - g4 (cL + cR) < 0 without vacuum, so bisect [0, hi] to the
 * same relative change test.  *nb: halvings done. */
static double bisect_pstar(const gconst *c, double dl, double pl, double cl, double dr,
                           double pr, double cr, double du, double tol, int *nb) {
    double lo = 0.0, hi = pl > pr ? pl : pr, fl, fld, fr, frd;
    for (int k = 0; k < 200; ++k) {
        prefun(c, hi, dl, pl, cl, &fl, &fld);
        prefun(c, hi, dr, pr, cr, &fr, &frd);
        if (fl + fr + du >= 0.0) break;
        lo = hi;
        hi *= 2.0;
    }
    *nb = 0;
    while (*nb < 200) {
        ++*nb;
        double mid = 0.5 * (lo + hi);
        prefun(c, mid, dl, pl, cl, &fl, &fld);
        prefun(c, mid, dr, pr, cr, &fr, &frd);
        if (fl + fr + du < 0.0) lo = mid;
        else hi = mid;
        if (2.0 * (hi - lo) <= tol * (hi + lo)) break;
    }
    return 0.5 * (lo + hi);
}

/* star pressure of the exact solver (0 for vacuum states), for tests */
double orc_riemann_pstar(double gamma, double dl, double ul, double pl, double dr, double ur,
                         double pr, double tol, int max_iter, int *iters) {
    gconst c = mk(gamma);
    *iters = 0;
    if (!(dl > 0.0) || !(dr > 0.0)) return 0.0;
    double cl = sqrt(gamma * pl / dl), cr = sqrt(gamma * pr / dr), du = ur - ul;
    if (c.g4 * (cl + cr) <= du) return 0.0;
    double tiny = tol * 0.5 * (pl + pr);
    double pold = guess(&c, dl, ul, pl, cl, dr, ur, pr, cr, tiny);
    double p = pold, fl, fld, fr, frd;
    int it;
    for (it = 1; it <= max_iter; ++it) {
        prefun(&c, pold, dl, pl, cl, &fl, &fld);
        prefun(&c, pold, dr, pr, cr, &fr, &frd);
        p = pold - (fl + fr + du) / (fld + frd);
        double change = 2.0 * fabs((p - pold) / (p + pold));
        if (change <= tol) break;
        if (p < 0.0) p = tiny;
        pold = p;
    }
    *iters = it;
    if (it > max_iter) {
        int nb = 0;
        p = bisect_pstar(&c, dl, pl, cl, dr, pr, cr, du, tol, &nb);
        *iters = max_iter + 1 + nb;
    }
    return p;
}

int orc_riemann_solve(double gamma, double dl, double ul, double pl,
                      double dr, double ur, double pr, double tol,
                      int max_iter, double *d_s, double *u_s, double *p_s,
                      int *iters) {
    gconst c = mk(gamma);
    *iters = 0;
    if (!(dl > 0.0) || !(dr > 0.0)) {
        double cl = dl > 0.0 ? sqrt(gamma * pl / dl) : 0.0;
        double cr = dr > 0.0 ? sqrt(gamma * pr / dr) : 0.0;
        return sample_vacuum(&c, dl, ul, pl, cl, dr, ur, pr, cr, d_s, u_s, p_s);
    }
    double cl = sqrt(gamma * pl / dl);
    double cr = sqrt(gamma * pr / dr);
    double du = ur - ul;
    if (c.g4 * (cl + cr) <= du) /* pressure positivity condition (4.40) */
        return sample_vacuum(&c, dl, ul, pl, cl, dr, ur, pr, cr, d_s, u_s, p_s);

    double tiny = tol * 0.5 * (pl + pr);
    double pold = guess(&c, dl, ul, pl, cl, dr, ur, pr, cr, tiny);
    double p = pold, fl, fld, fr, frd;
    int it;
    for (it = 1; it <= max_iter; ++it) {
        prefun(&c, pold, dl, pl, cl, &fl, &fld);
        prefun(&c, pold, dr, pr, cr, &fr, &frd);
        p = pold - (fl + fr + du) / (fld + frd);
        double change = 2.0 * fabs((p - pold) / (p + pold));
        if (change <= tol) break;
        if (p < 0.0) p = tiny;
        pold = p;
    }
    *iters = it;
    if (it > max_iter) {
        int nb = 0;
        p = bisect_pstar(&c, dl, pl, cl, dr, pr, cr, du, tol, &nb);
        *iters = max_iter + 1 + nb;
    }
    prefun(&c, p, dl, pl, cl, &fl, &fld);
    prefun(&c, p, dr, pr, cr, &fr, &frd);
    double um = 0.5 * (ul + ur + fr - fl);
    double pm = p;

    /* sample at S = 0 (Toro §4.5, routine SAMPLE) */
    const double s = 0.0;
    if (s <= um) { /* left of the contact */
        if (pm <= pl) { /* left rarefaction */
            double shl = ul - cl;
            if (s <= shl) { *d_s = dl; *u_s = ul; *p_s = pl; return -1; }
            double cml = cl * pow(pm / pl, c.g1);
            double stl = um - cml;
            if (s > stl) {
                *d_s = dl * pow(pm / pl, 1.0 / gamma);
                *u_s = um; *p_s = pm;
                return -1;
            }
            double cc = c.g5 * (cl + c.g7 * (ul - s));
            *u_s = c.g5 * (cl + c.g7 * ul + s);
            *d_s = dl * pow(cc / cl, c.g4);
            *p_s = pl * pow(cc / cl, c.g3);
            return -1;
        } else { /* left shock */
            double pml = pm / pl;
            double sl = ul - cl * sqrt(c.g2 * pml + c.g1);
            if (s <= sl) { *d_s = dl; *u_s = ul; *p_s = pl; return -1; }
            *d_s = dl * (pml + c.g6) / (pml * c.g6 + 1.0);
            *u_s = um; *p_s = pm;
            return -1;
        }
    } else { /* right of the contact */
        if (pm > pr) { /* right shock */
            double pmr = pm / pr;
            double sr = ur + cr * sqrt(c.g2 * pmr + c.g1);
            if (s >= sr) { *d_s = dr; *u_s = ur; *p_s = pr; return 1; }
            *d_s = dr * (pmr + c.g6) / (pmr * c.g6 + 1.0);
            *u_s = um; *p_s = pm;
            return 1;
        } else { /* right rarefaction */
            double shr = ur + cr;
            if (s >= shr) { *d_s = dr; *u_s = ur; *p_s = pr; return 1; }
            double cmr = cr * pow(pm / pr, c.g1);
            double str = um + cmr;
            if (s <= str) {
                *d_s = dr * pow(pm / pr, 1.0 / gamma);
                *u_s = um; *p_s = pm;
                return 1;
            }
            double cc = c.g5 * (cr - c.g7 * (ur - s));
            *u_s = c.g5 * (-cr + c.g7 * ur + s);
            *d_s = dr * pow(cc / cr, c.g4);
            *p_s = pr * pow(cc / cr, c.g3);
            return 1;
        }
    }
}

void orc_riemann_batch(int64_t n, double gamma, const double *rhoL,
                       const double *uL, const double *PL, const double *rhoR,
                       const double *uR, const double *PR, double tol,
                       int max_iter, double *rho_s, double *u_s, double *P_s,
                       int32_t *flag, int32_t *iters) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        int it = 0;
        int f = orc_riemann_solve(gamma, rhoL[k], uL[k], PL[k], rhoR[k], uR[k],
                                  PR[k], tol, max_iter, &rho_s[k], &u_s[k],
                                  &P_s[k], &it);
        if (flag) flag[k] = f;
        if (iters) iters[k] = it;
    }
}
