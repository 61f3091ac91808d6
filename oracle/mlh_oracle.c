/*
 * mlh_oracle.c -- CPU FP64 oracle of the MFV time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library (see
 * mlh_oracle.h).  Nothing in the product path loads this file.
 *
 * Each stage follows the reference function cited beside it
 * (jammartin/meshlessHydro, demonstrator/src/Particles.cpp unless noted), in
 * the reference's order of floating-point operations where that order
 * decides a result (neighbour distances, kernel, ghost coordinates,
 * limiter).  Reference behaviour kept on purpose:
 *   - grid cells are not periodic; periodic neighbours come only from ghost
 *     particles, created with max+(x-min) / min-(max-x) and found by an
 *     exhaustive distance test (ghostNNS, :2237-2260);
 *   - omega = sum_real W + W(0) + sum_ghost W (:1168-1183, :2273-2287);
 *   - DBL_MIN as the start of the limiter maxima and of vSig (:1344-1346,
 *     :1450);
 *   - limiter |.| through int abs(int) unless limiter_abs == FABS;
 *   - a pair is solved from the lower index, the other side negated
 *     (:1838-1858, ghosts: parent(g) >= i, :1886-1897);
 *   - the time step uses real neighbours only (:1446-1485);
 *   - the drift uses the pre-update velocity (:2082-2087).
 * Where the reference has no 3D periodic path (it exits, :2185-2188) the 3D
 * choices are: ghost images per imaged axis subset in descending axis-mask
 * order (reduces to the 2D order corner, y-image, x-image); 3D ghost pairs
 * are reconstructed like real pairs; the z momentum of ghost pairs is
 * summed; the 3D face frame is the basis-free projection
 * v* = u* n + (v_up - (v_up.n) n), equal to the Rodrigues rotation of
 * Helper.cpp:48-77 wherever that one is defined.
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "mlh_oracle.h"

#define MAXD 3

struct orc {
    mlh_params p;
    int N, D;
    double *x[MAXD], *v[MAXD], *m, *u;
    double *rho, *P, *omega;
    double *grad; /* [(D+2)][N][D] */
    double *B;    /* [N][D*D] */
    /* domain / grid (Domain.cpp:9-54) */
    double bmin[MAXD], bmax[MAXD], cs[MAXD];
    int nc[MAXD], ncells;
    int *cell, *cstart, *clist;
    /* ghosts in ghost-index order (Particles.cpp:2113-2191) */
    int NG;
    double *gx[MAXD];
    int *gpar, *gcode;
    /* neighbour lists, fixed stride like nnl / nnlGhosts */
    int maxr, maxg;
    int *nnl, *noi, *nnlg, *noig;
    /* per-pair arrays in CSR order (offsets from the counts) */
    int64_t *offr, *offg;
    double *psi, *psig, *A, *Ag, *F, *Fg;
    double *mF, *vF, *eF;
    int stage;
    int limited;  /* gradients of the last gradient stage are limited */
    int brute;
    mlh_diag diag;
};

/* ------------------------------------------------------------------ util */
static void *xcalloc(size_t n, size_t sz) {
    if (n == 0) n = 1;
    return calloc(n, sz);
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* Kernel::cubicSpline, Particles.cpp:7-24 (support radius h, uses h/2). */
double orc_kernel(int dim, double r, double h) {
    double h2 = h / 2.;
    double sigma = dim == 2 ? 10. / (7. * M_PI * h2 * h2)
                            : 1. / (M_PI * h2 * h2 * h2);
    double q = r / h2;
    if (0. <= q && q <= 1.) {
        return sigma * (1. - 3. / 2. * q * q * (1. - q / 2.));
    } else if (1. < q && q < 2.) {
        return sigma / 4. * pow(2. - q, 3.);
    }
    return 0.;
}

/* d^2 as pow(a,2)+pow(b,2)(+pow(c,2)), Particles.cpp:340-345 */
static inline double dsq(int D, const double *a, const double *b) {
    double d = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]);
    if (D == 3) d += (a[2] - b[2]) * (a[2] - b[2]);
    return d;
}

/* Helper::dotProduct, Helper.cpp:20-26 */
static inline double dot(int D, const double *a, const double *b) {
    double r = 0.;
    for (int k = 0; k < D; ++k) r += a[k] * b[k];
    return r;
}

/* abs() of the limiter: int abs(int) of the truncated double (x86
 * cvttsd2si gives INT_MIN for NaN / out of range, abs(INT_MIN) == INT_MIN)
 * or the intended fabs. */
static inline double lim_abs(const orc *o, double x) {
    if (o->p.limiter_abs == MLH_ABS_FABS) return fabs(x);
    int32_t t;
    if (isnan(x) || x >= 2147483648.0 || x <= -2147483649.0) {
        t = INT32_MIN;
    } else {
        t = (int32_t)x;
    }
    if (t != INT32_MIN && t < 0) t = -t;
    return (double)t;
}

static void particle_pos(const orc *o, int i, double *xi) {
    for (int a = 0; a < o->D; ++a) xi[a] = o->x[a][i];
}
static void ghost_pos(const orc *o, int g, double *xg) {
    for (int a = 0; a < o->D; ++a) xg[a] = o->gx[a][g];
}

/* --------------------------------------------------------------- create */
orc *orc_create(const mlh_params *p, int32_t n) {
    if (!p || n <= 0 || (p->dim != 2 && p->dim != 3)) return NULL;
    orc *o = (orc *)calloc(1, sizeof(orc));
    o->p = *p;
    o->N = n;
    o->D = p->dim;
    int D = o->D;
    for (int a = 0; a < D; ++a) {
        o->x[a] = xcalloc(n, sizeof(double));
        o->v[a] = xcalloc(n, sizeof(double));
    }
    o->m = xcalloc(n, sizeof(double));
    o->u = xcalloc(n, sizeof(double));
    o->rho = xcalloc(n, sizeof(double));
    o->P = xcalloc(n, sizeof(double));
    o->omega = xcalloc(n, sizeof(double));
    o->grad = xcalloc((size_t)(D + 2) * n * D, sizeof(double));
    o->B = xcalloc((size_t)n * D * D, sizeof(double));
    o->cell = xcalloc(n, sizeof(int));
    o->maxr = p->max_interactions;
    o->maxg = p->max_ghost_interactions;
    o->nnl = xcalloc((size_t)n * o->maxr, sizeof(int));
    o->noi = xcalloc(n, sizeof(int));
    o->nnlg = xcalloc((size_t)n * (o->maxg > 0 ? o->maxg : 1), sizeof(int));
    o->noig = xcalloc(n, sizeof(int));
    o->offr = xcalloc((size_t)n + 1, sizeof(int64_t));
    o->offg = xcalloc((size_t)n + 1, sizeof(int64_t));
    o->mF = xcalloc(n, sizeof(double));
    o->vF = xcalloc((size_t)n * D, sizeof(double));
    o->eF = xcalloc(n, sizeof(double));
    /* ghost capacity (2^D - 1) * N (at most 2^D-1 images per particle) */
    int capg = ((1 << D) - 1) * n;
    for (int a = 0; a < D; ++a) o->gx[a] = xcalloc(capg, sizeof(double));
    o->gpar = xcalloc(capg, sizeof(int));
    o->gcode = xcalloc(capg, sizeof(int));
    return o;
}

void orc_destroy(orc *o) {
    if (!o) return;
    for (int a = 0; a < o->D; ++a) {
        free(o->x[a]); free(o->v[a]); free(o->gx[a]);
    }
    free(o->m); free(o->u); free(o->rho); free(o->P); free(o->omega);
    free(o->grad); free(o->B); free(o->cell); free(o->cstart);
    free(o->clist); free(o->gpar); free(o->gcode); free(o->nnl);
    free(o->noi); free(o->nnlg); free(o->noig); free(o->offr); free(o->offg);
    free(o->psi); free(o->psig); free(o->A); free(o->Ag); free(o->F);
    free(o->Fg); free(o->mF); free(o->vF); free(o->eF);
    free(o);
}

void orc_set_bruteforce_ghosts(orc *o, int on) { o->brute = on; }

int orc_set_state(orc *o, const double *x, const double *y, const double *z,
                  const double *vx, const double *vy, const double *vz,
                  const double *m, const double *u) {
    size_t nb = sizeof(double) * o->N;
    memcpy(o->x[0], x, nb); memcpy(o->x[1], y, nb);
    memcpy(o->v[0], vx, nb); memcpy(o->v[1], vy, nb);
    if (o->D == 3) { memcpy(o->x[2], z, nb); memcpy(o->v[2], vz, nb); }
    memcpy(o->m, m, nb); memcpy(o->u, u, nb);
    o->stage = 0;
    return 0;
}

int orc_get_state(orc *o, double *x, double *y, double *z, double *vx,
                  double *vy, double *vz, double *m, double *u) {
    size_t nb = sizeof(double) * o->N;
    if (x) memcpy(x, o->x[0], nb);
    if (y) memcpy(y, o->x[1], nb);
    if (vx) memcpy(vx, o->v[0], nb);
    if (vy) memcpy(vy, o->v[1], nb);
    if (o->D == 3) {
        if (z) memcpy(z, o->x[2], nb);
        if (vz) memcpy(vz, o->v[2], nb);
    }
    if (m) memcpy(m, o->m, nb);
    if (u) memcpy(u, o->u, nb);
    return 0;
}

/* ---------------------------------------------------------- neighbours */
/* Particles::getDomainLimits, Particles.cpp:227-268 (non-periodic runs):
 * running min with `else if` max, maxima start at DBL_MIN. */
static void domain_limits(orc *o) {
    for (int a = 0; a < o->D; ++a) {
        double mn = DBL_MAX, mx = DBL_MIN;
        for (int i = 0; i < o->N; ++i) {
            double xv = o->x[a][i];
            if (xv < mn) mn = xv;
            else if (xv > mx) mx = xv;
        }
        o->bmin[a] = mn;
        o->bmax[a] = mx;
    }
}

/* Domain::createGrid, Domain.cpp:9-54 */
static int create_grid(orc *o) {
    double h = o->p.kernel_size;
    o->ncells = 1;
    for (int a = 0; a < o->D; ++a) {
        o->nc[a] = (int)floor((o->bmax[a] - o->bmin[a]) / h);
        if (o->nc[a] < 1) return MLH_ERR_GRID;
        if (o->p.periodic && o->nc[a] < 3) return MLH_ERR_GRID;
        o->cs[a] = (o->bmax[a] - o->bmin[a]) / (double)o->nc[a];
        o->ncells *= o->nc[a];
    }
    free(o->cstart);
    o->cstart = xcalloc((size_t)o->ncells + 1, sizeof(int));
    return 0;
}

/* Particles::assignParticlesAndCells, Particles.cpp:270-322 */
static void assign_cells(orc *o) {
    int D = o->D;
    for (int i = 0; i < o->N; ++i) {
        int f[MAXD];
        for (int a = 0; a < D; ++a) {
            f[a] = (int)floor((o->x[a][i] - o->bmin[a]) / o->cs[a]);
            if (f[a] == o->nc[a]) f[a] -= 1;
            /* outside the box the reference indexes out of bounds; clamp */
            if (f[a] < 0) f[a] = 0;
            if (f[a] > o->nc[a] - 1) f[a] = o->nc[a] - 1;
        }
        int c = f[0] + f[1] * o->nc[0];
        if (D == 3) c += f[2] * o->nc[0] * o->nc[1];
        o->cell[i] = c;
    }
    /* cell lists in ascending particle index (push_back order) */
    memset(o->cstart, 0, sizeof(int) * ((size_t)o->ncells + 1));
    for (int i = 0; i < o->N; ++i) o->cstart[o->cell[i] + 1]++;
    for (int c = 0; c < o->ncells; ++c) o->cstart[c + 1] += o->cstart[c];
    free(o->clist);
    o->clist = xcalloc(o->N, sizeof(int));
    int *fill = xcalloc((size_t)o->ncells, sizeof(int));
    for (int i = 0; i < o->N; ++i) {
        int c = o->cell[i];
        o->clist[o->cstart[c] + fill[c]++] = i;
    }
    free(fill);
}

/* Particles::createGhostParticles, Particles.cpp:2113-2191, generalised to
 * D axes: per axis an image when x <= min+h (to max+(x-min)) else when
 * max-h < x (to min-(max-x)); one ghost per non-empty subset of the imaged
 * axes, subsets in descending bit order (2D: corner, y-image, x-image). */
static void create_ghosts(orc *o) {
    int D = o->D;
    double h = o->p.kernel_size;
    int ng = 0;
    for (int i = 0; i < o->N; ++i) {
        int s[MAXD] = {0, 0, 0};
        double img[MAXD];
        int avail = 0;
        for (int a = 0; a < D; ++a) {
            double xa = o->x[a][i];
            if (xa <= o->bmin[a] + h) {
                img[a] = o->bmax[a] + (xa - o->bmin[a]);
                s[a] = +1;
                avail |= 1 << a;
            } else if (o->bmax[a] - h < xa) {
                img[a] = o->bmin[a] - (o->bmax[a] - xa);
                s[a] = -1;
                avail |= 1 << a;
            } else {
                img[a] = xa;
            }
        }
        for (int mask = (1 << D) - 1; mask >= 1; --mask) {
            if ((mask & avail) != mask) continue;
            int code = 0, mul = 1;
            for (int a = 0; a < 3; ++a) {
                int sa = (a < D && (mask >> a & 1)) ? s[a] : 0;
                if (a < D) o->gx[a][ng] = (mask >> a & 1) ? img[a] : o->x[a][i];
                code += (sa + 1) * mul;
                mul *= 3;
            }
            o->gpar[ng] = i;
            o->gcode[ng] = code;
            ng++;
        }
    }
    o->NG = ng;
}

/* Domain::getNeighborCells, Domain.cpp:83-118: x outer, y, z inner; -1
 * outside the grid (no wrap). */
static int stencil(const orc *o, int c, int *out) {
    int D = o->D;
    int ix = c % o->nc[0];
    int iy = (c / o->nc[0]) % o->nc[1];
    int iz = D == 3 ? c / (o->nc[0] * o->nc[1]) : 0;
    int n = 0;
    for (int k = ix - 1; k <= ix + 1; ++k)
        for (int l = iy - 1; l <= iy + 1; ++l) {
            if (D == 2) {
                out[n++] = (k < 0 || k >= o->nc[0] || l < 0 || l >= o->nc[1])
                               ? -1 : k + l * o->nc[0];
            } else {
                for (int mm = iz - 1; mm <= iz + 1; ++mm) {
                    out[n++] = (k < 0 || k >= o->nc[0] || l < 0 ||
                                l >= o->nc[1] || mm < 0 || mm >= o->nc[2])
                                   ? -1
                                   : k + l * o->nc[0] + mm * o->nc[0] * o->nc[1];
                }
            }
        }
    return n;
}

/* Particles::gridNNS, Particles.cpp:324-365 */
static int grid_nns(orc *o) {
    int D = o->D, err = 0;
    const double hsq = o->p.kernel_size * o->p.kernel_size;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : err)
    for (int i = 0; i < o->N; ++i) {
        int cells[27];
        int nsc = stencil(o, o->cell[i], cells);
        int cnt = 0;
        double xi[MAXD], xj[MAXD];
        particle_pos(o, i, xi);
        for (int s = 0; s < nsc; ++s) {
            if (cells[s] < 0) continue;
            for (int q = o->cstart[cells[s]]; q < o->cstart[cells[s] + 1]; ++q) {
                int j = o->clist[q];
                if (j == i) continue;
                particle_pos(o, j, xj);
                double d = dsq(D, xj, xi);
                if (d < hsq) {
                    if (cnt >= o->maxr) { err |= 1; break; }
                    o->nnl[(size_t)i * o->maxr + cnt++] = j;
                }
            }
        }
        o->noi[i] = cnt;
    }
    return err ? MLH_ERR_NEIGHBOUR_OVERFLOW : 0;
}

/* Particles::ghostNNS, Particles.cpp:2237-2260: every ghost is tested, hits
 * kept in ghost-index order.  The default path bins ghosts by extended cell
 * (floor((g-min)/cs), clamped to [-2, n+1]) and tests the 5^D window around
 * the particle's cell, then orders hits by ghost index; a ghost within h of
 * x_i differs from it by less than one cell size per axis, so the window
 * holds every hit.  The exhaustive loop stays available (orc_set_bruteforce
 * _ghosts) and the tests compare both. */
static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static int ghost_nns(orc *o) {
    int D = o->D, err = 0;
    const double hsq = o->p.kernel_size * o->p.kernel_size;
    if (o->brute) {
#pragma omp parallel for schedule(dynamic, 64) reduction(| : err)
        for (int i = 0; i < o->N; ++i) {
            double xi[MAXD], xg[MAXD];
            particle_pos(o, i, xi);
            int cnt = 0;
            for (int g = 0; g < o->NG; ++g) {
                ghost_pos(o, g, xg);
                double d = dsq(D, xg, xi);
                if (d < hsq) {
                    if (cnt >= o->maxg) { err |= 1; break; }
                    o->nnlg[(size_t)i * o->maxg + cnt++] = g;
                }
            }
            o->noig[i] = cnt;
        }
        return err ? MLH_ERR_GHOST_OVERFLOW : 0;
    }
    /* binned search over an extended grid of (n+4)^D cells */
    int ext[MAXD], next = 1;
    for (int a = 0; a < D; ++a) { ext[a] = o->nc[a] + 4; next *= ext[a]; }
    int *gstart = xcalloc((size_t)next + 1, sizeof(int));
    int *gcell = xcalloc((size_t)(o->NG > 0 ? o->NG : 1), sizeof(int));
    int *glist = xcalloc((size_t)(o->NG > 0 ? o->NG : 1), sizeof(int));
    for (int g = 0; g < o->NG; ++g) {
        int c = 0, mul = 1;
        for (int a = 0; a < D; ++a) {
            int f = (int)floor((o->gx[a][g] - o->bmin[a]) / o->cs[a]);
            if (f < -2) f = -2;
            if (f > o->nc[a] + 1) f = o->nc[a] + 1;
            c += (f + 2) * mul;
            mul *= ext[a];
        }
        gcell[g] = c;
        gstart[c + 1]++;
    }
    for (int c = 0; c < next; ++c) gstart[c + 1] += gstart[c];
    int *fill = xcalloc((size_t)next, sizeof(int));
    for (int g = 0; g < o->NG; ++g) glist[gstart[gcell[g]] + fill[gcell[g]]++] = g;
    free(fill);
#pragma omp parallel for schedule(dynamic, 64) reduction(| : err)
    for (int i = 0; i < o->N; ++i) {
        double xi[MAXD], xg[MAXD];
        particle_pos(o, i, xi);
        int ci[MAXD];
        for (int a = 0; a < D; ++a) {
            int f = (int)floor((xi[a] - o->bmin[a]) / o->cs[a]);
            if (f < -2) f = -2;
            if (f > o->nc[a] + 1) f = o->nc[a] + 1;
            ci[a] = f;
        }
        int cnt = 0;
        int *row = &o->nnlg[(size_t)i * o->maxg];
        int lo[MAXD] = {0, 0, 0}, hi[MAXD] = {0, 0, 0};
        for (int a = 0; a < D; ++a) {
            lo[a] = ci[a] - 2 < -2 ? -2 : ci[a] - 2;
            hi[a] = ci[a] + 2 > o->nc[a] + 1 ? o->nc[a] + 1 : ci[a] + 2;
        }
        for (int cz = (D == 3 ? lo[2] : 0); cz <= (D == 3 ? hi[2] : 0); ++cz)
            for (int cy = lo[1]; cy <= hi[1]; ++cy)
                for (int cx = lo[0]; cx <= hi[0]; ++cx) {
                    int c = (cx + 2) + (cy + 2) * ext[0];
                    if (D == 3) c += (cz + 2) * ext[0] * ext[1];
                    for (int q = gstart[c]; q < gstart[c + 1]; ++q) {
                        int g = glist[q];
                        ghost_pos(o, g, xg);
                        double d = dsq(D, xg, xi);
                        if (d < hsq) {
                            if (cnt >= o->maxg) { err |= 1; goto done; }
                            row[cnt++] = g;
                        }
                    }
                }
    done:
        qsort(row, cnt, sizeof(int), cmp_int);
        o->noig[i] = cnt;
    }
    free(gstart); free(gcell); free(glist);
    return err ? MLH_ERR_GHOST_OVERFLOW : 0;
}

static void build_offsets(orc *o) {
    o->offr[0] = 0;
    o->offg[0] = 0;
    for (int i = 0; i < o->N; ++i) {
        o->offr[i + 1] = o->offr[i] + o->noi[i];
        o->offg[i + 1] = o->offg[i] + (o->p.periodic ? o->noig[i] : 0);
    }
    int D = o->D;
    size_t nr = (size_t)o->offr[o->N], ng = (size_t)o->offg[o->N];
    free(o->psi); free(o->psig); free(o->A); free(o->Ag); free(o->F); free(o->Fg);
    o->psi = xcalloc(nr * D, sizeof(double));
    o->A = xcalloc(nr * D, sizeof(double));
    o->F = xcalloc(nr * (D + 2), sizeof(double));
    o->psig = xcalloc(ng * D, sizeof(double));
    o->Ag = xcalloc(ng * D, sizeof(double));
    o->Fg = xcalloc(ng * (D + 2), sizeof(double));
}

int orc_neighbours(orc *o) {
    int D = o->D;
    if (o->p.periodic) {
        for (int a = 0; a < D; ++a) {
            o->bmin[a] = o->p.box_min[a];
            o->bmax[a] = o->p.box_max[a];
        }
    } else {
        domain_limits(o);
    }
    int st = create_grid(o);
    if (st) return st;
    assign_cells(o);
    o->NG = 0;
    if (o->p.periodic) create_ghosts(o);
    st = grid_nns(o);
    if (st) return st;
    if (o->p.periodic) {
        st = ghost_nns(o);
        if (st) return st;
    } else {
        memset(o->noig, 0, sizeof(int) * o->N);
    }
    int mr = 0, mg = 0;
    for (int i = 0; i < o->N; ++i) {
        if (o->noi[i] > mr) mr = o->noi[i];
        if (o->noig[i] > mg) mg = o->noig[i];
    }
    o->diag.max_neighbours = mr;
    o->diag.max_ghost_neighbours = mg;
    build_offsets(o);
    o->stage = 1;
    return 0;
}

int orc_get_neighbours(orc *o, int32_t *noi, int32_t *noi_ghost,
                       int32_t *nbr) {
    if (o->stage < 1) return MLH_ERR_CALL_ORDER;
    int stride = o->maxr + o->maxg;
    for (int i = 0; i < o->N; ++i) {
        noi[i] = o->noi[i];
        noi_ghost[i] = o->p.periodic ? o->noig[i] : 0;
        int32_t *row = &nbr[(size_t)i * stride];
        /* packed as the C ABI's entries: index | image code << 27, formed
         * in uint32 (codes up to 26 exceed INT32_MAX once shifted) */
        for (int n = 0; n < o->noi[i]; ++n)
            row[n] = (int32_t)((uint32_t)o->nnl[(size_t)i * o->maxr + n] | (13u << 27));
        if (o->p.periodic)
            for (int n = 0; n < o->noig[i]; ++n) {
                int g = o->nnlg[(size_t)i * o->maxg + n];
                row[o->noi[i] + n] =
                    (int32_t)((uint32_t)o->gpar[g] | ((uint32_t)o->gcode[g] << 27));
            }
    }
    return 0;
}

/* ---------------------------------------------------------------- density */
/* compDensity / compOmega (Particles.cpp:1151-1184), ghost part
 * (:2262-2290), compPressure (:1272-1288) */
int orc_density(orc *o) {
    if (o->stage < 1) return MLH_ERR_CALL_ORDER;
    int D = o->D;
    double h = o->p.kernel_size;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->N; ++i) {
        double xi[MAXD], xj[MAXD];
        particle_pos(o, i, xi);
        double omg = 0.;
        for (int n = 0; n < o->noi[i]; ++n) {
            int j = o->nnl[(size_t)i * o->maxr + n];
            particle_pos(o, j, xj);
            double r = sqrt(dsq(D, xi, xj));
            omg += orc_kernel(D, r, h);
        }
        double om = omg + orc_kernel(D, 0., h);
        if (o->p.periodic) {
            double og = om;
            for (int n = 0; n < o->noig[i]; ++n) {
                ghost_pos(o, o->nnlg[(size_t)i * o->maxg + n], xj);
                double r = sqrt(dsq(D, xi, xj));
                og += orc_kernel(D, r, h);
            }
            om = og;
        }
        o->omega[i] = om;
        o->rho[i] = o->m[i] * om;
        o->P[i] = (o->p.gamma - 1.) * o->rho[i] * o->u[i];
    }
    o->stage = 2;
    return 0;
}

int orc_get_density(orc *o, double *rho, double *P, double *omega) {
    if (o->stage < 2) return MLH_ERR_CALL_ORDER;
    size_t nb = sizeof(double) * o->N;
    if (rho) memcpy(rho, o->rho, nb);
    if (P) memcpy(P, o->P, nb);
    if (omega) memcpy(omega, o->omega, nb);
    return 0;
}

/* compGlobalTimestep, Particles.cpp:1446-1485 (real neighbours only) */
int orc_timestep(orc *o, double *dt_out) {
    if (o->stage < 2) return MLH_ERR_CALL_ORDER;
    int D = o->D;
    double g = o->p.gamma;
    double dtmin = DBL_MAX;
#pragma omp parallel for schedule(static) reduction(min : dtmin)
    for (int i = 0; i < o->N; ++i) {
        double vsig = DBL_MIN;
        double ci = sqrt(g * o->P[i] / o->rho[i]);
        for (int n = 0; n < o->noi[i]; ++n) {
            int j = o->nnl[(size_t)i * o->maxr + n];
            double cj = sqrt(g * o->P[j] / o->rho[j]);
            double xij[MAXD], vij[MAXD];
            for (int a = 0; a < D; ++a) {
                xij[a] = o->x[a][i] - o->x[a][j];
                vij[a] = o->v[a][i] - o->v[a][j];
            }
            double vx = dot(D, vij, xij) / sqrt(dot(D, xij, xij));
            vx = vx < 0. ? vx : 0.;
            double vs = ci + cj - vx;
            vsig = vs > vsig ? vs : vsig;
        }
        double dt = o->p.cfl * o->p.kernel_size / vsig;
        dtmin = dt < dtmin ? dt : dtmin;
    }
    *dt_out = dtmin;
    return 0;
}

/* ------------------------------------------------------------- gradients */
static inline double field(const orc *o, int f, int i) {
    if (f == 0) return o->rho[i];
    if (f == o->D + 1) return o->P[i];
    return o->v[f - 1][i];
}
static inline double *gradp(orc *o, int f, int i) {
    return &o->grad[((size_t)f * o->N + i) * o->D];
}

/* Helper::inverseMatrix (Helper.cpp:7-18): LAPACK dgetrf_ + dgetri_ on a
 * DIMxDIM matrix.  Restated as Gauss-Jordan elimination with the same
 * partial pivoting (largest |a_kj| row); results agree with LU to ~1 ulp. */
static void invert(int D, double *A) {
    double M[MAXD][2 * MAXD];
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < 2 * D; ++c)
            M[r][c] = c < D ? A[r * D + c] : (c - D == r ? 1. : 0.);
    for (int k = 0; k < D; ++k) {
        int piv = k;
        for (int r = k + 1; r < D; ++r)
            if (fabs(M[r][k]) > fabs(M[piv][k])) piv = r;
        if (piv != k)
            for (int c = 0; c < 2 * D; ++c) {
                double t = M[k][c]; M[k][c] = M[piv][c]; M[piv][c] = t;
            }
        double d = M[k][k];
        for (int c = 0; c < 2 * D; ++c) M[k][c] /= d;
        for (int r = 0; r < D; ++r) {
            if (r == k) continue;
            double f = M[r][k];
            for (int c = 0; c < 2 * D; ++c) M[r][c] -= f * M[k][c];
        }
    }
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) A[r * D + c] = M[r][D + c];
}

/* one neighbour (real j >= 0 or ghost g) as position + parent index */
static inline int nb_get(const orc *o, int i, int n, double *xj) {
    if (n < o->noi[i]) {
        int j = o->nnl[(size_t)i * o->maxr + n];
        particle_pos(o, j, xj);
        return j;
    }
    int g = o->nnlg[(size_t)i * o->maxg + (n - o->noi[i])];
    ghost_pos(o, g, xj);
    return o->gpar[g];
}
static inline int nb_count(const orc *o, int i) {
    return o->noi[i] + (o->p.periodic ? o->noig[i] : 0);
}
static inline double *psi_at(orc *o, int i, int n) {
    return n < o->noi[i] ? &o->psi[(o->offr[i] + n) * o->D]
                         : &o->psig[(o->offg[i] + n - o->noi[i]) * o->D];
}

/* slopeLimiter for one field, Particles.cpp:1339-1443 */
static void limit_field(orc *o, int f, int i) {
    int D = o->D;
    double h = o->p.kernel_size;
    double fi = field(o, f, i);
    double *gi = gradp(o, f, i);
    double maxNgb = DBL_MIN, minNgb = DBL_MAX, maxMid = DBL_MIN, minMid = DBL_MAX;
    double xi[MAXD], xj[MAXD];
    particle_pos(o, i, xi);
    int cnt = nb_count(o, i);
    for (int n = 0; n < cnt; ++n) {
        int j = nb_get(o, i, n, xj);
        double xijxi[MAXD];
        for (int a = 0; a < D; ++a) {
            double xij = o->p.first_order_quad_point
                             ? (xi[a] + xj[a]) / 2.
                             : xi[a] + h / 4. * (xj[a] - xi[a]);
            xijxi[a] = xij - xi[a];
        }
        double fj = field(o, f, j);
        if (maxNgb < fj) maxNgb = fj;
        if (minNgb > fj) minNgb = fj;
        double fij = fi + dot(D, gi, xijxi);
        if (maxMid < fij) maxMid = fij;
        if (minMid > fij) minMid = fij;
    }
    double aMax = lim_abs(o, (maxNgb - fi) / (maxMid - fi));
    double aMin = lim_abs(o, (fi - minNgb) / (fi - minMid));
    double beta = o->p.beta;
    if (aMin <= aMax && beta * aMin < 1.) {
        for (int a = 0; a < D; ++a) gi[a] *= aMin;
    } else if (aMax <= aMin && beta * aMax < 1.) {
        for (int a = 0; a < D; ++a) gi[a] *= aMax;
    }
}

/* compPsijTilde (Particles.cpp:1186-1254; ghosts :2292-2455), gradient
 * (:1257-1270; :2457-2502), slopeLimiter (:1313-1444) when limit */
static int gradients_impl(orc *o, int limit) {
    if (o->stage < 2) return MLH_ERR_CALL_ORDER;
    int D = o->D, NF = D + 2;
    double h = o->p.kernel_size;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->N; ++i) {
        double xi[MAXD], xj[MAXD];
        particle_pos(o, i, xi);
        double E[MAXD * MAXD] = {0};
        int cnt = nb_count(o, i);
        for (int n = 0; n < cnt; ++n) {
            nb_get(o, i, n, xj);
            double r = sqrt(dsq(D, xi, xj));
            double w = orc_kernel(D, r, h) / o->omega[i];
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b)
                    E[D * a + b] += (xj[a] - xi[a]) * (xj[b] - xi[b]) * w;
        }
        invert(D, E);
        memcpy(&o->B[(size_t)i * D * D], E, sizeof(double) * D * D);
        for (int n = 0; n < cnt; ++n) {
            nb_get(o, i, n, xj);
            double r = sqrt(dsq(D, xi, xj));
            double w = orc_kernel(D, r, h) / o->omega[i];
            double *ps = psi_at(o, i, n);
            for (int a = 0; a < D; ++a) {
                ps[a] = 0.;
                for (int b = 0; b < D; ++b) ps[a] += E[a * D + b] * (xj[b] - xi[b]) * w;
            }
        }
        for (int f = 0; f < NF; ++f) {
            double *g = gradp(o, f, i);
            for (int a = 0; a < D; ++a) g[a] = 0;
            double fi = field(o, f, i);
            for (int n = 0; n < cnt; ++n) {
                int j = nb_get(o, i, n, xj);
                double *ps = psi_at(o, i, n);
                for (int a = 0; a < D; ++a) g[a] += (field(o, f, j) - fi) * ps[a];
            }
        }
    }
    if (limit) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < o->N; ++i)
            for (int f = 0; f < NF; ++f) limit_field(o, f, i);
    }
    o->stage = 3;
    o->limited = limit;
    return 0;
}

int orc_gradients(orc *o) { return gradients_impl(o, o->p.slope_limiting); }

/* compPsijTilde + gradient without slopeLimiter */
int orc_psi_gradients(orc *o) { return gradients_impl(o, 0); }

/* slopeLimiter (Particles.cpp:1313-1444) on unlimited gradients */
int orc_limit(orc *o) {
    if (o->stage < 3 || o->limited) return MLH_ERR_CALL_ORDER;
    int NF = o->D + 2;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->N; ++i)
        for (int f = 0; f < NF; ++f) limit_field(o, f, i);
    o->limited = 1;
    return 0;
}

/* gradient(f, grad) (Particles.cpp:1257-1270; with ghosts :2457-2502, the
 * ghost's value being its parent's, updateGhostState :2193-2206) */
int orc_gradient_field(orc *o, const double *f, double *grad) {
    if (o->stage < 3) return MLH_ERR_CALL_ORDER;
    int D = o->D;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->N; ++i) {
        double xj[MAXD];
        double *g = &grad[(size_t)i * D];
        for (int a = 0; a < D; ++a) g[a] = 0.;
        int cnt = nb_count(o, i);
        for (int n = 0; n < cnt; ++n) {
            int j = nb_get(o, i, n, xj);
            const double *ps = psi_at(o, i, n);
            for (int a = 0; a < D; ++a) g[a] += (f[j] - f[i]) * ps[a];
        }
    }
    return 0;
}

int orc_get_gradients(orc *o, double *grads, double *B) {
    if (o->stage < 3) return MLH_ERR_CALL_ORDER;
    int D = o->D;
    if (grads) memcpy(grads, o->grad, sizeof(double) * (D + 2) * o->N * D);
    if (B) memcpy(B, o->B, sizeof(double) * o->N * D * D);
    return 0;
}

/* ------------------------------------------------------------------ faces */
/* pairwiseLimiter, Particles.cpp:1735-1785 */
static double pairwise_limiter(const orc *o, double phi0, double phi_i,
                               double phi_j, double xijxi_abs,
                               double xjxi_abs) {
    double phi_ = phi_i;
    double phi_ij = phi_i + xijxi_abs / xjxi_abs * (phi_j - phi_i);
    double phiMin, phiMax;
    if (phi_i < phi_j) { phiMin = phi_i; phiMax = phi_j; }
    else { phiMin = phi_j; phiMax = phi_i; }
    double delta1 = o->p.psi1 * lim_abs(o, phi_i - phi_j);
    double delta2 = o->p.psi2 * lim_abs(o, phi_i - phi_j);
    double phiMinus, phiPlus;
    if ((phiMax + delta1 >= 0. && phiMax >= 0.) || (phiMax + delta1 < 0. && phiMax < 0.))
        phiPlus = phiMax + delta1;
    else
        phiPlus = phiMax / (1. + delta1 / lim_abs(o, phiMax));
    if ((phiMin - delta1 >= 0. && phiMin >= 0.) || (phiMin - delta1 < 0. && phiMin < 0.))
        phiMinus = phiMin - delta1;
    else
        phiMinus = phiMin / (1. + delta1 / lim_abs(o, phiMin));
    if (phi_i < phi_j) {
        double minPhiD2 = phi_ij + delta2 < phi0 ? phi_ij + delta2 : phi0;
        phi_ = phiMinus > minPhiD2 ? phiMinus : minPhiD2;
    } else if (phi_i > phi_j) {
        double maxPhiD2 = phi_ij - delta2 > phi0 ? phi_ij - delta2 : phi0;
        phi_ = phiPlus < maxPhiD2 ? phiPlus : maxPhiD2;
    }
    return phi_;
}

/* compRiemannStatesLR for one directed pair i -> neighbour at xj with
 * parent j (Particles.cpp:1488-1733 real, :2533-2683 ghosts).
 * WR = i-state, WL = j-state, component order {rho, P, v...}. */
static void face_states(const orc *o, int i, int j, const double *xj,
                        int is_ghost, double dt, double *WR, double *WL,
                        double *vFrame) {
    int D = o->D;
    double h = o->p.kernel_size, g = o->p.gamma;
    double xi[MAXD], xijxi[MAXD], xijxj[MAXD], xjxi[MAXD];
    particle_pos(o, i, xi);
    for (int a = 0; a < D; ++a) {
        xjxi[a] = xj[a] - xi[a];
        if (o->p.first_order_quad_point) {
            xijxj[a] = .5 * (xi[a] - xj[a]);
            xijxi[a] = .5 * (xj[a] - xi[a]);
        } else {
            double xij = xi[a] + h / 4. * xjxi[a];
            xijxi[a] = xij - xi[a];
            xijxj[a] = xij - xj[a];
        }
    }
    if (!o->p.move_particles) {
        for (int a = 0; a < D; ++a) vFrame[a] = 0.;
    } else if (o->p.first_order_quad_point) {
        for (int a = 0; a < D; ++a) vFrame[a] = (o->v[a][i] + o->v[a][j]) / 2.;
    } else {
        double dp = dot(D, xijxi, xjxi), ds = dot(D, xjxi, xjxi);
        for (int a = 0; a < D; ++a)
            vFrame[a] = o->v[a][i] + (o->v[a][j] - o->v[a][i]) * dp / ds;
    }
    const double *gr_i = gradp((orc *)o, 0, i), *gr_j = gradp((orc *)o, 0, j);
    const double *gP_i = gradp((orc *)o, D + 1, i), *gP_j = gradp((orc *)o, D + 1, j);
    WR[0] = o->rho[i]; WL[0] = o->rho[j];
    WR[1] = o->P[i];   WL[1] = o->P[j];
    for (int a = 0; a < D; ++a) {
        WR[2 + a] = o->v[a][i] - vFrame[a];
        WL[2 + a] = o->v[a][j] - vFrame[a];
    }
    double WRb[MAXD + 2], WLb[MAXD + 2];
    memcpy(WRb, WR, sizeof(double) * (D + 2));
    memcpy(WLb, WL, sizeof(double) * (D + 2));
    WR[0] += dot(D, gr_i, xijxi); WL[0] += dot(D, gr_j, xijxj);
    WR[1] += dot(D, gP_i, xijxi); WL[1] += dot(D, gP_j, xijxj);
    for (int a = 0; a < D; ++a) {
        WR[2 + a] += dot(D, gradp((orc *)o, 1 + a, i), xijxi);
        WL[2 + a] += dot(D, gradp((orc *)o, 1 + a, j), xijxj);
    }
    if (o->p.pairwise_limiter && !is_ghost) {
        double xijxi_abs = 0., xijxj_abs = 0., xjxi_abs = 0.;
        for (int a = 0; a < D; ++a) {
            xijxi_abs += xijxi[a] * xijxi[a];
            xijxj_abs += xijxj[a] * xijxj[a];
            xjxi_abs += xjxi[a] * xjxi[a];
        }
        xijxi_abs = sqrt(xijxi_abs); xijxj_abs = sqrt(xijxj_abs);
        xjxi_abs = sqrt(xjxi_abs);
        for (int k = 0; k < D + 2; ++k) {
            WR[k] = pairwise_limiter(o, WR[k], WRb[k], WLb[k], xijxi_abs, xjxi_abs);
            WL[k] = pairwise_limiter(o, WL[k], WLb[k], WRb[k], xijxj_abs, xjxi_abs);
        }
    }
    /* half-step predictor */
    const double *gvx_i = gradp((orc *)o, 1, i), *gvx_j = gradp((orc *)o, 1, j);
    const double *gvy_i = gradp((orc *)o, 2, i), *gvy_j = gradp((orc *)o, 2, j);
    double viDiv = gvx_i[0] + gvy_i[1];
    double vjDiv = gvx_j[0] + gvy_j[1];
    if (D == 3) {
        viDiv += gradp((orc *)o, 3, i)[2];
        vjDiv += gradp((orc *)o, 3, j)[2];
    }
    double vxi = o->v[0][i] - vFrame[0], vyi = o->v[1][i] - vFrame[1];
    double vxj = o->v[0][j] - vFrame[0], vyj = o->v[1][j] - vFrame[1];
    double rhoi = o->rho[i], rhoj = o->rho[j], Pi = o->P[i], Pj = o->P[j];
    WR[0] -= dt / 2. * (rhoi * viDiv + vxi * gr_i[0] + vyi * gr_i[1]);
    WL[0] -= dt / 2. * (rhoj * vjDiv + vxj * gr_j[0] + vyj * gr_j[1]);
    WR[1] -= dt / 2. * (g * Pi * viDiv + vxi * gP_i[0] + vyi * gP_i[1]);
    WL[1] -= dt / 2. * (g * Pj * vjDiv + vxj * gP_j[0] + vyj * gP_j[1]);
    WR[2] -= dt / 2. * (gP_i[0] / rhoi + vxi * gvx_i[0] + vyi * gvx_i[1]);
    WL[2] -= dt / 2. * (gP_j[0] / rhoj + vxj * gvx_j[0] + vyj * gvx_j[1]);
    WR[3] -= dt / 2. * (gP_i[1] / rhoi + vxi * gvy_i[0] + vyi * gvy_i[1]);
    WL[3] -= dt / 2. * (gP_j[1] / rhoj + vxj * gvy_j[0] + vyj * gvy_j[1]);
    if (D == 3) {
        const double *gvz_i = gradp((orc *)o, 3, i), *gvz_j = gradp((orc *)o, 3, j);
        double vzi = o->v[2][i] - vFrame[2], vzj = o->v[2][j] - vFrame[2];
        /* Particles.cpp:1717-1719 uses vz[i] on the j side; kept behind a
         * switch, the default uses vz[j] */
        double vzjq = o->p.quirk_3d_predictor ? vzi : vzj;
        WR[0] -= dt / 2. * vzi * gr_i[2];
        WL[0] -= dt / 2. * vzj * gr_j[2];
        WR[1] -= dt / 2. * vzi * gP_i[2];
        WL[1] -= dt / 2. * vzj * gP_j[2];
        WR[2] -= dt / 2. * vzi * gvx_i[2];
        WL[2] -= dt / 2. * vzjq * gvx_j[2];
        WR[3] -= dt / 2. * vzi * gvy_i[2];
        WL[3] -= dt / 2. * vzjq * gvy_j[2];
        WR[4] -= dt / 2. * (gP_i[2] / rhoi + vxi * gvz_i[0] + vyi * gvz_i[1] + vzi * gvz_i[2]);
        WL[4] -= dt / 2. * (gP_j[2] / rhoj + vxj * gvz_j[0] + vyj * gvz_j[1] + vzj * gvz_j[2]);
    }
}

/* Test access for the oracle's own pins (tests/test_oracle_mfv.py): the
 * face states of the pair (i, j), with xj the neighbour's position and
 * rho / P [N], grads [(D+2) fields][N][D] installed in place of the density
 * and gradient stages' results.  Velocities and positions are the state. */
int orc_test_face_states(orc *o, int i, int j, const double *xj, int is_ghost, double dt,
                         const double *rho, const double *P, const double *grads,
                         double *WR, double *WL, double *vFrame) {
    if (!o || i < 0 || j < 0 || i >= o->N || j >= o->N) return MLH_ERR_INVALID_ARGUMENT;
    memcpy(o->rho, rho, sizeof(double) * o->N);
    memcpy(o->P, P, sizeof(double) * o->N);
    memcpy(o->grad, grads, sizeof(double) * (o->D + 2) * o->N * o->D);
    face_states(o, i, j, xj, is_ghost, dt, WR, WL, vFrame);
    return 0;
}

/* pairwiseLimiter alone, with the psi / abs() settings of p */
double orc_test_pairwise_limiter(const mlh_params *p, double phi0, double phi_i, double phi_j,
                                 double xijxi_abs, double xjxi_abs) {
    orc tmp;
    memset(&tmp, 0, sizeof(tmp));
    tmp.p = *p;
    return pairwise_limiter(&tmp, phi0, phi_i, phi_j, xijxi_abs, xjxi_abs);
}

/* Riemann::Riemann + Riemann::exact + rotateAndProjectFluxes2D/3D
 * (Riemann.cpp:7-229).  Wi = i-state (the solver's left), Wj = j-state. */
static int riemann_flux(orc *o, const double *Wi_in, const double *Wj_in,
                        const double *vFrame, const double *A, double *F) {
    int D = o->D;
    double g = o->p.gamma;
    double Wi[MAXD + 2], Wj[MAXD + 2];
    memcpy(Wi, Wi_in, sizeof(double) * (D + 2));
    memcpy(Wj, Wj_in, sizeof(double) * (D + 2));
    double An = sqrt(dot(D, A, A));
    double hA[MAXD];
    for (int a = 0; a < D; ++a) hA[a] = 1. / An * A[a];
    double rs, us, Ps, vS[MAXD];
    int it = 0, flag;
    if (D == 2) {
        /* rotationMatrix2D(hatA, e_x), Helper.cpp:39-45 */
        double L0 = hA[0] * 1. + hA[1] * 0.;
        double L1 = -(hA[0] * 0. - hA[1] * 1.);
        double L2 = -L1, L3 = L0;
        double bi0 = Wi[2], bi1 = Wi[3], bj0 = Wj[2], bj1 = Wj[3];
        Wi[2] = L0 * bi0 + L1 * bi1; Wi[3] = L2 * bi0 + L3 * bi1;
        Wj[2] = L0 * bj0 + L1 * bj1; Wj[3] = L2 * bj0 + L3 * bj1;
        flag = orc_riemann_solve(g, Wi[0], Wi[2], Wi[1], Wj[0], Wj[2], Wj[1],
                                 o->p.riemann_tol, o->p.riemann_max_iter, &rs,
                                 &us, &Ps, &it);
        vS[0] = us;
        if (flag == 1) vS[1] = Wj[3];
        else if (flag == -1) vS[1] = Wi[3];
        else vS[1] = 0.;
        /* rotationMatrix2D(e_x, hatA) */
        double K0 = 1. * hA[0] + 0. * hA[1];
        double K1 = -(1. * hA[1] - 0. * hA[0]);
        double K2 = -K1, K3 = K0;
        double s0 = vS[0], s1 = vS[1];
        vS[0] = K0 * s0 + K1 * s1;
        vS[1] = K2 * s0 + K3 * s1;
    } else {
        double ui = dot(D, hA, &Wi[2]), uj = dot(D, hA, &Wj[2]);
        flag = orc_riemann_solve(g, Wi[0], ui, Wi[1], Wj[0], uj, Wj[1],
                                 o->p.riemann_tol, o->p.riemann_max_iter, &rs,
                                 &us, &Ps, &it);
        const double *vup = flag == 1 ? &Wj[2] : &Wi[2];
        double un = flag == 1 ? uj : ui;
        for (int a = 0; a < D; ++a)
            vS[a] = flag == 0 ? us * hA[a] : us * hA[a] + (vup[a] - un * hA[a]);
    }
#pragma omp atomic
    o->diag.newton_iterations += it;
#pragma omp atomic
    o->diag.faces_solved += 1;
    if (flag == 0) {
#pragma omp atomic
        o->diag.vacuum_faces += 1;
    }
    if (it > o->p.riemann_max_iter) {
#pragma omp atomic
        o->diag.nonconverged_faces += 1;
    }
    /* projection onto A, Riemann.cpp:164-181 / :204-227 */
    double vLab[MAXD];
    F[0] = 0.;
    for (int a = 0; a < D; ++a) F[0] += A[a] * rs * vS[a];
    for (int a = 0; a < D; ++a) vLab[a] = vS[a] + vFrame[a];
    if (o->p.mfm)
        for (int a = 0; a < D; ++a) vS[a] = 0.;
    for (int c = 0; c < D; ++c) {
        double s = 0.;
        for (int b = 0; b < D; ++b)
            s += b == c ? A[b] * (rs * vLab[c] * vS[b] + Ps)
                        : A[b] * rs * vLab[c] * vS[b];
        F[2 + c] = s;
    }
    double ekin = dot(D, vLab, vLab);
    double e = 0.;
    for (int b = 0; b < D; ++b)
        e += A[b] * (vS[b] * (Ps / (g - 1.) + rs * .5 * ekin) + Ps * vLab[b]);
    F[1] = e;
    return flag;
}

/* Riemann::Riemann + Riemann::exact (Riemann.cpp:7-229) on n given faces:
 * w_left / w_right = {rho, P, v[D]} of the solver's left (particle i) and
 * right (particle j) states in the frame moving with v_frame. */
void orc_riemann_flux_batch(orc *o, int64_t n, const double *w_left, const double *w_right,
                            const double *v_frame, const double *a_ij, double *f_ij,
                            int32_t *flag) {
    int D = o->D;
    for (int64_t k = 0; k < n; ++k) {
        int fl = riemann_flux(o, &w_left[k * (D + 2)], &w_right[k * (D + 2)], &v_frame[k * D],
                              &a_ij[k * D], &f_ij[k * (D + 2)]);
        if (flag) flag[k] = fl;
    }
}

/* compEffectiveFace (:1290-1311, ghosts :2504-2531), compRiemannStatesLR,
 * solveRiemannProblems (:1787-1911), collectFluxes (:1913-2011) */
int orc_fluxes(orc *o, double dt) {
    if (o->stage < 3) return MLH_ERR_CALL_ORDER;
    int D = o->D, NF = D + 2;
    int per = o->p.periodic;
    int sym = o->p.enforce_flux_sym;
    int64_t neg_real = 0, neg_ghost = 0, missing = 0;
    /* effective faces */
#pragma omp parallel for schedule(static) reduction(+ : missing)
    for (int i = 0; i < o->N; ++i) {
        for (int n = 0; n < o->noi[i]; ++n) {
            int j = o->nnl[(size_t)i * o->maxr + n];
            int ij;
            for (ij = 0; ij < o->noi[j]; ++ij)
                if (o->nnl[(size_t)j * o->maxr + ij] == i) break;
            double *A = &o->A[(o->offr[i] + n) * D];
            const double *pi = &o->psi[(o->offr[i] + n) * D];
            if (ij == o->noi[j]) {
                missing++;
                for (int a = 0; a < D; ++a) A[a] = 1. / o->omega[i] * pi[a];
                continue;
            }
            const double *pj = &o->psi[(o->offr[j] + ij) * D];
            for (int a = 0; a < D; ++a)
                A[a] = 1. / o->omega[i] * pi[a] - 1. / o->omega[j] * pj[a];
        }
        if (!per) continue;
        for (int n = 0; n < o->noig[i]; ++n) {
            int g = o->nnlg[(size_t)i * o->maxg + n];
            int p = o->gpar[g];
            int ij;
            for (ij = 0; ij < o->noig[p]; ++ij)
                if (o->gpar[o->nnlg[(size_t)p * o->maxg + ij]] == i) break;
            double *A = &o->Ag[(o->offg[i] + n) * D];
            const double *pi = &o->psig[(o->offg[i] + n) * D];
            if (ij == o->noig[p]) {
                missing++;
                for (int a = 0; a < D; ++a) A[a] = 1. / o->omega[i] * pi[a];
                continue;
            }
            const double *pj = &o->psig[(o->offg[p] + ij) * D];
            for (int a = 0; a < D; ++a)
                A[a] = 1. / o->omega[i] * pi[a] - 1. / o->omega[p] * pj[a];
        }
    }
    o->diag.missing_mirror_pairs += missing;
    /* faces solved from the lower index (real: j > i; ghost: parent >= i) */
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : neg_real, neg_ghost)
    for (int i = 0; i < o->N; ++i) {
        double xj[MAXD], WR[MAXD + 2], WL[MAXD + 2], vF[MAXD];
        for (int n = 0; n < o->noi[i]; ++n) {
            int j = o->nnl[(size_t)i * o->maxr + n];
            particle_pos(o, j, xj);
            face_states(o, i, j, xj, 0, dt, WR, WL, vF);
            if (WR[1] < 0. || WL[1] < 0.) neg_real++;
            if (sym && j < i) continue;
            riemann_flux(o, WR, WL, vF, &o->A[(o->offr[i] + n) * D],
                         &o->F[(o->offr[i] + n) * NF]);
        }
        if (!per) continue;
        for (int n = 0; n < o->noig[i]; ++n) {
            int g = o->nnlg[(size_t)i * o->maxg + n];
            int p = o->gpar[g];
            ghost_pos(o, g, xj);
            face_states(o, i, p, xj, 1, dt, WR, WL, vF);
            if (WR[1] < 0. || WL[1] < 0.) neg_ghost++;
            if (sym && p < i) continue;
            riemann_flux(o, WR, WL, vF, &o->Ag[(o->offg[i] + n) * D],
                         &o->Fg[(o->offg[i] + n) * NF]);
        }
    }
    o->diag.neg_face_pressure_real += neg_real;
    o->diag.neg_face_pressure_ghost += neg_ghost;
    /* the other side: F_ij = -F_ji via reverse lookup */
    if (sym) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < o->N; ++i) {
            for (int n = 0; n < o->noi[i]; ++n) {
                int j = o->nnl[(size_t)i * o->maxr + n];
                if (j > i) continue;
                int ij;
                for (ij = 0; ij < o->noi[j]; ++ij)
                    if (o->nnl[(size_t)j * o->maxr + ij] == i) break;
                double *F = &o->F[(o->offr[i] + n) * NF];
                if (ij == o->noi[j]) { for (int k = 0; k < NF; ++k) F[k] = 0.; continue; }
                const double *Fr = &o->F[(o->offr[j] + ij) * NF];
                for (int k = 0; k < NF; ++k) F[k] = -Fr[k];
            }
            if (!per) continue;
            for (int n = 0; n < o->noig[i]; ++n) {
                int g = o->nnlg[(size_t)i * o->maxg + n];
                int p = o->gpar[g];
                if (p >= i) continue;
                int ij;
                for (ij = 0; ij < o->noig[p]; ++ij)
                    if (o->gpar[o->nnlg[(size_t)p * o->maxg + ij]] == i) break;
                double *F = &o->Fg[(o->offg[i] + n) * NF];
                if (ij == o->noig[p]) { for (int k = 0; k < NF; ++k) F[k] = 0.; continue; }
                const double *Fr = &o->Fg[(o->offg[p] + ij) * NF];
                for (int k = 0; k < NF; ++k) F[k] = -Fr[k];
            }
        }
    }
    /* collectFluxes */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->N; ++i) {
        double mF = 0., eF = 0., vF[MAXD] = {0., 0., 0.};
        for (int n = 0; n < o->noi[i]; ++n) {
            const double *F = &o->F[(o->offr[i] + n) * NF];
            mF += F[0];
            for (int a = 0; a < D; ++a) vF[a] += F[2 + a];
            eF += F[1];
        }
        if (per)
            for (int n = 0; n < o->noig[i]; ++n) {
                const double *F = &o->Fg[(o->offg[i] + n) * NF];
                mF += F[0];
                for (int a = 0; a < D; ++a) vF[a] += F[2 + a];
                eF += F[1];
            }
        o->mF[i] = mF;
        o->eF[i] = eF;
        for (int a = 0; a < D; ++a) o->vF[(size_t)i * D + a] = vF[a];
    }
    o->stage = 4;
    if (o->p.debug_level >= 1 && neg_ghost > 0) return MLH_ERR_NEGATIVE_FACE_PRESSURE;
    if (o->p.debug_level > 1 && neg_real > 0) return MLH_ERR_NEGATIVE_FACE_PRESSURE;
    return 0;
}

/* The per-pair arrays of the face stages in the C ABI's row layout
 * ([N][maxr + maxg][components], real entries then ghosts): Aij / AijGhosts
 * and Fij / FijGhosts as orc_fluxes stored them, WijR / WijL / vFrame of
 * compRiemannStatesLR with step dt.  Any output may be NULL. */
int orc_get_face_arrays(orc *o, double dt, double *A, double *WR, double *WL, double *vF,
                        double *F) {
    if (o->stage < 4) return MLH_ERR_CALL_ORDER;
    int D = o->D, NF = D + 2;
    size_t S = (size_t)o->maxr + o->maxg;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->N; ++i) {
        double xj[MAXD], wr[MAXD + 2], wl[MAXD + 2], vf[MAXD];
        int cnt = nb_count(o, i);
        for (int n = 0; n < cnt; ++n) {
            int ghost = n >= o->noi[i];
            int j = nb_get(o, i, n, xj);
            size_t e = (size_t)i * S + n;
            const double *Ap = ghost ? &o->Ag[(o->offg[i] + n - o->noi[i]) * D]
                                     : &o->A[(o->offr[i] + n) * D];
            const double *Fp = ghost ? &o->Fg[(o->offg[i] + n - o->noi[i]) * NF]
                                     : &o->F[(o->offr[i] + n) * NF];
            if (A) memcpy(&A[e * D], Ap, sizeof(double) * D);
            if (F) memcpy(&F[e * NF], Fp, sizeof(double) * NF);
            if (WR || WL || vF) {
                face_states(o, i, j, xj, ghost, dt, wr, wl, vf);
                if (WR) memcpy(&WR[e * NF], wr, sizeof(double) * NF);
                if (WL) memcpy(&WL[e * NF], wl, sizeof(double) * NF);
                if (vF) memcpy(&vF[e * D], vf, sizeof(double) * D);
            }
        }
    }
    return 0;
}

/* the ghost particles of the last neighbour search (createGhostParticles):
 * parents, image codes and coordinates in ghost-index order; up to cap */
int orc_get_ghosts(orc *o, int32_t cap, int32_t *parent, int32_t *code, double *gx, double *gy,
                   double *gz, int32_t *n) {
    if (o->stage < 1 || !o->p.periodic) return MLH_ERR_CALL_ORDER;
    *n = o->NG;
    if (o->NG > cap) return MLH_ERR_INVALID_ARGUMENT;
    double *out[3] = {gx, gy, gz};
    for (int k = 0; k < o->NG; ++k) {
        if (parent) parent[k] = o->gpar[k];
        if (code) code[k] = o->gcode[k];
        for (int a = 0; a < o->D; ++a)
            if (out[a]) out[a][k] = o->gx[a][k];
    }
    return 0;
}

/* checkFluxSymmetry (Particles.cpp:2888-2976): per pair, the reverse entry
 * and the one-sided test F_ij[c] + F_ji[c] > tol (ghost pairs: c = 0..3, as
 * the reference's ghost branch).  A missing reverse entry (the reference
 * then reads past the partner's list) is counted apart. */
int orc_check_flux_symmetry(orc *o, double tol, int64_t *asym, int64_t *missing) {
    if (o->stage < 4) return MLH_ERR_CALL_ORDER;
    int D = o->D, NF = D + 2;
    int64_t na = 0, nm = 0;
#pragma omp parallel for schedule(static) reduction(+ : na, nm)
    for (int i = 0; i < o->N; ++i) {
        for (int n = 0; n < o->noi[i]; ++n) {
            int j = o->nnl[(size_t)i * o->maxr + n];
            int ij;
            for (ij = 0; ij < o->noi[j]; ++ij)
                if (o->nnl[(size_t)j * o->maxr + ij] == i) break;
            if (ij == o->noi[j]) { nm++; continue; }
            const double *a = &o->F[(o->offr[i] + n) * NF], *b = &o->F[(o->offr[j] + ij) * NF];
            int bad = 0;
            for (int c = 0; c < NF; ++c) bad |= b[c] + a[c] > tol;
            na += bad;
        }
        if (!o->p.periodic) continue;
        for (int n = 0; n < o->noig[i]; ++n) {
            int p = o->gpar[o->nnlg[(size_t)i * o->maxg + n]];
            int ij;
            for (ij = 0; ij < o->noig[p]; ++ij)
                if (o->gpar[o->nnlg[(size_t)p * o->maxg + ij]] == i) break;
            if (ij == o->noig[p]) { nm++; continue; }
            const double *a = &o->Fg[(o->offg[i] + n) * NF], *b = &o->Fg[(o->offg[p] + ij) * NF];
            int bad = 0;
            for (int c = 0; c < 4; ++c) bad |= b[c] + a[c] > tol;
            na += bad;
        }
    }
    if (asym) *asym = na;
    if (missing) *missing = nm;
    return 0;
}

int orc_get_fluxes(orc *o, double *mF, double *vF, double *eF) {
    if (o->stage < 4) return MLH_ERR_CALL_ORDER;
    size_t nb = sizeof(double) * o->N;
    if (mF) memcpy(mF, o->mF, nb);
    if (vF) memcpy(vF, o->vF, nb * o->D);
    if (eF) memcpy(eF, o->eF, nb);
    return 0;
}

/* updateStateAndPosition, Particles.cpp:2013-2110 */
int orc_update(orc *o, double dt) {
    if (o->stage < 4) return MLH_ERR_CALL_ORDER;
    int D = o->D;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o->N; ++i) {
        double vo[MAXD], Q[MAXD + 1];
        for (int a = 0; a < D; ++a) vo[a] = o->v[a][i];
        double ke = vo[0] * vo[0] + vo[1] * vo[1];
        if (D == 3) ke += vo[2] * vo[2];
        Q[0] = o->m[i] * (o->u[i] + .5 * ke);
        for (int a = 0; a < D; ++a) Q[1 + a] = o->m[i] * vo[a];
        if (!o->p.mfm) o->m[i] -= dt * o->mF[i];
        for (int a = 0; a < D; ++a) Q[1 + a] -= dt * o->vF[(size_t)i * D + a];
        for (int a = 0; a < D; ++a) o->v[a][i] = Q[1 + a] / o->m[i];
        Q[0] -= dt * o->eF[i];
        double kn = o->v[0][i] * o->v[0][i] + o->v[1][i] * o->v[1][i];
        if (D == 3) kn += o->v[2][i] * o->v[2][i];
        o->u[i] = Q[0] / o->m[i] - .5 * kn;
        if (o->p.move_particles) {
            for (int a = 0; a < D; ++a) {
                o->x[a][i] += vo[a] * dt;
                if (o->p.periodic) {
                    double mn = o->p.box_min[a], mx = o->p.box_max[a];
                    if (o->x[a][i] < mn) o->x[a][i] = mx - (mn - o->x[a][i]);
                    else if (mx <= o->x[a][i]) o->x[a][i] = mn + (o->x[a][i] - mx);
                }
            }
        }
    }
    o->stage = 0;
    return 0;
}

int orc_step(orc *o, double dt_fixed, double dt_max, double *dt_used) {
    int st;
    if ((st = orc_neighbours(o))) return st;
    if ((st = orc_density(o))) return st;
    double dt = dt_fixed;
    if (!(dt_fixed > 0.)) {
        if ((st = orc_timestep(o, &dt))) return st;
        if (dt_max > 0. && dt > dt_max) dt = dt_max;
    }
    if ((st = orc_gradients(o))) return st;
    if ((st = orc_fluxes(o, dt))) return st;
    if ((st = orc_update(o, dt))) return st;
    if (dt_used) *dt_used = dt;
    return 0;
}

/* sumVolume, sumMass, sumEnergy, sumMomentumX/Y/Z (Particles.cpp:2830-2886) */
int orc_conserved_sums(orc *o, double *out) {
    int D = o->D;
    double V = 0., M = 0., E = 0., px = 0., py = 0., pz = 0.;
    for (int i = 0; i < o->N; ++i) {
        V += o->stage >= 2 ? 1. / o->omega[i] : 0.;
        M += o->m[i];
        double ke = o->v[0][i] * o->v[0][i] + o->v[1][i] * o->v[1][i];
        if (D == 3) ke += o->v[2][i] * o->v[2][i];
        E += o->m[i] * (o->u[i] + .5 * ke);
        px += o->m[i] * o->v[0][i];
        py += o->m[i] * o->v[1][i];
        if (D == 3) pz += o->m[i] * o->v[2][i];
    }
    out[0] = V; out[1] = M; out[2] = E; out[3] = px; out[4] = py; out[5] = pz;
    return 0;
}

int orc_get_diag(orc *o, mlh_diag *d) {
    *d = o->diag;
    return 0;
}
