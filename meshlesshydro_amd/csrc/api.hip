// api.hip -- the C ABI (include/mlh.h): context, device buffers, stage
// sequencing, readbacks.  Host code; every kernel runs on the context's one
// HIP stream.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>
#include "../../include/mlh.h"
#include "kernels.h"
#include "mlp_api.h"

using namespace mlh;

static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIPCHK(expr)                                                                \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                       \
            return fail(MLH_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

struct mlh_ctx {
    mlh_params p;
    int N = 0, D = 2, prec = 0;
    size_t tsz = 8;  // bytes of the derived-data type
    int stride = 0;  // maxr + maxg
    hipStream_t s = nullptr;
    // periodic domains: the image search's stream and its fork / join events
    hipStream_t s_img = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int *img_cnt = nullptr;
    int img_base = 0;
    Geom g{};
    double *x[3] = {}, *v[3] = {}, *m = nullptr, *u = nullptr;
    double *xn[3] = {}, *vn[3] = {}, *mn = nullptr, *un = nullptr;
    double4 *pos = nullptr;    // cell order
    double4 *pos_w = nullptr;  // caller order (work order = caller order)
    int order = 0;             // resolved work order (MLH_ORDER_CELL / _CALLER), 0 = not yet
    uint32_t *cell_key = nullptr, *sorted_idx = nullptr, *sorted_key = nullptr,
             *sort_scratch = nullptr;
    uint32_t *inv_idx = nullptr;  // original index -> cell-order position
    int *cstart = nullptr, *cend = nullptr;
    int cell_cap = 0;                // cells cstart / cend can hold
    double *lim_scratch = nullptr;   // domain limits (non-periodic)
    double *lim_out = nullptr;
    uint32_t *nbr = nullptr;
    int *noi = nullptr, *nog = nullptr;
    void *prim = nullptr, *frec = nullptr;
    double *dt_block = nullptr, *dt = nullptr, *dt_hist = nullptr;
    int dt_hist_cap = 0;
    double *mF = nullptr, *vF = nullptr, *eF = nullptr;
    unsigned long long *counters = nullptr;
    int *status = nullptr, *max_counts = nullptr;
    double *sums_block = nullptr;
    MlpState mlp;
    int stage = 0;  // 0 state, 1 neighbours, 2 density, 3 gradients, 4 fluxes
    // per-method stages: 1 gradients unlimited (mlh_psi_gradients), 2 limited
    int grad_state = 0;
    // per-entry face arrays (allocated on first use); bits of face_stage:
    // 1 A_ij, 2 W_R / W_L / vFrame, 4 F_ij
    FaceArrays fa{};
    int face_stage = 0;
    double *field_in = nullptr, *field_grad = nullptr;  // mlh_gradient_field
    unsigned long long *sym_counts = nullptr;            // mlh_check_flux_symmetry
    int img_window = 2;
    std::vector<hipEvent_t> evpool;
    float timings[4] = {0, 0, 0, 0};
    std::vector<void *> allocs;
    int64_t step_count = 0;
};

extern "C" {

void mlh_params_default(mlh_params *p) {
    memset(p, 0, sizeof(*p));
    p->dim = 2;
    p->periodic = 1;
    for (int a = 0; a < 3; ++a) { p->box_min[a] = 0.0; p->box_max[a] = 1.0; }
    p->kernel_size = 0.04;
    p->gamma = 5.0 / 3.0;
    p->cfl = 0.4;
    p->max_interactions = 400;
    p->max_ghost_interactions = 300;
    p->slope_limiting = 1;
    p->beta = 4.0;
    p->pairwise_limiter = 0;
    p->psi1 = 0.5;
    p->psi2 = 0.25;
    p->mfm = 0;
    p->enforce_flux_sym = 1;
    p->move_particles = 1;
    p->first_order_quad_point = 1;
    p->debug_level = 1;
    p->limiter_abs = MLH_ABS_INT_X86;
    p->quirk_3d_predictor = 0;
    p->precision = MLH_FP64;
    p->riemann_seed = MLH_SEED_STANDARD;
    p->riemann_tol = 1e-6;
    p->riemann_max_iter = 50;
    p->learned_iters = 3;
    p->work_order = MLH_ORDER_AUTO;
}

int mlh_version(void) { return MLH_ABI_VERSION; }

const char *mlh_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

template <typename P>
static int dalloc(mlh_ctx *c, P **ptr, size_t bytes) {
    void *q = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
        return fail(MLH_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->allocs.push_back(q);
    *ptr = (P *)q;
    return 0;
}

// Domain::createGrid (Domain.cpp:9-54) over [lo, hi]: the periodic box of
// the parameters, or the particles' limits of this step (non-periodic)
static int make_geom(mlh_ctx *c, const double *lo, const double *hi) {
    const mlh_params &p = c->p;
    Geom &g = c->g;
    g.periodic = p.periodic;
    g.h = p.kernel_size;
    g.hsq = p.kernel_size * p.kernel_size;
    double h2 = p.kernel_size / 2.;
    g.sigma = c->D == 2 ? 10. / (7. * M_PI * h2 * h2) : 1. / (M_PI * h2 * h2 * h2);
    long long ncells = 1;
    for (int a = 0; a < 3; ++a) {
        g.bmin[a] = a < c->D ? lo[a] : 0.0;
        g.bmax[a] = a < c->D ? hi[a] : 1.0;
        g.nc[a] = 1;
        g.cs[a] = 1.0;
    }
    for (int a = 0; a < c->D; ++a) {
        double ext = g.bmax[a] - g.bmin[a];
        int n = (int)floor(ext / p.kernel_size);
        if (p.periodic && n < 3)
            return fail(MLH_ERR_GRID, "fewer than 3 search cells per axis (box / kernelSize)");
        if (!(n >= 1))
            return fail(MLH_ERR_GRID, "particle extent below one kernel size on an axis");
        g.nc[a] = n;
        g.cs[a] = ext / (double)n;
        ncells *= n;
    }
    if (ncells > (1ll << 27)) return fail(MLH_ERR_GRID, "too many search cells");
    g.ncells = (int)ncells;
    // Morton keys when their table (2^(D bits) entries) stays within 2^24
    // entries and 8x the cell count; linear keys otherwise
    int nmax = 1;
    for (int a = 0; a < c->D; ++a) nmax = g.nc[a] > nmax ? g.nc[a] : nmax;
    int mb = 0;
    while ((1 << mb) < nmax) ++mb;
    long long mkeys = 1ll << (c->D * mb);
    const int maxbits = c->D == 2 ? 16 : 10;
    if (c->p.work_order != MLH_ORDER_CALLER && mb <= maxbits && mkeys <= (1ll << 24) &&
        mkeys <= 8 * ncells) {
        g.mbits = mb;
        g.nkeys = (int)mkeys;
    } else {
        g.mbits = 0;
        g.nkeys = g.ncells;
    }
    // Two points with computed d^2 < h^2 differ by less than h per axis.
    // Their computed cell indices differ by at most 1 when the cell size
    // exceeds h by more than the rounding of floor((x-min)/cs) and of the
    // image coordinates (~1e-14 relative), so a +-1 window of virtual cells
    // finds every periodic image; otherwise (cs == h up to rounding, e.g. a
    // box of exactly 25 kernel sizes) a +-2 window is used.
    c->img_window = 1;
    for (int a = 0; a < c->D; ++a)
        if (!(g.cs[a] > p.kernel_size * (1.0 + 1e-10))) c->img_window = 2;
    return 0;
}

static HydroArgs hydro_args(mlh_ctx *c) {
    HydroArgs a{};
    const mlh_params &p = c->p;
    a.g = c->g;
    a.N = c->N;
    a.gamma = p.gamma; a.cfl = p.cfl; a.beta = p.beta; a.psi1 = p.psi1; a.psi2 = p.psi2;
    a.riemann_tol = p.riemann_tol;
    a.riemann_max_iter = p.riemann_max_iter;
    a.learned_iters = p.learned_iters;
    a.limiter_abs = p.limiter_abs;
    a.slope_limiting = p.slope_limiting;
    a.pairwise_limiter = p.pairwise_limiter;
    a.mfm = p.mfm;
    a.enforce_flux_sym = p.enforce_flux_sym;
    a.move_particles = p.move_particles;
    a.foqp = p.first_order_quad_point;
    a.quirk3d = p.quirk_3d_predictor;
    a.write_fluxes = 0;
    a.learned = p.riemann_seed == MLH_SEED_LEARNED;
    a.debug = p.debug_level >= 1 ? 1 : 0;
    const bool co = c->order != MLH_ORDER_CELL;
    a.sidx = co ? nullptr : c->sorted_idx;
    a.pos = co ? c->pos_w : c->pos;
    a.nbr = c->nbr;
    a.noi = c->noi;
    a.nog = c->nog;
    for (int k = 0; k < 3; ++k) {
        a.x[k] = c->x[k]; a.v[k] = c->v[k];
        a.xn[k] = c->xn[k]; a.vn[k] = c->vn[k];
    }
    a.m = c->m; a.u = c->u; a.mn = c->mn; a.un = c->un;
    a.prim = c->prim;
    a.frec = c->frec;
    a.dt_block = c->dt_block;
    a.dt = c->dt;
    a.mF = c->mF; a.vF = c->vF; a.eF = c->eF;
    a.counters = c->counters;
    a.status = c->status;
    a.mlp_frags = c->mlp.frags;
    a.collect = c->mlp.collect_stride > 0 ? c->mlp.collect : nullptr;
    a.collect_n = c->mlp.collect_n;
    a.collect_scap = c->mlp.collect_scap;
    a.collect_smask = c->mlp.collect_stripes > 0 ? c->mlp.collect_stripes - 1 : 0;
    a.collect_stride = c->mlp.collect_stride > 0 ? c->mlp.collect_stride : 1;
    return a;
}

extern "C" {

int mlh_create(const mlh_params *p, int32_t n, mlh_ctx **out) {
    if (!p || !out || n <= 0) return fail(MLH_ERR_INVALID_ARGUMENT, "null argument or n <= 0");
    if (p->dim != 2 && p->dim != 3) return fail(MLH_ERR_INVALID_ARGUMENT, "dim must be 2 or 3");
    if (n > (int32_t)kIdxMask) return fail(MLH_ERR_INVALID_ARGUMENT, "n exceeds 2^27-1");
    if (p->max_interactions <= 0 || p->max_ghost_interactions < 0)
        return fail(MLH_ERR_INVALID_ARGUMENT, "bad interaction capacities");
    if (p->precision != MLH_FP64 && p->precision != MLH_FP32)
        return fail(MLH_ERR_INVALID_ARGUMENT, "bad precision");
    if (p->work_order < MLH_ORDER_AUTO || p->work_order > MLH_ORDER_CALLER)
        return fail(MLH_ERR_INVALID_ARGUMENT, "bad work_order");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(MLH_ERR_DEVICE, "no HIP device");
    mlh_ctx *c = new mlh_ctx();
    c->p = *p;
    c->N = n;
    c->D = p->dim;
    c->prec = p->precision;
    c->tsz = c->prec == MLH_FP64 ? 8 : 4;
    c->stride = p->max_interactions + p->max_ghost_interactions;
    int st = p->periodic ? make_geom(c, p->box_min, p->box_max) : MLH_OK;
    if (st) { delete c; return st; }
    if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(MLH_ERR_DEVICE, "hipStreamCreate failed");
    }
    if (p->periodic &&
        (hipStreamCreateWithFlags(&c->s_img, hipStreamNonBlocking) != hipSuccess ||
         hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
         hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)) {
        mlh_destroy(c);
        return fail(MLH_ERR_DEVICE, "hipStreamCreate / hipEventCreate failed");
    }
    c->img_base = image_base(p->max_interactions);
    size_t N = (size_t)n, dsz = sizeof(double) * N;
    int D = c->D;
    int frn = D == 2 ? FRec<2>::N : FRec<3>::N;
#define A_(ptr, bytes) if ((st = dalloc(c, ptr, bytes))) { mlh_destroy(c); return st; }
    for (int a = 0; a < D; ++a) {
        A_(&c->x[a], dsz); A_(&c->v[a], dsz); A_(&c->xn[a], dsz); A_(&c->vn[a], dsz);
    }
    A_(&c->m, dsz); A_(&c->u, dsz); A_(&c->mn, dsz); A_(&c->un, dsz);
    A_(&c->pos, sizeof(double4) * N);
    A_(&c->pos_w, sizeof(double4) * N);
    A_(&c->cell_key, 4 * N); A_(&c->sorted_idx, 4 * N); A_(&c->sorted_key, 4 * N);
    A_(&c->inv_idx, 4 * N);
    A_(&c->sort_scratch, 4 * (2 * N + radix_sort_scratch_words(n)));
    if (p->periodic) {
        A_(&c->cstart, 4 * (size_t)c->g.nkeys); A_(&c->cend, 4 * (size_t)c->g.nkeys);
        c->cell_cap = c->g.nkeys;
    } else {
        A_(&c->lim_scratch, sizeof(double) * 3 * D * (size_t)limits_blocks(n));
        A_(&c->lim_out, sizeof(double) * 2 * D);
    }
    // four guard rows past the list: the search stores up to four hits past
    // MAX_NUM_INTERACTIONS before it detects the overflow; periodic domains
    // also hold the image search's entries from img_base on
    {
        int rows = c->stride + 4;
        if (p->periodic && c->img_base + p->max_ghost_interactions > rows)
            rows = c->img_base + p->max_ghost_interactions;
        A_(&c->nbr, 4 * nb_words((int)N, rows));
    }
    if (p->periodic) A_(&c->img_cnt, 4 * N);
    A_(&c->noi, 4 * N); A_(&c->nog, 4 * N);
    A_(&c->prim, c->tsz * kPrimN * N);
    A_(&c->frec, c->tsz * frn * N);
    A_(&c->dt_block, sizeof(double) * dt_blocks(n));
    A_(&c->dt, sizeof(double) * 4);
    A_(&c->mF, dsz); A_(&c->vF, dsz * D); A_(&c->eF, dsz);
    A_(&c->counters, sizeof(unsigned long long) * kNumCounters);
    A_(&c->status, sizeof(int) * 2);
    A_(&c->max_counts, sizeof(int) * 2);
    A_(&c->sums_block, sizeof(double) * 6 * (size_t)dt_blocks(n));
#undef A_
    hipMemsetAsync(c->counters, 0, sizeof(unsigned long long) * kNumCounters, c->s);
    hipMemsetAsync(c->status, 0, sizeof(int), c->s);
    hipMemsetAsync(c->status + 1, 0xff, sizeof(int), c->s);
    hipMemsetAsync(c->max_counts, 0, sizeof(int) * 2, c->s);
    hipMemsetAsync(c->noi, 0, 4 * N, c->s);
    hipMemsetAsync(c->nog, 0, 4 * N, c->s);
    // the cell order starts as the identity permutation, so that every
    // kernel indexing through it stays in bounds before the first sort
    std::vector<uint32_t> iota(N);
    for (size_t q = 0; q < N; ++q) iota[q] = (uint32_t)q;
    hipMemcpy(c->sorted_idx, iota.data(), 4 * N, hipMemcpyHostToDevice);
    if (hipStreamSynchronize(c->s) != hipSuccess) {
        mlh_destroy(c);
        return fail(MLH_ERR_DEVICE, "initialisation failed");
    }
    *out = c;
    return MLH_OK;
}

int mlh_destroy(mlh_ctx *c) {
    if (!c) return MLH_OK;
    if (c->s) hipStreamSynchronize(c->s);
    for (void *q : c->allocs) hipFree(q);
    if (!c->p.periodic) {  // grown per step, not in allocs
        if (c->cstart) hipFree(c->cstart);
        if (c->cend) hipFree(c->cend);
    }
    mlp_free(c->mlp);
    for (hipEvent_t e : c->evpool) hipEventDestroy(e);
    if (c->dt_hist) hipFree(c->dt_hist);
    if (c->s_img) hipStreamSynchronize(c->s_img);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->s_img) hipStreamDestroy(c->s_img);
    if (c->s) hipStreamDestroy(c->s);
    delete c;
    return MLH_OK;
}

int mlh_upload_state(mlh_ctx *c, const double *x, const double *y, const double *z,
                     const double *vx, const double *vy, const double *vz, const double *m,
                     const double *u) {
    if (!c || !x || !y || !vx || !vy || !m || !u || (c->D == 3 && (!z || !vz)))
        return fail(MLH_ERR_INVALID_ARGUMENT, "null state array");
    size_t b = sizeof(double) * c->N;
    const double *xs[3] = {x, y, z}, *vs[3] = {vx, vy, vz};
    for (int a = 0; a < c->D; ++a) {
        HIPCHK(hipMemcpyAsync(c->x[a], xs[a], b, hipMemcpyHostToDevice, c->s));
        HIPCHK(hipMemcpyAsync(c->v[a], vs[a], b, hipMemcpyHostToDevice, c->s));
    }
    HIPCHK(hipMemcpyAsync(c->m, m, b, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->u, u, b, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->stage = 0;
    if (c->p.work_order == MLH_ORDER_AUTO) c->order = 0;  // settled at the next search
    return MLH_OK;
}

int mlh_download_state(mlh_ctx *c, double *x, double *y, double *z, double *vx, double *vy,
                       double *vz, double *m, double *u) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    size_t b = sizeof(double) * c->N;
    double *xs[3] = {x, y, z}, *vs[3] = {vx, vy, vz};
    for (int a = 0; a < c->D; ++a) {
        if (xs[a]) HIPCHK(hipMemcpyAsync(xs[a], c->x[a], b, hipMemcpyDeviceToHost, c->s));
        if (vs[a]) HIPCHK(hipMemcpyAsync(vs[a], c->v[a], b, hipMemcpyDeviceToHost, c->s));
    }
    if (m) HIPCHK(hipMemcpyAsync(m, c->m, b, hipMemcpyDeviceToHost, c->s));
    if (u) HIPCHK(hipMemcpyAsync(u, c->u, b, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return MLH_OK;
}

}  // extern "C"

// ---- stage launchers (asynchronous) ----------------------------------------
// MLH_ORDER_AUTO resolves to cell order: with Morton-ordered cell keys it
// is the faster order on every input measured (lattices in their natural
// order included) and it does not depend on the caller's particle order
static void resolve_order(mlh_ctx *c) {
    c->order = c->p.work_order == MLH_ORDER_AUTO ? MLH_ORDER_CELL : c->p.work_order;
}

// non-periodic: the domain limits of this step (one synchronisation), the
// grid over them, and cell-range buffers of that size
static int prepare_open_grid(mlh_ctx *c) {
    double lim[6];
    launch_domain_limits(c->x[0], c->x[1], c->D == 3 ? c->x[2] : c->x[0], c->D, c->N,
                         c->lim_scratch, c->lim_out, c->s);
    HIPCHK(hipMemcpyAsync(lim, c->lim_out, sizeof(double) * 2 * c->D, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    double lo[3], hi[3];
    for (int a = 0; a < c->D; ++a) {
        lo[a] = lim[2 * a];
        hi[a] = lim[2 * a + 1];
    }
    int st = make_geom(c, lo, hi);
    if (st) return st;
    if (c->g.nkeys > c->cell_cap) {
        if (c->cstart) hipFree(c->cstart);
        if (c->cend) hipFree(c->cend);
        c->cstart = c->cend = nullptr;
        HIPCHK(hipMalloc(&c->cstart, 4 * (size_t)c->g.nkeys));
        HIPCHK(hipMalloc(&c->cend, 4 * (size_t)c->g.nkeys));
        c->cell_cap = c->g.nkeys;
    }
    return MLH_OK;
}

// fuse_density: the step's compDensity + compPressure run inside the search
// (the same records, bitwise, as mlh_density's k_density)
static int enqueue_neighbours(mlh_ctx *c, bool fuse_density = false) {
    if (!c->p.periodic) {
        int st = prepare_open_grid(c);
        if (st) return st;
    }
    launch_cells(c->g, c->D, c->x[0], c->x[1], c->D == 3 ? c->x[2] : c->x[0], c->N, c->pos,
                 c->cell_key, c->sorted_idx, c->sorted_key, c->sort_scratch, c->cstart, c->cend,
                 c->inv_idx, c->s);
    if (c->order == 0) resolve_order(c);
    const bool co = c->order == MLH_ORDER_CALLER;
    if (co)
        launch_pack_pos(c->x[0], c->x[1], c->D == 3 ? c->x[2] : c->x[0], c->D, c->N, c->pos_w,
                        c->s);
    NbrArgs a{};
    a.g = c->g;
    a.N = c->N;
    a.maxr = c->p.max_interactions;
    a.maxg = c->p.max_ghost_interactions;
    a.pos_cell = c->pos;
    a.pos_w = co ? c->pos_w : c->pos;
    a.ck_w = co ? c->cell_key : c->sorted_key;
    a.caller_order = co ? 1 : 0;
    a.cstart = c->cstart;
    a.cend = c->cend;
    a.inv = c->inv_idx;
    a.nbr = c->nbr;
    a.noi = c->noi;
    a.nog = c->nog;
    a.status = c->status;
    a.max_counts = c->max_counts;
    a.img_window = c->img_window;
    a.img_cnt = c->img_cnt;
    a.img_base = c->img_base;
    a.dens = fuse_density ? (c->prec == MLH_FP32 ? 4 : 8) : 0;
    if (fuse_density) {
        a.prim = c->prim;
        a.m = c->m;
        a.u = c->u;
        for (int k = 0; k < 3; ++k) a.v[k] = c->v[k];
        a.gamma = c->p.gamma;
    }
    const bool fused = launch_neighbours(a, c->D, c->s, c->s_img, c->ev_fork, c->ev_join);
    if (fuse_density && !fused) launch_density(hydro_args(c), c->D, c->prec, c->s);
    c->face_stage = 0;
    c->grad_state = 0;
    // DEBUG_LVL >= 1: every entry's partner must list the particle (the face
    // kernel's F_ji = -F_ij relies on it); misses go to missing_mirror_pairs
    if (c->p.debug_level >= 1) launch_mirror_check(hydro_args(c), c->s);
    return MLH_OK;
}

// work position -> caller's particle index
static int read_order(mlh_ctx *c, std::vector<uint32_t> &perm) {
    perm.resize(c->N);
    if (c->order != MLH_ORDER_CELL) {
        for (int q = 0; q < c->N; ++q) perm[q] = (uint32_t)q;
        return MLH_OK;
    }
    HIPCHK(hipMemcpy(perm.data(), c->sorted_idx, 4 * (size_t)c->N, hipMemcpyDeviceToHost));
    return MLH_OK;
}

// status words of a step or run: cleared here, then error bits are sticky
// until the next call (a run does not lose an overflow of an early step)
static int clear_status(mlh_ctx *c) {
    HIPCHK(hipMemsetAsync(c->status, 0, sizeof(int), c->s));
    HIPCHK(hipMemsetAsync(c->status + 1, 0xff, sizeof(int), c->s));
    return MLH_OK;
}

static int check_status(mlh_ctx *c) {
    int st[2];
    HIPCHK(hipMemcpyAsync(st, c->status, sizeof(st), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipGetLastError());
    // a run names the step (the state is the one at its start: the face
    // kernel of that step and of every later one keeps the state unchanged)
    const std::string at =
        st[1] >= 0 ? " in step " + std::to_string(st[1]) + " of the run (state kept from its start)"
                   : "";
    if (st[0] & (1 << 1))
        return fail(MLH_ERR_NEIGHBOUR_OVERFLOW, "MAX_NUM_INTERACTIONS exceeded" + at);
    if (st[0] & (1 << 3))
        return fail(MLH_ERR_GHOST_OVERFLOW, "MAX_NUM_GHOST_INTERACTIONS exceeded" + at);
    return MLH_OK;
}

static void swap_state(mlh_ctx *c) {
    for (int a = 0; a < 3; ++a) {
        std::swap(c->x[a], c->xn[a]);
        std::swap(c->v[a], c->vn[a]);
    }
    std::swap(c->m, c->mn);
    std::swap(c->u, c->un);
}

static int read_counters(mlh_ctx *c, unsigned long long *out) {
    HIPCHK(hipMemcpyAsync(out, c->counters, sizeof(unsigned long long) * kNumCounters,
                          hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return MLH_OK;
}

// negative face pressure aborts as in the reference (Particles.cpp:1824-1827,
// 1878-1881)
static int face_pressure_status(mlh_ctx *c, const unsigned long long *before,
                                const unsigned long long *after) {
    unsigned long long nr = after[kCntNegReal] - before[kCntNegReal];
    unsigned long long ng = after[kCntNegGhost] - before[kCntNegGhost];
    if ((c->p.debug_level >= 1 && ng > 0) || (c->p.debug_level > 1 && nr > 0))
        return fail(MLH_ERR_NEGATIVE_FACE_PRESSURE, "negative reconstructed face pressure");
    return MLH_OK;
}

static int check_learned(mlh_ctx *c) {
    if (c->p.riemann_seed == MLH_SEED_LEARNED && !c->mlp.frags)
        return fail(MLH_ERR_NOT_TRAINED, "learned seed requested but no network parameters set");
    return MLH_OK;
}

static int ensure_events(mlh_ctx *c, int nsteps) {
    size_t need = (size_t)5 * (nsteps > 0 ? nsteps : 1);
    while (c->evpool.size() < need) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c->evpool.push_back(e);
    }
    return MLH_OK;
}

static int enqueue_step(mlh_ctx *c, double dt_fixed, double t_now, double t_clip,
                         const double *dt_in, double *dt_hist, int step, hipEvent_t *ev) {
    if (ev) hipEventRecord(ev[0], c->s);
    // the search with the density fused in (compDensity + compPressure)
    int st = enqueue_neighbours(c, true);
    if (st) return st;
    // after the search: it settles the work order (and, without periodic
    // boundaries, the grid) the arguments depend on
    HydroArgs a = hydro_args(c);
    if (ev) hipEventRecord(ev[1], c->s);
    if (ev) hipEventRecord(ev[2], c->s);
    launch_gradients(a, c->D, c->prec, c->s);
    launch_dt_finalize(a, dt_fixed, t_now, t_clip, dt_in, dt_hist, step, c->s);
    if (ev) hipEventRecord(ev[3], c->s);
    launch_faces(a, c->D, c->prec, c->s);
    if (ev) hipEventRecord(ev[4], c->s);
    swap_state(c);
    return MLH_OK;
}

// average device time per step of each kernel group over nsteps steps
static void collect_timings(mlh_ctx *c, int nsteps) {
    for (int k = 0; k < 4; ++k) c->timings[k] = 0.f;
    for (int st = 0; st < nsteps; ++st)
        for (int k = 0; k < 4; ++k) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, c->evpool[5 * st + k], c->evpool[5 * st + k + 1]);
            c->timings[k] += ms / (float)nsteps;
        }
}

extern "C" {

int mlh_step(mlh_ctx *c, double dt_fixed, double t_now, double t_clip, double *dt_used,
             int32_t *clipped) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    int st = check_learned(c);
    if (st) return st;
    unsigned long long before[kNumCounters], after[kNumCounters];
    if ((st = read_counters(c, before))) return st;
    if ((st = ensure_events(c, 1))) return st;
    if ((st = clear_status(c))) return st;
    if ((st = enqueue_step(c, dt_fixed, t_now, t_clip, nullptr, nullptr, 0, c->evpool.data())))
        return st;
    double dtv[3];
    HIPCHK(hipMemcpyAsync(dtv, c->dt, sizeof(dtv), hipMemcpyDeviceToHost, c->s));
    if ((st = check_status(c))) { c->stage = 0; return st; }
    collect_timings(c, 1);
    if (dt_used) *dt_used = dtv[0];
    if (clipped) *clipped = dtv[2] != 0.;
    c->stage = 0;
    c->step_count++;
    if ((st = read_counters(c, after))) return st;
    return face_pressure_status(c, before, after);
}

int mlh_run(mlh_ctx *c, int32_t n_steps, double dt_fixed, const double *dts_in,
            double *dts_out) {
    if (!c || n_steps < 0) return fail(MLH_ERR_INVALID_ARGUMENT, "bad arguments");
    int st = check_learned(c);
    if (st) return st;
    if (c->dt_hist_cap < n_steps) {
        if (c->dt_hist) hipFree(c->dt_hist);
        c->dt_hist = nullptr;
        HIPCHK(hipMalloc(&c->dt_hist, sizeof(double) * 2 * (size_t)(n_steps > 0 ? n_steps : 1)));
        c->dt_hist_cap = n_steps;
    }
    double *dev_in = nullptr;
    if (dts_in && n_steps > 0) {
        dev_in = c->dt_hist + c->dt_hist_cap;
        HIPCHK(hipMemcpyAsync(dev_in, dts_in, sizeof(double) * n_steps, hipMemcpyHostToDevice, c->s));
    }
    if ((st = ensure_events(c, n_steps))) return st;
    unsigned long long before[kNumCounters], after[kNumCounters];
    if ((st = read_counters(c, before))) return st;
    if ((st = clear_status(c))) return st;
    for (int k = 0; k < n_steps; ++k)
        if ((st = enqueue_step(c, dt_fixed, -1.0, -1.0, dev_in, c->dt_hist, k, &c->evpool[5 * k])))
            return st;
    if ((st = check_status(c))) { c->stage = 0; return st; }
    if (n_steps > 0) collect_timings(c, n_steps);
    if (dts_out && n_steps > 0)
        HIPCHK(hipMemcpy(dts_out, c->dt_hist, sizeof(double) * n_steps, hipMemcpyDeviceToHost));
    c->stage = 0;
    c->step_count += n_steps;
    if ((st = read_counters(c, after))) return st;
    return face_pressure_status(c, before, after);
}

int mlh_neighbours(mlh_ctx *c) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    int st = clear_status(c);
    if (!st) st = enqueue_neighbours(c);
    if (!st) st = check_status(c);
    c->stage = st ? 0 : 1;
    return st;
}

int mlh_density(mlh_ctx *c) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if (c->stage < 1) return fail(MLH_ERR_CALL_ORDER, "mlh_neighbours first");
    HydroArgs a = hydro_args(c);
    launch_density(a, c->D, c->prec, c->s);
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipGetLastError());
    c->stage = 2;
    return MLH_OK;
}

static int gradients_stage(mlh_ctx *c, bool limit) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if (c->stage < 2) return fail(MLH_ERR_CALL_ORDER, "mlh_density first");
    HydroArgs a = hydro_args(c);
    if (!limit) a.slope_limiting = 0;
    launch_gradients(a, c->D, c->prec, c->s);
    launch_dt_finalize(a, -1.0, -1.0, -1.0, nullptr, nullptr, 0, c->s);
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipGetLastError());
    c->stage = 3;
    c->grad_state = limit ? 2 : 1;
    c->face_stage = 0;
    return MLH_OK;
}

int mlh_gradients(mlh_ctx *c) { return gradients_stage(c, true); }

int mlh_psi_gradients(mlh_ctx *c) { return gradients_stage(c, false); }

int mlh_limit(mlh_ctx *c) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if (c->stage < 3 || c->grad_state != 1)
        return fail(MLH_ERR_CALL_ORDER, "mlh_psi_gradients first (gradients not yet limited)");
    launch_stage(hydro_args(c), c->D, c->prec, kOpLimit, 0, c->fa, nullptr, nullptr, c->s);
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipGetLastError());
    c->grad_state = 2;
    c->face_stage = 0;
    return MLH_OK;
}

int mlh_timestep(mlh_ctx *c, double *dt) {
    if (!c || !dt) return fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    if (c->stage < 2) return fail(MLH_ERR_CALL_ORDER, "mlh_density first");
    // the CFL step is pass 1 of the gradient kernel: run it unless a
    // gradient stage (limited or not) already has
    if (c->stage < 3) {
        int st = mlh_gradients(c);
        if (st) return st;
    }
    double dtv[3];
    HIPCHK(hipMemcpy(dtv, c->dt, sizeof(dtv), hipMemcpyDeviceToHost));
    *dt = dtv[1];
    return MLH_OK;
}

int mlh_fluxes(mlh_ctx *c, double dt) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if (c->stage < 3) return fail(MLH_ERR_CALL_ORDER, "mlh_gradients first");
    int st = check_learned(c);
    if (st) return st;
    unsigned long long before[kNumCounters], after[kNumCounters];
    if ((st = read_counters(c, before))) return st;
    HIPCHK(hipMemcpyAsync(c->dt, &dt, sizeof(double), hipMemcpyHostToDevice, c->s));
    HydroArgs a = hydro_args(c);
    a.write_fluxes = 1;
    launch_faces(a, c->D, c->prec, c->s);
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipGetLastError());
    c->stage = 4;
    if ((st = read_counters(c, after))) return st;
    return face_pressure_status(c, before, after);
}

int mlh_update(mlh_ctx *c) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if (c->stage < 4) return fail(MLH_ERR_CALL_ORDER, "mlh_fluxes first");
    swap_state(c);
    c->stage = 0;
    return MLH_OK;
}

}  // extern "C"

// ---- the per-method face stages ---------------------------------------------
// per-entry FP64 arrays, allocated on first use: A, vF [D], W_R, W_L, F [D+2]
// per row entry
static int ensure_face_arrays(mlh_ctx *c) {
    if (c->fa.A) return MLH_OK;
    const size_t ent = (size_t)c->stride * c->N * sizeof(double);
    const int D = c->D;
    int st;
    if ((st = dalloc(c, &c->fa.A, ent * D)) || (st = dalloc(c, &c->fa.vF, ent * D)) ||
        (st = dalloc(c, &c->fa.WR, ent * (D + 2))) || (st = dalloc(c, &c->fa.WL, ent * (D + 2))) ||
        (st = dalloc(c, &c->fa.F, ent * (D + 2)))) {
        c->fa = FaceArrays{};
        return st;
    }
    c->fa.rowlen = c->stride;
    return MLH_OK;
}

static int face_stage_run(mlh_ctx *c, int op, int what, const HydroArgs &a) {
    launch_stage(a, c->D, c->prec, op, what, c->fa, nullptr, nullptr, c->s);
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipGetLastError());
    return MLH_OK;
}

extern "C" {

int mlh_gradient_field(mlh_ctx *c, const double *f, double *grad) {
    if (!c || !f || !grad) return fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    if (c->stage < 3) return fail(MLH_ERR_CALL_ORDER, "a gradient stage first (psi-tilde)");
    if (!c->field_in) {
        int st;
        if ((st = dalloc(c, &c->field_in, sizeof(double) * c->N)) ||
            (st = dalloc(c, &c->field_grad, sizeof(double) * c->N * c->D)))
            return st;
    }
    HIPCHK(hipMemcpyAsync(c->field_in, f, sizeof(double) * c->N, hipMemcpyHostToDevice, c->s));
    launch_stage(hydro_args(c), c->D, c->prec, kOpGradField, 0, c->fa, c->field_in, c->field_grad,
                 c->s);
    HIPCHK(hipMemcpyAsync(grad, c->field_grad, sizeof(double) * c->N * c->D,
                          hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipGetLastError());
    return MLH_OK;
}

int mlh_effective_faces(mlh_ctx *c) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if (c->stage < 3) return fail(MLH_ERR_CALL_ORDER, "a gradient stage first (psi-tilde)");
    int st = ensure_face_arrays(c);
    if (!st) st = face_stage_run(c, kOpFaceStates, 1, hydro_args(c));
    if (st) return st;
    c->face_stage = (c->face_stage & 2) | 1;
    return MLH_OK;
}

int mlh_riemann_states(mlh_ctx *c, double dt) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if (c->stage < 3) return fail(MLH_ERR_CALL_ORDER, "the gradient stages first");
    int st = ensure_face_arrays(c);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(c->dt, &dt, sizeof(double), hipMemcpyHostToDevice, c->s));
    if ((st = face_stage_run(c, kOpFaceStates, 2, hydro_args(c)))) return st;
    c->face_stage = (c->face_stage & 1) | 2;
    return MLH_OK;
}

int mlh_solve_faces(mlh_ctx *c) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if ((c->face_stage & 3) != 3)
        return fail(MLH_ERR_CALL_ORDER, "mlh_effective_faces and mlh_riemann_states first");
    unsigned long long before[kNumCounters], after[kNumCounters];
    int st = read_counters(c, before);
    const HydroArgs a = hydro_args(c);
    if (!st) st = face_stage_run(c, kOpSolveOwn, 0, a);
    if (!st) st = face_stage_run(c, kOpSolveMirror, 0, a);
    if (!st) st = read_counters(c, after);
    if (st) return st;
    c->face_stage |= 4;
    return face_pressure_status(c, before, after);
}

int mlh_collect_fluxes(mlh_ctx *c) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if (!(c->face_stage & 4)) return fail(MLH_ERR_CALL_ORDER, "mlh_solve_faces first");
    HydroArgs a = hydro_args(c);
    a.write_fluxes = 1;
    int st = face_stage_run(c, kOpCollect, 0, a);
    if (st) return st;
    c->stage = 4;
    return MLH_OK;
}

int mlh_check_flux_symmetry(mlh_ctx *c, double tol, int64_t *asymmetric, int64_t *missing) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if (!(c->face_stage & 4)) return fail(MLH_ERR_CALL_ORDER, "mlh_solve_faces first");
    if (!c->sym_counts) {
        int st = dalloc(c, &c->sym_counts, sizeof(unsigned long long) * 2);
        if (st) return st;
    }
    HIPCHK(hipMemsetAsync(c->sym_counts, 0, sizeof(unsigned long long) * 2, c->s));
    launch_flux_check(hydro_args(c), c->D, c->fa, tol, c->sym_counts, c->s);
    unsigned long long h[2];
    HIPCHK(hipMemcpyAsync(h, c->sym_counts, sizeof(h), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipGetLastError());
    if (asymmetric) *asymmetric = (int64_t)h[0];
    if (missing) *missing = (int64_t)h[1];
    return MLH_OK;
}

int mlh_get_neighbours(mlh_ctx *c, int32_t *noi, int32_t *noi_ghost, int32_t *nbr) {
    if (!c || !noi || !noi_ghost || !nbr) return fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    if (c->stage < 1) return fail(MLH_ERR_CALL_ORDER, "no neighbour lists yet");
    size_t N = c->N;
    // device rows and entries are in cell order: map both back to the
    // caller's particle indices
    std::vector<uint32_t> perm;
    int st = read_order(c, perm);
    if (st) return st;
    std::vector<int32_t> cn(N), cg(N);
    HIPCHK(hipMemcpy(cn.data(), c->noi, 4 * N, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cg.data(), c->nog, 4 * N, hipMemcpyDeviceToHost));
    std::vector<uint32_t> t(nb_words((int)N, c->stride));
    HIPCHK(hipMemcpy(t.data(), c->nbr, 4 * t.size(), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < N; ++q) {
        size_t i = perm[q];
        noi[i] = cn[q];
        noi_ghost[i] = cg[q];
        int cnt = cn[q] + cg[q];
        for (int n = 0; n < c->stride; ++n) {
            uint32_t e = t[nb_slot((int)N, (int)q, n)];
            nbr[i * c->stride + n] =
                n < cnt ? (int32_t)(perm[nb_index(e)] | (nb_code(e) << kIdxBits)) : 0;
        }
    }
    return MLH_OK;
}

}  // extern "C"

template <typename T>
static void to_double(const void *src, double *dst, size_t n) {
    const T *s = (const T *)src;
    for (size_t k = 0; k < n; ++k) dst[k] = (double)s[k];
}

static int read_derived(mlh_ctx *c, void *dev, size_t count, std::vector<double> &out) {
    std::vector<unsigned char> raw(count * c->tsz);
    HIPCHK(hipMemcpy(raw.data(), dev, raw.size(), hipMemcpyDeviceToHost));
    out.resize(count);
    if (c->tsz == 8) to_double<double>(raw.data(), out.data(), count);
    else to_double<float>(raw.data(), out.data(), count);
    return MLH_OK;
}

extern "C" {

// one per-entry array [nc][stride][N] (work order) -> [N][stride][nc] by the
// caller's particle index, zero past each row's entries
static int read_entry_array(mlh_ctx *c, const double *dev, int nc, double *out,
                            const std::vector<uint32_t> &perm, const std::vector<int32_t> &cnt) {
    const size_t N = c->N, S = c->stride;
    std::vector<double> h(nc * S * N);
    HIPCHK(hipMemcpy(h.data(), dev, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < N; ++q) {
        const size_t i = perm[q];
        for (size_t n = 0; n < S; ++n)
            for (int k = 0; k < nc; ++k)
                out[(i * S + n) * nc + k] = (int)n < cnt[q] ? h[((size_t)k * S + n) * N + q] : 0.;
    }
    return MLH_OK;
}

int mlh_get_face_arrays(mlh_ctx *c, double *Aij, double *WR, double *WL, double *vFrame,
                        double *Fij) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if ((Aij && !(c->face_stage & 1)) || ((WR || WL || vFrame) && !(c->face_stage & 2)) ||
        (Fij && !(c->face_stage & 4)))
        return fail(MLH_ERR_CALL_ORDER, "the face stage of a requested array has not run");
    std::vector<uint32_t> perm;
    int st = read_order(c, perm);
    if (st) return st;
    const size_t N = c->N;
    std::vector<int32_t> cn(N), cg(N);
    HIPCHK(hipMemcpy(cn.data(), c->noi, 4 * N, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cg.data(), c->nog, 4 * N, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < N; ++q) cn[q] += cg[q];
    const int D = c->D;
    if (Aij && (st = read_entry_array(c, c->fa.A, D, Aij, perm, cn))) return st;
    if (WR && (st = read_entry_array(c, c->fa.WR, D + 2, WR, perm, cn))) return st;
    if (WL && (st = read_entry_array(c, c->fa.WL, D + 2, WL, perm, cn))) return st;
    if (vFrame && (st = read_entry_array(c, c->fa.vF, D, vFrame, perm, cn))) return st;
    if (Fij && (st = read_entry_array(c, c->fa.F, D + 2, Fij, perm, cn))) return st;
    return MLH_OK;
}

int mlh_get_density(mlh_ctx *c, double *rho, double *P, double *omega) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if (c->stage < 2) return fail(MLH_ERR_CALL_ORDER, "no density yet");
    std::vector<double> pr;
    int st = read_derived(c, c->prim, (size_t)c->N * kPrimN, pr);
    if (st) return st;
    std::vector<uint32_t> perm;
    if ((st = read_order(c, perm))) return st;
    for (int q = 0; q < c->N; ++q) {
        size_t i = perm[q];
        if (rho) rho[i] = pr[(size_t)q * kPrimN + kPrimRho];
        if (P) P[i] = pr[(size_t)q * kPrimN + kPrimP];
        if (omega) omega[i] = pr[(size_t)q * kPrimN + kPrimOmega];
    }
    return MLH_OK;
}

int mlh_get_gradients(mlh_ctx *c, double *grads, double *B) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if (c->stage < 3) return fail(MLH_ERR_CALL_ORDER, "no gradients yet");
    int D = c->D, NF = D + 2, frn = D == 2 ? FRec<2>::N : FRec<3>::N;
    std::vector<double> fr;
    int st = read_derived(c, c->frec, (size_t)c->N * frn, fr);
    if (st) return st;
    std::vector<uint32_t> perm;
    if ((st = read_order(c, perm))) return st;
    size_t N = c->N;
    for (size_t q = 0; q < N; ++q) {
        const double *r = &fr[q * frn];
        size_t i = perm[q];
        if (grads)
            for (int f = 0; f < NF; ++f)
                for (int a = 0; a < D; ++a) grads[((size_t)f * N + i) * D + a] = r[f * D + a];
        if (B)
            for (int k = 0; k < D * D; ++k) B[i * D * D + k] = r[NF * D + k];
    }
    return MLH_OK;
}

int mlh_get_fluxes(mlh_ctx *c, double *mF, double *vF, double *eF) {
    if (!c) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx");
    if (c->stage < 4) return fail(MLH_ERR_CALL_ORDER, "no fluxes yet");
    size_t b = sizeof(double) * c->N;
    if (mF) HIPCHK(hipMemcpy(mF, c->mF, b, hipMemcpyDeviceToHost));
    if (vF) HIPCHK(hipMemcpy(vF, c->vF, b * c->D, hipMemcpyDeviceToHost));
    if (eF) HIPCHK(hipMemcpy(eF, c->eF, b, hipMemcpyDeviceToHost));
    return MLH_OK;
}

int mlh_get_diagnostics(mlh_ctx *c, mlh_diag *d) {
    if (!c || !d) return fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    unsigned long long k[kNumCounters];
    int st = read_counters(c, k);
    if (st) return st;
    int mc[2];
    HIPCHK(hipMemcpy(mc, c->max_counts, sizeof(mc), hipMemcpyDeviceToHost));
    d->neg_face_pressure_real = (int64_t)k[kCntNegReal];
    d->neg_face_pressure_ghost = (int64_t)k[kCntNegGhost];
    d->vacuum_faces = (int64_t)k[kCntVacuum];
    d->nonconverged_faces = (int64_t)k[kCntNonconv];
    d->missing_mirror_pairs = (int64_t)k[kCntMissing];
    d->learned_fallbacks = (int64_t)k[kCntFallback];
    d->newton_iterations = (int64_t)k[kCntIters];
    d->faces_solved = (int64_t)k[kCntFaces];
    d->learned_seeded = (int64_t)k[kCntSeeded];
    d->singular_matrices = (int64_t)k[kCntSingular];
    d->nonfinite_particles = (int64_t)k[kCntNonfinite];
    d->acoustic_faces = (int64_t)k[kCntAcoustic];
    d->max_neighbours = mc[0];
    d->max_ghost_neighbours = mc[1];
    return MLH_OK;
}

int mlh_conserved_sums(mlh_ctx *c, double *out) {
    if (!c || !out) return fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    HydroArgs a = hydro_args(c);
    launch_sums(a, c->D, c->prec, c->sums_block, c->s);
    int nb = dt_blocks(c->N);
    std::vector<double> h((size_t)nb * 6);
    HIPCHK(hipMemcpyAsync(h.data(), c->sums_block, sizeof(double) * h.size(),
                          hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    for (int k = 0; k < 6; ++k) out[k] = 0.;
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < 6; ++k) out[k] += h[(size_t)b * 6 + k];
    if (c->stage < 2) out[0] = 0.;
    return MLH_OK;
}

int mlh_riemann_flux_batch(mlh_ctx *c, int64_t n, const double *w_left, const double *w_right,
                           const double *v_frame, const double *a_ij, double *f_ij,
                           int32_t *flag) {
    if (!c || n < 0) return fail(MLH_ERR_INVALID_ARGUMENT, "null ctx or n < 0");
    if (n == 0) return MLH_OK;
    if (!w_left || !w_right || !v_frame || !a_ij || !f_ij || !flag)
        return fail(MLH_ERR_INVALID_ARGUMENT, "null array");
    const size_t D = c->D, nw = (size_t)n * (D + 2), nv = (size_t)n * D;
    double *buf = nullptr;
    int *fl = nullptr;
    HIPCHK(hipMalloc(&buf, sizeof(double) * (3 * nw + 2 * nv)));
    if (hipMalloc(&fl, sizeof(int) * (size_t)n) != hipSuccess) {
        hipFree(buf);
        return fail(MLH_ERR_DEVICE, "hipMalloc failed");
    }
    double *dwl = buf, *dwr = buf + nw, *dvf = buf + 2 * nw, *da = dvf + nv, *df = da + nv;
    hipError_t e = hipMemcpyAsync(dwl, w_left, sizeof(double) * nw, hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) e = hipMemcpyAsync(dwr, w_right, sizeof(double) * nw, hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) e = hipMemcpyAsync(dvf, v_frame, sizeof(double) * nv, hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) e = hipMemcpyAsync(da, a_ij, sizeof(double) * nv, hipMemcpyHostToDevice, c->s);
    if (e == hipSuccess) {
        HydroArgs a = hydro_args(c);
        launch_riemann_flux(a, c->D, (long long)n, dwl, dwr, dvf, da, df, fl, c->s);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(f_ij, df, sizeof(double) * nw, hipMemcpyDeviceToHost, c->s);
    if (e == hipSuccess) e = hipMemcpyAsync(flag, fl, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, c->s);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s);
    hipFree(buf);
    hipFree(fl);
    if (e != hipSuccess) return fail(MLH_ERR_DEVICE, hipGetErrorString(e));
    return MLH_OK;
}

int mlh_set_face_collection(mlh_ctx *c, int64_t cap, int32_t particle_stride) {
    if (!c || cap < 0) return fail(MLH_ERR_INVALID_ARGUMENT, "bad arguments");
    MlpState &m = c->mlp;
    if (m.collect) { hipFree(m.collect); m.collect = nullptr; }
    if (!m.collect_n) HIPCHK(hipMalloc(&m.collect_n, sizeof(unsigned long long) * kCollectStripes));
    HIPCHK(hipMemset(m.collect_n, 0, sizeof(unsigned long long) * kCollectStripes));
    m.collect_cap = m.collect_scap = 0;
    m.collect_stripes = 0;
    m.collect_stride = 0;
    if (particle_stride <= 0 || cap == 0) return MLH_OK;
    HIPCHK(hipMalloc(&m.collect, sizeof(double) * 6 * (size_t)cap));
    // as many stripes as fit (a power of two, at most kCollectStripes), each
    // of cap / stripes rows; the rows past stripes * scap stay unused
    int stripes = 1;
    while (stripes * 2 <= kCollectStripes && stripes * 2 <= cap) stripes *= 2;
    m.collect_cap = cap;
    m.collect_stripes = stripes;
    m.collect_scap = cap / stripes;
    m.collect_stride = particle_stride;
    return MLH_OK;
}

int mlh_get_collected_faces(mlh_ctx *c, double *states6, int64_t *n_out) {
    if (!c || !n_out) return fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    MlpState &m = c->mlp;
    *n_out = 0;
    if (!m.collect || !m.collect_n) return MLH_OK;
    unsigned long long n[kCollectStripes] = {};
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipMemcpy(n, m.collect_n, sizeof(unsigned long long) * m.collect_stripes,
                     hipMemcpyDeviceToHost));
    // each stripe's filled prefix, stripe after stripe
    int64_t out = 0;
    for (int s = 0; s < m.collect_stripes; ++s) {
        const int64_t k = n[s] < (unsigned long long)m.collect_scap ? (int64_t)n[s] : m.collect_scap;
        if (states6 && k)
            HIPCHK(hipMemcpy(states6 + 6 * out, m.collect + 6 * (size_t)s * m.collect_scap,
                             sizeof(double) * 6 * k, hipMemcpyDeviceToHost));
        out += k;
    }
    HIPCHK(hipMemset(m.collect_n, 0, sizeof(unsigned long long) * kCollectStripes));
    *n_out = out;
    return MLH_OK;
}

int mlh_get_timings(mlh_ctx *c, float *ms4) {
    if (!c || !ms4) return fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    for (int k = 0; k < 4; ++k) ms4[k] = c->timings[k];
    return MLH_OK;
}

}  // extern "C"

// internal accessors for the MLP module
MlpState &ctx_mlp(mlh_ctx *c) { return c->mlp; }
hipStream_t ctx_stream(mlh_ctx *c) { return c ? c->s : nullptr; }
double ctx_gamma(mlh_ctx *c) { return c->p.gamma; }
int ctx_fail(int code, const char *msg) { return fail(code, msg); }
