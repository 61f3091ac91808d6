// Particles.cpp -- host mirror of Particles over the C ABI (see Particles.h).
#include "Particles.h"

#include <cmath>

#include "H5IO.h"
#include "Logger.h"

namespace mlh_host {

void check(int st, const char *where) {
    if (st == MLH_OK) return;
    std::string msg = std::string(where) + ": " + mlh_last_error();
    throw MlhError(st, msg);
}

Particles::Particles(int numParticles, const mlh_params &params)
    : N(numParticles), dim(params.dim), p_(params) {
    size_t n = (size_t)N, nd = n * (size_t)dim;
    matId.assign(n, 0);
    noi.assign(n, 0);
    noiGhosts.assign(n, 0);
    for (auto *v : {&m, &u, &x, &y, &z, &vx, &vy, &vz, &rho, &P}) v->assign(n, 0.0);
    for (auto *v : {&rhoGrad, &vxGrad, &vyGrad, &vzGrad, &PGrad}) v->assign(nd, 0.0);
    check(mlh_create(&p_, N, &ctx_), "mlh_create");
}

mlh_params &Particles::defaultParams() {
    static mlh_params p = [] {
        mlh_params d;
        mlh_params_default(&d);
        return d;
    }();
    return p;
}

Particles::Particles(int numParticles, bool ghostSet)
    : N(numParticles), dim(defaultParams().dim), ghosts(ghostSet), p_(defaultParams()) {
    if (!ghosts) {
        size_t n = (size_t)N, nd = n * (size_t)dim;
        matId.assign(n, 0);
        noi.assign(n, 0);
        noiGhosts.assign(n, 0);
        for (auto *v : {&m, &u, &x, &y, &z, &vx, &vy, &vz, &rho, &P}) v->assign(n, 0.0);
        for (auto *v : {&rhoGrad, &vxGrad, &vyGrad, &vzGrad, &PGrad}) v->assign(nd, 0.0);
        check(mlh_create(&p_, N, &ctx_), "mlh_create");
        return;
    }
    // the image container: room for dim * numParticles images (as the
    // reference allocates), no device context, N = 0 until filled
    const size_t cap = (size_t)dim * (size_t)numParticles;
    for (auto *v : {&x, &y, &z, &vx, &vy, &vz, &rho, &P}) v->reserve(cap);
    parent.reserve(cap);
    N = 0;
}

Particles::~Particles() {
    if (ctx_) mlh_destroy(ctx_);
}

void Particles::upload() {
    check(mlh_upload_state(ctx_, x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(),
                           m.data(), u.data()),
          "mlh_upload_state");
}

void Particles::download() {
    check(mlh_download_state(ctx_, x.data(), y.data(), dim == 3 ? z.data() : nullptr, vx.data(),
                             vy.data(), dim == 3 ? vz.data() : nullptr, m.data(), u.data()),
          "mlh_download_state");
}

void Particles::gridNNS(double kernelSize) {
    if (kernelSize != p_.kernel_size)
        throw MlhError(MLH_ERR_INVALID_ARGUMENT, "gridNNS: kernelSize differs from the context's");
    check(mlh_neighbours(ctx_), "gridNNS");
}

void Particles::compDensity(double kernelSize) {
    if (kernelSize != p_.kernel_size)
        throw MlhError(MLH_ERR_INVALID_ARGUMENT, "compDensity: kernelSize differs from the context's");
    check(mlh_density(ctx_), "compDensity");
}

void Particles::compPressure(double gamma) {
    if (gamma != p_.gamma)
        throw MlhError(MLH_ERR_INVALID_ARGUMENT, "compPressure: gamma differs from the context's");
}

double Particles::compGlobalTimestep(double gamma, double kernelSize) {
    if (gamma != p_.gamma || kernelSize != p_.kernel_size)
        throw MlhError(MLH_ERR_INVALID_ARGUMENT, "compGlobalTimestep: gamma / kernelSize differ");
    double dt = 0;
    check(mlh_timestep(ctx_, &dt), "compGlobalTimestep");
    return dt;
}

void Particles::compGradients() { check(mlh_gradients(ctx_), "compGradients"); }

void Particles::solveRiemannProblems(double dt) { check(mlh_fluxes(ctx_, dt), "solveRiemannProblems"); }

void Particles::updateStateAndPosition(double) { check(mlh_update(ctx_), "updateStateAndPosition"); }

void Particles::compPsijTilde(double kernelSize) {
    if (kernelSize != p_.kernel_size)
        throw MlhError(MLH_ERR_INVALID_ARGUMENT, "compPsijTilde: kernelSize differs from the context's");
    check(mlh_psi_gradients(ctx_), "compPsijTilde");
}

void Particles::gradient(const double *f, double *grad) {
    check(mlh_gradient_field(ctx_, f, grad), "gradient");
}

void Particles::slopeLimiter(double kernelSize, Particles *) {
    if (kernelSize != p_.kernel_size)
        throw MlhError(MLH_ERR_INVALID_ARGUMENT, "slopeLimiter: kernelSize differs from the context's");
    check(mlh_limit(ctx_), "slopeLimiter");
}

void Particles::compEffectiveFace() { check(mlh_effective_faces(ctx_), "compEffectiveFace"); }

void Particles::compRiemannStatesLR(double dt, double kernelSize, double gamma) {
    if (kernelSize != p_.kernel_size || gamma != p_.gamma)
        throw MlhError(MLH_ERR_INVALID_ARGUMENT, "compRiemannStatesLR: kernelSize / gamma differ");
    check(mlh_riemann_states(ctx_, dt), "compRiemannStatesLR");
}

void Particles::solveRiemannProblems(double gamma, const Particles &) {
    if (gamma != p_.gamma)
        throw MlhError(MLH_ERR_INVALID_ARGUMENT, "solveRiemannProblems: gamma differs");
    check(mlh_solve_faces(ctx_), "solveRiemannProblems");
}

void Particles::collectFluxes() { check(mlh_collect_fluxes(ctx_), "collectFluxes"); }

std::pair<int64_t, int64_t> Particles::checkFluxSymmetry(Particles *) {
    int64_t asym = 0, missing = 0;
    check(mlh_check_flux_symmetry(ctx_, 1e-20, &asym, &missing), "checkFluxSymmetry");
    if (asym)
        Logger(WARN) << "  > Fluxes are NOT symmetric for " << asym << " interactions";
    if (missing)
        Logger(WARN) << "  > " << missing << " interactions have no reverse interaction";
    return {asym, missing};
}

void ghost_images(const mlh_params &p, int n, const double *const xs[3], double kernelSize,
                  std::vector<int> &parent, std::vector<double> *gx[3]) {
    const int dim = p.dim;
    parent.clear();
    for (int a = 0; a < 3; ++a) gx[a]->clear();
    for (int i = 0; i < n; ++i) {
        // per axis: the image coordinate when x is within kernelSize of a
        // face (lower test first, Particles.cpp:2123-2141)
        double img[3] = {0, 0, 0};
        int near = 0;
        for (int a = 0; a < dim; ++a) {
            const double xa = xs[a][i], mn = p.box_min[a], mx = p.box_max[a];
            if (xa <= mn + kernelSize) {
                img[a] = mx + (xa - mn);
                near |= 1 << a;
            } else if (mx - kernelSize < xa) {
                img[a] = mn - (mx - xa);
                near |= 1 << a;
            }
        }
        // every non-empty subset of the near axes, largest axis mask first
        // (2D: corner, then y only, then x only, :2145-2178)
        for (int mask = near; mask > 0; mask = (mask - 1) & near) {
            for (int a = 0; a < 3; ++a)
                gx[a]->push_back(a < dim ? ((mask >> a) & 1 ? img[a] : xs[a][i]) : 0.0);
            parent.push_back(i);
        }
    }
}

void Particles::createGhostParticles(Particles &g, double kernelSize) const {
    if (!p_.periodic) throw MlhError(MLH_ERR_INVALID_ARGUMENT, "createGhostParticles: no periodic box");
    const double *xs[3] = {x.data(), y.data(), dim == 3 ? z.data() : nullptr};
    std::vector<double> *gx[3] = {&g.x, &g.y, &g.z};
    ghost_images(p_, N, xs, kernelSize, g.parent, gx);
    g.N = (int)g.parent.size();
    g.dim = dim;
    for (auto *v : {&g.vx, &g.vy, &g.vz, &g.rho, &g.P}) v->assign(g.N, 0.0);
}

void Particles::updateGhostState(Particles &g) const {
    for (int k = 0; k < g.N; ++k) {
        const int i = g.parent[k];
        g.rho[k] = rho[i];
        g.P[k] = P[i];
        g.vx[k] = vx[i];
        g.vy[k] = vy[i];
        if (dim == 3) g.vz[k] = vz[i];
    }
}

double Particles::step(double dt_fixed, double t_now, double t_clip, bool *clipped) {
    double dt = 0;
    int32_t c = 0;
    check(mlh_step(ctx_, dt_fixed, t_now, t_clip, &dt, &c), "step");
    if (clipped) *clipped = c != 0;
    return dt;
}

void Particles::fetchDensity() {
    check(mlh_get_density(ctx_, rho.data(), P.data(), nullptr), "fetchDensity");
}

void Particles::fetchGradients() {
    size_t nd = (size_t)N * dim;
    std::vector<double> g((size_t)(dim + 2) * nd);
    check(mlh_get_gradients(ctx_, g.data(), nullptr), "fetchGradients");
    // field order of the library: rho, v..., P
    std::vector<double> *dst[5] = {&rhoGrad, &vxGrad, &vyGrad, dim == 3 ? &vzGrad : &PGrad, &PGrad};
    for (int f = 0; f < dim + 2; ++f) {
        std::vector<double> &d = *dst[f];
        for (size_t k = 0; k < nd; ++k) d[k] = g[(size_t)f * nd + k];
    }
}

void Particles::fetchNeighbourCounts() {
    std::vector<int32_t> nbr((size_t)N * (p_.max_interactions + p_.max_ghost_interactions));
    check(mlh_get_neighbours(ctx_, noi.data(), noiGhosts.data(), nbr.data()), "fetchNeighbourCounts");
}

void Particles::sums(double out[6]) { check(mlh_conserved_sums(ctx_, out), "sums"); }

double Particles::sumVolume() { double s[6]; sums(s); return s[0]; }
double Particles::sumMass() { double s[6]; sums(s); return s[1]; }
double Particles::sumEnergy() { double s[6]; sums(s); return s[2]; }
double Particles::sumMomentumX() { double s[6]; sums(s); return s[3]; }
double Particles::sumMomentumY() { double s[6]; sums(s); return s[4]; }
double Particles::sumMomentumZ() { double s[6]; sums(s); return s[5]; }

mlh_diag Particles::diagnostics() {
    mlh_diag d;
    check(mlh_get_diagnostics(ctx_, &d), "diagnostics");
    return d;
}

void Particles::dump2file(const std::string &filename, double simTime) {
    download();
    fetchDensity();
    fetchGradients();
    fetchNeighbourCounts();
    double s[6];
    sums(s);
    Snapshot sn;
    sn.N = N;
    sn.D = dim;
    sn.time = simTime;
    sn.totalMass = s[1];
    sn.energy = s[2];
    sn.mom[0] = s[3];
    sn.mom[1] = s[4];
    sn.mom[2] = s[5];
    sn.rho = rho;
    sn.m = m;
    sn.u = u;
    sn.P = P;
    sn.noi = noi;
    size_t nd = (size_t)N * dim;
    sn.x.resize(nd);
    sn.v.resize(nd);
    sn.rhoGrad = rhoGrad;
    for (int i = 0; i < N; ++i) {
        sn.x[(size_t)i * dim + 0] = x[i];
        sn.x[(size_t)i * dim + 1] = y[i];
        sn.v[(size_t)i * dim + 0] = vx[i];
        sn.v[(size_t)i * dim + 1] = vy[i];
        if (dim == 3) {
            sn.x[(size_t)i * dim + 2] = z[i];
            sn.v[(size_t)i * dim + 2] = vz[i];
        }
    }
    write_snapshot(filename, sn);
}

}  // namespace mlh_host
