// Particles.cpp -- host mirror of Particles over the C ABI (see Particles.h).
#include "Particles.h"

#include <cmath>

#include "H5IO.h"

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
