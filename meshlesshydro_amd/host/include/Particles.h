// Particles.h -- host mirror of the reference's Particles class over the
// C ABI of include/mlh.h.
//
// The reference keeps every per-particle and per-pair array on the host and
// runs each stage as a member function (demonstrator/include/Particles.h:
// 30-196).  Here the device context owns the state and the stages; this
// class keeps the reference's public surface:
//   * public SoA host arrays N, matId, m, u, x, y, z, vx, vy, vz, rho, P and
//     the gradients rhoGrad, vxGrad, vyGrad, vzGrad, PGrad ([N][dim] each,
//     flattened; the reference's double(*)[DIM]);
//   * the stage functions under the reference's names, each one call of the
//     C ABI on the device state;
//   * diagnostics sumVolume ... sumMomentumZ and dump2file.
// Host arrays are copies: upload() sends x..u to the device (the reference's
// "state" after InitialDistribution::getAllParticles), download() and the
// fetch* functions bring device results back.  Errors of the C ABI are
// thrown as MlhError carrying the reference's exit code (1, 3, 6, 9) or the
// library's status.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mlh.h"

namespace mlh_host {

struct MlhError : std::runtime_error {
    int status;
    MlhError(int st, const std::string &what) : std::runtime_error(what), status(st) {}
};
// throws MlhError for status != MLH_OK
void check(int status, const char *where);

class Particles {
public:
    // Particles(int numParticles, bool ghosts=false), Particles.h:30; the
    // params fix dim, box, h, gamma and the algorithm flags for the run
    Particles(int numParticles, const mlh_params &params);
    ~Particles();
    Particles(const Particles &) = delete;
    Particles &operator=(const Particles &) = delete;

    int N;
    int dim;
    std::vector<int> matId, noi, noiGhosts;
    std::vector<double> m, u, x, y, z, vx, vy, vz, rho, P;
    std::vector<double> rhoGrad, vxGrad, vyGrad, vzGrad, PGrad;  // [N][dim]

    // host arrays -> device state / device state -> host arrays
    void upload();
    void download();

    // assignParticlesAndCells + createGhostParticles + gridNNS + ghostNNS
    // (Particles.cpp:270-365, :2113-2260): cell list and neighbour lists
    void gridNNS(double kernelSize);
    // compDensity incl. the periodic images + compPressure
    // (Particles.cpp:1151-1184, :2262-2290, :1272-1288)
    void compDensity(double kernelSize);
    void compPressure(double gamma);  // part of compDensity on the device
    // compGlobalTimestep (Particles.cpp:1446-1485): CFL step, real pairs
    double compGlobalTimestep(double gamma, double kernelSize);
    // compPsijTilde + gradient x (D+2) + slopeLimiter
    // (Particles.cpp:1186-1270, :1313-1444 and the image twins)
    void compGradients();
    // compEffectiveFace + compRiemannStatesLR + solveRiemannProblems +
    // collectFluxes (Particles.cpp:1290-2011 and the image twins)
    void solveRiemannProblems(double dt);
    // updateStateAndPosition (Particles.cpp:2013-2110)
    void updateStateAndPosition(double dt);

    // the whole step on the device (one status check); dt_fixed > 0 uses
    // that step, else the CFL step clipped at t_clip (MeshlessScheme.cpp:
    // 91-105).  Returns the step taken; *clipped reports the clip.
    double step(double dt_fixed, double t_now, double t_clip, bool *clipped);

    // results of the last stages
    void fetchDensity();        // rho, P
    void fetchGradients();      // rhoGrad ... PGrad
    void fetchNeighbourCounts();  // noi, noiGhosts

    double sumVolume();
    double sumMass();
    double sumEnergy();
    double sumMomentumX();
    double sumMomentumY();
    double sumMomentumZ();

    // Particles::dump2file (Particles.cpp:2975-3075): state, rho, P, rhoGrad,
    // noi and the conserved sums of the current step
    void dump2file(const std::string &filename, double simTime);

    mlh_diag diagnostics();
    mlh_ctx *context() { return ctx_; }
    const mlh_params &params() const { return p_; }

private:
    mlh_params p_;
    mlh_ctx *ctx_ = nullptr;
    void sums(double out[6]);
};

}  // namespace mlh_host
