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
#include <utility>
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
    // The params fix dim, box, h, gamma and the algorithm flags for the run.
    Particles(int numParticles, const mlh_params &params);
    // Particles(int numParticles, bool ghosts=false), Particles.h:30, with
    // the process-wide parameters of defaultParams() (the reference fixes
    // them at compile time in parameter.h).  ghosts = true makes the
    // container of periodic images that MeshlessScheme owns
    // (MeshlessScheme.h:40): host arrays only, room for dim * numParticles
    // images, filled by createGhostParticles; N is then the image count.
    explicit Particles(int numParticles, bool ghosts = false);
    ~Particles();
    Particles(const Particles &) = delete;
    Particles &operator=(const Particles &) = delete;

    // parameters that Particles(int, bool) uses; mlh_params_default() until
    // the caller changes them (e.g. with a parameter.h / config.info read)
    static mlh_params &defaultParams();

    int N;
    int dim;
    bool ghosts = false;
    std::vector<int> matId, noi, noiGhosts;
    std::vector<double> m, u, x, y, z, vx, vy, vz, rho, P;
    std::vector<double> rhoGrad, vxGrad, vyGrad, vzGrad, PGrad;  // [N][dim]
    // image container only: the particle each image copies
    std::vector<int> parent;

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

    // ---- the reference's methods one at a time (Particles.h:137-151) ----
    // compPsijTilde + gradient of rho, v, P, unlimited (Particles.cpp:
    // 1186-1270 and the image twins); with periodic boundaries the
    // ghostParticles argument of the reference is implicit
    void compPsijTilde(double kernelSize);
    // gradient(f, grad) (Particles.cpp:1257-1270): f [N], grad [N][dim]
    void gradient(const double *f, double *grad);
    // slopeLimiter (Particles.cpp:1313-1444) on the gradients of
    // compPsijTilde
    void slopeLimiter(double kernelSize, Particles *ghostParticles = nullptr);
    // compEffectiveFace (Particles.cpp:1290-1311, 2504-2531)
    void compEffectiveFace();
    // compRiemannStatesLR (Particles.cpp:1488-1733, 2533-2683)
    void compRiemannStatesLR(double dt, double kernelSize, double gamma);
    // solveRiemannProblems (Particles.cpp:1787-1911) on the stored states
    void solveRiemannProblems(double gamma, const Particles &ghostParticles);
    // collectFluxes (Particles.cpp:1913-2011); updateStateAndPosition then
    // commits the step
    void collectFluxes();
    // checkFluxSymmetry (Particles.cpp:2888-2976): a WARN line per failing
    // pair count, as the reference warns per pair; returns the counts
    // (asymmetric entries, entries without a mirror)
    std::pair<int64_t, int64_t> checkFluxSymmetry(Particles *ghostParticles = nullptr);

    // createGhostParticles (Particles.cpp:2113-2191) on the host arrays x, y
    // (z) of this set: the images of the particles within kernelSize of a
    // periodic box face, in the reference's ghost order (per particle the
    // image shifted along every axis it is near first, 2D: corner, y, x;
    // 3D the same rule over the seven axis subsets).  The device search
    // forms the same images implicitly; this fills the container for
    // callers that inspect them (e.g. dumping them).
    void createGhostParticles(Particles &ghostParticles, double kernelSize) const;
    // updateGhostState (Particles.cpp:2193-2206): rho, P, v of every image
    // from its parent's host arrays
    void updateGhostState(Particles &ghostParticles) const;

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

// The periodic images of n particles at xs[0..dim) (createGhostParticles,
// Particles.cpp:2113-2191, over dim axes): parents and image coordinates in
// the reference's ghost order.
void ghost_images(const mlh_params &p, int n, const double *const xs[3], double kernelSize,
                  std::vector<int> &parent, std::vector<double> *gx[3]);

}  // namespace mlh_host
