// MeshlessScheme.h -- host mirror of the reference's time-integration driver
// (demonstrator/include/MeshlessScheme.h, src/MeshlessScheme.cpp:21-254).
//
// The loop keeps the reference's order and dump semantics:
//   * ADAPTIVE_TIMESTEP: dump times i * timeStep * h5DumpInterval; the CFL
//     step is clipped to the next dump time, and the state at that time is
//     dumped at the start of the following step (after the gradients, before
//     the update), as outDir/%06d.h5 numbered by dump;
//   * fixed step: dump every h5DumpInterval steps, numbered by step;
//   * the loop stops when t >= timeEnd (checked after a dump) and runs while
//     t < timeEnd + timeStep; running out of dump times is exit code 9.
// Steps without a dump run as one device call (mlh_step); dump steps run the
// stages one by one so that the dump sees rho, P, rhoGrad and noi of that
// step, as the reference's does.
#pragma once
#include <string>

#include "Particles.h"

namespace mlh_host {

class MeshlessScheme {
public:
    struct Configuration {
        std::string initFile;
        std::string outDir;
        double timeStep = 0;
        double timeEnd = 0;
        int h5DumpInterval = 1;
        double periodicBoxLimits[6] = {0, 0, 0, 1, 1, 1};  // lowerX..Z, upperX..Z
        double kernelSize = 0;
        double gamma = 0;
    };

    // perMethod: every step calls the Particles methods one at a time in
    // the reference's order (gridNNS, compDensity, compPressure,
    // compGlobalTimestep, compPsijTilde, slopeLimiter, compEffectiveFace,
    // compRiemannStatesLR, solveRiemannProblems, checkFluxSymmetry at
    // DEBUG_LVL >= 1, collectFluxes, updateStateAndPosition,
    // MeshlessScheme.cpp:53-224); the results are bitwise those of the
    // default path
    MeshlessScheme(Configuration config, Particles *particles, bool adaptiveTimestep,
                   bool perMethod = false);
    void run();

    int steps() const { return step_; }
    int dumps() const { return dumps_; }
    double time() const { return t_; }

private:
    Configuration config_;
    Particles *particles_;
    bool adaptive_;
    bool per_method_;
    int step_ = 0, dumps_ = 0;
    double t_ = 0;
};

}  // namespace mlh_host
