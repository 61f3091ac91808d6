// MeshlessScheme.cpp -- time-integration driver (see MeshlessScheme.h).
#include "MeshlessScheme.h"

#include <cstdio>
#include <limits>
#include <vector>

#include "Logger.h"

namespace mlh_host {

MeshlessScheme::MeshlessScheme(Configuration config, Particles *particles, bool adaptiveTimestep,
                               bool perMethod)
    : config_(std::move(config)), particles_(particles), adaptive_(adaptiveTimestep),
      per_method_(perMethod) {
    if (config_.h5DumpInterval <= 0)
        throw MlhError(MLH_ERR_INVALID_ARGUMENT, "h5DumpInterval must be positive");
    if (config_.timeStep <= 0)
        throw MlhError(MLH_ERR_INVALID_ARGUMENT, "timeStep must be positive");
}

void MeshlessScheme::run() {
    Particles &P = *particles_;
    const double h = config_.kernelSize, gamma = config_.gamma;
    double t = 0;
    int step = 0;
    // MeshlessScheme.cpp:24-35
    int numDumpTimes = 0;
    std::vector<double> dumpTimes;
    if (adaptive_) {
        numDumpTimes = (int)(config_.timeEnd / config_.timeStep) / config_.h5DumpInterval + 1;
        for (int k = 0; k < numDumpTimes; ++k)
            dumpTimes.push_back(k * config_.timeStep * config_.h5DumpInterval);
    }
    int dumpStep = 0;
    bool dump = true, dumpNext = false;
    do {
        Logger(INFO) << "  > TIME: " << t << ", STEP: " << step;
        const bool dumpNow = adaptive_ ? dump : (step % config_.h5DumpInterval == 0);
        double tclip = -1.0;
        if (adaptive_) {
            // MeshlessScheme.cpp:93-101; dumpTimes[numDumpTimes] (read past the
            // end by the reference on its last dump) is taken as "no clip"
            if (dumpStep >= numDumpTimes)
                throw MlhError(MLH_ERR_DUMP_OVERRUN,
                               "Simulation did not abort after reaching timeEnd. Exiting.");
            tclip = dumpStep + 1 < numDumpTimes ? dumpTimes[dumpStep + 1]
                                                : std::numeric_limits<double>::infinity();
        }
        if (!dumpNow && t >= config_.timeEnd) {
            Logger(INFO) << "    > t = " << t << " -> FINISHED!";
            break;
        }
        double dt;
        if (per_method_) {
            // the reference's sequence of Particles methods
            P.gridNNS(h);
            P.compDensity(h);
            P.compPressure(gamma);
            if (adaptive_) {
                dt = P.compGlobalTimestep(gamma, h);
                if (t + dt >= tclip) {
                    dumpNext = true;
                    dt = tclip - t;
                }
            } else {
                dt = config_.timeStep;
            }
            P.compPsijTilde(h);
            if (P.params().slope_limiting) P.slopeLimiter(h);
            P.compEffectiveFace();
            P.compRiemannStatesLR(dt, h, gamma);
            if (dumpNow) {
                char name[32];
                std::snprintf(name, sizeof(name), "%06d", adaptive_ ? dumpStep : step);
                Logger(INFO) << "      > Dump particles to file";
                P.dump2file(config_.outDir + "/" + name + ".h5", t);
                ++dumpStep;
                ++dumps_;
                if (adaptive_) dump = false;
                if (t >= config_.timeEnd) {
                    Logger(INFO) << "    > t = " << t << " -> FINISHED!";
                    break;
                }
            }
            P.solveRiemannProblems(gamma, P);
            if (P.params().debug_level >= 1) P.checkFluxSymmetry();
            P.collectFluxes();
            P.updateStateAndPosition(dt);
        } else if (dumpNow) {
            P.gridNNS(h);
            P.compDensity(h);
            if (adaptive_) {
                dt = P.compGlobalTimestep(gamma, h);  // also the gradients
                if (t + dt >= tclip) {
                    dumpNext = true;
                    dt = tclip - t;
                }
            } else {
                dt = config_.timeStep;
                P.compGradients();
            }
            char name[32];
            std::snprintf(name, sizeof(name), "%06d", adaptive_ ? dumpStep : step);
            Logger(INFO) << "      > Dump particles to file";
            P.dump2file(config_.outDir + "/" + name + ".h5", t);
            ++dumpStep;
            ++dumps_;
            if (adaptive_) dump = false;
            if (t >= config_.timeEnd) {
                Logger(INFO) << "    > t = " << t << " -> FINISHED!";
                break;
            }
            P.solveRiemannProblems(dt);
            P.updateStateAndPosition(dt);
        } else {
            bool clipped = false;
            dt = P.step(adaptive_ ? -1.0 : config_.timeStep, t,
                        adaptive_ && tclip < std::numeric_limits<double>::infinity() ? tclip : -1.0,
                        &clipped);
            if (clipped) dumpNext = true;
        }
        Logger(INFO) << "Time  > dt = " << dt << " selected.";
        t += dt;
        ++step;
        if (adaptive_ && dumpNext) {
            dump = true;
            dumpNext = false;
        }
        t_ = t;
        step_ = step;
    } while (t < config_.timeEnd + config_.timeStep);
    t_ = t;
    step_ = step;
}

}  // namespace mlh_host
