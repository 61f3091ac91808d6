// Driver.h -- config.info + parameter.h -> a configured run; shared by the
// mlh executable and the C entry point mlh_host_run.
#pragma once
#include <string>

#include "H5IO.h"
#include "MeshlessScheme.h"
#include "Parameters.h"
#include "Particles.h"

namespace mlh_host {

struct Run {
    Parameters prm;
    MeshlessScheme::Configuration cfg;
};

// main.cpp:54-80 of the reference: the config values, the periodic box and
// the parameter.h flags; relative paths are taken relative to the config
Run load_run(const std::string &config, const std::string &parameter_h);
// InitialDistribution::getAllParticles (InitialDistribution.cpp:33-63)
void fill_particles(Particles &P, const InitialDistribution &ic);

}  // namespace mlh_host
