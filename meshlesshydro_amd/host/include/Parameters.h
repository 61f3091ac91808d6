// Parameters.h -- the reference's compile-time parameter.h, read at run time.
//
// The reference selects its algorithm with #defines in parameter.h
// (demonstrator/include/parameter.h; testcases/*/parameter*.h).  Here the
// same file is read as text and mapped onto mlh_params, so a testcase's
// parameter.h configures a run without recompiling:
//   DIM, PERIODIC_BOUNDARIES, CFL, MAX_NUM_INTERACTIONS,
//   MAX_NUM_GHOST_INTERACTIONS, SLOPE_LIMITING, BETA, PAIRWISE_LIMITER,
//   PSI_1, PSI_2, MESHLESS_FINITE_MASS, ENFORCE_FLUX_SYM, MOVE_PARTICLES,
//   DEBUG_LVL, FIRST_ORDER_QUAD_POINT  -> the mlh_params fields;
//   ADAPTIVE_TIMESTEP                  -> Parameters::adaptive_timestep;
//   RUNSPH = 1                         -> rejected (the SPH solver is not
//                                         part of this library).
// Names this library adds (MLH_PRECISION, MLH_RIEMANN_SEED, MLH_LIMITER_ABS,
// MLH_QUIRK_3D_PREDICTOR, MLH_WORK_ORDER) may be set the same way.
// Other defines (ARTVISC, VERBOSITY_PARTICLES, FLUX_SYM_TOL, ...) are
// recorded in `ignored`.
#pragma once
#include <string>
#include <vector>

#include "../../../include/mlh.h"

namespace mlh_host {

struct Parameters {
    mlh_params params;
    bool adaptive_timestep = true;
    std::vector<std::string> ignored;
};

// defaults: mlh_params_default() and ADAPTIVE_TIMESTEP 1
Parameters default_parameters();
// read "#define NAME value" lines over the defaults; throws on a malformed
// value or an unsupported setting
Parameters read_parameter_header(const std::string &path);
void apply_define(Parameters &p, const std::string &name, const std::string &value);

}  // namespace mlh_host
