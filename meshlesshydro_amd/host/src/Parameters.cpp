// Parameters.cpp -- parameter.h at run time (see Parameters.h).
#include "Parameters.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <stdexcept>

namespace mlh_host {

namespace {

std::string strip_comments(const std::string &s) {
    std::string t = s;
    size_t c = t.find("//");
    if (c != std::string::npos) t = t.substr(0, c);
    size_t b = t.find("/*");
    if (b != std::string::npos) t = t.substr(0, b);
    return t;
}

double to_double(const std::string &name, const std::string &v) {
    char *end = nullptr;
    double d = std::strtod(v.c_str(), &end);
    if (end == v.c_str() || *end != '\0')
        throw std::runtime_error("parameter.h: " + name + " = '" + v + "' is not a number");
    return d;
}

int to_int(const std::string &name, const std::string &v) {
    double d = to_double(name, v);
    if (d != (double)(int)d) throw std::runtime_error("parameter.h: " + name + " must be an integer");
    return (int)d;
}

}  // namespace

Parameters default_parameters() {
    Parameters p;
    mlh_params_default(&p.params);
    p.adaptive_timestep = true;
    return p;
}

void apply_define(Parameters &p, const std::string &n, const std::string &v) {
    mlh_params &q = p.params;
    if (n == "DIM") {
        q.dim = to_int(n, v);
        if (q.dim != 2 && q.dim != 3)
            throw std::runtime_error("parameter.h: DIM must be 2 or 3, not '" + v + "'");
    }
    else if (n == "PERIODIC_BOUNDARIES") q.periodic = to_int(n, v);
    else if (n == "ADAPTIVE_TIMESTEP") p.adaptive_timestep = to_int(n, v) != 0;
    else if (n == "CFL") q.cfl = to_double(n, v);
    else if (n == "MAX_NUM_INTERACTIONS") q.max_interactions = to_int(n, v);
    else if (n == "MAX_NUM_GHOST_INTERACTIONS") q.max_ghost_interactions = to_int(n, v);
    else if (n == "SLOPE_LIMITING") q.slope_limiting = to_int(n, v);
    else if (n == "BETA") q.beta = to_double(n, v);
    else if (n == "PAIRWISE_LIMITER") q.pairwise_limiter = to_int(n, v);
    else if (n == "PSI_1") q.psi1 = to_double(n, v);
    else if (n == "PSI_2") q.psi2 = to_double(n, v);
    else if (n == "MESHLESS_FINITE_MASS") q.mfm = to_int(n, v);
    else if (n == "ENFORCE_FLUX_SYM") q.enforce_flux_sym = to_int(n, v);
    else if (n == "MOVE_PARTICLES") q.move_particles = to_int(n, v);
    else if (n == "DEBUG_LVL") q.debug_level = to_int(n, v);
    else if (n == "FIRST_ORDER_QUAD_POINT") q.first_order_quad_point = to_int(n, v);
    else if (n == "RUNSPH") {
        if (to_int(n, v) != 0)
            throw std::runtime_error("parameter.h: RUNSPH = 1 selects the SPH solver, which this "
                                     "library does not provide (MFV / MFM only)");
    } else if (n == "MLH_PRECISION") q.precision = to_int(n, v);
    else if (n == "MLH_RIEMANN_SEED") q.riemann_seed = to_int(n, v);
    else if (n == "MLH_LIMITER_ABS") q.limiter_abs = to_int(n, v);
    else if (n == "MLH_QUIRK_3D_PREDICTOR") q.quirk_3d_predictor = to_int(n, v);
    else if (n == "MLH_WORK_ORDER") q.work_order = to_int(n, v);
    else p.ignored.push_back(n);
}

Parameters read_parameter_header(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("parameter.h: cannot open '" + path + "'");
    Parameters p = default_parameters();
    std::string line;
    bool in_block_comment = false;
    while (std::getline(in, line)) {
        if (in_block_comment) {
            size_t e = line.find("*/");
            if (e == std::string::npos) continue;
            line = line.substr(e + 2);
            in_block_comment = false;
        }
        size_t b = line.find("/*");
        if (b != std::string::npos && line.find("*/", b) == std::string::npos) {
            in_block_comment = true;
            line = line.substr(0, b);
        }
        std::string t = strip_comments(line);
        size_t k = 0;
        while (k < t.size() && std::isspace((unsigned char)t[k])) ++k;
        if (t.compare(k, 7, "#define") != 0) continue;
        k += 7;
        while (k < t.size() && std::isspace((unsigned char)t[k])) ++k;
        size_t a = k;
        while (k < t.size() && !std::isspace((unsigned char)t[k])) ++k;
        std::string name = t.substr(a, k - a);
        while (k < t.size() && std::isspace((unsigned char)t[k])) ++k;
        size_t e = t.size();
        while (e > k && std::isspace((unsigned char)t[e - 1])) --e;
        std::string value = t.substr(k, e - k);
        if (value.empty()) continue;  // include guards and flag-only defines
        apply_define(p, name, value);
    }
    return p;
}

}  // namespace mlh_host
