// capi.cpp -- C entry points of the host layer (for tools and tests in other
// languages): parameter.h and config.info readers and the whole driver.
#include <stdint.h>

#include <cstring>
#include <fstream>
#include <string>

#include "ConfigParser.h"
#include "Driver.h"
#include "Logger.h"

namespace mlh_host {

LogConfig &log_config() {
    static LogConfig c;
    return c;
}

Run load_run(const std::string &config, const std::string &parameter_h) {
    Run r;
    r.prm = parameter_h.empty() ? default_parameters() : read_parameter_header(parameter_h);
    ConfigParser cp(config);
    r.cfg.initFile = cp.getVal<std::string>("initFile");
    r.cfg.outDir = cp.getVal<std::string>("outDir");
    r.cfg.timeStep = cp.getVal<double>("timeStep");
    r.cfg.timeEnd = cp.getVal<double>("timeEnd");
    r.cfg.h5DumpInterval = cp.getVal<int>("h5DumpInterval");
    r.cfg.kernelSize = cp.getVal<double>("kernelSize");
    r.cfg.gamma = cp.getVal<double>("gamma");
    mlh_params &p = r.prm.params;
    p.kernel_size = r.cfg.kernelSize;
    p.gamma = r.cfg.gamma;
    if (p.periodic) {  // main.cpp:71-80
        ConfigParser b = cp.getObj("periodicBoxLimits");
        const char *lo[3] = {"lowerX", "lowerY", "lowerZ"}, *hi[3] = {"upperX", "upperY", "upperZ"};
        for (int a = 0; a < p.dim; ++a) {
            p.box_min[a] = r.cfg.periodicBoxLimits[a] = b.getVal<double>(lo[a]);
            p.box_max[a] = r.cfg.periodicBoxLimits[3 + a] = b.getVal<double>(hi[a]);
        }
    }
    // relative initFile / outDir are taken relative to the config file
    std::string dir = config.substr(0, config.find_last_of('/') + 1);
    if (!r.cfg.initFile.empty() && r.cfg.initFile[0] != '/') r.cfg.initFile = dir + r.cfg.initFile;
    if (!r.cfg.outDir.empty() && r.cfg.outDir[0] != '/') r.cfg.outDir = dir + r.cfg.outDir;
    return r;
}

void fill_particles(Particles &P, const InitialDistribution &ic) {
    for (int i = 0; i < ic.N; ++i) {  // InitialDistribution::getAllParticles
        P.m[i] = ic.m[i];
        P.u[i] = ic.u[i];
        P.matId[i] = ic.matId[i];
        P.x[i] = ic.x[(size_t)i * ic.D + 0];
        P.y[i] = ic.x[(size_t)i * ic.D + 1];
        P.vx[i] = ic.v[(size_t)i * ic.D + 0];
        P.vy[i] = ic.v[(size_t)i * ic.D + 1];
        if (ic.D == 3) {
            P.z[i] = ic.x[(size_t)i * ic.D + 2];
            P.vz[i] = ic.v[(size_t)i * ic.D + 2];
        }
    }
}

}  // namespace mlh_host

namespace {
thread_local std::string g_err2;
}

extern "C" {

const char *mlh_host_error(void) { return g_err2.c_str(); }

int mlh_host_read_parameters(const char *path, mlh_params *out, int32_t *adaptive) {
    try {
        mlh_host::Parameters p = mlh_host::read_parameter_header(path);
        *out = p.params;
        if (adaptive) *adaptive = p.adaptive_timestep ? 1 : 0;
        return 0;
    } catch (const std::exception &e) {
        g_err2 = e.what();
        return -1;
    }
}

int mlh_host_config_value(const char *path, const char *key, char *out, int32_t cap) {
    try {
        mlh_host::ConfigParser cp(path);
        std::string v = cp.getVal<std::string>(key);
        if ((int32_t)v.size() + 1 > cap) throw std::runtime_error("buffer too small");
        std::memcpy(out, v.c_str(), v.size() + 1);
        return 0;
    } catch (const std::exception &e) {
        g_err2 = e.what();
        return -1;
    }
}

// the whole driver: config.info + parameter.h (may be NULL: defaults) ->
// snapshots in outDir; returns 0 or the status of the failing call
// (1, 3, 6, 9 as the reference's exit codes), -1 for other errors
int mlh_host_run_methods(const char *config, const char *parameter_h, int32_t log_level,
                         int32_t per_method, int64_t *steps, double *t_end, int32_t *dumps);

int mlh_host_run(const char *config, const char *parameter_h, int32_t log_level,
                 int64_t *steps, double *t_end, int32_t *dumps) {
    return mlh_host_run_methods(config, parameter_h, log_level, 0, steps, t_end, dumps);
}

// the periodic images of n particles (host restatement of
// createGhostParticles): up to cap images into parent / gx, gy, gz
int mlh_host_ghost_images(const mlh_params *p, int32_t n, const double *x, const double *y,
                          const double *z, int32_t cap, int32_t *parent, double *gx,
                          double *gy, double *gz, int32_t *n_out) {
    try {
        const double *xs[3] = {x, y, z};
        std::vector<int> par;
        std::vector<double> g[3];
        std::vector<double> *gp[3] = {&g[0], &g[1], &g[2]};
        mlh_host::ghost_images(*p, n, xs, p->kernel_size, par, gp);
        *n_out = (int32_t)par.size();
        if ((int64_t)par.size() > cap) throw std::runtime_error("image capacity exceeded");
        double *out[3] = {gx, gy, gz};
        for (size_t k = 0; k < par.size(); ++k) {
            parent[k] = par[k];
            for (int a = 0; a < 3; ++a)
                if (out[a]) out[a][k] = g[a][k];
        }
        return 0;
    } catch (const std::exception &e) {
        g_err2 = e.what();
        return -1;
    }
}

int mlh_host_run_methods(const char *config, const char *parameter_h, int32_t log_level,
                         int32_t per_method, int64_t *steps, double *t_end, int32_t *dumps) {
    try {
        mlh_host::log_config().level = (mlh_host::LogLevel)log_level;
        mlh_host::Run r = mlh_host::load_run(config, parameter_h ? parameter_h : "");
        mlh_host::InitialDistribution ic = mlh_host::read_initial_distribution(r.cfg.initFile);
        if (ic.D != r.prm.params.dim) throw std::runtime_error("initial distribution dimension != DIM");
        mlh_host::Particles P(ic.N, r.prm.params);
        mlh_host::fill_particles(P, ic);
        P.upload();
        mlh_host::MeshlessScheme s(r.cfg, &P, r.prm.adaptive_timestep, per_method != 0);
        s.run();
        if (steps) *steps = s.steps();
        if (t_end) *t_end = s.time();
        if (dumps) *dumps = s.dumps();
        return 0;
    } catch (const mlh_host::MlhError &e) {
        g_err2 = e.what();
        return e.status;
    } catch (const std::exception &e) {
        g_err2 = e.what();
        return -1;
    }
}

}  // extern "C"
