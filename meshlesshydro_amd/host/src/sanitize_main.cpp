// sanitize_main.cpp -- runs the host layer's CPU paths for its
// AddressSanitizer / UndefinedBehaviorSanitizer build (make sanitize):
// the parameter.h and config.info readers (valid and malformed input), the
// HDF5 initial conditions and snapshots, and the periodic ghost images in
// 2D and 3D.  No device work.  Exits 1 on a failed check; the sanitizers
// abort on the first error they find.
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

#include <cmath>
#include <fstream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "ConfigParser.h"
#include "H5IO.h"
#include "Parameters.h"
#include "Particles.h"

using namespace mlh_host;

static int g_fail = 0;
#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            ++g_fail;                                                            \
        }                                                                        \
    } while (0)

template <typename F>
static bool throws(F f) {
    try {
        f();
    } catch (const std::exception &) {
        return true;
    }
    return false;
}

static void write_file(const std::string &path, const std::string &text) {
    std::ofstream(path) << text;
}

static void parameters(const std::string &dir) {
    const std::string ph = dir + "/parameter.h";
    write_file(ph,
               "#ifndef P_H\n#define P_H\n#define DIM 3\n#define PERIODIC_BOUNDARIES 1\n"
               "#define ADAPTIVE_TIMESTEP 0\n#define CFL .3 // comment\n"
               "#define MAX_NUM_INTERACTIONS 128\n/* #define MAX_NUM_INTERACTIONS 7 */\n"
               "#define MAX_NUM_GHOST_INTERACTIONS 96\n#define BETA 2.\n#define PSI_1 .4\n"
               "#define ARTVISC 1\n#endif\n");
    Parameters p = read_parameter_header(ph);
    CHECK(p.params.dim == 3 && p.params.periodic == 1 && !p.adaptive_timestep);
    CHECK(p.params.max_interactions == 128 && p.params.max_ghost_interactions == 96);
    CHECK(std::fabs(p.params.cfl - 0.3) < 1e-15 && std::fabs(p.params.psi1 - 0.4) < 1e-15);
    CHECK(!p.ignored.empty());
    CHECK(throws([&] { apply_define(p, "DIM", "4"); }));
    CHECK(throws([&] { apply_define(p, "CFL", "x"); }));
    CHECK(throws([&] { apply_define(p, "RUNSPH", "1"); }));
    CHECK(throws([&] { read_parameter_header(dir + "/missing.h"); }));
    fprintf(stderr, "[sanitize] parameter.h reader: done\n");
}

static void config(const std::string &dir) {
    const std::string ci = dir + "/config.info";
    write_file(ci,
               "; comment\ninitFile ic.h5\noutDir \"out dir\"\ntimeStep .005\n"
               "periodicBoxLimits {\n  lowerX 0.\n  lowerY -1.\n  upperX 1.\n  upperY 2.\n}\n"
               "list {\n  a 1\n  b 2\n  c 3\n}\nkernelSize .025 ; trailing\n");
    ConfigParser c(ci);
    CHECK(c.getVal<std::string>("initFile") == "ic.h5");
    CHECK(c.getVal<std::string>("outDir") == "out dir");
    CHECK(c.getVal<double>("timeStep") == 0.005);
    CHECK(c.getObj("periodicBoxLimits").getVal<double>("upperY") == 2.0);
    CHECK(c.getVal<double>("periodicBoxLimits.lowerY") == -1.0);
    auto l = c.getList<int>("list");
    CHECK(l.size() == 3 && l.front() == 1 && l.back() == 3);
    CHECK(c.has("kernelSize") && !c.has("nothing"));
    CHECK(!c.flatten().empty());
    CHECK(throws([&] { c.getVal<double>("nothing"); }));
    CHECK(throws([&] { c.getVal<double>("initFile"); }));
    write_file(dir + "/bad.info", "block {\n  a 1\n");  // unterminated block
    CHECK(throws([&] { ConfigParser bad(dir + "/bad.info"); }));
    CHECK(throws([&] { ConfigParser none(dir + "/none.info"); }));
    fprintf(stderr, "[sanitize] config.info reader: done\n");
}

static void hdf5(const std::string &dir) {
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    InitialDistribution ic;
    ic.N = 257;
    ic.D = 3;
    for (int k = 0; k < ic.N; ++k) {
        for (int a = 0; a < ic.D; ++a) {
            ic.x.push_back(U(rng));
            ic.v.push_back(U(rng) - 0.5);
        }
        ic.m.push_back(1.0 / ic.N);
        ic.u.push_back(1.5 + U(rng));
        ic.matId.push_back(k % 3);
    }
    const std::string f = dir + "/ic.h5";
    write_initial_distribution(f, ic);
    InitialDistribution r = read_initial_distribution(f);
    CHECK(r.N == ic.N && r.D == ic.D && r.x == ic.x && r.v == ic.v && r.m == ic.m &&
          r.u == ic.u && r.matId == ic.matId);
    Snapshot s;
    s.N = ic.N;
    s.D = ic.D;
    s.time = 0.25;
    s.rho.assign(s.N, 1.0);
    s.P.assign(s.N, 0.6);
    s.m = ic.m;
    s.u = ic.u;
    s.x = ic.x;
    s.v = ic.v;
    s.rhoGrad.assign((size_t)s.N * s.D, 0.0);
    s.noi.assign(s.N, 32);
    s.totalMass = 1.0;
    const std::string sf = dir + "/000001.h5";
    write_snapshot(sf, s);
    std::vector<size_t> dims;
    CHECK(read_dataset_f64(sf, "/x", &dims) == s.x && dims.size() == 2 && dims[1] == 3);
    CHECK(read_dataset_i32(sf, "/noi") == s.noi);
    CHECK(read_dataset_f64(sf, "/time")[0] == 0.25);
    CHECK(throws([&] { read_dataset_f64(sf, "/missing"); }));
    CHECK(throws([&] { read_initial_distribution(dir + "/none.h5"); }));
    fprintf(stderr, "[sanitize] HDF5 I/O: done\n");
}

static void ghosts() {
    for (int dim = 2; dim <= 3; ++dim) {
        mlh_params p;
        mlh_params_default(&p);
        p.dim = dim;
        p.periodic = 1;
        const double h = 0.2;
        std::mt19937_64 rng(11 + dim);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        const int n = 300;
        std::vector<double> c[3];
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < n; ++k) c[a].push_back(a < dim ? U(rng) : 0.0);
        c[0][0] = 0.0;  // on the boundary
        c[1][1] = 1.0 - 1e-12;
        const double *xs[3] = {c[0].data(), c[1].data(), c[2].data()};
        std::vector<int> parent;
        std::vector<double> g[3];
        std::vector<double> *gx[3] = {&g[0], &g[1], &g[2]};
        ghost_images(p, n, xs, h, parent, gx);
        CHECK(!parent.empty());
        for (size_t k = 0; k < parent.size(); ++k) {
            CHECK(parent[k] >= 0 && parent[k] < n);
            bool outside = false;
            for (int a = 0; a < dim; ++a) {
                const double x = (*gx[a])[k];
                CHECK(x > -h - 1e-12 && x < 1.0 + h + 1e-12);
                outside = outside || x <= 0.0 || x >= 1.0;
            }
            CHECK(outside);
        }
    }
    fprintf(stderr, "[sanitize] ghost images: done\n");
}

int main() {
    char tmpl[] = "/tmp/mlh_sanitize_XXXXXX";
    const char *dir = mkdtemp(tmpl);
    if (!dir) return 2;
    try {
        parameters(dir);
        config(dir);
        hdf5(dir);
        ghosts();
    } catch (const std::exception &e) {
        fprintf(stderr, "unexpected exception: %s\n", e.what());
        ++g_fail;
    }
    const std::string cmd = std::string("rm -rf ") + dir;
    if (system(cmd.c_str()) != 0) ++g_fail;
    return g_fail ? 1 : 0;
}
