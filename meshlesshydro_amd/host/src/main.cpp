// main.cpp -- the `mlh` driver (demonstrator/src/main.cpp): reads
// config.info and parameter.h, the initial distribution, and runs the MFV
// scheme on the GPU, writing snapshots as the reference does.
//
//   mlh -c config.info [-p parameter.h] [-v | -s] [--mlp network.f32]
//
// The process exits with the reference's codes on its failures (1 neighbour
// overflow, 3 ghost overflow, 6 negative face pressure, 9 dump overrun) and
// with the library's status otherwise.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "Driver.h"
#include "Logger.h"

using namespace mlh_host;

static void usage() {
    std::printf(
        "mlh - MFV / MFM meshless hydrodynamics on one AMD Instinct GPU\n"
        "  -c, --config FILE     config.info (default config.info)\n"
        "  -p, --parameters FILE parameter.h read at run time (default: built-in\n"
        "                        Kelvin-Helmholtz long-run set)\n"
        "      --mlp FILE        trained p* seed network (4929 float32 values)\n"
        "      --per-method      call the Particles methods one at a time, in the\n"
        "                        reference's order (same results)\n"
        "  -v, --verbose         more printouts\n"
        "  -s, --silent          warnings and errors only\n"
        "  -h, --help            this help\n");
}

int main(int argc, char **argv) {
    std::string config = "config.info", params, mlp;
    bool verbose = false, silent = false, per_method = false;
    for (int k = 1; k < argc; ++k) {
        std::string a = argv[k];
        auto next = [&](const char *what) -> std::string {
            if (k + 1 >= argc) {
                std::fprintf(stderr, "%s needs a value\n", what);
                std::exit(2);
            }
            return argv[++k];
        };
        if (a == "-c" || a == "--config") config = next("-c");
        else if (a == "-p" || a == "--parameters") params = next("-p");
        else if (a == "--mlp") mlp = next("--mlp");
        else if (a == "--per-method") per_method = true;
        else if (a == "-v" || a == "--verbose") verbose = true;
        else if (a == "-s" || a == "--silent") silent = true;
        else if (a == "-h" || a == "--help") { usage(); return 0; }
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); usage(); return 2; }
    }
    if (verbose && silent) {
        std::fprintf(stderr, "Command line options -s and -v are incompatible\n");
        return 2;
    }
    log_config().level = verbose ? DEBUG : (silent ? WARN : INFO);
    try {
        Logger(INFO) << "Reading configuration ... ";
        Run r = load_run(config, params);
        for (const auto &n : r.prm.ignored) Logger(DEBUG) << "    > parameter.h: " << n << " not used";
        Logger(INFO) << "    > Initial distribution: " << r.cfg.initFile;
        Logger(INFO) << "    > Output directory: " << r.cfg.outDir;
        Logger(INFO) << "    > Using global kernel size h = " << r.cfg.kernelSize;
        InitialDistribution ic = read_initial_distribution(r.cfg.initFile);
        if (ic.D != r.prm.params.dim)
            throw MlhError(MLH_ERR_INVALID_ARGUMENT, "initial distribution has D != DIM");
        Logger(INFO) << "    > N = " << ic.N;
        Particles P(ic.N, r.prm.params);
        fill_particles(P, ic);
        P.upload();
        if (!mlp.empty()) {
            std::vector<float> w(MLH_MLP_NPARAM);
            std::ifstream f(mlp, std::ios::binary);
            if (!f.read((char *)w.data(), sizeof(float) * w.size()))
                throw MlhError(MLH_ERR_INVALID_ARGUMENT, "cannot read " + mlp);
            check(mlh_mlp_set_params(P.context(), w.data()), "mlh_mlp_set_params");
        }
        MeshlessScheme s(r.cfg, &P, r.prm.adaptive_timestep, per_method);
        Logger(INFO) << "Starting time integration ...";
        s.run();
        mlh_diag d = P.diagnostics();
        Logger(INFO) << "... done: " << s.steps() << " steps, t = " << s.time() << ", "
                     << s.dumps() << " snapshots, " << d.faces_solved << " faces solved";
        return 0;
    } catch (const MlhError &e) {
        Logger(ERROR) << e.what();
        return e.status;
    } catch (const std::exception &e) {
        Logger(ERROR) << e.what();
        return 1;
    }
}
