// H5IO.h -- HDF5 initial conditions and snapshots in the reference's layout,
// written with the HDF5 C library.
//
// Initial conditions (InitialDistribution.cpp:8-29): datasets /m [N],
// /x [N][D], /v [N][D], /u [N], /materialId [N] (int).
// Snapshots (Particles::dump2file, Particles.cpp:2975-3075), file
// outDir/%06d.h5: /time [1], /totalMass [1], /energy [1], /xMomentum [1],
// /yMomentum [1], /zMomentum [1] (3D), /rho [N], /m [N], /u [N], /x [N][D],
// /v [N][D], /rhoGrad [N][D], /P [N], /noi [N] (int).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace mlh_host {

struct InitialDistribution {
    int N = 0, D = 0;
    std::vector<double> x, v;  // [N][D]
    std::vector<double> m, u;  // [N]
    std::vector<int> matId;    // [N]
};

// throws std::runtime_error when a dataset is missing or lengths mismatch
// (the reference throws std::length_error, InitialDistribution.cpp:28)
InitialDistribution read_initial_distribution(const std::string &file);
void write_initial_distribution(const std::string &file, const InitialDistribution &ic);

struct Snapshot {
    int N = 0, D = 0;
    double time = 0, totalMass = 0, energy = 0, mom[3] = {0, 0, 0};
    std::vector<double> rho, m, u, P;      // [N]
    std::vector<double> x, v, rhoGrad;     // [N][D]
    std::vector<int> noi;                  // [N]
};
void write_snapshot(const std::string &file, const Snapshot &s);

// 1-D or 2-D dataset read into doubles / ints; dims receives the shape
std::vector<double> read_dataset_f64(const std::string &file, const std::string &name,
                                     std::vector<size_t> *dims = nullptr);
std::vector<int> read_dataset_i32(const std::string &file, const std::string &name,
                                  std::vector<size_t> *dims = nullptr);

}  // namespace mlh_host

// C entry points of the same I/O (for tools and tests in other languages).
// Return 0 on success, -1 on failure (message in mlh_host_last_error()).
extern "C" {
const char *mlh_host_last_error(void);
int mlh_h5_write_ic(const char *file, int32_t n, int32_t dim, const double *x, const double *v,
                    const double *m, const double *u, const int32_t *material_id);
/* one snapshot in the dump2file layout; x, v, rho_grad are [n][dim],
   momentum has dim entries */
int mlh_h5_write_snapshot(const char *file, int32_t n, int32_t dim, double time,
                          double total_mass, double energy, const double *momentum,
                          const double *rho, const double *m, const double *u, const double *x,
                          const double *v, const double *rho_grad, const double *P,
                          const int32_t *noi);
/* shape of dataset `name`: ndims <= 2 into dims[0..1] */
int mlh_h5_dataset_shape(const char *file, const char *name, int32_t *ndims, int64_t *dims);
int mlh_h5_read_f64(const char *file, const char *name, double *out, int64_t count);
int mlh_h5_read_i32(const char *file, const char *name, int32_t *out, int64_t count);
}
