// H5IO.cpp -- HDF5 I/O in the reference's layout (see H5IO.h).
#include "H5IO.h"

#include <hdf5.h>

#include <stdexcept>
#include <string>

namespace mlh_host {

namespace {

struct H5Quiet {  // no HDF5 error stack printing; errors surface as exceptions
    H5Quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};
const H5Quiet kQuiet;

struct Handle {
    hid_t id;
    herr_t (*close)(hid_t);
    Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Handle() {
        if (id >= 0) close(id);
    }
    Handle(const Handle &) = delete;
    Handle &operator=(const Handle &) = delete;
};

void write_ds(hid_t f, const char *name, hid_t type, int rank, const hsize_t *dims,
              const void *data) {
    Handle sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    if (sp.id < 0) throw std::runtime_error(std::string("h5: dataspace for ") + name);
    Handle ds(H5Dcreate2(f, name, type, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (ds.id < 0) throw std::runtime_error(std::string("h5: cannot create ") + name);
    if (H5Dwrite(ds.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
        throw std::runtime_error(std::string("h5: cannot write ") + name);
}

void shape(hid_t f, const std::string &name, std::vector<size_t> &dims) {
    Handle ds(H5Dopen2(f, name.c_str(), H5P_DEFAULT), H5Dclose);
    if (ds.id < 0) throw std::runtime_error("h5: no dataset '" + name + "'");
    Handle sp(H5Dget_space(ds.id), H5Sclose);
    int r = H5Sget_simple_extent_ndims(sp.id);
    if (r < 0 || r > 2) throw std::runtime_error("h5: dataset '" + name + "' has rank > 2");
    hsize_t d[2] = {0, 0};
    H5Sget_simple_extent_dims(sp.id, d, nullptr);
    dims.assign(d, d + r);
}

template <typename T>
std::vector<T> read_ds(hid_t f, const std::string &name, hid_t mem, std::vector<size_t> *dims) {
    std::vector<size_t> d;
    shape(f, name, d);
    size_t n = 1;
    for (size_t k : d) n *= k;
    std::vector<T> out(n);
    Handle ds(H5Dopen2(f, name.c_str(), H5P_DEFAULT), H5Dclose);
    if (n && H5Dread(ds.id, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) < 0)
        throw std::runtime_error("h5: cannot read '" + name + "'");
    if (dims) *dims = d;
    return out;
}

hid_t open_ro(const std::string &file) {
    hid_t f = H5Fopen(file.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) throw std::runtime_error("h5: cannot open '" + file + "'");
    return f;
}

}  // namespace

InitialDistribution read_initial_distribution(const std::string &file) {
    Handle f(open_ro(file), H5Fclose);
    InitialDistribution ic;
    std::vector<size_t> dx, dv, dm, du, di;
    ic.m = read_ds<double>(f.id, "/m", H5T_NATIVE_DOUBLE, &dm);
    ic.x = read_ds<double>(f.id, "/x", H5T_NATIVE_DOUBLE, &dx);
    ic.v = read_ds<double>(f.id, "/v", H5T_NATIVE_DOUBLE, &dv);
    ic.u = read_ds<double>(f.id, "/u", H5T_NATIVE_DOUBLE, &du);
    ic.matId = read_ds<int>(f.id, "/materialId", H5T_NATIVE_INT, &di);
    if (dx.size() != 2 || dv.size() != 2 || dx[1] != dv[1])
        throw std::runtime_error("h5: /x and /v must be [N][D] with the same D");
    size_t N = dx[0];
    if (dv[0] != N || ic.m.size() != N || ic.u.size() != N || ic.matId.size() != N)
        throw std::runtime_error("Length mismatch between mass, position and/or velocity vectors.");
    ic.N = (int)N;
    ic.D = (int)dx[1];
    return ic;
}

void write_initial_distribution(const std::string &file, const InitialDistribution &ic) {
    Handle f(H5Fcreate(file.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) throw std::runtime_error("h5: cannot create '" + file + "'");
    hsize_t n[1] = {(hsize_t)ic.N}, nd[2] = {(hsize_t)ic.N, (hsize_t)ic.D};
    write_ds(f.id, "/m", H5T_NATIVE_DOUBLE, 1, n, ic.m.data());
    write_ds(f.id, "/x", H5T_NATIVE_DOUBLE, 2, nd, ic.x.data());
    write_ds(f.id, "/v", H5T_NATIVE_DOUBLE, 2, nd, ic.v.data());
    write_ds(f.id, "/u", H5T_NATIVE_DOUBLE, 1, n, ic.u.data());
    write_ds(f.id, "/materialId", H5T_NATIVE_INT, 1, n, ic.matId.data());
}

void write_snapshot(const std::string &file, const Snapshot &s) {
    Handle f(H5Fcreate(file.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) throw std::runtime_error("h5: cannot create '" + file + "'");
    hsize_t one[1] = {1}, n[1] = {(hsize_t)s.N}, nd[2] = {(hsize_t)s.N, (hsize_t)s.D};
    write_ds(f.id, "/time", H5T_NATIVE_DOUBLE, 1, one, &s.time);
    write_ds(f.id, "/totalMass", H5T_NATIVE_DOUBLE, 1, one, &s.totalMass);
    write_ds(f.id, "/energy", H5T_NATIVE_DOUBLE, 1, one, &s.energy);
    write_ds(f.id, "/xMomentum", H5T_NATIVE_DOUBLE, 1, one, &s.mom[0]);
    write_ds(f.id, "/yMomentum", H5T_NATIVE_DOUBLE, 1, one, &s.mom[1]);
    if (s.D == 3) write_ds(f.id, "/zMomentum", H5T_NATIVE_DOUBLE, 1, one, &s.mom[2]);
    write_ds(f.id, "/rho", H5T_NATIVE_DOUBLE, 1, n, s.rho.data());
    write_ds(f.id, "/m", H5T_NATIVE_DOUBLE, 1, n, s.m.data());
    write_ds(f.id, "/u", H5T_NATIVE_DOUBLE, 1, n, s.u.data());
    write_ds(f.id, "/x", H5T_NATIVE_DOUBLE, 2, nd, s.x.data());
    write_ds(f.id, "/v", H5T_NATIVE_DOUBLE, 2, nd, s.v.data());
    write_ds(f.id, "/rhoGrad", H5T_NATIVE_DOUBLE, 2, nd, s.rhoGrad.data());
    write_ds(f.id, "/P", H5T_NATIVE_DOUBLE, 1, n, s.P.data());
    write_ds(f.id, "/noi", H5T_NATIVE_INT, 1, n, s.noi.data());
}

std::vector<double> read_dataset_f64(const std::string &file, const std::string &name,
                                     std::vector<size_t> *dims) {
    Handle f(open_ro(file), H5Fclose);
    return read_ds<double>(f.id, name, H5T_NATIVE_DOUBLE, dims);
}

std::vector<int> read_dataset_i32(const std::string &file, const std::string &name,
                                  std::vector<size_t> *dims) {
    Handle f(open_ro(file), H5Fclose);
    return read_ds<int>(f.id, name, H5T_NATIVE_INT, dims);
}

}  // namespace mlh_host

// ---------------------------------------------------------------- C entries
namespace {
thread_local std::string g_err;
template <typename F>
int guard(F &&fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}
}  // namespace

extern "C" {

const char *mlh_host_last_error(void) { return g_err.c_str(); }

int mlh_h5_write_ic(const char *file, int32_t n, int32_t dim, const double *x, const double *v,
                    const double *m, const double *u, const int32_t *material_id) {
    return guard([&] {
        if (!file || n < 0 || (dim != 2 && dim != 3) || !x || !v || !m || !u)
            throw std::runtime_error("mlh_h5_write_ic: bad arguments");
        mlh_host::InitialDistribution ic;
        ic.N = n;
        ic.D = dim;
        ic.x.assign(x, x + (size_t)n * dim);
        ic.v.assign(v, v + (size_t)n * dim);
        ic.m.assign(m, m + n);
        ic.u.assign(u, u + n);
        ic.matId.assign((size_t)n, 0);
        if (material_id) ic.matId.assign(material_id, material_id + n);
        mlh_host::write_initial_distribution(file, ic);
    });
}

int mlh_h5_write_snapshot(const char *file, int32_t n, int32_t dim, double time,
                          double total_mass, double energy, const double *momentum,
                          const double *rho, const double *m, const double *u, const double *x,
                          const double *v, const double *rho_grad, const double *P,
                          const int32_t *noi) {
    return guard([&] {
        if (!file || n < 0 || (dim != 2 && dim != 3) || !momentum || !rho || !m || !u || !x ||
            !v || !rho_grad || !P || !noi)
            throw std::runtime_error("mlh_h5_write_snapshot: bad arguments");
        mlh_host::Snapshot s;
        s.N = n;
        s.D = dim;
        s.time = time;
        s.totalMass = total_mass;
        s.energy = energy;
        for (int a = 0; a < dim; ++a) s.mom[a] = momentum[a];
        s.rho.assign(rho, rho + n);
        s.m.assign(m, m + n);
        s.u.assign(u, u + n);
        s.P.assign(P, P + n);
        s.x.assign(x, x + (size_t)n * dim);
        s.v.assign(v, v + (size_t)n * dim);
        s.rhoGrad.assign(rho_grad, rho_grad + (size_t)n * dim);
        s.noi.assign(noi, noi + n);
        mlh_host::write_snapshot(file, s);
    });
}

int mlh_h5_dataset_shape(const char *file, const char *name, int32_t *ndims, int64_t *dims) {
    return guard([&] {
        if (!file || !name || !ndims || !dims) throw std::runtime_error("null argument");
        std::vector<size_t> d;
        mlh_host::read_dataset_f64(file, name, &d);
        *ndims = (int32_t)d.size();
        for (size_t k = 0; k < d.size(); ++k) dims[k] = (int64_t)d[k];
    });
}

int mlh_h5_read_f64(const char *file, const char *name, double *out, int64_t count) {
    return guard([&] {
        auto v = mlh_host::read_dataset_f64(file, name);
        if ((int64_t)v.size() != count) throw std::runtime_error("size mismatch");
        for (size_t k = 0; k < v.size(); ++k) out[k] = v[k];
    });
}

int mlh_h5_read_i32(const char *file, const char *name, int32_t *out, int64_t count) {
    return guard([&] {
        auto v = mlh_host::read_dataset_i32(file, name);
        if ((int64_t)v.size() != count) throw std::runtime_error("size mismatch");
        for (size_t k = 0; k < v.size(); ++k) out[k] = v[k];
    });
}

}  // extern "C"
