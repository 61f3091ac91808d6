#pragma once
// hydro_impl.h -- MFV physics kernels: density, gradients + limiter + CFL
// step, faces + exact Riemann + flux sums + state update.  Included by
// hydro.hip (FP64 instantiations, common kernels, launchers) and
// hydro_fp32.hip (FP32 instantiations).
//
// One thread per particle, in work order (position t; the particle's
// original index is sidx[t], or t when the work order is the caller's);
// neighbour rows come from cells.hip in the reference's order (real
// neighbours in stencil order, then periodic images in ghost-index order),
// so every per-particle sum below runs in the reference's summation order.
// Per-particle records (pos, prim, frec) are read and written at work
// positions; the state, its successor and the flux sums at original
// indices.  Blocks are mapped XCD-aware (xcd_block).
//
// Faces are evaluated i-centric and without atomics: particle i visits each
// of its neighbours and solves the pair from the side the reference solves
// it (the lower index; for an image pair the reference's test parent >= i),
// i.e. when i is the higher index it recomputes the lower index's face with
// the same operands and subtracts it.  Both sides therefore hold bitwise
// opposite fluxes (F_ij = -F_ji) and the update is conservative to
// round-off and independent of scheduling.
#include <hip/hip_runtime.h>
#include "device_common.h"
#include "kernels.h"
#include "riemann.h"
#include "mlp_device.h"
#include "mlp_mfma.h"

// FP64 kernels (hydro.hip) are compiled without contraction, so that they
// round like the reference; the FP32 kernels (hydro_fp32.hip, MLH_FP32_TU)
// let the compiler fuse a * b + c within one expression into an FMA.  The
// FP64 limiter extrema keep separate rounding in both (see k_gradients).
#ifdef MLH_FP32_TU
#pragma clang fp contract(on)
#else
#pragma clang fp contract(off)
#endif

namespace mlh {

template <typename T> __device__ inline T tmin_pos();
template <> __device__ inline double tmin_pos<double>() { return DBL_MIN; }
template <> __device__ inline float tmin_pos<float>() { return FLT_MIN; }
template <typename T> __device__ inline T tmax_val();
template <> __device__ inline double tmax_val<double>() { return DBL_MAX; }
template <> __device__ inline float tmax_val<float>() { return FLT_MAX; }

__device__ inline double tsqrt(double x) { return sqrt(x); }
__device__ inline float tsqrt(float x) { return __builtin_sqrtf(x); }

template <int D>
__device__ inline double dist2(const double *a, const double *b) {
    double d = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]);
    if (D == 3) d += (a[2] - b[2]) * (a[2] - b[2]);
    return d;
}

template <int D, typename T>
__device__ inline T dotT(const T *a, const T *b) {
    T r = T(0);
    for (int k = 0; k < D; ++k) r += a[k] * b[k];
    return r;
}

// neighbour position (periodic image applied) and parent index
template <int D>
__device__ inline int nb_pos(const HydroArgs &a, uint32_t e, double *xj, int *s) {
    int j = (int)nb_index(e);
    uint32_t code = nb_code(e);
    double4 p = a.pos[j];
    xj[0] = p.x; xj[1] = p.y; xj[2] = p.z;
    s[0] = s[1] = s[2] = 0;
    if (code != kCodeReal) {
        for (int ax = 0; ax < D; ++ax) {
            s[ax] = code_shift(code, ax);
            xj[ax] = image_coord(xj[ax], s[ax], a.g.bmin[ax], a.g.bmax[ax]);
        }
    }
    return j;
}

// position of row entry e from its loaded cell-order record p (the periodic
// image applied for an image entry), as nb_pos gives it
template <int D>
__device__ inline void entry_pos(const HydroArgs &a, uint32_t e, const double4 &p, double *xj) {
    xj[0] = p.x; xj[1] = p.y; xj[2] = p.z;
    const uint32_t code = nb_code(e);
    if (code != kCodeReal)
        for (int ax = 0; ax < D; ++ax)
            xj[ax] = image_coord(xj[ax], code_shift(code, ax), a.g.bmin[ax], a.g.bmax[ax]);
}

template <typename T, int N>
__device__ inline void load_vec(T *dst, const T *__restrict__ src) {
    static_assert((N * sizeof(T)) % 16 == 0, "record size");
    const T *s = (const T *)__builtin_assume_aligned(src, 16);
    __builtin_memcpy(dst, s, N * sizeof(T));
}

__device__ inline unsigned long long wave_sum(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ------------------------------------------------------------------ density
// compDensity / compOmega (Particles.cpp:1151-1184), ghost part
// (:2262-2290), compPressure (:1272-1288); also c = sqrt(gamma P / rho)
// for the time step (:1453).
template <int D, typename T>
__global__ __launch_bounds__(256) void k_density(HydroArgs a) {
    int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= a.N) return;
    const int N = a.N;
    const int i = a.sidx ? (int)a.sidx[t] : t;
    double4 pi = a.pos[t];
    double xi[3] = {pi.x, pi.y, pi.z};
    const T h2 = T(a.g.h / 2.), sig = T(a.g.sigma);
    int nr = a.noi[t], ng = a.nog[t];
    T omg = T(0);
    // real neighbours (no image transform), eight entries at a time: the
    // eight row entries, then the eight positions, are in flight together;
    // the sum keeps the entry order
    int n = 0;
    for (; n + 8 <= nr; n += 8) {
        uint32_t e[8];
        nb_get8(a.nbr, N, t, n, e);
        double4 pj[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) pj[u] = a.pos[nb_index(e[u])];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double xj[3] = {pj[u].x, pj[u].y, pj[u].z};
            T r = tsqrt(T(dist2<D>(xi, xj)));
            omg += kernel_w<T>(r, h2, sig);
        }
    }
    for (; n < nr; ++n) {
        const double4 pj = a.pos[nb_index(nb_get(a.nbr, N, t, n))];
        const double xj[3] = {pj.x, pj.y, pj.z};
        T r = tsqrt(T(dist2<D>(xi, xj)));
        omg += kernel_w<T>(r, h2, sig);
    }
    T om = omg + kernel_w<T>(T(0), h2, sig);
    if (a.g.periodic) {
        // images, gathered eight at a time like the real neighbours
        T og = om;
        int n = nr;
        for (; n + 8 <= nr + ng; n += 8) {
            uint32_t e[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) e[u] = nb_get(a.nbr, N, t, n + u);
            double4 pj[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) pj[u] = a.pos[nb_index(e[u])];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                double xj[3];
                entry_pos<D>(a, e[u], pj[u], xj);
                T r = tsqrt(T(dist2<D>(xi, xj)));
                og += kernel_w<T>(r, h2, sig);
            }
        }
        for (; n < nr + ng; ++n) {
            double xj[3];
            int s[3];
            nb_pos<D>(a, nb_get(a.nbr, N, t, n), xj, s);
            T r = tsqrt(T(dist2<D>(xi, xj)));
            og += kernel_w<T>(r, h2, sig);
        }
        om = og;
    }
    write_prim<D, T>((T *)a.prim + (size_t)t * kPrimN, om, a.m[i], a.u[i], a.v[0][i], a.v[1][i],
                     D == 3 ? a.v[2][i] : 0., a.gamma);
}

// Helper::inverseMatrix (Helper.cpp:7-18, LAPACK dgetrf_/dgetri_): Gauss-
// Jordan with partial pivoting, FP64.
// Returns false when a pivot is zero or not finite (LAPACK dgetrf's INFO > 0,
// which the reference does not check: the inverse is then inf / NaN).
template <int D>
__device__ inline bool invert(double *A) {
    bool ok = true;
    double M[D][2 * D];
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < 2 * D; ++c) M[r][c] = c < D ? A[r * D + c] : (c - D == r ? 1. : 0.);
    for (int k = 0; k < D; ++k) {
        // first row with the largest |M[r][k]|, swapped with row k; rows are
        // addressed with compile-time indices only (no scratch memory)
        int piv = k;
        double best = fabs(M[k][k]);
        for (int r = k + 1; r < D; ++r) {
            double v = fabs(M[r][k]);
            if (v > best) { best = v; piv = r; }
        }
        for (int r = k + 1; r < D; ++r) {
            bool sw = piv == r;
            for (int c = 0; c < 2 * D; ++c) {
                double a0 = M[k][c], b0 = M[r][c];
                M[k][c] = sw ? b0 : a0;
                M[r][c] = sw ? a0 : b0;
            }
        }
        double d = M[k][k];
        ok &= d != 0. && isfinite(d);
        for (int c = 0; c < 2 * D; ++c) M[k][c] /= d;
        for (int r = 0; r < D; ++r) {
            if (r == k) continue;
            double f = M[r][k];
            for (int c = 0; c < 2 * D; ++c) M[r][c] -= f * M[k][c];
        }
    }
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) A[r * D + c] = M[r][D + c];
    return ok;
}

// ---------------------------------------------------------------- gradients
// slopeLimiter (Particles.cpp:1313-1444) for particle t (work position) with
// fi its fields, grad its gradients (limited in place) and ngbmax / ngbmin
// the extrema of its neighbours' fields.  The reconstructed extrema and the
// ratios are evaluated in FP64 in both precisions, with separate rounding
// as in the reference: with its int abs() a rounding tie f_ij == f_i turns
// into a 2^31 scaling, and FP32 reconstructions would create such ties
// where FP64 has none.
template <int D, typename T>
__device__ inline void slope_limit(const HydroArgs &a, int t, int cnt, const double *xi,
                                   const T *fi, T (&grad)[D + 2][D], const T *ngbmax,
                                   const T *ngbmin) {
#pragma clang fp contract(off)
    constexpr int NF = D + 2;
    const int N = a.N;
    // The extrema of f_ij = f_i + grad f . (x_ij - x_i) over the neighbours
    // are those of the dot products, f_i added once after the loop: rounding
    // is monotone, so max_j fl(f_i + d_j) = fl(f_i + max_j d_j) exactly (and
    // NaN products are skipped either way)
    double dmax[NF], dmin[NF];
    for (int f = 0; f < NF; ++f) {
        dmax[f] = -HUGE_VAL;
        dmin[f] = HUGE_VAL;
    }
    double gd[NF][D];
    for (int f = 0; f < NF; ++f)
        for (int ax = 0; ax < D; ++ax) gd[f][ax] = (double)grad[f][ax];
    auto visit = [&](const double *xj) {
        double xijxi[D];
        for (int ax = 0; ax < D; ++ax) {
            double xij = a.foqp ? (xi[ax] + xj[ax]) / 2.
                                : xi[ax] + a.g.h / 4. * (xj[ax] - xi[ax]);
            xijxi[ax] = xij - xi[ax];
        }
        for (int f = 0; f < NF; ++f) {
            const double dj = dotT<D, double>(gd[f], xijxi);
            if (dmax[f] < dj) dmax[f] = dj;
            if (dmin[f] > dj) dmin[f] = dj;
        }
    };
    // neighbours in batches of four: the four row entries (one 16-byte
    // load), then their positions, are in flight together; the batch is
    // visited in entry order
    int n = 0;
    for (; n + 4 <= cnt; n += 4) {
        uint32_t e[4];
        nb_get4(a.nbr, N, t, n, e);
        double4 pb[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) pb[u] = a.pos[nb_index(e[u])];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            double xj[3];
            entry_pos<D>(a, e[u], pb[u], xj);
            visit(xj);
        }
    }
    for (; n < cnt; ++n) {
        double xj[3];
        int s[3];
        nb_pos<D>(a, nb_get(a.nbr, N, t, n), xj, s);
        visit(xj);
    }
    // midmax / midmin start from the smallest positive and the largest
    // value of T, as in the reference
    double midmax[NF], midmin[NF];
    for (int f = 0; f < NF; ++f) {
        midmax[f] = tmin_pos<T>();
        midmin[f] = tmax_val<T>();
        const double hi = (double)fi[f] + dmax[f], lo = (double)fi[f] + dmin[f];
        if (midmax[f] < hi) midmax[f] = hi;
        if (midmin[f] > lo) midmin[f] = lo;
    }
    const double beta = a.beta;
    for (int f = 0; f < NF; ++f) {
        double ff = (double)fi[f];
        double amax = lim_abs<double>(((double)ngbmax[f] - ff) / (midmax[f] - ff), a.limiter_abs);
        double amin = lim_abs<double>((ff - (double)ngbmin[f]) / (ff - midmin[f]), a.limiter_abs);
        if (amin <= amax && beta * amin < 1.) {
            for (int ax = 0; ax < D; ++ax) grad[f][ax] = T((double)grad[f][ax] * amin);
        } else if (amax <= amin && beta * amax < 1.) {
            for (int ax = 0; ax < D; ++ax) grad[f][ax] = T((double)grad[f][ax] * amax);
        }
    }
}

#ifndef MLH_GRAD_BATCH
#define MLH_GRAD_BATCH 4
#endif
// real neighbours per gather batch of the FP32 gradient pass
constexpr int kGradBatch = MLH_GRAD_BATCH;
// compPsijTilde (Particles.cpp:1186-1254, ghosts :2292-2455): E_i in FP64,
// B_i = E_i^-1 in FP64; gradient (:1257-1270, :2457-2502) for rho, v, P;
// slopeLimiter (:1313-1444); compGlobalTimestep (:1446-1485) per particle.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_gradients(HydroArgs a) {
    constexpr int NF = D + 2;
    constexpr int FRN = FRec<D>::N;
    __shared__ double sdt[256];
    int t = xcd_block() * blockDim.x + threadIdx.x;
    double dt_i = DBL_MAX;
    bool singular = false;
    if (t < a.N) {
        const int N = a.N;
        const int i = a.sidx ? (int)a.sidx[t] : t;
        double4 pi = a.pos[t];
        double xi[3] = {pi.x, pi.y, pi.z};
        const T *prim = (const T *)a.prim;
        const T *pri = prim + (size_t)t * kPrimN;
        T om_i = pri[kPrimOmega];
        T fi[NF];
        for (int f = 0; f < NF; ++f) fi[f] = pri[field_slot<D>(f)];
        const T h2 = T(a.g.h / 2.), sig = T(a.g.sigma);
        int cnt = a.noi[t] + a.nog[t], nr = a.noi[t];
        // pass 1: E_i and the signal speed (real neighbours only).  FP32
        // mode also accumulates G_f = sum_j (f_j - f_i)(x_j - x_i) W/omega_i
        // and the neighbour extrema here: grad f = sum_j (f_j - f_i) psi_j
        // = B G_f, so pass 2 is not needed.  FP64 mode keeps the reference's
        // per-neighbour psi_j and its summation order (pass 2).
        constexpr bool kFused = sizeof(T) == 4;
        double E[D * D];
        for (int k = 0; k < D * D; ++k) E[k] = 0.;
        double vsig = DBL_MIN;
        double ci = (double)pri[kPrimC];
        double vi[3] = {a.v[0][i], a.v[1][i], D == 3 ? a.v[2][i] : 0.};
        T grad[NF][D];
        T ngbmax[NF], ngbmin[NF];
        T G[kFused ? NF : 1][D];
        for (int f = 0; f < NF; ++f) {
            for (int ax = 0; ax < D; ++ax) grad[f][ax] = T(0);
            ngbmax[f] = tmin_pos<T>();
            ngbmin[f] = tmax_val<T>();
        }
        for (int f = 0; f < (kFused ? NF : 1); ++f)
            for (int ax = 0; ax < D; ++ax) G[f][ax] = T(0);
        // one neighbour of pass 1: n = entry, j = its work position, xj its
        // position (image applied), prj its record (FP32 mode)
        auto visit = [&](int n, int j, const double *xj, const T *prj) {
            // one set of differences dd = x_j - x_i serves the distance, E,
            // the FP32 gradient sums and the signal speed; (-a)(-a) = a a and
            // -(a + b) = (-a) + (-b) exactly, so |dd|^2 is dist2(x_i, x_j)
            // and -(v_ij . dd) is v_ij . x_ij bit for bit
            double dd[3] = {0., 0., 0.};
            for (int p = 0; p < D; ++p) dd[p] = xj[p] - xi[p];
            const double d2 = dotT<D, double>(dd, dd);
            T r = tsqrt(T(d2));
            T wt = kernel_w<T>(r, h2, sig) / om_i;
            double w = (double)wt;
            // E is symmetric and a*b*w == b*a*w in IEEE arithmetic: the
            // upper triangle, mirrored after the loop
            for (int p = 0; p < D; ++p)
                for (int q = p; q < D; ++q) E[D * p + q] += dd[p] * dd[q] * w;
            if (kFused) {
                T dx[D];
                for (int p = 0; p < D; ++p) dx[p] = T(dd[p]) * wt;
                for (int f = 0; f < (kFused ? NF : 1); ++f) {
                    T fj = prj[field_slot<D>(f)];
                    T df = fj - fi[f];
                    for (int p = 0; p < D; ++p) G[f][p] += df * dx[p];
                    if (ngbmax[f] < fj) ngbmax[f] = fj;
                    if (ngbmin[f] > fj) ngbmin[f] = fj;
                }
            }
            if (n < nr) {
                const T *pj = prim + (size_t)j * kPrimN;
                double cj = kFused ? (double)prj[kPrimC] : (double)pj[kPrimC];
                double vij[3];
                for (int ax = 0; ax < D; ++ax)
                    vij[ax] = vi[ax] - (kFused ? (double)prj[kPrimV + ax] : (double)pj[kPrimV + ax]);
                // vx = min(v_ij.x_ij / |x_ij|, 0) with x_ij = x_i - x_j = -dd:
                // only approaching pairs (v_ij.x_ij < 0) need the divide and
                // square root; for the others (including NaN) the reference's
                // value is ci + cj
                double dv = -dotT<D, double>(vij, dd);
                double vs = ci + cj;
                if constexpr (kFused) {
                    // FP32 mode: 1/|x_ij| as a refined reciprocal square
                    // root (error ~1e-14, no divide) for every pair, then
                    // the same selection, without a divergent branch
                    const double vx = dv * rsqrt_nodiv(d2);
                    vs = dv < 0. ? vs - (vx < 0. ? vx : 0.) : vs;
                } else if (dv < 0.) {
                    double vx = dv / sqrt(d2);
                    vx = vx < 0. ? vx : 0.;
                    vs = vs - vx;
                }
                vsig = vs > vsig ? vs : vsig;
            }
        };
        int n = 0;
        {
            // neighbours (real, then images) in batches: the batch's row
            // entries, then its positions (and in FP32 its records), are in
            // flight together; the batch is visited in entry order, so E and
            // the sums keep the reference order
            constexpr int BG = kGradBatch;
            for (; n + BG <= cnt; n += BG) {
                uint32_t eb[BG];
                if constexpr (BG == 4) {
                    nb_get4(a.nbr, N, t, n, eb);
                } else {
#pragma unroll
                    for (int u = 0; u < BG; ++u) eb[u] = nb_get(a.nbr, N, t, n + u);
                }
                double4 pb[BG];
                T rb[kFused ? BG : 1][kPrimN];
#pragma unroll
                for (int u = 0; u < BG; ++u) {
                    const int j = (int)nb_index(eb[u]);
                    pb[u] = a.pos[j];
                    if constexpr (kFused) load_vec<T, kPrimN>(rb[u], prim + (size_t)j * kPrimN);
                }
#pragma unroll
                for (int u = 0; u < BG; ++u) {
                    double xj[3];
                    entry_pos<D>(a, eb[u], pb[u], xj);
                    visit(n + u, (int)nb_index(eb[u]), xj, rb[kFused ? u : 0]);
                }
            }
        }
        for (; n < cnt; ++n) {
            double xj[3];
            int s[3];
            int j = nb_pos<D>(a, nb_get(a.nbr, N, t, n), xj, s);
            T prj[kPrimN];
            if (kFused) load_vec<T, kPrimN>(prj, prim + (size_t)j * kPrimN);
            visit(n, j, xj, prj);
        }
        for (int p = 1; p < D; ++p)
            for (int q = 0; q < p; ++q) E[D * p + q] = E[D * q + p];
        const bool regular = invert<D>(E);
        singular = a.debug && !regular;
        T B[D * D];
        for (int k = 0; k < D * D; ++k) B[k] = T(E[k]);
        if (kFused) {
            // grad f = B G_f, the product in FP64
            for (int f = 0; f < (kFused ? NF : 1); ++f)
                for (int p = 0; p < D; ++p) {
                    double g = 0.;
                    for (int q = 0; q < D; ++q) g += E[p * D + q] * (double)G[f][q];
                    grad[f][p] = T(g);
                }
        } else {
            // pass 2: gradients and neighbour extrema, the neighbours in
            // batches of four (entries, positions and records in flight
            // together), visited in entry order
            auto visit2 = [&](const double *xj, const T *prj) {
                T r = tsqrt(T(dist2<D>(xi, xj)));
                T w = kernel_w<T>(r, h2, sig) / om_i;
                T psi[D];
                for (int p = 0; p < D; ++p) {
                    psi[p] = T(0);
                    for (int q = 0; q < D; ++q) psi[p] += B[p * D + q] * T(xj[q] - xi[q]) * w;
                }
                for (int f = 0; f < NF; ++f) {
                    T fj = prj[field_slot<D>(f)];
                    for (int p = 0; p < D; ++p) grad[f][p] += (fj - fi[f]) * psi[p];
                    if (ngbmax[f] < fj) ngbmax[f] = fj;
                    if (ngbmin[f] > fj) ngbmin[f] = fj;
                }
            };
            int n2 = 0;
            for (; n2 + 4 <= cnt; n2 += 4) {
                uint32_t eb[4];
                nb_get4(a.nbr, N, t, n2, eb);
                double4 pb[4];
                T rb[4][kPrimN];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = (int)nb_index(eb[u]);
                    pb[u] = a.pos[j];
                    load_vec<T, kPrimN>(rb[u], prim + (size_t)j * kPrimN);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    double xj[3];
                    entry_pos<D>(a, eb[u], pb[u], xj);
                    visit2(xj, rb[u]);
                }
            }
            for (; n2 < cnt; ++n2) {
                double xj[3];
                int s[3];
                int j = nb_pos<D>(a, nb_get(a.nbr, N, t, n2), xj, s);
                T prj[kPrimN];
                load_vec<T, kPrimN>(prj, prim + (size_t)j * kPrimN);
                visit2(xj, prj);
            }
        }
        // pass 3: slope limiter
        if (a.slope_limiting) slope_limit<D, T>(a, t, cnt, xi, fi, grad, ngbmax, ngbmin);
        T *fr = (T *)a.frec + (size_t)t * FRN;
        for (int f = 0; f < NF; ++f)
            for (int ax = 0; ax < D; ++ax) fr[FRec<D>::G + f * D + ax] = grad[f][ax];
        for (int k = 0; k < D * D; ++k) fr[FRec<D>::B + k] = B[k];
        dt_i = a.cfl * a.g.h / vsig;
    }
    // block minimum of the CFL step (min is order independent)
    sdt[threadIdx.x] = dt_i;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            double o = sdt[threadIdx.x + off];
            if (o < sdt[threadIdx.x]) sdt[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) a.dt_block[blockIdx.x] = sdt[0];
    if (a.debug) {
        // E_i with a zero pivot (DEBUG_LVL >= 1): one atomic per wavefront
        const unsigned long long ns = wave_sum((unsigned long long)(singular ? 1 : 0));
        if (lane_id() == 0 && ns) atomicAdd(&a.counters[kCntSingular], ns);
    }
}

#ifndef MLH_FP32_TU
// dt = min over blocks; fixed dt or clip to the next dump time
// (MeshlessScheme.cpp:93-101: if t + dt >= t_dump then dt = t_dump - t).
__global__ __launch_bounds__(256) void k_dt_finalize(const double *__restrict__ dt_block,
                                                     int nblocks, double dt_fixed,
                                                     double t_now, double t_clip,
                                                     const double *__restrict__ dt_in,
                                                     double *__restrict__ dt,
                                                     double *__restrict__ dt_hist, int step,
                                                     int *__restrict__ status) {
    __shared__ double sh[256];
    // eight block minima per thread in flight together; the minimum does not
    // depend on the order (a NaN is skipped either way)
    constexpr int U = 8;
    double v = DBL_MAX;
    for (int k0 = threadIdx.x; k0 < nblocks; k0 += 256 * U) {
        double o[U];
#pragma unroll
        for (int u = 0; u < U; ++u) o[u] = k0 + 256 * u < nblocks ? dt_block[k0 + 256 * u] : DBL_MAX;
#pragma unroll
        for (int u = 0; u < U; ++u) v = o[u] < v ? o[u] : v;
    }
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off && sh[threadIdx.x + off] < sh[threadIdx.x])
            sh[threadIdx.x] = sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double d = sh[0];
        double clipped = 0.;
        dt[1] = d;
        if (dt_in) {
            d = dt_in[step];
        } else if (dt_fixed > 0.) {
            d = dt_fixed;
        } else if (t_clip > 0. && t_now + d >= t_clip) {
            d = t_clip - t_now;
            clipped = 1.;
        }
        dt[0] = d;
        dt[2] = clipped;
        if (dt_hist) dt_hist[step] = d;
        // the first step of the run whose neighbour search overflowed
        if ((status[0] & kStatusErrMask) && status[1] < 0) status[1] = step;
    }
}
#endif

// -------------------------------------------------------------------- faces
// A particle's face inputs held in registers: position, primitive record
// and face record, each read with 16-byte vector loads (the records are
// 16-byte aligned).  The solving side is chosen by index, so no per-field
// selects are needed.
template <int D, typename T>
struct View {
    static constexpr int NP = kPrimN, NFR = FRec<D>::N;
    double x[3];
    T p[NP];
    T f[NFR];
};
template <int D, typename T>
__device__ inline View<D, T> make_view(const HydroArgs &a, int j) {
    View<D, T> v;
    double4 q = a.pos[j];
    v.x[0] = q.x; v.x[1] = q.y; v.x[2] = q.z;
    load_vec<T, View<D, T>::NP>(v.p, (const T *)a.prim + (size_t)j * kPrimN);
    load_vec<T, View<D, T>::NFR>(v.f, (const T *)a.frec + (size_t)j * FRec<D>::N);
    return v;
}
#define PD_RHO(d) (d).p[kPrimRho]
#define PD_P(d) (d).p[kPrimP]
#define PD_OM(d) (d).p[kPrimOmega]
#define PD_V(d, ax) (d).p[kPrimV + (ax)]
#define PD_G(d, fi_, ax) (d).f[(fi_) * D + (ax)]
#define PD_B(d, p_, q_) (d).f[FRec<D>::B + (p_) * D + (q_)]

struct FaceStats {
    int neg_real, neg_ghost, vacuum, seeded;
    int nonfinite;  // particles whose new state is not finite (DEBUG_LVL >= 1)
    SolveStats st;
};

// pairwiseLimiter, Particles.cpp:1735-1785
template <typename T>
__device__ inline T pairwise_limiter(const HydroArgs &a, T phi0, T phi_i, T phi_j,
                                     T xijxi_abs, T xjxi_abs) {
    T phi_ = phi_i;
    T phi_ij = phi_i + xijxi_abs / xjxi_abs * (phi_j - phi_i);
    T phiMin, phiMax;
    if (phi_i < phi_j) { phiMin = phi_i; phiMax = phi_j; }
    else { phiMin = phi_j; phiMax = phi_i; }
    T delta1 = T(a.psi1) * lim_abs<T>(phi_i - phi_j, a.limiter_abs);
    T delta2 = T(a.psi2) * lim_abs<T>(phi_i - phi_j, a.limiter_abs);
    T phiMinus, phiPlus;
    if ((phiMax + delta1 >= T(0) && phiMax >= T(0)) || (phiMax + delta1 < T(0) && phiMax < T(0)))
        phiPlus = phiMax + delta1;
    else
        phiPlus = phiMax / (T(1) + delta1 / lim_abs<T>(phiMax, a.limiter_abs));
    if ((phiMin - delta1 >= T(0) && phiMin >= T(0)) || (phiMin - delta1 < T(0) && phiMin < T(0)))
        phiMinus = phiMin - delta1;
    else
        phiMinus = phiMin / (T(1) + delta1 / lim_abs<T>(phiMin, a.limiter_abs));
    if (phi_i < phi_j) {
        T minPhiD2 = phi_ij + delta2 < phi0 ? phi_ij + delta2 : phi0;
        phi_ = phiMinus > minPhiD2 ? phiMinus : minPhiD2;
    } else if (phi_i > phi_j) {
        T maxPhiD2 = phi_ij - delta2 > phi0 ? phi_ij - delta2 : phi0;
        phi_ = phiPlus < maxPhiD2 ? phiPlus : maxPhiD2;
    }
    return phi_;
}

// The face between L (the side that solves it) and H seen from L with
// image shift s: effective face (compEffectiveFace, :1290-1311 / :2504-2531),
// reconstructed + predicted states (compRiemannStatesLR, :1488-1733 /
// :2533-2683) and the Riemann flux (Riemann.cpp:7-229).  F = {mass,
// energy, momentum...} as seen from L.
template <int D, typename T>
struct FacePrep {
    T A[D], hA[D], WR[D + 2], WL[D + 2], vF[D];
    T ui, uj, ti0, tj0;  // normal / transverse (2D) velocities in the face frame
};
template <int D, typename T>
__device__ inline void face_frame(FacePrep<D, T> &fp);

template <int D, typename T>
__device__ inline void face_prepare(const HydroArgs &a, const View<D, T> &L,
                                    const View<D, T> &H, const int *s, bool image,
                                    T dt, FacePrep<D, T> &fp, FaceStats &fs) {
    constexpr int NF = D + 2;
    const T h2 = T(a.g.h / 2.), sig = T(a.g.sigma);
    T *A = fp.A;
    T xijxi[D], xijxj[D], xjxi[D];
    if constexpr (sizeof(T) == 4) {
        // FP32: one FP64 difference d = x_H(image) - x_L, rounded once; the
        // geometry of both sides follows from d in FP32 (x_L(image) - x_H =
        // -d and one distance |d| for both kernel weights)
        T d[D];
        for (int ax = 0; ax < D; ++ax) {
            double g = image ? image_coord(H.x[ax], s[ax], a.g.bmin[ax], a.g.bmax[ax]) : H.x[ax];
            d[ax] = T(g - L.x[ax]);
        }
        T r = tsqrt(dotT<D, T>(d, d));
        T wk = kernel_w<T>(r, h2, sig);
        T iL = T(1) / PD_OM(L), iH = T(1) / PD_OM(H);
        T w1 = wk * iL, w2 = wk * iH;
        // effective face A = psi_H(x_L)/omega_L - psi_L(x_H)/omega_H
        // = (B_L d) w1 / omega_L + (B_H d) w2 / omega_H
        for (int p = 0; p < D; ++p) {
            T p1 = T(0), p2 = T(0);
            for (int q = 0; q < D; ++q) {
                p1 += PD_B(L, p, q) * d[q];
                p2 += PD_B(H, p, q) * d[q];
            }
            A[p] = iL * (p1 * w1) + iH * (p2 * w2);
        }
        const T fl = a.foqp ? T(.5) : T(a.g.h / 4.);
        for (int ax = 0; ax < D; ++ax) {
            xjxi[ax] = d[ax];
            xijxi[ax] = fl * d[ax];
            xijxj[ax] = (fl - T(1)) * d[ax];
        }
    } else {
        double g[3], gp[3];
        for (int ax = 0; ax < D; ++ax) {
            g[ax] = image_coord(H.x[ax], s[ax], a.g.bmin[ax], a.g.bmax[ax]);
            gp[ax] = image_coord(L.x[ax], -s[ax], a.g.bmin[ax], a.g.bmax[ax]);
        }
        // effective face A = psi_H(x_L)/omega_L - psi_L(x_H)/omega_H
        T r1 = tsqrt(T(dist2<D>(L.x, g)));
        T r2 = tsqrt(T(dist2<D>(H.x, gp)));
        T w1 = kernel_w<T>(r1, h2, sig) / PD_OM(L);
        T w2 = kernel_w<T>(r2, h2, sig) / PD_OM(H);
        for (int p = 0; p < D; ++p) {
            T p1 = T(0), p2 = T(0);
            for (int q = 0; q < D; ++q) {
                p1 += PD_B(L, p, q) * T(g[q] - L.x[q]) * w1;
                p2 += PD_B(H, p, q) * T(gp[q] - H.x[q]) * w2;
            }
            A[p] = T(1) / PD_OM(L) * p1 - T(1) / PD_OM(H) * p2;
        }
        // states at the face, L perspective: WR = L, WL = H
        for (int ax = 0; ax < D; ++ax) {
            xjxi[ax] = T(g[ax] - L.x[ax]);
            if (a.foqp) {
                xijxj[ax] = T(.5 * (L.x[ax] - g[ax]));
                xijxi[ax] = T(.5 * (g[ax] - L.x[ax]));
            } else {
                double xij = L.x[ax] + a.g.h / 4. * (g[ax] - L.x[ax]);
                xijxi[ax] = T(xij - L.x[ax]);
                xijxj[ax] = T(xij - g[ax]);
            }
        }
    }
    T *vF = fp.vF;
    if (!a.move_particles) {
        for (int ax = 0; ax < D; ++ax) vF[ax] = T(0);
    } else if (a.foqp) {
        for (int ax = 0; ax < D; ++ax) vF[ax] = (PD_V(L, ax) + PD_V(H, ax)) / T(2);
    } else {
        T dp = dotT<D, T>(xijxi, xjxi), ds = dotT<D, T>(xjxi, xjxi);
        for (int ax = 0; ax < D; ++ax) vF[ax] = PD_V(L, ax) + (PD_V(H, ax) - PD_V(L, ax)) * dp / ds;
    }
    T *WR = fp.WR, *WL = fp.WL;
    WR[0] = PD_RHO(L); WL[0] = PD_RHO(H);
    WR[1] = PD_P(L);   WL[1] = PD_P(H);
    for (int ax = 0; ax < D; ++ax) {
        WR[2 + ax] = PD_V(L, ax) - vF[ax];
        WL[2 + ax] = PD_V(H, ax) - vF[ax];
    }
    T WRb[NF], WLb[NF];
    for (int k = 0; k < NF; ++k) { WRb[k] = WR[k]; WLb[k] = WL[k]; }
    // gradient field order in the record: rho, v..., P
    WR[0] += dotT<D, T>(&PD_G(L, 0, 0), xijxi);
    WL[0] += dotT<D, T>(&PD_G(H, 0, 0), xijxj);
    WR[1] += dotT<D, T>(&PD_G(L, D + 1, 0), xijxi);
    WL[1] += dotT<D, T>(&PD_G(H, D + 1, 0), xijxj);
    for (int ax = 0; ax < D; ++ax) {
        WR[2 + ax] += dotT<D, T>(&PD_G(L, 1 + ax, 0), xijxi);
        WL[2 + ax] += dotT<D, T>(&PD_G(H, 1 + ax, 0), xijxj);
    }
    if (a.pairwise_limiter && !image) {
        T ai = T(0), aj = T(0), ad = T(0);
        for (int ax = 0; ax < D; ++ax) {
            ai += xijxi[ax] * xijxi[ax];
            aj += xijxj[ax] * xijxj[ax];
            ad += xjxi[ax] * xjxi[ax];
        }
        ai = tsqrt(ai); aj = tsqrt(aj); ad = tsqrt(ad);
        for (int k = 0; k < NF; ++k) {
            WR[k] = pairwise_limiter<T>(a, WR[k], WRb[k], WLb[k], ai, ad);
            WL[k] = pairwise_limiter<T>(a, WL[k], WLb[k], WRb[k], aj, ad);
        }
    }
    // half-step predictor
    const T gam = T(a.gamma), hdt = dt / T(2);
    T viDiv = PD_G(L, 1, 0) + PD_G(L, 2, 1);
    T vjDiv = PD_G(H, 1, 0) + PD_G(H, 2, 1);
    if (D == 3) { viDiv += PD_G(L, 3, 2); vjDiv += PD_G(H, 3, 2); }
    T vxi = PD_V(L, 0) - vF[0], vyi = PD_V(L, 1) - vF[1];
    T vxj = PD_V(H, 0) - vF[0], vyj = PD_V(H, 1) - vF[1];
    const int GP = D + 1;
    WR[0] -= hdt * (PD_RHO(L) * viDiv + vxi * PD_G(L, 0, 0) + vyi * PD_G(L, 0, 1));
    WL[0] -= hdt * (PD_RHO(H) * vjDiv + vxj * PD_G(H, 0, 0) + vyj * PD_G(H, 0, 1));
    WR[1] -= hdt * (gam * PD_P(L) * viDiv + vxi * PD_G(L, GP, 0) + vyi * PD_G(L, GP, 1));
    WL[1] -= hdt * (gam * PD_P(H) * vjDiv + vxj * PD_G(H, GP, 0) + vyj * PD_G(H, GP, 1));
    WR[2] -= hdt * (PD_G(L, GP, 0) / PD_RHO(L) + vxi * PD_G(L, 1, 0) + vyi * PD_G(L, 1, 1));
    WL[2] -= hdt * (PD_G(H, GP, 0) / PD_RHO(H) + vxj * PD_G(H, 1, 0) + vyj * PD_G(H, 1, 1));
    WR[3] -= hdt * (PD_G(L, GP, 1) / PD_RHO(L) + vxi * PD_G(L, 2, 0) + vyi * PD_G(L, 2, 1));
    WL[3] -= hdt * (PD_G(H, GP, 1) / PD_RHO(H) + vxj * PD_G(H, 2, 0) + vyj * PD_G(H, 2, 1));
    if (D == 3) {
        T vzi = PD_V(L, 2) - vF[2], vzj = PD_V(H, 2) - vF[2];
        T vzq = a.quirk3d ? vzi : vzj;
        WR[0] -= hdt * vzi * PD_G(L, 0, 2);
        WL[0] -= hdt * vzj * PD_G(H, 0, 2);
        WR[1] -= hdt * vzi * PD_G(L, GP, 2);
        WL[1] -= hdt * vzj * PD_G(H, GP, 2);
        WR[2] -= hdt * vzi * PD_G(L, 1, 2);
        WL[2] -= hdt * vzq * PD_G(H, 1, 2);
        WR[3] -= hdt * vzi * PD_G(L, 2, 2);
        WL[3] -= hdt * vzq * PD_G(H, 2, 2);
        WR[4] -= hdt * (PD_G(L, GP, 2) / PD_RHO(L) + vxi * PD_G(L, 3, 0) + vyi * PD_G(L, 3, 1) +
                        vzi * PD_G(L, 3, 2));
        WL[4] -= hdt * (PD_G(H, GP, 2) / PD_RHO(H) + vxj * PD_G(H, 3, 0) + vyj * PD_G(H, 3, 1) +
                        vzj * PD_G(H, 3, 2));
    }
    if (WR[1] < T(0) || WL[1] < T(0)) {
        if (image) fs.neg_ghost++;
        else fs.neg_real++;
    }
    face_frame<D, T>(fp);
}

// Riemann (Riemann.cpp:7-141): face frame along A/|A| of fp.A, normal
// (and 2D transverse) velocities of the solver's left (WR) and right (WL)
// states in it
template <int D, typename T>
__device__ inline void face_frame(FacePrep<D, T> &fp) {
    const T *A = fp.A, *WR = fp.WR, *WL = fp.WL;
    T An = tsqrt(dotT<D, T>(A, A));
    T *hA = fp.hA;
    for (int ax = 0; ax < D; ++ax) hA[ax] = T(1) / An * A[ax];
    T ui, uj;
    T ti0 = T(0), tj0 = T(0);
    if (D == 2) {
        // rotationMatrix2D(hatA, e_x), Helper.cpp:39-45
        T L0 = hA[0] * T(1) + hA[1] * T(0);
        T L1 = -(hA[0] * T(0) - hA[1] * T(1));
        T L2 = -L1, L3 = L0;
        ui = L0 * WR[2] + L1 * WR[3];
        ti0 = L2 * WR[2] + L3 * WR[3];
        uj = L0 * WL[2] + L1 * WL[3];
        tj0 = L2 * WL[2] + L3 * WL[3];
    } else {
        ui = dotT<D, T>(hA, &WR[2]);
        uj = dotT<D, T>(hA, &WL[2]);
    }
    fp.ui = ui; fp.uj = uj; fp.ti0 = ti0; fp.tj0 = tj0;
}

// FP32 faces: the first step of the mixed solver (an FP64 Newton step at
// the standard guess) on the face's 1D problem; valid = false for a face
// without area
template <int D>
__device__ inline MixedProbe face_probe(const HydroArgs &a, const FacePrep<D, float> &fp,
                                        SolveStats &st) {
    MixedProbe m{};
    if (!(dotT<D, float>(fp.A, fp.A) > 0.f)) return m;
    GasConst<float> gf(a.gamma);
    GasConst<double> gc(a.gamma);
    return mixed_probe(gf, gc, fp.WR[0], fp.ui, fp.WR[1], fp.WL[0], fp.uj, fp.WL[1],
                       a.riemann_tol, -1.f, st);
}

// exact Riemann solve from the FacePrep states and the flux projected on A
// (Riemann.cpp:83-229).  seed > 0: Newton starts there (learned seed).  FP32
// with a probe: the solve continues from face_probe, and seed > 0 restarts
// the FP32 iterations there.
template <int D, typename T>
__device__ inline int face_finish(const HydroArgs &a, const FacePrep<D, T> &fp, float seed,
                                  T *F, FaceStats &fs, const MixedProbe *probe = nullptr) {
    const T *A = fp.A, *hA = fp.hA, *WR = fp.WR, *WL = fp.WL, *vF = fp.vF;
    const T ui = fp.ui, uj = fp.uj, ti0 = fp.ti0, tj0 = fp.tj0;
    const T gam = T(a.gamma);
    // A = 0 (a neighbour at r = h exactly: W = 0 on both sides) has no
    // normal; the face carries no flux
    if (!(dotT<D, T>(A, A) > T(0))) {
        for (int k = 0; k < D + 2; ++k) F[k] = T(0);
        return 2;  // no face
    }
    T rs, us, Ps;
    int flag;
    if constexpr (sizeof(T) == 8) {
        GasConst<double> gc(a.gamma);
        RSol<double> r = solve_fp64(gc, WR[0], ui, WR[1], WL[0], uj, WL[1], a.riemann_tol,
                                    a.riemann_max_iter, (double)seed, fs.st);
        rs = r.d; us = r.u; Ps = r.p; flag = r.flag;
    } else {
        GasConst<float> gf(a.gamma);
        GasConst<double> gc(a.gamma);
        RSol<float> r =
            probe ? mixed_finish(gf, gc, WR[0], ui, WR[1], WL[0], uj, WL[1], a.riemann_tol,
                                 a.riemann_max_iter, a.learned_iters, seed, *probe, fs.st)
                  : solve_mixed(gf, gc, WR[0], ui, WR[1], WL[0], uj, WL[1], a.riemann_tol,
                                a.riemann_max_iter, a.learned_iters, seed, fs.st);
        rs = r.d; us = r.u; Ps = r.p; flag = r.flag;
    }
    if (flag == 0) fs.vacuum++;
    T vS[D];
    if (D == 2) {
        T t0 = flag == 1 ? tj0 : (flag == -1 ? ti0 : T(0));
        T K0 = T(1) * hA[0] + T(0) * hA[1];
        T K1 = -(T(1) * hA[1] - T(0) * hA[0]);
        T K2 = -K1, K3 = K0;
        vS[0] = K0 * us + K1 * t0;
        vS[1] = K2 * us + K3 * t0;
    } else {
        T un = flag == 1 ? uj : ui;
        for (int ax = 0; ax < D; ++ax) {
            T vup = flag == 1 ? WL[2 + ax] : WR[2 + ax];
            vS[ax] = flag == 0 ? us * hA[ax] : us * hA[ax] + (vup - un * hA[ax]);
        }
    }
    // fluxes projected on A (Riemann.cpp:164-181, 3D :204-227)
    T vLab[D];
    for (int ax = 0; ax < D; ++ax) vLab[ax] = vS[ax] + vF[ax];
    if constexpr (sizeof(T) == 4) {
        // FP32: the same fluxes with the sums over A factored,
        //   F_m = rho* (A.v*), F_p = vLab F_m + P* A,
        //   F_E = (A.v*)(P*/(gamma-1) + rho* |vLab|^2 / 2) + P* (A.vLab),
        // where MFM drops the advective A.v* of momentum and energy
        const T Av = dotT<D, T>(A, vS);
        F[0] = rs * Av;
        const T Avm = a.mfm ? T(0) : Av;
        const T mf = rs * Avm;
        for (int c = 0; c < D; ++c) F[2 + c] = vLab[c] * mf + Ps * A[c];
        const T ekin = dotT<D, T>(vLab, vLab);
        F[1] = Avm * (Ps / (gam - T(1)) + rs * T(.5) * ekin) + Ps * dotT<D, T>(A, vLab);
        return flag;
    }
    T f0 = T(0);
    for (int ax = 0; ax < D; ++ax) f0 += A[ax] * rs * vS[ax];
    F[0] = f0;
    if (a.mfm)
        for (int ax = 0; ax < D; ++ax) vS[ax] = T(0);
    for (int c = 0; c < D; ++c) {
        T sm = T(0);
        for (int b = 0; b < D; ++b)
            sm += b == c ? A[b] * (rs * vLab[c] * vS[b] + Ps) : A[b] * rs * vLab[c] * vS[b];
        F[2 + c] = sm;
    }
    T ekin = dotT<D, T>(vLab, vLab);
    T e = T(0);
    for (int b = 0; b < D; ++b)
        e += A[b] * (vS[b] * (Ps / (gam - T(1)) + rs * T(.5) * ekin) + Ps * vLab[b]);
    F[1] = e;
    return flag;
}

// Riemann::Riemann + Riemann::exact (Riemann.cpp:7-229) on given faces, FP64:
// wl / wr = {rho, P, v[D]} of the solver's left (particle i) and right
// (particle j) states in the frame moving with vf; F = {mass, energy,
// momentum[D]}; flag = sampled side (-1 left, +1 right, 0 vacuum, 2 A = 0).
template <int D>
__global__ __launch_bounds__(256) void k_riemann_flux(HydroArgs a, long long n,
                                                      const double *__restrict__ wl,
                                                      const double *__restrict__ wr,
                                                      const double *__restrict__ vf,
                                                      const double *__restrict__ aij,
                                                      double *__restrict__ f,
                                                      int *__restrict__ flag) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    FacePrep<D, double> fp;
    for (int c = 0; c < D + 2; ++c) {
        fp.WR[c] = wl[k * (D + 2) + c];
        fp.WL[c] = wr[k * (D + 2) + c];
    }
    for (int ax = 0; ax < D; ++ax) {
        fp.A[ax] = aij[k * D + ax];
        fp.vF[ax] = vf[k * D + ax];
    }
    double An = sqrt(dotT<D, double>(fp.A, fp.A));
    for (int ax = 0; ax < D; ++ax) fp.hA[ax] = 1. / An * fp.A[ax];
    fp.ti0 = fp.tj0 = 0.;
    if (D == 2) {
        double L0 = fp.hA[0] * 1. + fp.hA[1] * 0.;
        double L1 = -(fp.hA[0] * 0. - fp.hA[1] * 1.);
        double L2 = -L1, L3 = L0;
        fp.ui = L0 * fp.WR[2] + L1 * fp.WR[3];
        fp.ti0 = L2 * fp.WR[2] + L3 * fp.WR[3];
        fp.uj = L0 * fp.WL[2] + L1 * fp.WL[3];
        fp.tj0 = L2 * fp.WL[2] + L3 * fp.WL[3];
    } else {
        fp.ui = dotT<D, double>(fp.hA, &fp.WR[2]);
        fp.uj = dotT<D, double>(fp.hA, &fp.WL[2]);
    }
    FaceStats fs{};
    double F[D + 2];
    int fl = face_finish<D, double>(a, fp, -1.f, F, fs);
    for (int c = 0; c < D + 2; ++c) f[k * (D + 2) + c] = F[c];
    flag[k] = fl;
}


// The face of neighbour entry n of work position t (original index i), seen
// from the side that solves it: the lower original index (images: parent
// >= i).  FP32: also the first step of the mixed solver.
template <int D, typename T>
__device__ inline void face_of_entry(const HydroArgs &a, int t, int i, uint32_t e, int jo, T dt,
                                     FacePrep<D, T> &fp, MixedProbe &mp, bool &self_lo,
                                     FaceStats &fs, const View<D, T> *own = nullptr) {
    int j = (int)nb_index(e);
    uint32_t code = nb_code(e);
    bool image = code != kCodeReal;
    int s[3] = {0, 0, 0};
    if (image)
        for (int ax = 0; ax < D; ++ax) s[ax] = code_shift(code, ax);
    self_lo = !a.enforce_flux_sym || (image ? jo >= i : jo > i);
    View<D, T> L, H;
    if (own) {
        // the particle's own view is held in registers for its whole row:
        // only the neighbour's is gathered, and the sides are selected
        const View<D, T> nb = make_view<D, T>(a, j);
        L = self_lo ? *own : nb;
        H = self_lo ? nb : *own;
    } else {
        L = make_view<D, T>(a, self_lo ? t : j);
        H = make_view<D, T>(a, self_lo ? j : t);
    }
    int sl[3];
    for (int ax = 0; ax < 3; ++ax) sl[ax] = self_lo ? s[ax] : -s[ax];
    face_prepare<D, T>(a, L, H, sl, image, dt, fp, fs);
    if constexpr (sizeof(T) == 4) mp = face_probe<D>(a, fp, fs.st);
}

// the face of row entry n of work position t: its entry and the neighbour's
// original index (which decides the solving side) read here
template <int D, typename T>
__device__ inline void face_of_entry(const HydroArgs &a, int t, int i, int n, T dt,
                                     FacePrep<D, T> &fp, MixedProbe &mp, bool &self_lo,
                                     FaceStats &fs) {
    const uint32_t e = nb_get(a.nbr, a.N, t, n);
    face_of_entry<D, T>(a, t, i, e, (int)a.pos[nb_index(e)].w, dt, fp, mp, self_lo, fs);
}

// Network seed for this lane's face when want; every lane of the wavefront
// calls it together (BF16 MFMA over the wavefront's 64 faces).
template <int D, typename T>
__device__ inline float face_net_seed(const HydroArgs &a, bool want, const FacePrep<D, T> &fp,
                                      uint16_t *lds, FaceStats &fs) {
    float f[kMlpIn];
    float pref = 1.f;
    bool ok = want && fp.WR[0] > T(0) && fp.WL[0] > T(0) && fp.WR[1] > T(0) && fp.WL[1] > T(0);
    if (ok)
        pref = mlp_features_fast((float)a.gamma, (float)fp.WR[0], (float)fp.ui, (float)fp.WR[1],
                                 (float)fp.WL[0], (float)fp.uj, (float)fp.WL[1], f);
    else
        for (int q = 0; q < kMlpIn; ++q) f[q] = 0.f;
    float y = mlp_infer_wave_bf16_lean(a.mlp_frags, f, lds);
    fs.seeded += ok ? 1 : 0;
    return ok ? mlp_seed_from_output(y, f, pref) : -1.f;
}

struct FluxSums {
    double m, e, v[3];
};

// training data: the face states of the faces the solver needed Newton
// iterations for (left / right rho, normal velocity, P), for the lanes with
// `want`.  Called by the active lanes of a wavefront together: one atomic
// per wavefront for the slots, nothing once its stripe is full.  The
// wavefront writes to stripe (wavefront index) % stripes, each with its own
// counter: with one counter, the atomics of all wavefronts serialise on one
// address while the record fills.
template <int D, typename T>
__device__ inline void collect_faces(const HydroArgs &a, bool want, const FacePrep<D, T> &fp) {
    const unsigned long long b = __ballot(want);
    if (!b) return;
    const int s = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) & a.collect_smask;
    unsigned long long *cnt = a.collect_n + s;
    const unsigned long long cap = (unsigned long long)a.collect_scap;
    if (__atomic_load_n(cnt, __ATOMIC_RELAXED) >= cap) return;
    const int lane = lane_id();
    const int leader = __ffsll((long long)b) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(cnt, (unsigned long long)__popcll(b));
    base = __shfl(base, leader, 64);
    if (want) {
        const unsigned long long slot = base + __popcll(b & ((1ull << lane) - 1ull));
        if (slot < cap) {
            double *c = a.collect + ((unsigned long long)s * cap + slot) * 6;
            c[0] = fp.WR[0]; c[1] = fp.ui; c[2] = fp.WR[1];
            c[3] = fp.WL[0]; c[4] = fp.uj; c[5] = fp.WL[1];
        }
    }
}

// solve the face, add its flux (with the sign of this side) to the sums;
// face collection for the training data
template <int D, typename T>
__device__ inline void face_accumulate(const HydroArgs &a, const FacePrep<D, T> &fp,
                                       const MixedProbe &mp, float seed, bool self_lo, int it0,
                                       bool collect, FluxSums &acc, FaceStats &fs) {
    T F[D + 2];
    face_finish<D, T>(a, fp, seed, F, fs, sizeof(T) == 4 ? &mp : nullptr);
    if (a.collect) collect_faces<D, T>(a, collect && self_lo && fs.st.iters - it0 >= 2, fp);
    double sgn = self_lo ? 1. : -1.;
    acc.m += sgn * (double)F[0];
    for (int ax = 0; ax < D; ++ax) acc.v[ax] += sgn * (double)F[2 + ax];
    acc.e += sgn * (double)F[1];
}



// the state of particle i carried over unchanged (a step whose neighbour
// search overflowed)
template <int D>
__device__ inline void particle_keep(const HydroArgs &a, int i) {
    a.mn[i] = a.m[i];
    a.un[i] = a.u[i];
    for (int ax = 0; ax < D; ++ax) {
        a.vn[ax][i] = a.v[ax][i];
        a.xn[ax][i] = a.x[ax][i];
    }
}

// Flux sums and updateStateAndPosition (Particles.cpp:2013-2110) of the
// particle at work position t (original index i).
// Returns 1 when the debug counters are on (DEBUG_LVL >= 1) and the new
// state (m, u, v, x) is not finite, else 0.
template <int D>
__device__ inline int particle_finish(const HydroArgs &a, int i, const FluxSums acc) {
    const double dtd = a.dt[0];
    const double mF = acc.m, eF = acc.e;
    const double vFs[3] = {acc.v[0], acc.v[1], acc.v[2]};
    if (a.write_fluxes) {
        a.mF[i] = mF;
        a.eF[i] = eF;
        for (int ax = 0; ax < D; ++ax) a.vF[(size_t)i * D + ax] = vFs[ax];
    }
    double vo[3] = {a.v[0][i], a.v[1][i], D == 3 ? a.v[2][i] : 0.};
    double m0 = a.m[i], u0 = a.u[i];
    double ke = vo[0] * vo[0] + vo[1] * vo[1];
    if (D == 3) ke += vo[2] * vo[2];
    double Q0 = m0 * (u0 + .5 * ke);
    double Q[3];
    for (int ax = 0; ax < D; ++ax) Q[ax] = m0 * vo[ax];
    double m1 = a.mfm ? m0 : m0 - dtd * mF;
    double vn[3];
    for (int ax = 0; ax < D; ++ax) {
        Q[ax] -= dtd * vFs[ax];
        vn[ax] = Q[ax] / m1;
    }
    Q0 -= dtd * eF;
    double kn = vn[0] * vn[0] + vn[1] * vn[1];
    if (D == 3) kn += vn[2] * vn[2];
    double u1 = Q0 / m1 - .5 * kn;
    a.mn[i] = m1;
    a.un[i] = u1;
    bool fin = isfinite(m1) && isfinite(u1);
    for (int ax = 0; ax < D; ++ax) {
        a.vn[ax][i] = vn[ax];
        double x = a.x[ax][i];
        if (a.move_particles) {
            x += vo[ax] * dtd;
            if (a.g.periodic) {
                double mn = a.g.bmin[ax], mx = a.g.bmax[ax];
                if (x < mn) x = mx - (mn - x);
                else if (mx <= x) x = mn + (x - mx);
            }
        }
        a.xn[ax][i] = x;
        fin &= isfinite(vn[ax]) && isfinite(x);
    }
    return a.debug && !fin ? 1 : 0;
}

// diagnostics: one atomic per wavefront and counter
__device__ inline void add_counters(const HydroArgs &a, const FaceStats fs, int faces) {
    unsigned long long c[kNumCounters] = {
        (unsigned long long)fs.neg_real, (unsigned long long)fs.neg_ghost,
        (unsigned long long)fs.vacuum, (unsigned long long)fs.st.nonconv, 0ull,
        (unsigned long long)fs.st.fallback, (unsigned long long)fs.st.iters,
        (unsigned long long)faces, (unsigned long long)fs.seeded, 0ull,
        (unsigned long long)fs.nonfinite, (unsigned long long)fs.st.acoustic};
    for (int k = 0; k < kNumCounters; ++k) {
        unsigned long long v = wave_sum(c[k]);
        if (lane_id() == 0 && v) atomicAdd(&a.counters[k], v);
    }
}

// Waves per SIMD the face kernel is compiled for.  Two waves cap a lane at
// 256 registers: the FP32 learned kernel then spills a few registers in the
// network's branch but keeps the occupancy of the standard FP32 kernel.
#ifndef MLH_FACES_WAVES
#define MLH_FACES_WAVES(D, T, LEARNED) ((LEARNED && sizeof(T) == 4) ? 2 : 1)
#endif

// Faces of particle i.  LEARNED: the loop runs to the wavefront's largest
// neighbour count so that every lane takes part in the BF16 MFMA inference
// of the p* seed for the wavefront's 64 faces of that iteration.
template <int D, typename T, bool LEARNED>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MLH_FACES_WAVES(D, T, LEARNED)))) void k_faces(HydroArgs a) {
    constexpr int NF = D + 2;
    constexpr bool kFp32 = sizeof(T) == 4;
    // FP32 learned: faces whose probe fails are deferred to the end of the
    // row and seeded there by the network; the FP64 learned kernel runs the
    // network inside the face loop, over the wavefront's longest row
    constexpr bool kInLoopNet = LEARNED && !kFp32;
    constexpr bool kDefer = LEARNED && kFp32;
    __shared__ uint16_t lds[LEARNED ? 4 : 1][LEARNED ? 64 * 16 : 1];
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    const bool live = t < a.N;
    FaceStats fs{};
    const int N = a.N;
    const int i = live ? (a.sidx ? (int)a.sidx[t] : t) : 0;
    if (a.status[0] & kStatusErrMask) {
        // this step's neighbour search overflowed (the reference exits in
        // gridNNS / ghostNNS): the state is carried over unchanged
        if (live) particle_keep<D>(a, i);
        return;
    }
    const double dtd = a.dt[0];
    const T dt = T(dtd);
    int cnt = live ? a.noi[t] + a.nog[t] : 0;
    int lim = cnt;
    if (kInLoopNet) {
        for (int off = 32; off > 0; off >>= 1) {
            int o = __shfl_xor(lim, off, 64);
            lim = o > lim ? o : lim;
        }
    }
    const bool collect = a.collect && live && (i % a.collect_stride) == 0;
    // the particle's own view, read once for its row, where the registers
    // it takes cost no occupancy (the 2D FP64 kernel would drop from 2 to
    // 1 wavefront per SIMD and gathers both views per face instead)
    constexpr bool kOwnView = kFp32 || D == 3;
    const View<D, T> own = make_view<D, T>(a, live && kOwnView ? t : 0);
    FluxSums acc{0., 0., {0., 0., 0.}};
    uint64_t defer[2] = {0ull, 0ull};
    bool defer_far = false;
    // software pipeline of the face's first two loads: row entries two
    // faces ahead, the neighbour's original index one face ahead, so a face
    // waits only on its records
    uint32_t e0 = 0, e1 = 0;
    int jo0 = 0;
    if (cnt > 0) e0 = nb_get(a.nbr, N, t, 0);
    if (cnt > 1) e1 = nb_get(a.nbr, N, t, 1);
    if (cnt > 0) jo0 = (int)a.pos[nb_index(e0)].w;
    for (int n = 0; n < lim; ++n) {
        const bool act = n < cnt;
        const uint32_t e = e0;
        const int jo = jo0;
        if (n + 1 < cnt) jo0 = (int)a.pos[nb_index(e1)].w;
        e0 = e1;
        if (n + 2 < cnt) e1 = nb_get(a.nbr, N, t, n + 2);
        FacePrep<D, T> fp;
        MixedProbe mp{};
        bool self_lo = true;
        const int it0 = fs.st.iters;
        if (act) face_of_entry<D, T>(a, t, i, e, jo, dt, fp, mp, self_lo, fs, kOwnView ? &own : nullptr);
        float seed = -1.f;
        if constexpr (kInLoopNet) {
            // FP64: the network runs for faces with a real jump
            // (near-uniform faces keep the standard guess, which is converged
            // to O(jump^2) there): |dp| > k (pL + pR) or
            // du^2 rhoR rhoL > 2 k^2 gamma (pR rhoL + pL rhoR)
            bool strong = false;
            if (act) {
                T dp = fp.WR[1] - fp.WL[1], du = fp.uj - fp.ui;
                T psum = fp.WR[1] + fp.WL[1];
                T rr = fp.WR[0] * fp.WL[0];
                T c2r = T(a.gamma) * (fp.WR[1] * fp.WL[0] + fp.WL[1] * fp.WR[0]);
                strong = dp * dp > T(kSeedJump * kSeedJump) * psum * psum ||
                         du * du * rr > T(2 * kSeedJump * kSeedJump) * c2r;
            }
            if (__ballot(strong)) seed = face_net_seed<D, T>(a, strong, fp, lds[threadIdx.x >> 6], fs);
        }
        if (act) {
            if (kDefer && mp.valid && !mp.accepted && mp.chg > kSeedGate) {
                // the standard guess was far off: Newton from the network's
                // seed, after the row.  Entries 0..127 are marked in
                // registers; for later ones (rows longer than 128) a flag
                // has the row's tail searched again after the row (no
                // stores in the face loop).  Here the face goes on with
                // A = 0, for which face_finish returns zero flux at once:
                // the loop keeps one path for every face
                if (n < 128) defer[n >> 6] |= 1ull << (n & 63);
                else defer_far = true;
                for (int ax = 0; ax < D; ++ax) fp.A[ax] = T(0);
            }
            face_accumulate<D, T>(a, fp, mp, seed, self_lo, it0, collect, acc, fs);
        }
    }
    if constexpr (kDefer) {
        // the deferred faces, each lane's in ascending entry order after the
        // sums of its other faces, one face per lane and round: every lane
        // takes part in the round's BF16 MFMA inference of the seeds.
        // Entries 0..127 come from the masks; entries from 128 on (rows
        // longer than 128) are found again by the same probe.  Whether a
        // face is deferred depends on the face alone (the probe runs on the
        // solving side's operands), so both sides of a face defer it and
        // solve it identically, whatever its entry number.
        const bool dfr = live && ((defer[0] | defer[1]) != 0ull || defer_far);
        if (__ballot(dfr)) {
            int wd = 0, far_n = 128;
            unsigned long long m = dfr ? defer[0] : 0ull;
            for (;;) {
                FacePrep<D, float> fp;
                MixedProbe mp{};
                bool self_lo = true, act = false;
                if (dfr) {
                    // the probes were counted in the row loop
                    const FaceStats keep = fs;
                    for (;;) {
                        int n;
                        bool scan = false;
                        if (m != 0ull) {
                            n = wd * 64 + __builtin_ctzll(m);
                            m &= m - 1ull;
                        } else if (wd == 0) {
                            wd = 1;
                            m = defer[1];
                            continue;
                        } else if (defer_far && far_n < cnt) {
                            n = far_n++;
                            scan = true;
                        } else {
                            break;
                        }
                        face_of_entry<D, float>(a, t, i, n, dt, fp, mp, self_lo, fs);
                        if (!scan || (mp.valid && !mp.accepted && mp.chg > kSeedGate)) {
                            act = true;
                            break;
                        }
                    }
                    fs = keep;
                }
                if (!__ballot(act)) break;
                const float seed = face_net_seed<D, float>(a, act && mp.valid && !mp.accepted, fp,
                                                           lds[threadIdx.x >> 6], fs);
                if (act)
                    face_accumulate<D, float>(a, fp, mp, seed, self_lo, fs.st.iters, collect, acc,
                                              fs);
            }
        }
    }
    if (live) fs.nonfinite += particle_finish<D>(a, i, acc);
    add_counters(a, fs, cnt);
}


// conserved sums (Particles.cpp:2830-2886): V, M, E, px, py, pz per block
template <int D, typename T>
__global__ __launch_bounds__(256) void k_sums(HydroArgs a, double *out) {
    __shared__ double sh[6][256];
    int t = xcd_block() * blockDim.x + threadIdx.x;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (t < a.N) {
        const int i = a.sidx ? (int)a.sidx[t] : t;
        const T *pr = (const T *)a.prim + (size_t)t * kPrimN;
        double om = (double)pr[kPrimOmega];
        double m = a.m[i];
        double vx = a.v[0][i], vy = a.v[1][i], vz = D == 3 ? a.v[2][i] : 0.;
        double ke = vx * vx + vy * vy;
        if (D == 3) ke += vz * vz;
        v[0] = om > 0. ? 1. / om : 0.;
        v[1] = m;
        v[2] = m * (a.u[i] + .5 * ke);
        v[3] = m * vx;
        v[4] = m * vy;
        v[5] = m * vz;
    }
    for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) out[(size_t)blockIdx.x * 6 + k] = sh[k][0];
}

}  // namespace mlh
