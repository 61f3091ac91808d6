#pragma once
// stages_impl.h -- the reference's per-method stages as kernels of their
// own, for callers that drive a step one Particles method at a time:
//   gradient(f, grad)      any field's gradient from the psi-tilde of the
//                          gradient stage (Particles.cpp:1257-1270,
//                          2457-2502);
//   slopeLimiter           on the unlimited gradients (:1313-1444);
//   compEffectiveFace      A_ij of every row entry (:1290-1311, 2504-2531);
//   compRiemannStatesLR    W_R, W_L, vFrame of every row entry (:1488-1733,
//                          2533-2683);
//   solveRiemannProblems   F_ij of every row entry, solved from the lower
//                          index and negated for the other side through a
//                          linear search of the partner's row (:1787-1911);
//   collectFluxes          the flux sums in entry order (:1913-2011), and the
//                          update of the fused step prepared from them;
//   checkFluxSymmetry      F_ij + F_ji against FLUX_SYM_TOL per entry
//                          (:2888-2976);
// and the mirror check of the neighbour rows the fused step runs at
// DEBUG_LVL >= 1 (every entry's partner must list it: the fused face kernel
// relies on it for F_ji = -F_ij).
// The per-entry arithmetic is the fused kernels' own device functions
// (face_prepare, face_frame, face_finish, slope_limit, particle_finish), so
// both routes give the same bits.  Per-entry arrays are FP64, component-
// major: value c of row entry n of work position t at (c * rowlen + n) * N + t.
// Included by hydro.hip (T = double) and hydro_fp32.hip (T = float).

namespace mlh {

__device__ inline size_t fa_at(const FaceArrays &f, int c, int n, int N, int t) {
    return ((size_t)c * f.rowlen + n) * (size_t)N + t;
}

// original index of the particle at work position q
__device__ inline int orig_index(const HydroArgs &a, int q) {
    return a.sidx ? (int)a.sidx[q] : q;
}

// The entry of work position j's row that points back at work position t:
// the real entry t (rows hold each real neighbour once), or for an image
// entry the first image entry of j whose particle is t (the reference's
// search by parent, Particles.cpp:1886-1897).  -1 when j does not list t.
__device__ inline int mirror_entry(const HydroArgs &a, int t, int j, bool image) {
    const int N = a.N, nr = a.noi[j], ng = a.nog[j];
    const int lo = image ? nr : 0, hi = image ? nr + ng : nr;
    for (int m = lo; m < hi; ++m)
        if ((int)nb_index(nb_get(a.nbr, N, j, m)) == t) return m;
    return -1;
}

// slopeLimiter alone: the neighbour extrema of the fields, then the limiter
// on the gradients held in the face records
template <int D, typename T>
__global__ __launch_bounds__(256) void k_limit(HydroArgs a) {
    constexpr int NF = D + 2, FRN = FRec<D>::N;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.N) return;
    const int N = a.N;
    const T *prim = (const T *)a.prim;
    const T *pri = prim + (size_t)t * kPrimN;
    const double4 pi = a.pos[t];
    const double xi[3] = {pi.x, pi.y, pi.z};
    T fi[NF], ngbmax[NF], ngbmin[NF], grad[NF][D];
    T *fr = (T *)a.frec + (size_t)t * FRN;
    for (int f = 0; f < NF; ++f) {
        fi[f] = pri[field_slot<D>(f)];
        ngbmax[f] = tmin_pos<T>();
        ngbmin[f] = tmax_val<T>();
        for (int ax = 0; ax < D; ++ax) grad[f][ax] = fr[FRec<D>::G + f * D + ax];
    }
    const int cnt = a.noi[t] + a.nog[t];
    for (int n = 0; n < cnt; ++n) {
        const int j = (int)nb_index(nb_get(a.nbr, N, t, n));
        const T *prj = prim + (size_t)j * kPrimN;
        for (int f = 0; f < NF; ++f) {
            T fj = prj[field_slot<D>(f)];
            if (ngbmax[f] < fj) ngbmax[f] = fj;
            if (ngbmin[f] > fj) ngbmin[f] = fj;
        }
    }
    slope_limit<D, T>(a, t, cnt, xi, fi, grad, ngbmax, ngbmin);
    for (int f = 0; f < NF; ++f)
        for (int ax = 0; ax < D; ++ax) fr[FRec<D>::G + f * D + ax] = grad[f][ax];
}

// gradient(f, grad): grad_i = sum_j (f_j - f_i) psi_j(x_i) over the real
// neighbours, then the images with the parent's value (updateGhostState,
// Particles.cpp:2193-2206); psi_j(x_i) = B_i (x_j - x_i) W(r_ij) / omega_i as
// in the gradient stage.  f and grad ([N][D]) by original index.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_gradient_field(HydroArgs a, const double *__restrict__ f,
                                                        double *__restrict__ grad) {
    constexpr int FRN = FRec<D>::N;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.N) return;
    const int N = a.N, i = orig_index(a, t);
    const double4 pi = a.pos[t];
    const double xi[3] = {pi.x, pi.y, pi.z};
    const T om_i = ((const T *)a.prim)[(size_t)t * kPrimN + kPrimOmega];
    const T *B = (const T *)a.frec + (size_t)t * FRN + FRec<D>::B;
    const T h2 = T(a.g.h / 2.), sig = T(a.g.sigma);
    const double fi = f[i];
    double g[D];
    for (int p = 0; p < D; ++p) g[p] = 0.;
    const int cnt = a.noi[t] + a.nog[t];
    for (int n = 0; n < cnt; ++n) {
        double xj[3];
        int s[3];
        const int j = nb_pos<D>(a, nb_get(a.nbr, N, t, n), xj, s);
        T r = tsqrt(T(dist2<D>(xi, xj)));
        T w = kernel_w<T>(r, h2, sig) / om_i;
        const double df = f[orig_index(a, j)] - fi;
        for (int p = 0; p < D; ++p) {
            T psi = T(0);
            for (int q = 0; q < D; ++q) psi += B[p * D + q] * T(xj[q] - xi[q]) * w;
            g[p] += df * (double)psi;
        }
    }
    for (int p = 0; p < D; ++p) grad[(size_t)i * D + p] = g[p];
}

// compEffectiveFace (what & 1) and compRiemannStatesLR (what & 2) of every
// row entry, seen from the row's particle: A_ij, W_R = i-state, W_L =
// j-state, vFrame
template <int D, typename T>
__global__ __launch_bounds__(256) void k_face_states(HydroArgs a, FaceArrays fa, int what) {
    constexpr int NF = D + 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.N) return;
    const int N = a.N;
    const T dt = T(a.dt[0]);
    const View<D, T> L = make_view<D, T>(a, t);
    const int cnt = a.noi[t] + a.nog[t];
    for (int n = 0; n < cnt; ++n) {
        const uint32_t e = nb_get(a.nbr, N, t, n);
        const int j = (int)nb_index(e);
        const uint32_t code = nb_code(e);
        const bool image = code != kCodeReal;
        int s[3] = {0, 0, 0};
        if (image)
            for (int ax = 0; ax < D; ++ax) s[ax] = code_shift(code, ax);
        const View<D, T> H = make_view<D, T>(a, j);
        FacePrep<D, T> fp;
        FaceStats fs{};
        face_prepare<D, T>(a, L, H, s, image, dt, fp, fs);
        if (what & 1)
            for (int ax = 0; ax < D; ++ax) fa.A[fa_at(fa, ax, n, N, t)] = (double)fp.A[ax];
        if (what & 2) {
            for (int c = 0; c < NF; ++c) {
                fa.WR[fa_at(fa, c, n, N, t)] = (double)fp.WR[c];
                fa.WL[fa_at(fa, c, n, N, t)] = (double)fp.WL[c];
            }
            for (int ax = 0; ax < D; ++ax) fa.vF[fa_at(fa, ax, n, N, t)] = (double)fp.vF[ax];
        }
    }
}

// solveRiemannProblems.  pass 0: the entries the row's particle solves (the
// lower original index; images: parent >= i; every entry without
// ENFORCE_FLUX_SYM) from the stored A, W_R, W_L, vFrame, always from the
// standard guess; also the negative face pressure count of every entry.
// pass 1: the other entries, F_ij = -F_ji from the partner's mirror entry
// (0 and counted missing when the partner does not list the particle).
template <int D, typename T>
__global__ __launch_bounds__(256) void k_face_solve(HydroArgs a, FaceArrays fa, int pass) {
    constexpr int NF = D + 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    FaceStats fs{};
    int missing = 0, faces = 0;
    if (t < a.N) {
        const int N = a.N, i = orig_index(a, t);
        const int cnt = a.noi[t] + a.nog[t];
        for (int n = 0; n < cnt; ++n) {
            const uint32_t e = nb_get(a.nbr, N, t, n);
            const int j = (int)nb_index(e);
            const bool image = nb_code(e) != kCodeReal;
            const int jo = orig_index(a, j);
            const bool self_lo = !a.enforce_flux_sym || (image ? jo >= i : jo > i);
            if (pass == 0) {
                FacePrep<D, T> fp;
                for (int ax = 0; ax < D; ++ax) {
                    fp.A[ax] = T(fa.A[fa_at(fa, ax, n, N, t)]);
                    fp.vF[ax] = T(fa.vF[fa_at(fa, ax, n, N, t)]);
                }
                for (int c = 0; c < NF; ++c) {
                    fp.WR[c] = T(fa.WR[fa_at(fa, c, n, N, t)]);
                    fp.WL[c] = T(fa.WL[fa_at(fa, c, n, N, t)]);
                }
                if (fp.WR[1] < T(0) || fp.WL[1] < T(0)) {
                    if (image) fs.neg_ghost++;
                    else fs.neg_real++;
                }
                if (!self_lo) continue;
                face_frame<D, T>(fp);
                T F[NF];
                face_finish<D, T>(a, fp, -1.f, F, fs);
                ++faces;
                for (int c = 0; c < NF; ++c) fa.F[fa_at(fa, c, n, N, t)] = (double)F[c];
            } else if (!self_lo) {
                const int m = mirror_entry(a, t, j, image);
                if (m < 0) ++missing;
                for (int c = 0; c < NF; ++c)
                    fa.F[fa_at(fa, c, n, N, t)] = m < 0 ? 0. : -fa.F[fa_at(fa, c, m, N, j)];
            }
        }
    }
    add_counters(a, fs, faces);
    unsigned long long ms = wave_sum((unsigned long long)missing);
    if (lane_id() == 0 && ms) atomicAdd(&a.counters[kCntMissing], ms);
}

// collectFluxes: the stored F_ij summed in entry order (real neighbours,
// then images) in FP64, then the update of the fused step
template <int D, typename T>
__global__ __launch_bounds__(256) void k_collect(HydroArgs a, FaceArrays fa) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    FaceStats fs{};
    if (t < a.N) {
        const int N = a.N;
        FluxSums acc{0., 0., {0., 0., 0.}};
        const int cnt = a.noi[t] + a.nog[t];
        for (int n = 0; n < cnt; ++n) {
            acc.m += fa.F[fa_at(fa, 0, n, N, t)];
            for (int ax = 0; ax < D; ++ax) acc.v[ax] += fa.F[fa_at(fa, 2 + ax, n, N, t)];
            acc.e += fa.F[fa_at(fa, 1, n, N, t)];
        }
        fs.nonfinite = particle_finish<D>(a, orig_index(a, t), acc);
    }
    add_counters(a, fs, 0);
}

#ifndef MLH_FP32_TU
// checkFluxSymmetry (Particles.cpp:2888-2976): per entry, the partner's
// mirror entry and the reference's one-sided test F_ij[c] + F_ji[c] > tol.
// As in the reference, image entries test components 0..3 only (its ghost
// branch is the 2D one).  out[0]: entries failing the test, out[1]: entries
// without a mirror.
template <int D>
__global__ __launch_bounds__(256) void k_flux_check(HydroArgs a, FaceArrays fa, double tol,
                                                     unsigned long long *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int asym = 0, missing = 0;
    if (t < a.N) {
        const int N = a.N;
        const int cnt = a.noi[t] + a.nog[t];
        for (int n = 0; n < cnt; ++n) {
            const uint32_t e = nb_get(a.nbr, N, t, n);
            const int j = (int)nb_index(e);
            const bool image = nb_code(e) != kCodeReal;
            const int m = mirror_entry(a, t, j, image);
            if (m < 0) {
                ++missing;
                continue;
            }
            const int nc = image ? 4 : D + 2;
            bool bad = false;
            for (int c = 0; c < nc; ++c)
                bad |= fa.F[fa_at(fa, c, m, N, j)] + fa.F[fa_at(fa, c, n, N, t)] > tol;
            asym += bad ? 1 : 0;
        }
    }
    unsigned long long s0 = wave_sum((unsigned long long)asym);
    unsigned long long s1 = wave_sum((unsigned long long)missing);
    if (lane_id() == 0) {
        if (s0) atomicAdd(&out[0], s0);
        if (s1) atomicAdd(&out[1], s1);
    }
}

// Mirror check of the neighbour rows (the fused step at DEBUG_LVL >= 1):
// entries whose partner does not list the particle, into the
// missing_mirror_pairs counter.
__global__ __launch_bounds__(256) void k_mirror_check(HydroArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int missing = 0;
    if (t < a.N) {
        const int N = a.N;
        const int cnt = a.noi[t] + a.nog[t];
        for (int n = 0; n < cnt; ++n) {
            const uint32_t e = nb_get(a.nbr, N, t, n);
            if (mirror_entry(a, t, (int)nb_index(e), nb_code(e) != kCodeReal) < 0) ++missing;
        }
    }
    unsigned long long s = wave_sum((unsigned long long)missing);
    if (lane_id() == 0 && s) atomicAdd(&a.counters[kCntMissing], s);
}

#endif  // MLH_FP32_TU

}  // namespace mlh
