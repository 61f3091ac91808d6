// riemann.h -- exact ideal-gas Riemann solver, device side.
//
// Contract of the reference's external solver (src/Riemann.cpp:84,93-94):
// solve(rhoL,uL,PL, rhoR,uR,PR) -> (rho*, u*, P*) sampled at x/t = 0 and a
// flag (-1 left of the contact, +1 right, 0 vacuum).  The solver itself is
// not in the reference repository (fetched at build time from
// bwvdnbro/python_finite_volume_solver, unpinned); this is the published
// exact solver of Toro (3rd ed., ch. 4): Newton-Raphson on
// f(p) = f_L(p) + f_R(p) + du from the adaptive PVRS/TRRS/TSRS guess with
// the relative-change test 2|p1-p0|/(p1+p0) <= tol, sampled at S = 0.
//
// Precision variants:
//   solve_fp64   every operation in FP64 (the plain exact solver);
//   solve_mixed  mixed_probe then mixed_finish (below): the acoustic
//                solution for near-uniform faces, else an FP64 Newton step
//                at the seed or the standard guess, FP32 Newton only when
//                that step fails the exact solver's relative-change test,
//                an FP64 acceptance step, solve_fp64 as the fallback.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

namespace mlh {

// Fifth root in FP64: hardware log2/exp2 estimate (~1e-7), then two
// Newton steps y <- (4y + x/y^4)/5 (error ~1e-14, then round-off).
__device__ inline double root5(double x) {
    if (!(x > 1e-30) || !(x < 1e30)) return pow(x, 0.2);
    double y = (double)__builtin_amdgcn_exp2f(0.2f * __builtin_amdgcn_logf((float)x));
    for (int k = 0; k < 2; ++k) {
        double y2 = y * y;
        double y4 = y2 * y2;
        y = 0.8 * y + 0.2 * (x / y4);
    }
    return y;
}
__device__ inline float root5(float x) {
    return __builtin_amdgcn_exp2f(0.2f * __builtin_amdgcn_logf(x));
}

template <typename T>
struct GasConst {
    T g, g1, g2, g3, g4, g5, g6, g7, ginv;
    T sg5;     // sqrt(g5)
    bool g53;  // gamma == 5/3: g1 = 1/5, 1/gamma = 3/5, g3 = 5, g4 = 3
    __host__ __device__ explicit GasConst(double gam) {
        g53 = gam == 5.0 / 3.0;
        g = T(gam);
        g1 = T((gam - 1.0) / (2.0 * gam));
        g2 = T((gam + 1.0) / (2.0 * gam));
        g3 = T(2.0 * gam / (gam - 1.0));
        g4 = T(2.0 / (gam - 1.0));
        g5 = T(2.0 / (gam + 1.0));
        g6 = T((gam - 1.0) / (gam + 1.0));
        g7 = T((gam - 1.0) / 2.0);
        ginv = T(1.0 / gam);
        sg5 = T(sqrt(2.0 / (gam + 1.0)));
    }
};

__device__ inline double rpow(double x, double y) { return pow(x, y); }
// FP32 powers in the Newton loop use the hardware log2/exp2 pair
__device__ inline float rpow(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
// x^g1, x^(1/gamma), x^g4, x^g3 of the solver; closed forms for 5/3
template <typename T>
__device__ inline T pw_g1(const GasConst<T> &c, T x) { return c.g53 ? root5(x) : rpow(x, c.g1); }
template <typename T>
__device__ inline T pw_ginv(const GasConst<T> &c, T x) {
    if (!c.g53) return rpow(x, c.ginv);
    T r = root5(x);
    return r * r * r;
}
template <typename T>
__device__ inline T pw_g4(const GasConst<T> &c, T x) { return c.g53 ? x * x * x : rpow(x, c.g4); }
template <typename T>
__device__ inline T pw_g3(const GasConst<T> &c, T x) {
    if (!c.g53) return rpow(x, c.g3);
    T x2 = x * x;
    return x2 * x2 * x;
}

__device__ inline double rsqrt_(double x) { return sqrt(x); }
__device__ inline float rsqrt_(float x) { return __builtin_sqrtf(x); }

// Toro (4.6), (4.7), (4.37): f_K and f_K' for one side.
template <typename T>
__device__ inline void prefun(const GasConst<T> &c, T p, T rk, T pk, T ck,
                              T &f, T &fd) {
    if (p <= pk) {
        // (p/pk)^(-g2) = (p/pk)^g1 / (p/pk) since g2 = 1 - g1: one power
        T prat = p / pk;
        T y = pw_g1(c, prat);
        f = c.g4 * ck * (y - T(1));
        fd = (T(1) / (rk * ck)) * (y / prat);
    } else {
        T ak = c.g5 / rk;
        T bk = c.g6 * pk;
        T qrt = rsqrt_(ak / (bk + p));
        f = (p - pk) * qrt;
        fd = (T(1) - T(0.5) * (p - pk) / (bk + p)) * qrt;
    }
}

// The FP32 mode's FP64 helpers below start from the FP32 hardware estimate
// (relative error e ~ 1e-7) and take one FP64 Newton step, which squares
// the error: ~1e-14 for the reciprocal and its square root, ~1e-12 for the
// fifth root.  The FP64 residual needs to be far more accurate than the
// solver's relative-change test (1e-6) and than the FP32 p* it yields
// (6e-8), not rounding-exact.  Outside the FP32 normal range the exact
// operations.

// x^(1/5) without a division: the estimate of r = x^(-1/5) refined by
// r <- r (6 - x r^5) / 5 (error 3 e^2), then x^(1/5) = x r^4.
__device__ inline double root5_nodiv(double x) {
    if (!(x > 1e-30) || !(x < 1e30)) return pow(x, 0.2);
    double r = (double)__builtin_amdgcn_exp2f(-0.2f * __builtin_amdgcn_logf((float)x));
    double r2 = r * r;
    r = r * (1.2 - 0.2 * (x * (r2 * r2 * r)));
    r2 = r * r;
    return x * (r2 * r2);
}

// 1 / x without the divide expansion: r <- r (2 - x r) (error e^2)
__device__ inline double rcp_nodiv(double x) {
    if (!(fabs(x) > 1e-30) || !(fabs(x) < 1e30)) return 1.0 / x;
    double r = (double)__builtin_amdgcn_rcpf((float)x);
    return r * (2.0 - x * r);
}
// 1 / sqrt(x) without the square root: r <- r (3 - x r^2) / 2 (error
// 3 e^2 / 2)
__device__ inline double rsqrt_nodiv(double x) {
    if (!(x > 1e-30) || !(x < 1e30)) return 1.0 / sqrt(x);
    double r = (double)__builtin_amdgcn_rsqf((float)x);
    return r * (1.5 - 0.5 * x * (r * r));
}

// Residual of the exact solver for the FP32 mode's acceptance steps: f_K(p)
// in FP64 (it cancels against the other side and du at the root), its
// derivative in FP32 (it only scales the Newton correction; a 1e-7 relative
// error of the correction is far below the change test).  ckf is the FP32
// sound speed.  Divide-free, to FP64 rounding: p / p_K = p * (1 / p_K);
// both branches need one reciprocal square root, taken outside the branch:
//   rarefaction: c_K = sqrt(g p_K / rho_K) = g p_K / sqrt(g p_K rho_K),
//   shock:       sqrt(A_K / (B_K + p)) = sqrt(g5) / sqrt(rho_K (B_K + p)).
__device__ inline void prefun_res(const GasConst<double> &c, double p, double rk, double pk,
                                  float ckf, double &f, float &fd) {
    const bool rare = p <= pk;
    const double gp = c.g * pk, bk = c.g6 * pk;
    const double r = rsqrt_nodiv(rare ? gp * rk : rk * (bk + p));
    if (rare) {
        double prat = p * rcp_nodiv(pk);
        double y = c.g53 ? root5_nodiv(prat) : pow(prat, c.g1);
        f = c.g4 * (gp * r) * (y - 1.0);
        fd = (1.f / ((float)rk * ckf)) * ((float)y / (float)prat);
    } else {
        double qrt = c.sg5 * r;
        f = (p - pk) * qrt;
        fd = (1.f - 0.5f * (float)(p - pk) / (float)(bk + p)) * (float)qrt;
    }
}

// Adaptive initial guess (Toro 9.5.1): PVRS, else TRRS, else TSRS.
template <typename T>
__device__ inline T guess_p(const GasConst<T> &c, T dl, T ul, T pl, T cl, T dr,
                            T ur, T pr, T cr, T tiny) {
    T cup = T(0.25) * (dl + dr) * (cl + cr);
    T ppv = T(0.5) * (pl + pr) + T(0.5) * (ul - ur) * cup;
    if (ppv < T(0)) ppv = T(0);
    T pmin = pl < pr ? pl : pr;
    T pmax = pl > pr ? pl : pr;
    T qmax = pmax / pmin;
    T pm;
    if (qmax <= T(2) && pmin <= ppv && ppv <= pmax) {
        pm = ppv;
    } else if (ppv < pmin) {
        T pq = pw_g1(c, pl / pr);
        T um = (pq * ul / cl + ur / cr + c.g4 * (pq - T(1))) / (pq / cl + T(1) / cr);
        T ptl = T(1) + c.g7 * (ul - um) / cl;
        T ptr = T(1) + c.g7 * (um - ur) / cr;
        pm = T(0.5) * (pl * pw_g3(c, ptl) + pr * pw_g3(c, ptr));
    } else {
        T p0 = ppv > tiny ? ppv : tiny;
        T gel = rsqrt_((c.g5 / dl) / (c.g6 * pl + p0));
        T ger = rsqrt_((c.g5 / dr) / (c.g6 * pr + p0));
        pm = (gel * pl + ger * pr - (ur - ul)) / (gel + ger);
    }
    if (!(pm > tiny)) pm = tiny;
    return pm;
}

// Newton missed the change test within max_iter (near-vacuum states: f has
// an infinite slope at p = 0 and the iterates jump between the negative-
// pressure reset and overshoots).  f(p) = f_L + f_R + du increases with p
// and f(0) = du - g4 (c_L + c_R) < 0 when no vacuum forms, so bisect a
// bracket [0, hi] to the same relative change test.  nb: halvings done.
__device__ inline double bisect_pstar(const GasConst<double> &c, double dl, double pl,
                                      double cl, double dr, double pr, double cr, double du,
                                      double tol, int &nb) {
    double lo = 0.0, hi = pl > pr ? pl : pr, fl, fld, fr, frd;
    for (int k = 0; k < 200; ++k) {  // f grows like sqrt(p) on the shock branch
        prefun(c, hi, dl, pl, cl, fl, fld);
        prefun(c, hi, dr, pr, cr, fr, frd);
        if (fl + fr + du >= 0.0) break;
        lo = hi;
        hi *= 2.0;
    }
    nb = 0;
    while (nb < 200) {
        ++nb;
        double mid = 0.5 * (lo + hi);
        prefun(c, mid, dl, pl, cl, fl, fld);
        prefun(c, mid, dr, pr, cr, fr, frd);
        if (fl + fr + du < 0.0) lo = mid;
        else hi = mid;
        if (2.0 * (hi - lo) <= tol * (hi + lo)) break;
    }
    return 0.5 * (lo + hi);
}

template <typename T>
struct RSol {
    T d, u, p;
    int flag;
};

// Vacuum cases, Toro §4.6, sampled at S = 0: vacuum right (4.76)-(4.77),
// vacuum left (4.79)-(4.80), vacuum generated (4.82)-(4.84).
template <typename T>
__device__ inline RSol<T> sample_vacuum(const GasConst<T> &c, T dl, T ul, T pl,
                                        T cl, T dr, T ur, T pr, T cr) {
    const T s = T(0);
    const RSol<T> vac{T(0), T(0), T(0), 0};
    bool lv = !(dl > T(0)), rv = !(dr > T(0));
    if (lv && rv) return vac;
    if (!lv) {  // left gas present: left rarefaction into vacuum
        T shl = ul - cl, stl = ul + c.g4 * cl;
        if (s <= shl) return RSol<T>{dl, ul, pl, -1};
        if (s < stl) {
            T cc = c.g5 * (cl + c.g7 * (ul - s));
            return RSol<T>{dl * pw_g4(c, cc / cl), c.g5 * (cl + c.g7 * ul + s),
                           pl * pw_g3(c, cc / cl), -1};
        }
        if (rv) return vac;
    }
    // right gas present: right rarefaction into vacuum
    T shr = ur + cr, str = ur - c.g4 * cr;
    if (s >= shr) return RSol<T>{dr, ur, pr, 1};
    if (s > str) {
        T cc = c.g5 * (cr - c.g7 * (ur - s));
        return RSol<T>{dr * pw_g4(c, cc / cr), c.g5 * (-cr + c.g7 * ur + s),
                       pr * pw_g3(c, cc / cr), 1};
    }
    return vac;
}

// Sampling at S = 0 given p*, u* (Toro §4.5, SAMPLE).
template <typename T>
__device__ inline RSol<T> sample(const GasConst<T> &c, T pm, T um, T dl, T ul,
                                 T pl, T cl, T dr, T ur, T pr, T cr) {
    const T s = T(0);
    if (s <= um) {
        if (pm <= pl) {
            T shl = ul - cl;
            if (s <= shl) return RSol<T>{dl, ul, pl, -1};
            T cml = cl * pw_g1(c, pm / pl);
            T stl = um - cml;
            if (s > stl) return RSol<T>{dl * pw_ginv(c, pm / pl), um, pm, -1};
            T cc = c.g5 * (cl + c.g7 * (ul - s));
            return RSol<T>{dl * pw_g4(c, cc / cl), c.g5 * (cl + c.g7 * ul + s),
                           pl * pw_g3(c, cc / cl), -1};
        } else {
            T pml = pm / pl;
            T sl = ul - cl * rsqrt_(c.g2 * pml + c.g1);
            if (s <= sl) return RSol<T>{dl, ul, pl, -1};
            return RSol<T>{dl * (pml + c.g6) / (pml * c.g6 + T(1)), um, pm, -1};
        }
    } else {
        if (pm > pr) {
            T pmr = pm / pr;
            T sr = ur + cr * rsqrt_(c.g2 * pmr + c.g1);
            if (s >= sr) return RSol<T>{dr, ur, pr, 1};
            return RSol<T>{dr * (pmr + c.g6) / (pmr * c.g6 + T(1)), um, pm, 1};
        } else {
            T shr = ur + cr;
            if (s >= shr) return RSol<T>{dr, ur, pr, 1};
            T cmr = cr * pw_g1(c, pm / pr);
            T str = um + cmr;
            if (s <= str) return RSol<T>{dr * pw_ginv(c, pm / pr), um, pm, 1};
            T cc = c.g5 * (cr - c.g7 * (ur - s));
            return RSol<T>{dr * pw_g4(c, cc / cr), c.g5 * (-cr + c.g7 * ur + s),
                           pr * pw_g3(c, cc / cr), 1};
        }
    }
}

struct SolveStats {
    int iters;      // Newton iterations (FP32 + FP64)
    int fallback;   // learned / FP32 start rejected by the FP64 test
    int nonconv;    // FP64 iteration cap reached
    int acoustic;   // FP32: near-uniform faces given the acoustic solution
};

// Plain FP64 exact solver from the standard guess (or from p0 > 0).
__device__ inline RSol<double> solve_fp64(const GasConst<double> &c, double dl,
                                          double ul, double pl, double dr,
                                          double ur, double pr, double tol,
                                          int max_iter, double p0,
                                          SolveStats &st) {
    if (!(dl > 0.0) || !(dr > 0.0)) {
        double cl = dl > 0.0 ? sqrt(c.g * pl / dl) : 0.0;
        double cr = dr > 0.0 ? sqrt(c.g * pr / dr) : 0.0;
        return sample_vacuum(c, dl, ul, pl, cl, dr, ur, pr, cr);
    }
    double cl = sqrt(c.g * pl / dl), cr = sqrt(c.g * pr / dr);
    double du = ur - ul;
    if (c.g4 * (cl + cr) <= du) return sample_vacuum(c, dl, ul, pl, cl, dr, ur, pr, cr);
    double tiny = tol * 0.5 * (pl + pr);
    double pold = p0 > 0.0 ? p0 : guess_p(c, dl, ul, pl, cl, dr, ur, pr, cr, tiny);
    double p = pold, fl, fld, fr, frd;
    int it;
    for (it = 1; it <= max_iter; ++it) {
        prefun(c, pold, dl, pl, cl, fl, fld);
        prefun(c, pold, dr, pr, cr, fr, frd);
        p = pold - (fl + fr + du) / (fld + frd);
        double change = 2.0 * fabs((p - pold) / (p + pold));
        if (change <= tol) break;
        if (p < 0.0) p = tiny;
        pold = p;
    }
    if (it > max_iter) {
        st.nonconv += 1;
        int nb = 0;
        p = bisect_pstar(c, dl, pl, cl, dr, pr, cr, du, tol, nb);
        it = max_iter + 1 + nb;
    }
    st.iters += it;
    prefun(c, p, dl, pl, cl, fl, fld);
    prefun(c, p, dr, pr, cr, fr, frd);
    double um = 0.5 * (ul + ur + fr - fl);
    return sample(c, p, um, dl, ul, pl, cl, dr, ur, pr, cr);
}

// Converged star pressure p* of the plain FP64 solver (0 for vacuum).
__device__ inline double star_pressure_fp64(const GasConst<double> &c, double dl, double ul,
                                            double pl, double dr, double ur, double pr,
                                            double tol, int max_iter, SolveStats &st) {
    if (!(dl > 0.0) || !(dr > 0.0)) return 0.0;
    double cl = sqrt(c.g * pl / dl), cr = sqrt(c.g * pr / dr);
    double du = ur - ul;
    if (c.g4 * (cl + cr) <= du) return 0.0;
    double tiny = tol * 0.5 * (pl + pr);
    double pold = guess_p(c, dl, ul, pl, cl, dr, ur, pr, cr, tiny);
    double p = pold, fl, fld, fr, frd;
    int it;
    for (it = 1; it <= max_iter; ++it) {
        prefun(c, pold, dl, pl, cl, fl, fld);
        prefun(c, pold, dr, pr, cr, fr, frd);
        p = pold - (fl + fr + du) / (fld + frd);
        double change = 2.0 * fabs((p - pold) / (p + pold));
        if (change <= tol) break;
        if (p < 0.0) p = tiny;
        pold = p;
    }
    if (it > max_iter) {
        int nb = 0;
        p = bisect_pstar(c, dl, pl, cl, dr, pr, cr, du, tol, nb);
        it = max_iter + 1 + nb;
    }
    st.iters += it;
    return p;
}

// The mixed solver in two steps, so that work between them (the learned
// seed, for the faces that need one) can run on a whole wavefront:
//   mixed_probe   one FP64 Newton step at the start point (a seed, or the
//                 standard guess); in smooth flow the guess is exact to
//                 O(jump^2) and this step already passes the exact solver's
//                 relative-change test;
//   mixed_finish  otherwise FP32 Newton iterations from a restart point (a
//                 seed) or from the probe's FP64 step, then one FP64 step at
//                 the FP32 root; a face is accepted only when an FP64 step
//                 passes the test, else solve_fp64 runs from the standard
//                 guess (vacuum states and rejected faces share that call).
// Sampling uses the FP64-corrected p* and contact velocity.
struct MixedProbe {
    float p;          // FP32 iterate to continue from
    float cl, cr;     // FP32 sound speeds
    float chg;        // relative change of the FP64 step from the guess
    double pn, um;    // FP64 root and contact velocity once accepted
    bool valid;       // positive densities and no vacuum generated
    bool accepted;    // an FP64 step passed the change test
    bool acoustic;    // near-uniform face: pn, um from the acoustic solution
};

// Largest relative jump |drho|/rho, |dP|/P, |du|/c of a face that the FP32
// mode solves by the acoustic (linearised) solution, and the largest
// |u*| / c there.  The acoustic solution
//   P* = P + (u_L - u_R) Z / 2,  u* = u + (P_L - P_R) / (2 Z),
//   rho* = rho_K (1 + (P* - P_K) / (gamma P_K))  on the contact's upwind
// side K (L when u* >= 0), with the averages P, u, rho, c and Z = rho c
// (Z is the PVRS guess's impedance), differs from the exact solution by
// about 0.4 delta^2 relative (4e-11 at delta = 1e-5: far below FP32
// rounding).  |u*| <= c / 2 keeps x/t = 0 between the two acoustic waves
// (a transonic rarefaction fan is not linear there).
constexpr float kAcousticJump = 1e-5f;
constexpr float kAcousticMach = 0.5f;

__device__ inline MixedProbe mixed_probe(const GasConst<float> &cf, const GasConst<double> &cd,
                                         float dl, float ul, float pl, float dr, float ur,
                                         float pr, double tol, float p_start, SolveStats &st) {
    MixedProbe m{};
    bool ok = dl > 0.f && dr > 0.f;
    m.cl = ok ? __builtin_sqrtf(cf.g * pl / dl) : 0.f;
    m.cr = ok ? __builtin_sqrtf(cf.g * pr / dr) : 0.f;
    m.valid = ok && cf.g4 * (m.cl + m.cr) > ur - ul;
    if (!m.valid) return m;
    if (!(p_start > 0.f) && pl > 0.f && pr > 0.f) {
        // near-uniform face (a smooth or quiescent region): the acoustic
        // solution, no Newton step
        const float db = 0.5f * (dl + dr), pb = 0.5f * (pl + pr), cb = 0.5f * (m.cl + m.cr);
        if (__builtin_fabsf(dl - dr) <= kAcousticJump * db &&
            __builtin_fabsf(pl - pr) <= kAcousticJump * pb &&
            __builtin_fabsf(ul - ur) <= kAcousticJump * cb) {
            const float z = db * cb;
            const float ps = pb + 0.5f * (ul - ur) * z;
            const float us = 0.5f * (ul + ur) + 0.5f * (pl - pr) / z;
            if (__builtin_fabsf(us) <= kAcousticMach * cb) {
                m.accepted = m.acoustic = true;
                m.p = ps;
                m.pn = ps;
                m.um = us;
                st.acoustic += 1;
                return m;
            }
        }
    }
    float tiny = 1e-6f * 0.5f * (pl + pr);
    float p = p_start > 0.f ? p_start : guess_p(cf, dl, ul, pl, m.cl, dr, ur, pr, m.cr, tiny);
    const double dld = dl, pld = pl, drd = dr, prd = pr, dud = (double)ur - (double)ul;
    double fl, fr, pd = p;
    float fld, frd;
    prefun_res(cd, pd, dld, pld, m.cl, fl, fld);
    prefun_res(cd, pd, drd, prd, m.cr, fr, frd);
    double dp = (fl + fr + dud) * (double)(1.f / (fld + frd));
    double pn = pd - dp;
    st.iters += 1;
    m.chg = pn > 0.0 ? 2.f * __builtin_fabsf((float)dp / (float)(pn + pd)) : FLT_MAX;
    m.accepted = m.chg <= (float)tol && pn > 0.0;
    m.pn = pn;
    m.um = 0.5 * ((double)ul + (double)ur + (fr - fl) - (double)(frd - fld) * dp);
    m.p = pn > 0.0 ? (float)pn : p;
    return m;
}

__device__ inline RSol<float> mixed_finish(const GasConst<float> &cf, const GasConst<double> &cd,
                                           float dl, float ul, float pl, float dr, float ur,
                                           float pr, double tol, int max_iter, int iters32,
                                           float restart, const MixedProbe &m, SolveStats &st) {
    if (m.acoustic) {
        // the star state on the contact's upwind side (x/t = 0 lies between
        // the acoustic waves)
        const bool left = 0. <= m.um;
        const float dk = left ? dl : dr, pk = left ? pl : pr, ps = (float)m.pn;
        return RSol<float>{dk + dk * ((ps - pk) / (cf.g * pk)), (float)m.um, ps, left ? -1 : 1};
    }
    bool ok = m.valid && m.accepted;
    double pn = m.pn, um = m.um;
    if (m.valid && !m.accepted) {
        const float du = ur - ul, tiny = 1e-6f * 0.5f * (pl + pr);
        float p = restart > 0.f ? restart : m.p;
        int it = 0;
        for (; it < iters32; ++it) {
            float gl, gld, gr, grd;
            prefun(cf, p, dl, pl, m.cl, gl, gld);
            prefun(cf, p, dr, pr, m.cr, gr, grd);
            float q = p - (gl + gr + du) / (gld + grd);
            if (!(q > 0.f)) q = tiny;
            float ch = __builtin_fabsf(q - p) / (q + p);
            p = q;
            if (ch < 2e-6f) { ++it; break; }
        }
        st.iters += it;
        const double dld = dl, pld = pl, drd = dr, prd = pr, dud = (double)ur - (double)ul;
        double fl, fr, pd = p;
        float fld, frd;
        prefun_res(cd, pd, dld, pld, m.cl, fl, fld);
        prefun_res(cd, pd, drd, prd, m.cr, fr, frd);
        double dp = (fl + fr + dud) * (double)(1.f / (fld + frd));
        pn = pd - dp;
        st.iters += 1;
        ok = 2.f * __builtin_fabsf((float)dp / (float)(pn + pd)) <= (float)tol && pn > 0.0;
        if (!ok) st.fallback += 1;
        um = 0.5 * ((double)ul + (double)ur + (fr - fl) - (double)(frd - fld) * dp);
    }
    if (ok) return sample(cf, float(pn), float(um), dl, ul, pl, m.cl, dr, ur, pr, m.cr);
    // FP64 fallback, from the last FP64 step when it is positive (the FP32
    // root refined once: far closer than the standard guess), else from the
    // standard guess
    RSol<double> r = solve_fp64(cd, dl, ul, pl, dr, ur, pr, tol, max_iter,
                                m.valid && pn > 0.0 ? pn : -1.0, st);
    return RSol<float>{float(r.d), float(r.u), float(r.p), r.flag};
}

// probe at p_seed (the standard guess when <= 0), then finish
__device__ inline RSol<float> solve_mixed(const GasConst<float> &cf,
                                          const GasConst<double> &cd, float dl,
                                          float ul, float pl, float dr,
                                          float ur, float pr, double tol,
                                          int max_iter, int iters32,
                                          float p_seed, SolveStats &st) {
    MixedProbe m = mixed_probe(cf, cd, dl, ul, pl, dr, ur, pr, tol, p_seed, st);
    return mixed_finish(cf, cd, dl, ul, pl, dr, ur, pr, tol, max_iter, iters32, -1.f, m, st);
}

}  // namespace mlh
