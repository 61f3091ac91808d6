// mlp_device.h -- learned p* seed: face features and per-lane FP32 inference.
//
// Network: 10 -> 64 (tanh) -> 64 (tanh) -> 1, FP32 parameters laid out as
// W1[64][10], b1[64], W2[64][64], b2[64], W3[64], b3 (4929 floats).
// Input: features of the 1D face problem (rhoL,uL,PL | rhoR,uR,PR), gamma,
// with p_ref = (PL+PR)/2 and a = sound speed (all logs of max(x, 1e-30),
// clipped to [-40, 5]):
//   f0 = log(PL/p_ref)        f1 = log(PR/p_ref)
//   f2 = (uR-uL)/(aL+aR)      f3 = (aL-aR)/(aL+aR)
//   f4 = log(p_PVRS/p_ref)    f5 = log(p_TRRS/p_ref)
//   f6 = log(p_TSRS/p_ref)    f7 = log(p_guess/p_ref) (Toro's adaptive pick)
//   f8 = (gamma-1)/(gamma+1)  f9 = log(rhoL/rhoR)/4
// Output: y ~ log(p*/p_guess), the correction of Toro's adaptive guess;
// the seed is p_ref * exp(y + f7).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace mlh {

constexpr int kMlpIn = 10;
constexpr int kMlpH = 64;
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpH * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpH;
constexpr int kMlpB2 = kMlpW2 + kMlpH * kMlpH;
constexpr int kMlpW3 = kMlpB2 + kMlpH;
constexpr int kMlpB3 = kMlpW3 + kMlpH;
constexpr int kMlpN = kMlpB3 + 1;

__host__ __device__ inline float mlp_clog(float r) {
    float v = logf(r > 1e-30f ? r : 1e-30f);
    if (v < -40.f) v = -40.f;
    if (v > 5.f) v = 5.f;
    if (v != v) v = -40.f;
    return v;
}

// features in FP32 (the same formulas as the CPU oracle, which evaluates
// them in FP64); returns p_ref
__host__ __device__ inline float mlp_features(float gamma, float dl, float ul, float pl,
                                              float dr, float ur, float pr, float *f) {
    float g1 = (gamma - 1.f) / (2.f * gamma), g3 = 2.f * gamma / (gamma - 1.f);
    float g4 = 2.f / (gamma - 1.f), g5 = 2.f / (gamma + 1.f), g6 = (gamma - 1.f) / (gamma + 1.f),
          g7 = (gamma - 1.f) / 2.f;
    float cl = sqrtf(gamma * pl / dl), cr = sqrtf(gamma * pr / dr);
    float pref = 0.5f * (pl + pr);
    float du = ur - ul;
    float ppv = 0.5f * (pl + pr) - 0.125f * du * (dl + dr) * (cl + cr);
    if (ppv < 0.f) ppv = 0.f;
    float pq = powf(pl / pr, g1);
    float um = (pq * ul / cl + ur / cr + g4 * (pq - 1.f)) / (pq / cl + 1.f / cr);
    float ptl = 1.f + g7 * (ul - um) / cl;
    float ptr = 1.f + g7 * (um - ur) / cr;
    float ptrr = (ptl > 0.f && ptr > 0.f) ? 0.5f * (pl * powf(ptl, g3) + pr * powf(ptr, g3)) : 0.f;
    float tiny = 1e-6f * pref;
    float p0 = ppv > tiny ? ppv : tiny;
    float gel = sqrtf((g5 / dl) / (g6 * pl + p0));
    float ger = sqrtf((g5 / dr) / (g6 * pr + p0));
    float pts = (gel * pl + ger * pr - du) / (gel + ger);
    float pmin = pl < pr ? pl : pr, pmax = pl > pr ? pl : pr;
    float pg;
    if (pmax / pmin <= 2.f && pmin <= ppv && ppv <= pmax) pg = ppv;
    else if (ppv < pmin) pg = ptrr;
    else pg = pts;
    f[0] = mlp_clog(pl / pref);
    f[1] = mlp_clog(pr / pref);
    f[2] = du / (cl + cr);
    f[3] = (cl - cr) / (cl + cr);
    f[4] = mlp_clog(ppv / pref);
    f[5] = mlp_clog(ptrr / pref);
    f[6] = mlp_clog(pts / pref);
    f[7] = mlp_clog(pg / pref);
    f[8] = g6;
    f[9] = mlp_clog(dl / dr) / 4.f;
    return pref;
}

// The same features with the hardware log2 / exp2 / reciprocal (about
// 1e-6 relative), used inside the face kernel; BF16 inputs keep ~3
// significant digits anyway.
__device__ inline float fast_log(float x) { return __builtin_amdgcn_logf(x) * 0.69314718f; }
__device__ inline float fast_pow(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float clog_fast(float r) {
    float v = fast_log(r > 1e-30f ? r : 1e-30f);
    v = fminf(fmaxf(v, -40.f), 5.f);
    return v != v ? -40.f : v;
}

__device__ inline float mlp_features_fast(float gamma, float dl, float ul, float pl, float dr,
                                          float ur, float pr, float *f) {
    float g1 = (gamma - 1.f) / (2.f * gamma), g3 = 2.f * gamma / (gamma - 1.f);
    float g4 = 2.f / (gamma - 1.f), g5 = 2.f / (gamma + 1.f), g6 = (gamma - 1.f) / (gamma + 1.f),
          g7 = (gamma - 1.f) / 2.f;
    float cl = __builtin_sqrtf(gamma * pl * fast_rcp(dl)), cr = __builtin_sqrtf(gamma * pr * fast_rcp(dr));
    float pref = 0.5f * (pl + pr);
    float ipref = fast_rcp(pref);
    float du = ur - ul;
    float ppv = 0.5f * (pl + pr) - 0.125f * du * (dl + dr) * (cl + cr);
    if (ppv < 0.f) ppv = 0.f;
    float icl = fast_rcp(cl), icr = fast_rcp(cr);
    float pq = fast_pow(pl * fast_rcp(pr), g1);
    float um = (pq * ul * icl + ur * icr + g4 * (pq - 1.f)) * fast_rcp(pq * icl + icr);
    float ptl = 1.f + g7 * (ul - um) * icl;
    float ptr = 1.f + g7 * (um - ur) * icr;
    float ptrr = (ptl > 0.f && ptr > 0.f) ? 0.5f * (pl * fast_pow(ptl, g3) + pr * fast_pow(ptr, g3)) : 0.f;
    float tiny = 1e-6f * pref;
    float p0 = ppv > tiny ? ppv : tiny;
    float gel = __builtin_sqrtf((g5 * fast_rcp(dl)) * fast_rcp(g6 * pl + p0));
    float ger = __builtin_sqrtf((g5 * fast_rcp(dr)) * fast_rcp(g6 * pr + p0));
    float pts = (gel * pl + ger * pr - du) * fast_rcp(gel + ger);
    float pmin = pl < pr ? pl : pr, pmax = pl > pr ? pl : pr;
    float pg;
    if (pmax <= 2.f * pmin && pmin <= ppv && ppv <= pmax) pg = ppv;
    else if (ppv < pmin) pg = ptrr;
    else pg = pts;
    f[0] = clog_fast(pl * ipref);
    f[1] = clog_fast(pr * ipref);
    f[2] = du * fast_rcp(cl + cr);
    f[3] = (cl - cr) * fast_rcp(cl + cr);
    f[4] = clog_fast(ppv * ipref);
    f[5] = clog_fast(ptrr * ipref);
    f[6] = clog_fast(pts * ipref);
    f[7] = clog_fast(pg * ipref);
    f[8] = g6;
    f[9] = clog_fast(dl * fast_rcp(dr)) * 0.25f;
    return pref;
}

// relative pressure / velocity jump above which a face uses the learned seed
constexpr float kSeedJump = 1e-2f;
// FP32 mode: the network seeds a face only when the FP64 Newton step from
// the standard guess changed p by more than this fraction.  Below it the
// step's result (quadratically closer than the guess) is the better start
// for the remaining Newton iterations than the network's prediction.
constexpr float kSeedGate = 5e-2f;

__device__ inline float mlp_forward_scalar(const float *__restrict__ P, const float *x) {
    float h1[kMlpH];
    for (int o = 0; o < kMlpH; ++o) {
        float s = P[kMlpB1 + o];
        for (int k = 0; k < kMlpIn; ++k) s += P[kMlpW1 + o * kMlpIn + k] * x[k];
        h1[o] = tanhf(s);
    }
    float y = P[kMlpB3];
    for (int o = 0; o < kMlpH; ++o) {
        float s = P[kMlpB2 + o];
        for (int k = 0; k < kMlpH; ++k) s += P[kMlpW2 + o * kMlpH + k] * h1[k];
        y += P[kMlpW3 + o] * tanhf(s);
    }
    return y;
}

}  // namespace mlh
