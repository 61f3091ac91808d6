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
// Output: y ~ log(p*/p_ref); the seed is p_ref * exp(y).
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

// seed p0 for one face from the network (FP32 VALU inference)
__device__ inline float mlp_seed_scalar(const float *__restrict__ P, double gamma, double dl,
                                        double ul, double pl, double dr, double ur,
                                        double pr) {
    if (!P || !(dl > 0.0) || !(dr > 0.0) || !(pl > 0.0) || !(pr > 0.0)) return -1.f;
    float f[kMlpIn];
    float pref = mlp_features((float)gamma, (float)dl, (float)ul, (float)pl, (float)dr,
                              (float)ur, (float)pr, f);
    float y = mlp_forward_scalar(P, f);
    if (y < -40.f) y = -40.f;
    if (y > 5.f) y = 5.f;
    return pref * expf(y);
}

}  // namespace mlh
