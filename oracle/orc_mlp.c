/*
 * orc_mlp.c -- oracle for the learned p* seed: face features and the
 * 10-64-64-1 tanh MLP forward / backward in FP64.
 *
 * TEST INFRASTRUCTURE ONLY (see mlh_oracle.h).  The reference has no
 * learned component; this file is the CPU check for the HIP forward,
 * backward and inference kernels.  Gradients are plain reverse-mode
 * derivatives of loss = mean_k (y_k - t_k)^2 written out by hand (no
 * autograd), and the tests also compare them with central differences.
 *
 * Parameter layout (float32, MLH_MLP_NPARAM = 4929 values):
 *   W1[64][10], b1[64], W2[64][64], b2[64], W3[64], b3[1]
 *   h1 = tanh(W1 x + b1), h2 = tanh(W2 h1 + b2), y = W3 . h2 + b3.
 *
 * Features of one face (rhoL,uL,PL | rhoR,uR,PR), gamma; all logs are of
 * max(ratio, 1e-30) clipped to [-40, 5]; p_ref = (PL+PR)/2, a = sound speed:
 *   f0 = log(PL/p_ref)          f1 = log(PR/p_ref)
 *   f2 = (uR-uL)/(aL+aR)        f3 = (aL-aR)/(aL+aR)
 *   f4 = log(p_PVRS/p_ref)      f5 = log(p_TRRS/p_ref)
 *   f6 = log(p_TSRS/p_ref)      f7 = log(p_guess/p_ref) (adaptive choice)
 *   f8 = (gamma-1)/(gamma+1)    f9 = log(rhoL/rhoR)/4
 * Target: log(p_star / p_ref).  PVRS / TRRS / TSRS are Toro's approximate
 * solvers (ch. 9, eqs. 9.20, 9.32, 9.42).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "mlh_oracle.h"

#define NI MLH_MLP_IN
#define NH MLH_MLP_HIDDEN

static double clog_(double r) {
    double v = log(r > 1e-30 ? r : 1e-30);
    if (v < -40.) v = -40.;
    if (v > 5.) v = 5.;
    if (isnan(v)) v = -40.;
    return v;
}

void orc_mlp_features(int64_t n, double gamma, const double *rhoL,
                      const double *uL, const double *PL, const double *rhoR,
                      const double *uR, const double *PR, float *feats,
                      double *p_ref) {
    double g1 = (gamma - 1.) / (2. * gamma), g3 = 2. * gamma / (gamma - 1.);
    double g4 = 2. / (gamma - 1.), g5 = 2. / (gamma + 1.),
           g6 = (gamma - 1.) / (gamma + 1.), g7 = (gamma - 1.) / 2.;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double dl = rhoL[k], ul = uL[k], pl = PL[k];
        double dr = rhoR[k], ur = uR[k], pr = PR[k];
        double cl = sqrt(gamma * pl / dl), cr = sqrt(gamma * pr / dr);
        double pref = 0.5 * (pl + pr);
        double du = ur - ul;
        double ppv = 0.5 * (pl + pr) - 0.125 * du * (dl + dr) * (cl + cr);
        if (ppv < 0.) ppv = 0.;
        /* TRRS */
        double pq = pow(pl / pr, g1);
        double um = (pq * ul / cl + ur / cr + g4 * (pq - 1.)) / (pq / cl + 1. / cr);
        double ptl = 1. + g7 * (ul - um) / cl;
        double ptr = 1. + g7 * (um - ur) / cr;
        double ptr_ = (ptl > 0. && ptr > 0.)
                          ? 0.5 * (pl * pow(ptl, g3) + pr * pow(ptr, g3))
                          : 0.;
        /* TSRS */
        double tiny = 1e-6 * pref;
        double p0 = ppv > tiny ? ppv : tiny;
        double gel = sqrt((g5 / dl) / (g6 * pl + p0));
        double ger = sqrt((g5 / dr) / (g6 * pr + p0));
        double pts = (gel * pl + ger * pr - du) / (gel + ger);
        /* adaptive choice (Toro 9.5.1) */
        double pmin = pl < pr ? pl : pr, pmax = pl > pr ? pl : pr;
        double pg;
        if (pmax / pmin <= 2. && pmin <= ppv && ppv <= pmax) pg = ppv;
        else if (ppv < pmin) pg = ptr_;
        else pg = pts;
        float *f = &feats[k * NI];
        f[0] = (float)clog_(pl / pref);
        f[1] = (float)clog_(pr / pref);
        f[2] = (float)(du / (cl + cr));
        f[3] = (float)((cl - cr) / (cl + cr));
        f[4] = (float)clog_(ppv / pref);
        f[5] = (float)clog_(ptr_ / pref);
        f[6] = (float)clog_(pts / pref);
        f[7] = (float)clog_(pg / pref);
        f[8] = (float)g6;
        f[9] = (float)(clog_(dl / dr) / 4.);
        if (p_ref) p_ref[k] = pref;
    }
}

typedef struct {
    const float *W1, *b1, *W2, *b2, *W3, *b3;
} mlp_view;

static mlp_view view(const float *p) {
    mlp_view v;
    v.W1 = p;
    v.b1 = v.W1 + NH * NI;
    v.W2 = v.b1 + NH;
    v.b2 = v.W2 + NH * NH;
    v.W3 = v.b2 + NH;
    v.b3 = v.W3 + NH;
    return v;
}

static double fwd1(const mlp_view *v, const float *x, double *h1, double *h2) {
    for (int o = 0; o < NH; ++o) {
        double s = v->b1[o];
        for (int i = 0; i < NI; ++i) s += (double)v->W1[o * NI + i] * x[i];
        h1[o] = tanh(s);
    }
    for (int o = 0; o < NH; ++o) {
        double s = v->b2[o];
        for (int i = 0; i < NH; ++i) s += (double)v->W2[o * NH + i] * h1[i];
        h2[o] = tanh(s);
    }
    double y = v->b3[0];
    for (int i = 0; i < NH; ++i) y += (double)v->W3[i] * h2[i];
    return y;
}

void orc_mlp_forward(int64_t n, const float *params, const float *feats,
                     double *out) {
    mlp_view v = view(params);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double h1[NH], h2[NH];
        out[k] = fwd1(&v, &feats[k * NI], h1, h2);
    }
}

double orc_mlp_loss_grad(int64_t n, const float *params, const float *feats,
                         const float *target, double *grad) {
    mlp_view v = view(params);
    memset(grad, 0, sizeof(double) * MLH_MLP_NPARAM);
    double *gW1 = grad, *gb1 = gW1 + NH * NI, *gW2 = gb1 + NH,
           *gb2 = gW2 + NH * NH, *gW3 = gb2 + NH, *gb3 = gW3 + NH;
    double loss = 0.;
    for (int64_t k = 0; k < n; ++k) {
        const float *x = &feats[k * NI];
        double h1[NH], h2[NH];
        double y = fwd1(&v, x, h1, h2);
        double e = y - (double)target[k];
        loss += e * e;
        double dy = 2. * e / (double)n;
        gb3[0] += dy;
        double d2[NH], d1[NH];
        for (int i = 0; i < NH; ++i) {
            gW3[i] += dy * h2[i];
            d2[i] = dy * (double)v.W3[i] * (1. - h2[i] * h2[i]);
        }
        for (int o = 0; o < NH; ++o) {
            gb2[o] += d2[o];
            for (int i = 0; i < NH; ++i) gW2[o * NH + i] += d2[o] * h1[i];
        }
        for (int i = 0; i < NH; ++i) {
            double s = 0.;
            for (int o = 0; o < NH; ++o) s += (double)v.W2[o * NH + i] * d2[o];
            d1[i] = s * (1. - h1[i] * h1[i]);
        }
        for (int o = 0; o < NH; ++o) {
            gb1[o] += d1[o];
            for (int i = 0; i < NI; ++i) gW1[o * NI + i] += d1[o] * x[i];
        }
    }
    return loss / (double)n;
}
