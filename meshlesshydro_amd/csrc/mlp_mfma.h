// mlp_mfma.h -- one-wavefront BF16 inference of the 10-64-64-1 seed network
// on the CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Layout of v_mfma_f32_16x16x32_bf16 used here (lane l, g = l/16, c = l%16):
//   A (16 x 32): lane l holds row c, k slots 8g..8g+7;
//   B (32 x 16): lane l holds column c, the same k slots;
//   C/D (16 x 16 f32): lane l holds column c, rows 4g..4g+3.
// With Z^T = W X^T, rows are neurons and columns are faces, so the D
// fragment of a layer (neurons 4g+r of face c) becomes the next layer's B
// operand after relabelling K; mfma_h1_neuron is that relabelling.
#pragma once
#include <hip/hip_runtime.h>
#include "device_common.h"
#include "mlp_api.h"
#include "mlp_device.h"

namespace mlh {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// input neuron of layer 2 that sits in k slot j of lane group g, k step ks
__host__ __device__ inline int mfma_h1_neuron(int ks, int g, int j) {
    return 32 * ks + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

struct MlpRegs {
    bf16x8_t w1[4];
    bf16x8_t w2[4][2];
    f32x4_t b1[4], b2[4], w3[4];
    float b3;
};

__device__ inline void mlp_load_regs(const MlpFrags *__restrict__ F, MlpRegs &R) {
    int l = __lane_id();
    for (int mt = 0; mt < 4; ++mt) {
        R.w1[mt] = *(const bf16x8_t *)&F->w1[mt][l][0];
        R.w2[mt][0] = *(const bf16x8_t *)&F->w2[mt][0][l][0];
        R.w2[mt][1] = *(const bf16x8_t *)&F->w2[mt][1][l][0];
        R.b1[mt] = *(const f32x4_t *)&F->b1[mt][l][0];
        R.b2[mt] = *(const f32x4_t *)&F->b2[mt][l][0];
        R.w3[mt] = *(const f32x4_t *)&F->w3[mt][l][0];
    }
    R.b3 = F->b3;
}

// tanh(x) = 1 - 2 / (exp(2x) + 1) with the hardware exp2 / reciprocal
__device__ inline float tanh_fast(float x) {
    float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
    return 1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f);
}

// All 64 lanes must call this together.  x: the calling lane's 10 features.
// lds: 64*16 uint16 of this wavefront's LDS.  Returns this lane's output.
__device__ inline float mlp_infer_wave_bf16(const MlpRegs &R, const float *x, uint16_t *lds) {
    const int l = __lane_id(), g = l >> 4, c = l & 15;
    // features -> LDS as BF16 rows of 16 (slots 10..15 zero)
    {
        bf16x8_t lo, hi;
        for (int j = 0; j < 8; ++j) lo[j] = (__bf16)x[j];
        hi[0] = (__bf16)x[8];
        hi[1] = (__bf16)x[9];
        for (int j = 2; j < 8; ++j) hi[j] = (__bf16)0.f;
        *(bf16x8_t *)&lds[l * 16] = lo;
        *(bf16x8_t *)&lds[l * 16 + 8] = hi;
    }
    wave_lds_sync();
    f32x4_t part[4];
    f32x4_t h1[4][4];  // [M tile][N tile]
    for (int nt = 0; nt < 4; ++nt) {
        bf16x8_t bx;
        if (g < 2) {
            bx = *(const bf16x8_t *)&lds[(16 * nt + c) * 16 + 8 * g];
        } else {
            for (int j = 0; j < 8; ++j) bx[j] = (__bf16)0.f;
        }
        for (int mt = 0; mt < 4; ++mt) {
            f32x4_t acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w1[mt], bx, R.b1[mt], 0, 0, 0);
            for (int r = 0; r < 4; ++r) acc[r] = tanh_fast(acc[r]);
            h1[mt][nt] = acc;
        }
    }
    wave_lds_sync();
    for (int nt = 0; nt < 4; ++nt) {
        bf16x8_t b0, b1;
        for (int r = 0; r < 4; ++r) {
            b0[r] = (__bf16)h1[0][nt][r];
            b0[4 + r] = (__bf16)h1[1][nt][r];
            b1[r] = (__bf16)h1[2][nt][r];
            b1[4 + r] = (__bf16)h1[3][nt][r];
        }
        f32x4_t p = {0.f, 0.f, 0.f, 0.f};
        for (int mo = 0; mo < 4; ++mo) {
            f32x4_t acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w2[mo][0], b0, R.b2[mo], 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w2[mo][1], b1, acc, 0, 0, 0);
            for (int r = 0; r < 4; ++r) p[r] += R.w3[mo][r] * tanh_fast(acc[r]);
        }
        part[nt] = p;
    }
    // sum the 16 neurons held by this lane, then over the 4 lane groups
    float y[4];
    for (int nt = 0; nt < 4; ++nt) {
        float v = part[nt][0] + part[nt][1] + part[nt][2] + part[nt][3];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        y[nt] = v;
    }
    float out = g == 0 ? y[0] : (g == 1 ? y[1] : (g == 2 ? y[2] : y[3]));
    return out + R.b3;
}

// Same as mlp_infer_wave_bf16 with the fragments read from memory (L1 /
// L2 resident, 384 bytes per lane) at each call, for kernels that cannot
// keep ~100 VGPRs of weights live.
__device__ inline float mlp_infer_wave_bf16_mem(const MlpFrags *__restrict__ F, const float *x,
                                                uint16_t *lds) {
    MlpRegs R;
    mlp_load_regs(F, R);
    return mlp_infer_wave_bf16(R, x, lds);
}

// Register-lean variant for the face kernel: one face tile (16 faces) at a
// time, layer 1 straight into the packed BF16 operand of layer 2 (the lane
// that computes neurons 4g..4g+3 of face c is the lane that needs them), and
// each weight fragment read from memory (L1 resident) where it is used.
// About 30 VGPRs live instead of ~170; the arithmetic is that of
// mlp_infer_wave_bf16 except for the order of the output-layer sums.
__device__ inline float mlp_infer_wave_bf16_lean(const MlpFrags *__restrict__ F,
                                                 const float *x, uint16_t *lds) {
    const int l = __lane_id(), g = l >> 4, c = l & 15;
    {
        bf16x8_t lo, hi;
        for (int j = 0; j < 8; ++j) lo[j] = (__bf16)x[j];
        hi[0] = (__bf16)x[8];
        hi[1] = (__bf16)x[9];
        for (int j = 2; j < 8; ++j) hi[j] = (__bf16)0.f;
        *(bf16x8_t *)&lds[l * 16] = lo;
        *(bf16x8_t *)&lds[l * 16 + 8] = hi;
    }
    wave_lds_sync();
    float out = 0.f;
#pragma unroll 1
    for (int nt = 0; nt < 4; ++nt) {
        // keeps the fragment loads inside the loop (no hoisting into registers)
        asm volatile("" ::: "memory");
        bf16x8_t bx;
        if (g < 2) {
            bx = *(const bf16x8_t *)&lds[(16 * nt + c) * 16 + 8 * g];
        } else {
            for (int j = 0; j < 8; ++j) bx[j] = (__bf16)0.f;
        }
        bf16x8_t b0, b1;
        for (int mt = 0; mt < 4; ++mt) {
            bf16x8_t w = *(const bf16x8_t *)&F->w1[mt][l][0];
            f32x4_t bias = *(const f32x4_t *)&F->b1[mt][l][0];
            f32x4_t acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, bx, bias, 0, 0, 0);
            for (int r = 0; r < 4; ++r) {
                __bf16 h = (__bf16)tanh_fast(acc[r]);
                if (mt == 0) b0[r] = h;
                else if (mt == 1) b0[4 + r] = h;
                else if (mt == 2) b1[r] = h;
                else b1[4 + r] = h;
            }
        }
        f32x4_t p = {0.f, 0.f, 0.f, 0.f};
        for (int mo = 0; mo < 4; ++mo) {
            bf16x8_t wa = *(const bf16x8_t *)&F->w2[mo][0][l][0];
            bf16x8_t wb = *(const bf16x8_t *)&F->w2[mo][1][l][0];
            f32x4_t bias = *(const f32x4_t *)&F->b2[mo][l][0];
            f32x4_t w3 = *(const f32x4_t *)&F->w3[mo][l][0];
            f32x4_t acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, b0, bias, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wb, b1, acc, 0, 0, 0);
            for (int r = 0; r < 4; ++r) p[r] += w3[r] * tanh_fast(acc[r]);
        }
        float v = p[0] + p[1] + p[2] + p[3];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (g == nt) out = v;
    }
    return out + F->b3;
}

// The network predicts the correction of log p* over Toro's adaptive guess
// (feature 7): seed = p_ref * exp(y + f7), clamped like the features.
__device__ inline float mlp_seed_from_output(float y, const float *f, float pref) {
    float l = y + f[7];
    l = fminf(fmaxf(l, -40.f), 5.f);
    return pref * __expf(l);
}

}  // namespace mlh
