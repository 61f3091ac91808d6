// mlp.hip -- learned p* seed: training (FP32 forward/backward + Adam),
// inference (FP32 per lane, and BF16 on the matrix cores), face collection,
// and the batched exact Riemann entry point.
//
// BF16 inference runs one wavefront per 64 faces with
// v_mfma_f32_16x16x32_bf16 in the transposed form Z^T = W X^T, so that a
// layer's accumulator fragments are, lane for lane, the next layer's B
// operand (the K index inside one MFMA is only a label: W2's columns are
// permuted to match once, when the fragments are packed).  Biases ride in
// the accumulator initialisation; the last 64 -> 1 layer is a per-lane dot
// product plus two cross-lane adds.  Activations never leave registers;
// only the features go through 2 KB of LDS per wavefront.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <mutex>
#include <vector>
#include "../../include/mlh.h"
#include "device_common.h"
#include "kernels.h"
#include "mlp_api.h"
#include "mlp_device.h"
#include "mlp_mfma.h"
#include "riemann.h"

using namespace mlh;

namespace mlh {

void mlp_free(MlpState &m) {
    if (m.params) hipFree(m.params);
    if (m.frags) hipFree(m.frags);
    if (m.adam_m) hipFree(m.adam_m);
    if (m.adam_v) hipFree(m.adam_v);
    if (m.grad) hipFree(m.grad);
    if (m.collect) hipFree(m.collect);
    if (m.collect_n) hipFree(m.collect_n);
    m = MlpState{};
}

// ---- packing of the MFMA fragments ------------------------------------------
__global__ void k_mlp_pack(const float *__restrict__ P, MlpFrags *__restrict__ F) {
    int l = threadIdx.x;  // 64 threads = one lane each
    int g = l >> 4, c = l & 15;
    for (int mt = 0; mt < 4; ++mt) {
        for (int j = 0; j < 8; ++j) {
            int k = 8 * g + j;
            float w = k < kMlpIn ? P[kMlpW1 + (16 * mt + c) * kMlpIn + k] : 0.f;
            __bf16 b = (__bf16)w;
            F->w1[mt][l][j] = *(uint16_t *)&b;
        }
        for (int r = 0; r < 4; ++r) {
            int n = 16 * mt + 4 * g + r;
            F->b1[mt][l][r] = P[kMlpB1 + n];
            F->b2[mt][l][r] = P[kMlpB2 + n];
            F->w3[mt][l][r] = P[kMlpW3 + n];
        }
        for (int ks = 0; ks < 2; ++ks)
            for (int j = 0; j < 8; ++j) {
                int in = mfma_h1_neuron(ks, g, j);
                __bf16 b = (__bf16)P[kMlpW2 + (16 * mt + c) * kMlpH + in];
                F->w2[mt][ks][l][j] = *(uint16_t *)&b;
            }
    }
    if (l == 0) F->b3 = P[kMlpB3];
}

// ---- training --------------------------------------------------------------
// One block of four wavefronts per chunk of 64 samples of the batch.  The
// GEMM-shaped products run on the matrix cores in FP32
// (v_mfma_f32_16x16x4f32; lane l, g = l / 16, j = l % 16: A holds row j,
// k slot g; B holds column j, k slot g; D holds column j, rows 4g..4g+3).
// K is a label inside one MFMA, so step ks of lane group g takes the k
// index 16g + ks (4g + ks for the 16-wide input layer): a lane's A
// operands of one tile are then 16 consecutive floats of an LDS row.
// Wavefront w owns 16 neurons of each product:
//   forward    Z1 = X W1^T + b1, Z2 = H1 W2^T + b2    (samples x neurons)
//   backward   dW2 = dz2^T H1, dz1 = (dz2 W2)(1 - h1^2), [dW1 | db1] =
//              dz1^T [X | 1]                          (the ones column)
// and the small vector terms (the output layer, dy = 2 e / B, dz2 = dy w3
// (1 - h2^2), dW3, db2, db3) run on the VALU.  Each chunk stores its
// partial gradient; k_mlp_reduce sums the chunks' partials in chunk order
// and applies Adam: the result does not depend on scheduling and needs no
// atomics.
constexpr int kTrainChunk = 64;
constexpr int kTrainLd = kMlpH + 4;   // LDS row pitch: rows 16-byte aligned
constexpr int kTrainXLd = 16;         // input tile: 10 features, 1, zeros
constexpr int kPartialN = kMlpN + 1;  // gradient + the chunk's sum of e^2

__device__ inline f32x4_t mfma4(float a, float b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__global__ __launch_bounds__(256) void k_mlp_chunk_grad(const float *__restrict__ P,
                                                        const float *__restrict__ X,
                                                        const float *__restrict__ T,
                                                        const int *__restrict__ idx, int B,
                                                        float *__restrict__ partial) {
    __shared__ __align__(16) float XS[kTrainChunk][kTrainXLd];
    __shared__ __align__(16) float H1[kTrainChunk][kTrainLd];
    __shared__ __align__(16) float H2[kTrainChunk][kTrainLd];  // h2, then dz1
    __shared__ __align__(16) float DZ2[kTrainChunk][kTrainLd];
    __shared__ float YP[4][kTrainChunk];
    __shared__ float DY[kTrainChunk], E2[kTrainChunk];
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wavefront-uniform
    const int g = l >> 4, j = l & 15;
    const int c0 = blockIdx.x * kTrainChunk;
    // the input tile: features, a ones column (db1), zeros; samples past
    // the batch end are zero rows and contribute nothing (dy = 0)
    for (int e = tid; e < kTrainChunk * kTrainXLd; e += 256) {
        const int q = e >> 4, k = e & 15, b = c0 + q;
        float v = 0.f;
        if (b < B) {
            const int s = idx ? idx[b] : b;
            v = k < kMlpIn ? X[(size_t)s * kMlpIn + k] : (k == kMlpIn ? 1.f : 0.f);
        }
        XS[q][k] = v;
    }
    // W2 operands (registers): rows 16w + j for the forward product, column
    // 16w + j for dz1, each at k = 16g + ks
    float w2r[16], w2c[16];
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
        w2r[ks] = P[kMlpW2 + (16 * w + j) * kMlpH + 16 * g + ks];
        w2c[ks] = P[kMlpW2 + (16 * g + ks) * kMlpH + 16 * w + j];
    }
    __syncthreads();
    // layer 1: Z1[samples][16w..16w+15], K = 16 (k = 4g + ks)
    {
        const f32x4_t bias = {P[kMlpB1 + 16 * w + j], P[kMlpB1 + 16 * w + j],
                              P[kMlpB1 + 16 * w + j], P[kMlpB1 + 16 * w + j]};
        float w1[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int k = 4 * g + ks;
            w1[ks] = k < kMlpIn ? P[kMlpW1 + (16 * w + j) * kMlpIn + k] : 0.f;
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4_t acc = bias;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) acc = mfma4(XS[16 * mt + j][4 * g + ks], w1[ks], acc);
#pragma unroll
            for (int r = 0; r < 4; ++r) H1[16 * mt + 4 * g + r][16 * w + j] = tanhf(acc[r]);
        }
    }
    __syncthreads();
    // layer 2: Z2[samples][16w..16w+15], K = 64
    {
        const float b2 = P[kMlpB2 + 16 * w + j];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const f32x4_t *row = (const f32x4_t *)&H1[16 * mt + j][16 * g];
            f32x4_t a4[4] = {row[0], row[1], row[2], row[3]};
            f32x4_t acc = {b2, b2, b2, b2};
#pragma unroll
            for (int ks = 0; ks < 16; ++ks) acc = mfma4(a4[ks >> 2][ks & 3], w2r[ks], acc);
#pragma unroll
            for (int r = 0; r < 4; ++r) H2[16 * mt + 4 * g + r][16 * w + j] = tanhf(acc[r]);
        }
    }
    __syncthreads();
    // output layer, loss term, dz2 (thread: sample l, neurons 16w..16w+15)
    const float *w3 = P + kMlpW3 + 16 * w;
    {
        float yp = 0.f;
        for (int oo = 0; oo < 16; ++oo) yp += w3[oo] * H2[l][16 * w + oo];
        YP[w][l] = yp;
    }
    __syncthreads();
    {
        const int b = c0 + l;
        const bool act = b < B;
        const float y = P[kMlpB3] + ((YP[0][l] + YP[1][l]) + (YP[2][l] + YP[3][l]));
        const float e = act ? y - T[idx ? idx[b] : b] : 0.f;
        const float dy = 2.f * e / (float)B;
        if (w == 0) {
            DY[l] = dy;
            E2[l] = e * e;
        }
        for (int oo = 0; oo < 16; ++oo) {
            const float h2 = H2[l][16 * w + oo];
            DZ2[l][16 * w + oo] = dy * w3[oo] * (1.f - h2 * h2);
        }
    }
    __syncthreads();
    float *gr = partial + (size_t)blockIdx.x * kPartialN;
    // dW2[16w..16w+15][0..63] = dz2^T H1, K = samples (s = 16g + ks)
    {
        f32x4_t acc[4] = {};
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) {
            const int q = 16 * g + ks;
            const float a = DZ2[q][16 * w + j];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[nt] = mfma4(a, H1[q][16 * nt + j], acc[nt]);
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                gr[kMlpW2 + (16 * w + 4 * g + r) * kMlpH + 16 * nt + j] = acc[nt][r];
    }
    // dz1 before the h1 factor: (dz2 W2)[samples][16w..16w+15], K = 64
    f32x4_t dz1[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const f32x4_t *row = (const f32x4_t *)&DZ2[16 * mt + j][16 * g];
        f32x4_t a4[4] = {row[0], row[1], row[2], row[3]};
        f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) acc = mfma4(a4[ks >> 2][ks & 3], w2c[ks], acc);
        dz1[mt] = acc;
    }
    // dW3, db2 (threads 0-63, sample order), db3 and the sum of e^2
    if (tid < kMlpH) {
        float s3 = 0.f, bb = 0.f;
        for (int q = 0; q < kTrainChunk; ++q) {
            s3 += DY[q] * H2[q][tid];
            bb += DZ2[q][tid];
        }
        gr[kMlpW3 + tid] = s3;
        gr[kMlpB2 + tid] = bb;
    } else if (tid == kMlpH) {
        float b3 = 0.f, se = 0.f;
        for (int q = 0; q < kTrainChunk; ++q) {
            b3 += DY[q];
            se += E2[q];
        }
        gr[kMlpB3] = b3;
        gr[kMlpN] = se;
    }
    __syncthreads();  // H2 is read above, dz1 goes there below
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float h1 = H1[16 * mt + 4 * g + r][16 * w + j];
            H2[16 * mt + 4 * g + r][16 * w + j] = dz1[mt][r] * (1.f - h1 * h1);
        }
    __syncthreads();
    // [dW1 | db1][16w..16w+15][0..15] = dz1^T [X | 1], K = samples
    {
        f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) {
            const int q = 16 * g + ks;
            acc = mfma4(H2[q][16 * w + j], XS[q][j], acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int o = 16 * w + 4 * g + r;
            if (j < kMlpIn) gr[kMlpW1 + o * kMlpIn + j] = acc[r];
            else if (j == kMlpIn) gr[kMlpB1 + o] = acc[r];
        }
    }
}

// Sum of the chunks' partials, eight lanes per entry: lane q of entry p
// sums chunks q, q + 8, q + 16, ... in order, and the eight sums combine
// in a fixed butterfly (every lane gets the same bits: IEEE addition
// commutes).  Entry p < kMlpN: gradient entry p (stored in G, and with
// adam the Adam step of parameter p); entry kMlpN: the loss, mean of e^2
// over the batch.
constexpr int kReduceLanes = 8;
constexpr int kReducePerBlock = 256 / kReduceLanes;
__global__ __launch_bounds__(256) void k_mlp_reduce(const float *__restrict__ partial,
                                                    int nchunks, int B, float *__restrict__ G,
                                                    float *__restrict__ loss, int adam,
                                                    float *__restrict__ P, float *__restrict__ M,
                                                    float *__restrict__ V, float lr, float b1c,
                                                    float b2c) {
    const int p = blockIdx.x * kReducePerBlock + (threadIdx.x >> 3);
    const int q = threadIdx.x & (kReduceLanes - 1);
    float g = 0.f;
    if (p <= kMlpN)
        for (int c = q; c < nchunks; c += kReduceLanes) g += partial[(size_t)c * kPartialN + p];
    g += __shfl_xor(g, 1, 64);
    g += __shfl_xor(g, 2, 64);
    g += __shfl_xor(g, 4, 64);
    if (p > kMlpN || q != 0) return;
    if (p == kMlpN) {
        *loss = g / (float)B;
        return;
    }
    G[p] = g;
    if (!adam) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float m = b1 * M[p] + (1.f - b1) * g;
    const float v = b2 * V[p] + (1.f - b2) * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / b1c) / (sqrtf(v / b2c) + eps);
}

// ---- inference ---------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mlp_infer_f32(const float *__restrict__ P,
                                                       const float *__restrict__ X, int64_t n,
                                                       float *__restrict__ Y) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[kMlpIn];
    for (int k = 0; k < kMlpIn; ++k) x[k] = X[s * kMlpIn + k];
    Y[s] = mlp_forward_scalar(P, x);
}

__global__ __launch_bounds__(256) void k_mlp_infer_bf16(const MlpFrags *__restrict__ F,
                                                        const float *__restrict__ X, int64_t n,
                                                        float *__restrict__ Y) {
    __shared__ uint16_t lds[4][64 * 16];
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int w = threadIdx.x >> 6;
    bool active = s < n;
    float x[kMlpIn];
    for (int k = 0; k < kMlpIn; ++k) x[k] = active ? X[s * kMlpIn + k] : 0.f;
    MlpRegs R;
    mlp_load_regs(F, R);
    float y = mlp_infer_wave_bf16(R, x, lds[w]);
    if (active) Y[s] = y;
}

// ---- batched exact Riemann solver -------------------------------------------
// Faces the FP64 solver finished: added to fallbacks[0] (atomics), or, with
// block_counts, block b's count stored in block_counts[b] (every block
// stores its slot, zero included, so the slots need no clearing).
__global__ __launch_bounds__(256) void k_riemann_batch(
    int64_t n, double gamma, const double *__restrict__ dl, const double *__restrict__ ul,
    const double *__restrict__ pl, const double *__restrict__ dr, const double *__restrict__ ur,
    const double *__restrict__ pr, double *__restrict__ ds, double *__restrict__ us,
    double *__restrict__ ps, int *__restrict__ flag, int *__restrict__ iters, int precision,
    const MlpFrags *__restrict__ F, double tol, int max_iter, int iters32,
    unsigned long long *__restrict__ fallbacks, double *__restrict__ pstar,
    unsigned long long *__restrict__ block_counts) {
    __shared__ unsigned long long wcount[4];
    __shared__ uint16_t lds[4][64 * 16];
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool active = k < n;
    double a0 = active ? dl[k] : 1., a1 = active ? ul[k] : 0., a2 = active ? pl[k] : 1.;
    double b0 = active ? dr[k] : 1., b1 = active ? ur[k] : 0., b2 = active ? pr[k] : 1.;
    float seed = -1.f;
    if (F) {
        // every lane of the wavefront takes part in the MFMA inference
        float f[kMlpIn];
        bool ok = a0 > 0. && b0 > 0. && a2 > 0. && b2 > 0.;
        float pref = ok ? mlp_features_fast((float)gamma, (float)a0, (float)a1, (float)a2,
                                            (float)b0, (float)b1, (float)b2, f)
                        : 1.f;
        if (!ok)
            for (int q = 0; q < kMlpIn; ++q) f[q] = 0.f;
        MlpRegs R;
        mlp_load_regs(F, R);
        float y = mlp_infer_wave_bf16(R, f, lds[threadIdx.x >> 6]);
        seed = ok ? mlp_seed_from_output(y, f, pref) : -1.f;
    }
    SolveStats st{};
    if (pstar) {
        if (!active) return;
        GasConst<double> gc(gamma);
        pstar[k] = star_pressure_fp64(gc, a0, a1, a2, b0, b1, b2, tol, max_iter, st);
        if (iters) iters[k] = st.iters;
        return;
    }
    if (active) {
        double rd, ru, rp;
        int fl;
        if (precision == MLH_FP64) {
            GasConst<double> gc(gamma);
            RSol<double> r = solve_fp64(gc, a0, a1, a2, b0, b1, b2, tol, max_iter, (double)seed, st);
            rd = r.d; ru = r.u; rp = r.p; fl = r.flag;
        } else {
            GasConst<float> gf(gamma);
            GasConst<double> gc(gamma);
            RSol<float> r = solve_mixed(gf, gc, (float)a0, (float)a1, (float)a2, (float)b0,
                                        (float)b1, (float)b2, tol, max_iter, iters32, seed, st);
            rd = r.d; ru = r.u; rp = r.p; fl = r.flag;
        }
        ds[k] = rd; us[k] = ru; ps[k] = rp;
        if (flag) flag[k] = fl;
        if (iters) iters[k] = st.iters;
    }
    if (!block_counts) {
        if (st.fallback) atomicAdd(fallbacks, (unsigned long long)st.fallback);
        return;
    }
    // every thread of the block is here: wavefront sums, then the block's
    unsigned long long v = (unsigned long long)st.fallback;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) wcount[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        block_counts[blockIdx.x] = wcount[0] + wcount[1] + wcount[2] + wcount[3];
}

// out[0] = sum of in[0..n) (one block of 256 threads)
__global__ __launch_bounds__(256) void k_sum_u64(const unsigned long long *__restrict__ in,
                                                  int n, unsigned long long *__restrict__ out) {
    __shared__ unsigned long long sh[256];
    unsigned long long v = 0;
    for (int k = threadIdx.x; k < n; k += 256) v += in[k];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// ---- training sets built on the device -------------------------------------
// The learned seed's training set is built from face problems (rhoL, uL,
// PL, rhoR, uR, PR) already on the device, the run's record or an uploaded
// array, without a round trip through the host:
//   1. k_ds_prep: per face, the FP64 exact p* (tol 1e-12, at most 50
//      iterations, as mlh_riemann_pstar_batch), the FP32 features, the target
//      log(p* / p_ref) - f7 in FP32 and a class: dropped (a non-positive or
//      non-finite input state, or p* = 0), trivial (|target| < 1e-5: the
//      standard guess is converged) or non-trivial.
//   2. k_ds_select keeps every non-trivial face and each trivial one with
//      probability keep / n_trivial, keep = min(n_trivial, max(keep_trivial
//      * max(n_nontrivial, 1), 1024)), decided by a hash of the face.
//   3. The kept faces are sorted by their content hash h (by h >> 33, then
//      by the low 32 bits), which fixes the set's order whatever order the
//      record was filled in; the first max(n / 10, 64) are the validation
//      faces, the rest the training faces.
//   4. Each epoch visits the training faces in the order of a hash of
//      (seed, epoch, position), sorted the same way.
// Every step is deterministic, so a retraining gives the same weights
// whenever the same faces were recorded.

// splitmix64's output function (Steele, Lea and Flood, OOPSLA 2014)
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
constexpr uint64_t kSelectSalt = 0x5851f42d4c957f2dull;
constexpr uint64_t kEpochSalt = 0xa0761d6478bd642full;
constexpr float kTrivialTarget = 1e-5f;
constexpr int kMinTrainFaces = 256;

// face slot k of the input: counts null -> every slot [0, nslots) holds a
// face; else slot k is row r = k % scap of stripe s = k / scap, filled when
// r < counts[s].  Launched with 256 threads per block.
__global__ __launch_bounds__(256) void k_ds_prep(const double *__restrict__ F6,
                                                  const unsigned long long *__restrict__ counts,
                                                  int64_t scap, int64_t nslots, double gamma,
                                                  float *__restrict__ Xall,
                                                  float *__restrict__ Tall,
                                                  uint32_t *__restrict__ hlo,
                                                  uint32_t *__restrict__ hhi,
                                                  uint8_t *__restrict__ cls,
                                                  unsigned long long *__restrict__ ncls) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool used = k < nslots;
    if (used && counts) used = (unsigned long long)(k % scap) < counts[k / scap];
    uint8_t cl = 0;
    uint64_t h = ~0ull;
    if (used) {
        double a[6];
        h = 0;
        for (int q = 0; q < 6; ++q) {
            a[q] = F6[6 * k + q];
            h = mix64(h ^ (uint64_t)__double_as_longlong(a[q]));
        }
        bool ok = a[0] > 0. && a[3] > 0. && a[2] > 0. && a[5] > 0.;
        for (int q = 0; q < 6; ++q) ok = ok && isfinite(a[q]);
        if (ok) {
            GasConst<double> gc(gamma);
            SolveStats st{};
            const double ps = star_pressure_fp64(gc, a[0], a[1], a[2], a[3], a[4], a[5], 1e-12,
                                                 50, st);
            if (ps > 0.) {
                float f[kMlpIn];
                const float pref = mlp_features((float)gamma, (float)a[0], (float)a[1],
                                                (float)a[2], (float)a[3], (float)a[4],
                                                (float)a[5], f);
                const float t = (float)(log(ps / (double)pref) - (double)f[7]);
                cl = fabsf(t) < kTrivialTarget ? 1 : 2;
                for (int q = 0; q < kMlpIn; ++q) Xall[(size_t)k * kMlpIn + q] = f[q];
                Tall[k] = t;
            }
        }
    }
    if (k < nslots) {
        cls[k] = cl;
        hlo[k] = (uint32_t)h;
        hhi[k] = (uint32_t)(h >> 32);
    }
    // the block's class counts: ncls[block] trivial, ncls[nblocks + block]
    // non-trivial (every block stores both)
    __shared__ unsigned long long wc[2][4];
    const unsigned long long n1 = __popcll(__ballot(cl == 1)), n2 = __popcll(__ballot(cl == 2));
    if ((threadIdx.x & 63) == 0) {
        wc[0][threadIdx.x >> 6] = n1;
        wc[1][threadIdx.x >> 6] = n2;
    }
    __syncthreads();
    if (threadIdx.x < 2)
        ncls[threadIdx.x * gridDim.x + blockIdx.x] =
            wc[threadIdx.x][0] + wc[threadIdx.x][1] + wc[threadIdx.x][2] + wc[threadIdx.x][3];
}

// selection (trivial faces kept when the top 53 bits of a second hash are
// below thr53) and the high sort key: (h >> 33), with bit 31 set for the
// faces that are not kept.  Launched with 256 threads per block.
__global__ __launch_bounds__(256) void k_ds_select(const uint8_t *__restrict__ cls,
                                                    const uint32_t *__restrict__ hlo,
                                                    const uint32_t *__restrict__ hhi,
                                                    int64_t nslots, uint64_t thr53,
                                                    uint32_t *__restrict__ khi,
                                                    unsigned long long *__restrict__ nsel) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool sel = false;
    if (k < nslots) {
        const uint8_t cl = cls[k];
        const uint64_t h = (uint64_t)hlo[k] | ((uint64_t)hhi[k] << 32);
        sel = cl == 2 || (cl == 1 && (mix64(h ^ kSelectSalt) >> 11) < thr53);
        khi[k] = (hhi[k] >> 1) | (sel ? 0u : 0x80000000u);
    }
    // the block's count of kept faces in nsel[block]
    __shared__ unsigned long long wc[4];
    const unsigned long long n = __popcll(__ballot(sel));
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) nsel[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ __launch_bounds__(256) void k_iota_u32(uint32_t *__restrict__ v, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) v[k] = (uint32_t)k;
}

// out[k] = in[idx[k]]
__global__ __launch_bounds__(256) void k_gather_u32(const uint32_t *__restrict__ in,
                                                     const uint32_t *__restrict__ idx, int64_t n,
                                                     uint32_t *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = in[idx[k]];
}

// the kept faces' features and targets in sorted order
__global__ __launch_bounds__(256) void k_ds_gather(const uint32_t *__restrict__ order, int64_t n,
                                                    const float *__restrict__ Xall,
                                                    const float *__restrict__ Tall,
                                                    float *__restrict__ X, float *__restrict__ T) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const size_t s = order[k];
    for (int q = 0; q < kMlpIn; ++q) X[(size_t)k * kMlpIn + q] = Xall[s * kMlpIn + q];
    T[k] = Tall[s];
}

// epoch permutation keys: key[i] = top 32 bits of a hash of (seed, epoch,
// i), val[i] = first + i
__global__ __launch_bounds__(256) void k_epoch_keys(uint32_t *__restrict__ key,
                                                     uint32_t *__restrict__ val, int64_t n,
                                                     uint32_t first, uint64_t epoch_hash) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    key[k] = (uint32_t)(mix64(epoch_hash ^ (uint64_t)k) >> 32);
    val[k] = first + (uint32_t)k;
}

}  // namespace mlh

// ---- C ABI --------------------------------------------------------------------
#define CHK(expr)                                                                    \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) return ctx_fail(MLH_ERR_DEVICE, hipGetErrorString(e_)); \
    } while (0)

template <typename P>
static hipError_t dmalloc(P **p, size_t bytes) {
    return hipMalloc((void **)p, bytes ? bytes : 16);
}

// device buffer freed at scope exit
template <typename P>
struct DevBuf {
    P *p = nullptr;
    ~DevBuf() {
        if (p) hipFree(p);
    }
};

static int ensure_mlp(mlh_ctx *c) {
    MlpState &m = ctx_mlp(c);
    if (!m.params) {
        CHK(dmalloc(&m.params, sizeof(float) * kMlpN));
        CHK(dmalloc(&m.frags, sizeof(MlpFrags)));
        CHK(dmalloc(&m.adam_m, sizeof(float) * kMlpN));
        CHK(dmalloc(&m.adam_v, sizeof(float) * kMlpN));
        CHK(dmalloc(&m.grad, sizeof(float) * kMlpN));
        CHK(hipMemset(m.adam_m, 0, sizeof(float) * kMlpN));
        CHK(hipMemset(m.adam_v, 0, sizeof(float) * kMlpN));
        m.adam_t = 0;
    }
    return MLH_OK;
}

static void repack(mlh_ctx *c) {
    MlpState &m = ctx_mlp(c);
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, ctx_stream(c), m.params, m.frags);
}

extern "C" {

int mlh_mlp_features(int64_t n, double gamma, const double *rhoL, const double *uL,
                     const double *PL, const double *rhoR, const double *uR, const double *PR,
                     float *feats, double *p_ref) {
    if (n < 0 || !rhoL || !uL || !PL || !rhoR || !uR || !PR || !feats)
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    for (int64_t k = 0; k < n; ++k) {
        float pr = mlp_features((float)gamma, (float)rhoL[k], (float)uL[k], (float)PL[k],
                                (float)rhoR[k], (float)uR[k], (float)PR[k], &feats[k * kMlpIn]);
        if (p_ref) p_ref[k] = pr;
    }
    return MLH_OK;
}

int mlh_mlp_set_params(mlh_ctx *c, const float *params) {
    if (!c || !params) return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    int st = ensure_mlp(c);
    if (st) return st;
    MlpState &m = ctx_mlp(c);
    CHK(hipMemcpyAsync(m.params, params, sizeof(float) * kMlpN, hipMemcpyHostToDevice,
                       ctx_stream(c)));
    CHK(hipMemsetAsync(m.adam_m, 0, sizeof(float) * kMlpN, ctx_stream(c)));
    CHK(hipMemsetAsync(m.adam_v, 0, sizeof(float) * kMlpN, ctx_stream(c)));
    m.adam_t = 0;
    repack(c);
    CHK(hipStreamSynchronize(ctx_stream(c)));
    return MLH_OK;
}

int mlh_mlp_get_params(mlh_ctx *c, float *params) {
    if (!c || !params) return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    MlpState &m = ctx_mlp(c);
    if (!m.params) return ctx_fail(MLH_ERR_NOT_TRAINED, "no network parameters");
    CHK(hipMemcpy(params, m.params, sizeof(float) * kMlpN, hipMemcpyDeviceToHost));
    return MLH_OK;
}

struct TrainBufs {
    float *X = nullptr, *T = nullptr, *partial = nullptr, *loss = nullptr;
    int *idx = nullptr;
    // epoch permutation: sort keys and the radix sort's ping-pong / scratch
    uint32_t *pk = nullptr, *pk1 = nullptr, *pv1 = nullptr, *scr = nullptr;
    ~TrainBufs() {
        float *f[] = {X, T, partial, loss};
        for (float *p : f)
            if (p) hipFree(p);
        uint32_t *u[] = {pk, pk1, pv1, scr};
        for (uint32_t *p : u)
            if (p) hipFree(p);
        if (idx) hipFree(idx);
    }
};

static int train_chunks(int batch) { return (batch + kTrainChunk - 1) / kTrainChunk; }

// n samples; partial sums for batches of up to max_batch samples
static int alloc_train(TrainBufs &b, int64_t n, int max_batch) {
    CHK(dmalloc(&b.X, sizeof(float) * n * kMlpIn));
    CHK(dmalloc(&b.T, sizeof(float) * n));
    CHK(dmalloc(&b.partial, sizeof(float) * (size_t)train_chunks(max_batch) * kPartialN));
    CHK(dmalloc(&b.loss, sizeof(float)));
    CHK(dmalloc(&b.idx, sizeof(int) * n));
    CHK(dmalloc(&b.pk, sizeof(uint32_t) * n));
    CHK(dmalloc(&b.pk1, sizeof(uint32_t) * n));
    CHK(dmalloc(&b.pv1, sizeof(uint32_t) * n));
    CHK(dmalloc(&b.scr, sizeof(uint32_t) * radix_sort_scratch_words((int)n)));
    return MLH_OK;
}

// gradient (and loss) of the batch idx[0..B) (idx null: samples 0..B-1);
// with adam, the Adam step at bias corrections b1c / b2c
static void enqueue_grad(mlh_ctx *c, TrainBufs &b, const int *idx, int B, bool adam = false,
                         float lr = 0.f, float b1c = 1.f, float b2c = 1.f) {
    MlpState &m = ctx_mlp(c);
    hipStream_t s = ctx_stream(c);
    const int nch = train_chunks(B);
    hipLaunchKernelGGL(k_mlp_chunk_grad, dim3(nch), dim3(256), 0, s, m.params, b.X, b.T, idx, B,
                       b.partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kPartialN + kReducePerBlock - 1) / kReducePerBlock),
                       dim3(256), 0, s, b.partial,
                       nch, B, m.grad, b.loss, adam ? 1 : 0, m.params, m.adam_m, m.adam_v, lr,
                       b1c, b2c);
}

// Adam epochs over the samples [first, first + n) of b.X / b.T in batches
// of `batch` (a partial last batch is left out), each epoch in the order of
// k_epoch_keys; *last = the last batch's loss (0 without epochs).  Repacks
// the BF16 fragments and synchronises.
static int train_epochs(mlh_ctx *c, TrainBufs &b, int64_t first, int64_t n, int n_epochs,
                        int batch, float lr, uint64_t seed, float *last) {
    MlpState &m = ctx_mlp(c);
    hipStream_t s = ctx_stream(c);
    const int64_t nb = (n + 255) / 256;
    *last = 0.f;
    for (int ep = 0; ep < n_epochs; ++ep) {
        const uint64_t eh = mix64(mix64(seed ^ kEpochSalt) + (uint64_t)ep);
        hipLaunchKernelGGL(k_epoch_keys, dim3(nb), dim3(256), 0, s, b.pk, (uint32_t *)b.idx, n,
                           (uint32_t)first, eh);
        radix_sort_pairs(b.pk, (uint32_t *)b.idx, b.pk1, b.pv1, (int)n, 32, b.scr, s);
        for (int64_t off = 0; off + batch <= n; off += batch) {
            m.adam_t++;
            float b1c = 1.f - powf(0.9f, (float)m.adam_t);
            float b2c = 1.f - powf(0.999f, (float)m.adam_t);
            enqueue_grad(c, b, b.idx + off, batch, true, lr, b1c, b2c);
        }
    }
    CHK(hipGetLastError());
    if (n_epochs > 0) CHK(hipMemcpyAsync(last, b.loss, sizeof(float), hipMemcpyDeviceToHost, s));
    repack(c);
    CHK(hipStreamSynchronize(s));
    return MLH_OK;
}

struct DsCounts {
    int64_t faces = 0, trivial = 0, nontrivial = 0, nsel = 0;
};

// Builds the training set of the faces at F6 (slots as k_ds_prep) into b,
// whose partial sums hold batches of up to max(max_batch, 65536) samples.
// slots (optional): the input slot of each kept face, in the set's order.
static int build_set(mlh_ctx *c, const double *F6, const unsigned long long *counts,
                     int64_t scap, int64_t nslots, double gamma, double keep_trivial,
                     int max_batch, TrainBufs &b, DsCounts &dc,
                     std::vector<uint32_t> *slots = nullptr) {
    if (nslots <= 0 || nslots > INT32_MAX)
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "face count out of range");
    hipStream_t s = ctx_stream(c);
    DevBuf<float> Xall, Tall;
    DevBuf<uint32_t> hlo, hhi, khi, ord, k1, v1, scr;
    DevBuf<uint8_t> cls;
    DevBuf<unsigned long long> cnt;  // trivial, non-trivial, kept
    DevBuf<unsigned long long> blk;  // per-block counts
    CHK(dmalloc(&Xall.p, sizeof(float) * kMlpIn * nslots));
    CHK(dmalloc(&Tall.p, sizeof(float) * nslots));
    uint32_t **u[] = {&hlo.p, &hhi.p, &khi.p, &ord.p, &k1.p, &v1.p};
    for (uint32_t **q : u) CHK(dmalloc(q, sizeof(uint32_t) * nslots));
    CHK(dmalloc(&scr.p, sizeof(uint32_t) * radix_sort_scratch_words((int)nslots)));
    CHK(dmalloc(&cls.p, nslots));
    const int64_t nb = (nslots + 255) / 256;
    CHK(dmalloc(&cnt.p, sizeof(unsigned long long) * 3));
    CHK(dmalloc(&blk.p, sizeof(unsigned long long) * 2 * nb));
    hipLaunchKernelGGL(k_ds_prep, dim3(nb), dim3(256), 0, s, F6, counts, scap, nslots, gamma,
                       Xall.p, Tall.p, hlo.p, hhi.p, cls.p, blk.p);
    hipLaunchKernelGGL(k_sum_u64, dim3(1), dim3(256), 0, s, blk.p, (int)nb, cnt.p);
    hipLaunchKernelGGL(k_sum_u64, dim3(1), dim3(256), 0, s, blk.p + nb, (int)nb, cnt.p + 1);
    unsigned long long h[3] = {};
    CHK(hipMemcpyAsync(h, cnt.p, sizeof(unsigned long long) * 2, hipMemcpyDeviceToHost, s));
    CHK(hipStreamSynchronize(s));
    dc.trivial = (int64_t)h[0];
    dc.nontrivial = (int64_t)h[1];
    if (counts) {
        unsigned long long cn[kCollectStripes] = {};
        const int ns = (int)(nslots / scap);
        CHK(hipMemcpy(cn, counts, sizeof(unsigned long long) * ns, hipMemcpyDeviceToHost));
        dc.faces = 0;
        for (int q = 0; q < ns; ++q)
            dc.faces += cn[q] < (unsigned long long)scap ? (int64_t)cn[q] : scap;
    } else {
        dc.faces = nslots;
    }
    // trivial faces kept: keep = min(n_trivial, max(keep_trivial *
    // max(n_nontrivial, 1), 1024)), each with probability keep / n_trivial
    int64_t keep = (int64_t)(keep_trivial * (double)(dc.nontrivial > 1 ? dc.nontrivial : 1));
    if (keep < 1024) keep = 1024;
    if (keep > dc.trivial) keep = dc.trivial;
    uint64_t thr53 = 0;
    if (dc.trivial > 0)
        thr53 = keep >= dc.trivial ? (1ull << 53)
                                   : (uint64_t)((double)keep / (double)dc.trivial * 9007199254740992.0);
    hipLaunchKernelGGL(k_ds_select, dim3(nb), dim3(256), 0, s, cls.p, hlo.p, hhi.p, nslots, thr53,
                       khi.p, blk.p);
    hipLaunchKernelGGL(k_sum_u64, dim3(1), dim3(256), 0, s, blk.p, (int)nb, cnt.p + 2);
    // sort by (khi, hlo): the low word first, then stably by the high one
    hipLaunchKernelGGL(k_iota_u32, dim3(nb), dim3(256), 0, s, ord.p, nslots);
    radix_sort_pairs(hlo.p, ord.p, k1.p, v1.p, (int)nslots, 32, scr.p, s);
    hipLaunchKernelGGL(k_gather_u32, dim3(nb), dim3(256), 0, s, khi.p, ord.p, nslots, hhi.p);
    radix_sort_pairs(hhi.p, ord.p, k1.p, v1.p, (int)nslots, 32, scr.p, s);
    CHK(hipMemcpyAsync(h + 2, cnt.p + 2, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    CHK(hipStreamSynchronize(s));
    dc.nsel = (int64_t)h[2];
    int st = alloc_train(b, dc.nsel, max_batch > 65536 ? max_batch : 65536);
    if (st) return st;
    if (dc.nsel > 0)
        hipLaunchKernelGGL(k_ds_gather, dim3((dc.nsel + 255) / 256), dim3(256), 0, s, ord.p,
                           dc.nsel, Xall.p, Tall.p, b.X, b.T);
    CHK(hipGetLastError());
    if (slots) {
        slots->resize(dc.nsel);
        if (dc.nsel > 0)
            CHK(hipMemcpyAsync(slots->data(), ord.p, sizeof(uint32_t) * dc.nsel,
                               hipMemcpyDeviceToHost, s));
    }
    CHK(hipStreamSynchronize(s));  // before the temporaries are freed
    return MLH_OK;
}

// Loss over the first min(n, 65536) faces of the set, Adam epochs over the
// training faces, and the validation MSE of the BF16 network against the
// standard guess (target 0) on the validation faces.  Fewer than
// kMinTrainFaces faces in the set: nothing is trained, train_faces = 0.
static int train_on_set(mlh_ctx *c, TrainBufs &b, const DsCounts &dc, int n_epochs, int batch,
                        float lr, uint64_t seed, mlh_train_info *info) {
    MlpState &m = ctx_mlp(c);
    hipStream_t s = ctx_stream(c);
    *info = mlh_train_info{};
    info->faces = dc.faces;
    info->valid = dc.trivial + dc.nontrivial;
    info->nontrivial = dc.nontrivial;
    info->trivial_kept = dc.nsel - dc.nontrivial;
    info->epochs = n_epochs;
    const int64_t n = dc.nsel;
    if (n < kMinTrainFaces) return MLH_OK;
    const int64_t nval = n / 10 > 64 ? n / 10 : 64, ntrain = n - nval;
    if (batch > ntrain) batch = (int)ntrain;
    const int ninit = (int)(n < 65536 ? n : 65536);
    enqueue_grad(c, b, nullptr, ninit);
    float loss0 = 0.f, last = 0.f;
    CHK(hipMemcpyAsync(&loss0, b.loss, sizeof(float), hipMemcpyDeviceToHost, s));
    int st = train_epochs(c, b, nval, ntrain, n_epochs, batch, lr, seed, &last);
    if (st) return st;
    DevBuf<float> Y;
    CHK(dmalloc(&Y.p, sizeof(float) * nval));
    hipLaunchKernelGGL(k_mlp_infer_bf16, dim3((nval + 255) / 256), dim3(256), 0, s, m.frags, b.X,
                       nval, Y.p);
    std::vector<float> y(nval), t(nval);
    CHK(hipMemcpyAsync(y.data(), Y.p, sizeof(float) * nval, hipMemcpyDeviceToHost, s));
    CHK(hipMemcpyAsync(t.data(), b.T, sizeof(float) * nval, hipMemcpyDeviceToHost, s));
    CHK(hipStreamSynchronize(s));
    double e_net = 0., e_std = 0.;
    for (int64_t k = 0; k < nval; ++k) {
        const double d = (double)y[k] - (double)t[k];
        e_net += d * d;
        e_std += (double)t[k] * (double)t[k];
    }
    info->train_faces = ntrain;
    info->val_faces = nval;
    info->batch = batch;
    info->loss_init = loss0;
    info->loss_final = last;
    info->val_mse_log_pstar = e_net / (double)nval;
    info->val_mse_standard_guess = e_std / (double)nval;
    return MLH_OK;
}

int mlh_mlp_loss_grad(mlh_ctx *c, int64_t n, const float *feats, const float *target,
                      float *loss, float *grad) {
    if (!c || n <= 0 || n > (1 << 20) || !feats || !target)
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "bad arguments");
    MlpState &m = ctx_mlp(c);
    if (!m.params) return ctx_fail(MLH_ERR_NOT_TRAINED, "no network parameters");
    TrainBufs b;
    int st = alloc_train(b, n, (int)n);
    if (st) return st;
    hipStream_t s = ctx_stream(c);
    CHK(hipMemcpyAsync(b.X, feats, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice, s));
    CHK(hipMemcpyAsync(b.T, target, sizeof(float) * n, hipMemcpyHostToDevice, s));
    enqueue_grad(c, b, nullptr, (int)n);
    if (loss) CHK(hipMemcpyAsync(loss, b.loss, sizeof(float), hipMemcpyDeviceToHost, s));
    if (grad) CHK(hipMemcpyAsync(grad, m.grad, sizeof(float) * kMlpN, hipMemcpyDeviceToHost, s));
    CHK(hipStreamSynchronize(s));
    return MLH_OK;
}

int mlh_mlp_train(mlh_ctx *c, int64_t n, const float *feats, const float *target,
                  int32_t n_epochs, int32_t batch, float lr, uint64_t seed, float *final_loss) {
    if (!c || n <= 0 || n > INT32_MAX || !feats || !target || batch <= 0 || n_epochs < 0)
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "bad arguments");
    int st = ensure_mlp(c);
    if (st) return st;
    hipStream_t s = ctx_stream(c);
    if (batch > n) batch = (int)n;
    TrainBufs b;
    if ((st = alloc_train(b, n, batch))) return st;
    CHK(hipMemcpyAsync(b.X, feats, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice, s));
    CHK(hipMemcpyAsync(b.T, target, sizeof(float) * n, hipMemcpyHostToDevice, s));
    float last = 0.f;
    if ((st = train_epochs(c, b, 0, n, n_epochs, batch, lr, seed, &last))) return st;
    if (final_loss) *final_loss = last;
    return MLH_OK;
}

int mlh_mlp_train_faces(mlh_ctx *c, int64_t n, const double *states6, double gamma,
                        int32_t n_epochs, int32_t batch, float lr, uint64_t seed,
                        double keep_trivial, mlh_train_info *info) {
    if (!c || n <= 0 || n > INT32_MAX || !states6 || batch <= 0 || n_epochs < 0 || !info ||
        !(keep_trivial >= 0.))
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "bad arguments");
    MlpState &m = ctx_mlp(c);
    if (!m.params) return ctx_fail(MLH_ERR_NOT_TRAINED, "no network parameters to train from");
    hipStream_t s = ctx_stream(c);
    DevBuf<double> F6;
    CHK(dmalloc(&F6.p, sizeof(double) * 6 * n));
    CHK(hipMemcpyAsync(F6.p, states6, sizeof(double) * 6 * n, hipMemcpyHostToDevice, s));
    TrainBufs b;
    DsCounts dc;
    int st = build_set(c, F6.p, nullptr, n, n, gamma, keep_trivial, batch, b, dc);
    if (st) return st;
    return train_on_set(c, b, dc, n_epochs, batch, lr, seed, info);
}

int mlh_mlp_train_recorded(mlh_ctx *c, int32_t n_epochs, int32_t batch, float lr, uint64_t seed,
                           double keep_trivial, mlh_train_info *info) {
    if (!c || batch <= 0 || n_epochs < 0 || !info || !(keep_trivial >= 0.))
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "bad arguments");
    MlpState &m = ctx_mlp(c);
    if (!m.params) return ctx_fail(MLH_ERR_NOT_TRAINED, "no network parameters to train from");
    if (!m.collect || !m.collect_n || m.collect_stripes <= 0)
        return ctx_fail(MLH_ERR_CALL_ORDER, "no face record (mlh_set_face_collection)");
    hipStream_t s = ctx_stream(c);
    TrainBufs b;
    DsCounts dc;
    const int64_t nslots = (int64_t)m.collect_stripes * m.collect_scap;
    int st = build_set(c, m.collect, m.collect_n, m.collect_scap, nslots, ctx_gamma(c),
                       keep_trivial, batch, b, dc);
    // the record restarts whether or not the set was large enough
    CHK(hipMemsetAsync(m.collect_n, 0, sizeof(unsigned long long) * kCollectStripes, s));
    CHK(hipStreamSynchronize(s));
    if (st) return st;
    return train_on_set(c, b, dc, n_epochs, batch, lr, seed, info);
}

int mlh_mlp_dataset(mlh_ctx *c, int64_t n, const double *states6, double gamma,
                    double keep_trivial, int64_t cap, float *feats, float *target,
                    int64_t *rows, int64_t *n_out) {
    if (!c || n < 0 || n > INT32_MAX || (n > 0 && !states6) || cap < 0 || !n_out ||
        (cap > 0 && (!feats || !target)) || !(keep_trivial >= 0.))
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "bad arguments");
    *n_out = 0;
    MlpState &m = ctx_mlp(c);
    hipStream_t s = ctx_stream(c);
    DevBuf<double> F6;
    const double *src = nullptr;
    const unsigned long long *counts = nullptr;
    int64_t scap = n, nslots = n;
    if (states6) {
        CHK(dmalloc(&F6.p, sizeof(double) * 6 * n));
        CHK(hipMemcpyAsync(F6.p, states6, sizeof(double) * 6 * n, hipMemcpyHostToDevice, s));
        src = F6.p;
    } else {
        if (!m.collect || !m.collect_n || m.collect_stripes <= 0)
            return ctx_fail(MLH_ERR_CALL_ORDER, "no face record (mlh_set_face_collection)");
        src = m.collect;
        counts = m.collect_n;
        scap = m.collect_scap;
        nslots = (int64_t)m.collect_stripes * scap;
    }
    if (nslots == 0) return MLH_OK;
    TrainBufs b;
    DsCounts dc;
    std::vector<uint32_t> slots;
    int st = build_set(c, src, counts, scap, nslots, gamma, keep_trivial, 1, b, dc, &slots);
    if (st) return st;
    const int64_t k = dc.nsel < cap ? dc.nsel : cap;
    if (k > 0) {
        CHK(hipMemcpyAsync(feats, b.X, sizeof(float) * kMlpIn * k, hipMemcpyDeviceToHost, s));
        CHK(hipMemcpyAsync(target, b.T, sizeof(float) * k, hipMemcpyDeviceToHost, s));
    }
    CHK(hipStreamSynchronize(s));
    if (rows && k > 0) {
        // slot -> input row; the record's rows in mlh_get_collected_faces'
        // order (each stripe's filled prefix, stripe after stripe)
        std::vector<int64_t> first(1, 0);
        if (counts) {
            const int ns = (int)(nslots / scap);
            std::vector<unsigned long long> cn(ns);
            CHK(hipMemcpy(cn.data(), counts, sizeof(unsigned long long) * ns,
                          hipMemcpyDeviceToHost));
            first.resize(ns);
            int64_t o = 0;
            for (int q = 0; q < ns; ++q) {
                first[q] = o;
                o += cn[q] < (unsigned long long)scap ? (int64_t)cn[q] : scap;
            }
        }
        for (int64_t q = 0; q < k; ++q) {
            const int64_t sl = slots[q];
            rows[q] = counts ? first[sl / scap] + sl % scap : sl;
        }
    }
    *n_out = dc.nsel;
    return MLH_OK;
}

int mlh_mlp_infer(mlh_ctx *c, int64_t n, const float *feats, float *out, int32_t bf16) {
    if (!c || n <= 0 || !feats || !out) return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "bad arguments");
    MlpState &m = ctx_mlp(c);
    if (!m.params) return ctx_fail(MLH_ERR_NOT_TRAINED, "no network parameters");
    hipStream_t s = ctx_stream(c);
    float *X = nullptr, *Y = nullptr;
    CHK(dmalloc(&X, sizeof(float) * n * kMlpIn));
    CHK(dmalloc(&Y, sizeof(float) * n));
    hipMemcpyAsync(X, feats, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice, s);
    int64_t nb = (n + 255) / 256;
    if (bf16)
        hipLaunchKernelGGL(k_mlp_infer_bf16, dim3(nb), dim3(256), 0, s, m.frags, X, n, Y);
    else
        hipLaunchKernelGGL(k_mlp_infer_f32, dim3(nb), dim3(256), 0, s, m.params, X, n, Y);
    hipMemcpyAsync(out, Y, sizeof(float) * n, hipMemcpyDeviceToHost, s);
    hipError_t e = hipStreamSynchronize(s);
    hipFree(X);
    hipFree(Y);
    if (e != hipSuccess) return ctx_fail(MLH_ERR_DEVICE, hipGetErrorString(e));
    return MLH_OK;
}

int mlh_riemann_exact_batch(mlh_ctx *c, int64_t n, double gamma, const double *rhoL,
                            const double *uL, const double *PL, const double *rhoR,
                            const double *uR, const double *PR, double *rho_s, double *u_s,
                            double *P_s, int32_t *flag, int32_t *iters, int32_t precision,
                            int32_t seed, double tol) {
    if (n < 0 || !rhoL || !uL || !PL || !rhoR || !uR || !PR || !rho_s || !u_s || !P_s)
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    if (precision != MLH_FP64 && precision != MLH_FP32)
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "bad precision");
    const MlpFrags *F = nullptr;
    if (seed == MLH_SEED_LEARNED) {
        if (!c || !ctx_mlp(c).params) return ctx_fail(MLH_ERR_NOT_TRAINED, "no network parameters");
        F = ctx_mlp(c).frags;
    }
    if (n == 0) return MLH_OK;
    hipStream_t s = c ? ctx_stream(c) : nullptr;
    double *buf = nullptr;
    int *ib = nullptr;
    unsigned long long *fb = nullptr;
    CHK(dmalloc(&buf, sizeof(double) * n * 9));
    CHK(dmalloc(&ib, sizeof(int) * n * 2));
    CHK(dmalloc(&fb, sizeof(unsigned long long)));
    size_t b = sizeof(double) * n;
    const double *in[6] = {rhoL, uL, PL, rhoR, uR, PR};
    for (int k = 0; k < 6; ++k) hipMemcpyAsync(buf + k * n, in[k], b, hipMemcpyHostToDevice, s);
    hipMemsetAsync(fb, 0, sizeof(unsigned long long), s);
    int max_iter = 50;
    hipLaunchKernelGGL(k_riemann_batch, dim3((n + 255) / 256), dim3(256), 0, s, n, gamma, buf,
                       buf + n, buf + 2 * n, buf + 3 * n, buf + 4 * n, buf + 5 * n, buf + 6 * n,
                       buf + 7 * n, buf + 8 * n, ib, ib + n, precision, F, tol, max_iter, 3, fb,
                       nullptr, nullptr);
    hipMemcpyAsync(rho_s, buf + 6 * n, b, hipMemcpyDeviceToHost, s);
    hipMemcpyAsync(u_s, buf + 7 * n, b, hipMemcpyDeviceToHost, s);
    hipMemcpyAsync(P_s, buf + 8 * n, b, hipMemcpyDeviceToHost, s);
    if (flag) hipMemcpyAsync(flag, ib, sizeof(int) * n, hipMemcpyDeviceToHost, s);
    if (iters) hipMemcpyAsync(iters, ib + n, sizeof(int) * n, hipMemcpyDeviceToHost, s);
    hipError_t e = hipStreamSynchronize(s);
    hipFree(buf);
    hipFree(ib);
    hipFree(fb);
    if (e != hipSuccess) return ctx_fail(MLH_ERR_DEVICE, hipGetErrorString(e));
    return MLH_OK;
}

int mlh_device_alloc(int64_t bytes, void **ptr) {
    if (!ptr || bytes < 0) return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "bad arguments");
    *ptr = nullptr;
    if (bytes == 0) return MLH_OK;
    CHK(hipMalloc(ptr, (size_t)bytes));
    return MLH_OK;
}

int mlh_device_free(void *ptr) {
    if (ptr) CHK(hipFree(ptr));
    return MLH_OK;
}

int mlh_device_copy(void *dst, const void *src, int64_t bytes, int32_t kind) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src)) || kind < 1 || kind > 3)
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "bad arguments");
    if (bytes == 0) return MLH_OK;
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                          : kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    CHK(hipMemcpy(dst, src, (size_t)bytes, k));
    return MLH_OK;
}

int mlh_riemann_exact_batch_device(mlh_ctx *c, int64_t n, double gamma, const double *states,
                                   double *star, int32_t *flag, int32_t *iters, int32_t precision,
                                   int32_t seed, double tol, int64_t *fallbacks,
                                   float *kernel_ms) {
    if (n < 0 || (n > 0 && (!states || !star)))
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    if (precision != MLH_FP64 && precision != MLH_FP32)
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "bad precision");
    const MlpFrags *F = nullptr;
    if (seed == MLH_SEED_LEARNED) {
        if (!c || !ctx_mlp(c).params) return ctx_fail(MLH_ERR_NOT_TRAINED, "no network parameters");
        F = ctx_mlp(c).frags;
    }
    if (fallbacks) *fallbacks = 0;
    if (kernel_ms) *kernel_ms = 0.f;
    if (n == 0) return MLH_OK;
    hipStream_t s = c ? ctx_stream(c) : nullptr;
    unsigned long long *fb = nullptr;
    CHK(dmalloc(&fb, sizeof(unsigned long long)));
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipMemsetAsync(fb, 0, sizeof(unsigned long long), s);
    hipEventRecord(e0, s);
    hipLaunchKernelGGL(k_riemann_batch, dim3((n + 255) / 256), dim3(256), 0, s, n, gamma, states,
                       states + n, states + 2 * n, states + 3 * n, states + 4 * n, states + 5 * n,
                       star, star + n, star + 2 * n, flag, iters, precision, F, tol, 50, 3, fb,
                       nullptr, nullptr);
    hipEventRecord(e1, s);
    unsigned long long h_fb = 0;
    hipMemcpyAsync(&h_fb, fb, sizeof(h_fb), hipMemcpyDeviceToHost, s);
    hipError_t e = hipStreamSynchronize(s);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(fb);
    if (e != hipSuccess) return ctx_fail(MLH_ERR_DEVICE, hipGetErrorString(e));
    if (fallbacks) *fallbacks = (int64_t)h_fb;
    if (kernel_ms) *kernel_ms = ms;
    return MLH_OK;
}

}  // extern "C"

// ---- the allocation-free batch solve ---------------------------------------
// Scratch of mlh_riemann_exact_batch_device_noalloc: one fallback count per
// block of 256 faces, then their total.  A static device array: each GPU
// the library runs on holds its own copy until the process ends.
constexpr int kNoAllocBlocks = MLH_NOALLOC_MAX_FACES / 256;
static __device__ __align__(256) unsigned long long g_batch_scratch[kNoAllocBlocks + 32];

// The array's device address on each GPU, looked up once per device (with
// that device current) on the entry point's first call there.  In host code
// the array's name is not a device address.
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned long long *g_scratch_ptr[kMaxDevices];

static int batch_scratch(unsigned long long **out) {
    int dev = -1;
    CHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return ctx_fail(MLH_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        CHK(hipGetSymbolAddress(&p, HIP_SYMBOL(g_batch_scratch)));
        if (!p) return ctx_fail(MLH_ERR_DEVICE, "no device address for the batch scratch");
        g_scratch_ptr[dev] = (unsigned long long *)p;
    }
    *out = g_scratch_ptr[dev];
    return MLH_OK;
}

extern "C" {

int mlh_riemann_exact_batch_device_noalloc(mlh_ctx *c, int64_t n, double gamma,
                                           const double *states, double *star, int32_t *flag,
                                           int32_t *iters, int32_t precision, int32_t seed,
                                           double tol, int64_t *fallbacks) {
    if (n > MLH_NOALLOC_MAX_FACES)
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT,
                        "n exceeds MLH_NOALLOC_MAX_FACES (the static scratch holds one count per "
                        "256 faces for at most 2^26 faces); split the batch");
    if (n < 0 || (n > 0 && (!states || !star)))
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    if (precision != MLH_FP64 && precision != MLH_FP32)
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "bad precision");
    const MlpFrags *F = nullptr;
    if (seed == MLH_SEED_LEARNED) {
        if (!c || !ctx_mlp(c).params) return ctx_fail(MLH_ERR_NOT_TRAINED, "no network parameters");
        F = ctx_mlp(c).frags;
    }
    if (fallbacks) *fallbacks = 0;
    if (n == 0) return MLH_OK;
    unsigned long long *scr = nullptr;
    int st = batch_scratch(&scr);
    if (st) return st;
    hipStream_t s = c ? ctx_stream(c) : nullptr;
    const int nblocks = (int)((n + 255) / 256);  // <= kNoAllocBlocks
    hipLaunchKernelGGL(k_riemann_batch, dim3(nblocks), dim3(256), 0, s, n, gamma, states,
                       states + n, states + 2 * n, states + 3 * n, states + 4 * n, states + 5 * n,
                       star, star + n, star + 2 * n, flag, iters, precision, F, tol, 50, 3,
                       nullptr, nullptr, scr);
    hipLaunchKernelGGL(k_sum_u64, dim3(1), dim3(256), 0, s, scr, nblocks, scr + kNoAllocBlocks);
    unsigned long long total = 0;
    CHK(hipMemcpyAsync(&total, scr + kNoAllocBlocks, sizeof(total), hipMemcpyDeviceToHost, s));
    CHK(hipStreamSynchronize(s));
    CHK(hipGetLastError());
    if (fallbacks) *fallbacks = (int64_t)total;
    return MLH_OK;
}

int mlh_riemann_pstar_batch(int64_t n, double gamma, const double *rhoL, const double *uL,
                            const double *PL, const double *rhoR, const double *uR,
                            const double *PR, double tol, double *p_star, int32_t *iters) {
    if (n < 0 || !rhoL || !uL || !PL || !rhoR || !uR || !PR || !p_star)
        return ctx_fail(MLH_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) return MLH_OK;
    double *buf = nullptr;
    int *ib = nullptr;
    unsigned long long *fb = nullptr;
    CHK(dmalloc(&buf, sizeof(double) * n * 7));
    CHK(dmalloc(&ib, sizeof(int) * n));
    CHK(dmalloc(&fb, sizeof(unsigned long long)));
    size_t b = sizeof(double) * n;
    const double *in[6] = {rhoL, uL, PL, rhoR, uR, PR};
    for (int k = 0; k < 6; ++k) hipMemcpy(buf + k * n, in[k], b, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_riemann_batch, dim3((n + 255) / 256), dim3(256), 0, 0, n, gamma, buf,
                       buf + n, buf + 2 * n, buf + 3 * n, buf + 4 * n, buf + 5 * n, buf, buf, buf,
                       nullptr, ib, MLH_FP64, nullptr, tol, 50, 3, fb, buf + 6 * n, nullptr);
    hipMemcpy(p_star, buf + 6 * n, b, hipMemcpyDeviceToHost);
    if (iters) hipMemcpy(iters, ib, sizeof(int) * n, hipMemcpyDeviceToHost);
    hipError_t e = hipDeviceSynchronize();
    hipFree(buf);
    hipFree(ib);
    hipFree(fb);
    if (e != hipSuccess) return ctx_fail(MLH_ERR_DEVICE, hipGetErrorString(e));
    return MLH_OK;
}

}  // extern "C"
