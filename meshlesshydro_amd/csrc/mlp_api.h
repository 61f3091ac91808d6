// mlp_api.h -- device-side state of the learned p* seed network.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlh {

// MFMA fragment layout of the network (filled by mlp_pack from the FP32
// parameters): per lane of a wavefront, the A operands of
// v_mfma_f32_16x16x32_bf16 for W1 and W2, and the bias / output weights in
// the accumulator layout.
struct MlpFrags {
    uint16_t w1[4][64][8];      // [M tile][lane][k slot]   (bf16 bits)
    uint16_t w2[4][2][64][8];   // [M tile][k step][lane][k slot]
    float b1[4][64][4];         // [M tile][lane][acc reg]
    float b2[4][64][4];
    float w3[4][64][4];
    float b3;
};

struct MlpState {
    float *params = nullptr;    // FP32 parameters (device)
    MlpFrags *frags = nullptr;  // packed BF16 fragments (device)
    float *adam_m = nullptr, *adam_v = nullptr, *grad = nullptr;
    int64_t adam_t = 0;
    // face collection for (online) training: the record is split into
    // collect_stripes slices of collect_scap rows, each filled through its
    // own counter (a wavefront writes to slice wave % stripes)
    double *collect = nullptr;          // [cap][6] rhoL,uL,PL,rhoR,uR,PR
    unsigned long long *collect_n = nullptr;  // [kCollectStripes] fill counters
    int64_t collect_cap = 0;
    int64_t collect_scap = 0;
    int collect_stripes = 0;
    int collect_stride = 0;
};

// most stripes of the face record
constexpr int kCollectStripes = 64;

void mlp_free(MlpState &m);

}  // namespace mlh

struct mlh_ctx;
mlh::MlpState &ctx_mlp(mlh_ctx *c);
hipStream_t ctx_stream(mlh_ctx *c);
double ctx_gamma(mlh_ctx *c);  // the context's adiabatic index
int ctx_fail(int code, const char *msg);
