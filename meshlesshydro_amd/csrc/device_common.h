// device_common.h -- layouts and small device helpers shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

namespace mlh {

// Packed neighbour entry: particle index in bits 0..26, periodic image code
// s = (sx+1) + 3(sy+1) + 9(sz+1) in bits 27..31 (13 = no image).
constexpr int kIdxBits = 27;
constexpr uint32_t kIdxMask = (1u << kIdxBits) - 1u;
constexpr uint32_t kCodeReal = 13u;

__host__ __device__ inline uint32_t nb_index(uint32_t e) { return e & kIdxMask; }
__host__ __device__ inline uint32_t nb_code(uint32_t e) { return e >> kIdxBits; }
__host__ __device__ inline int code_shift(uint32_t code, int a) {
    int c = (int)code;
    if (a == 1) c /= 3;
    if (a == 2) c /= 9;
    return c % 3 - 1;
}
// bit a set when axis a carries an image shift
__host__ __device__ inline int code_mask(uint32_t code) {
    int m = 0;
    for (int a = 0; a < 3; ++a)
        if (code_shift(code, a) != 0) m |= 1 << a;
    return m;
}
// Neighbour rows, chunked: entry n of work position t sits at word
// ((n >> 3) * N + t) * 8 + (n & 7).  The eight entries 8c..8c+7 of one
// position are 32 contiguous bytes (one sector), so the search fills a
// position's sector before moving on, and a wavefront reads eight entries
// per lane with two 16-byte loads at consecutive sectors.
constexpr int kNbChunk = 8;
__host__ __device__ inline size_t nb_slot(int N, int t, int n) {
    return ((size_t)(n >> 3) * (size_t)N + (size_t)t) * kNbChunk + (size_t)(n & 7);
}
// words of a row array for S entries per position
__host__ __device__ inline size_t nb_words(int N, int S) {
    return (size_t)((S + kNbChunk - 1) / kNbChunk) * kNbChunk * (size_t)N;
}
__device__ inline uint32_t nb_get(const uint32_t *nbr, int N, int t, int n) {
    return nbr[nb_slot(N, t, n)];
}
// entries n..n+7, n a multiple of 8
__device__ inline void nb_get8(const uint32_t *nbr, int N, int t, int n, uint32_t *e) {
    const uint4 *p = (const uint4 *)(nbr + nb_slot(N, t, n));
    const uint4 a = p[0], b = p[1];
    e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = a.w;
    e[4] = b.x; e[5] = b.y; e[6] = b.z; e[7] = b.w;
}
// entries n..n+3, n a multiple of 4
__device__ inline void nb_get4(const uint32_t *nbr, int N, int t, int n, uint32_t *e) {
    const uint4 a = *(const uint4 *)(nbr + nb_slot(N, t, n));
    e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = a.w;
}

// ghost-index order of the reference: parent (original index) ascending,
// then the image subsets of one parent in descending axis-mask order
__host__ __device__ inline uint64_t ghost_key(uint32_t parent, uint32_t code) {
    return (uint64_t)parent * 8u + (uint64_t)(7 - code_mask(code));
}

// Search geometry (Domain::createGrid, Domain.cpp:9-54).  Cells are the
// reference's; their keys (the sort key and the index of the cell-range
// tables) follow a Morton curve when mbits > 0 (mbits bits per axis, nkeys =
// 2^(D mbits) table entries), so that consecutive particles of the cell
// order, and the blocks in flight on one XCD, are compact bricks of space.
// mbits = 0: linear keys x + nx (y + ny z), nkeys = ncells.
struct Geom {
    double bmin[3], bmax[3], cs[3];
    int nc[3];
    int ncells;
    int mbits, nkeys;
    double h, hsq;        // kernelSize and kernelSize*kernelSize
    double sigma;         // cubic-spline normalisation for h/2
    int periodic;
};

__host__ __device__ inline uint32_t spread2(uint32_t x) {  // 16 bits -> even bits
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    return (x | (x << 1)) & 0x55555555u;
}
__host__ __device__ inline uint32_t compact2(uint32_t x) {
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0f0f0f0fu;
    x = (x | (x >> 4)) & 0x00ff00ffu;
    return (x | (x >> 8)) & 0xffffu;
}
__host__ __device__ inline uint32_t spread3(uint32_t x) {  // 10 bits -> every third bit
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    return (x | (x << 2)) & 0x09249249u;
}
__host__ __device__ inline uint32_t compact3(uint32_t x) {
    x &= 0x09249249u;
    x = (x | (x >> 2)) & 0x030c30c3u;
    x = (x | (x >> 4)) & 0x0300f00fu;
    x = (x | (x >> 8)) & 0x030000ffu;
    return (x | (x >> 16)) & 0x3ffu;
}
// key of cell (cx, cy, cz)
template <int D>
__host__ __device__ inline uint32_t cell_key(const Geom &g, int cx, int cy, int cz) {
    if (g.mbits) {
        if (D == 2) return spread2((uint32_t)cx) | (spread2((uint32_t)cy) << 1);
        return spread3((uint32_t)cx) | (spread3((uint32_t)cy) << 1) | (spread3((uint32_t)cz) << 2);
    }
    return (uint32_t)(cx + cy * g.nc[0] + (D == 3 ? cz * g.nc[0] * g.nc[1] : 0));
}
template <int D>
__host__ __device__ inline void cell_coords(const Geom &g, uint32_t key, int *c) {
    if (g.mbits) {
        if (D == 2) {
            c[0] = (int)compact2(key);
            c[1] = (int)compact2(key >> 1);
            c[2] = 0;
        } else {
            c[0] = (int)compact3(key);
            c[1] = (int)compact3(key >> 1);
            c[2] = (int)compact3(key >> 2);
        }
        return;
    }
    c[0] = (int)(key % (uint32_t)g.nc[0]);
    c[1] = (int)((key / (uint32_t)g.nc[0]) % (uint32_t)g.nc[1]);
    c[2] = D == 3 ? (int)(key / (uint32_t)(g.nc[0] * g.nc[1])) : 0;
}

// Periodic image coordinate (createGhostParticles, Particles.cpp:2124-2138):
// s = +1 -> max + (x - min), s = -1 -> min - (max - x).
__device__ inline double image_coord(double x, int s, double mn, double mx) {
    return s == 0 ? x : (s > 0 ? mx + (x - mn) : mn - (mx - x));
}

// The cube of the kernel's outer branch, correctly rounded: a double-double
// cube rounded once (error of the unrounded value ~2^-106 relative).  glibc's
// pow(t, 3.) agrees except where its own result (within 0.52 ulp) is not the
// correctly rounded one: 2 of the 117 distinct arguments of a 2D lattice with
// h = 8 dx, none of a lattice with h = 4 dx; there the densities differ by
// one ulp.
__device__ inline double cube_cr(double t) {
    double hi = t * t;
    double lo = __fma_rn(t, t, -hi);
    double p = hi * t;
    double e = __fma_rn(hi, t, -p) + lo * t;
    return p + e;
}
__device__ inline float cube_cr(float t) { return t * t * t; }

// Kernel::cubicSpline (Particles.cpp:7-24); sigma precomputed on the host
// with the reference's expression.  The rounding does not depend on the
// translation unit: FP64 keeps the reference's separate roundings, FP32
// fuses a * b + c within an expression (as every FP32 kernel does).
template <typename T>
__device__ inline T kernel_w(T r, T h2, T sigma) {
    if constexpr (sizeof(T) == 4) {
#pragma clang fp contract(on)
        T q = r / h2;
        if (T(0) <= q && q <= T(1)) return sigma * (T(1) - T(3) / T(2) * q * q * (T(1) - q / T(2)));
        if (T(1) < q && q < T(2)) return sigma / T(4) * cube_cr(T(2) - q);
        return T(0);
    } else {
#pragma clang fp contract(off)
        T q = r / h2;
        if (T(0) <= q && q <= T(1)) return sigma * (T(1) - T(3) / T(2) * q * q * (T(1) - q / T(2)));
        if (T(1) < q && q < T(2)) return sigma / T(4) * cube_cr(T(2) - q);
        return T(0);
    }
}

// Per-particle primitive record (AoS, 8 values of T).
constexpr int kPrimRho = 0, kPrimV = 1, kPrimP = 4, kPrimC = 5, kPrimOmega = 6, kPrimN = 8;
template <int D>
__host__ __device__ constexpr int field_slot(int f) { return f == D + 1 ? kPrimP : f; }

// compDensity's rho = m omega and compPressure (Particles.cpp:1151-1184,
// 1272-1288), with c = sqrt(gamma P / rho) for the time step (:1453), into
// the particle's primitive record pr.  FP64 throughout (no a * b + c forms,
// so the rounding is the same in every translation unit).
template <int D, typename T>
__device__ inline void write_prim(T *pr, T om, double m, double u, double v0, double v1,
                                  double v2, double gamma) {
    double rho = m * (double)om;
    double P = (gamma - 1.) * rho * u;
    double c = sqrt(gamma * P / rho);
    T rec[kPrimN];
    rec[kPrimRho] = T(rho);
    rec[kPrimV + 0] = T(v0);
    rec[kPrimV + 1] = T(v1);
    rec[kPrimV + 2] = D == 3 ? T(v2) : T(0);
    rec[kPrimP] = T(P);
    rec[kPrimC] = T(c);
    rec[kPrimOmega] = om;
    rec[7] = T(0);
    for (int k = 0; k < kPrimN; ++k) pr[k] = rec[k];
}

// Face-input record: (D+2)*D limited gradients then D*D B = E^-1.
template <int D>
struct FRec {
    static constexpr int NF = D + 2;
    static constexpr int G = 0;
    static constexpr int B = NF * D;
    static constexpr int N = NF * D + D * D;   // 12 (2D) / 24 (3D)
};

__device__ inline int lane_id() { return __lane_id(); }

// LDS written by some lanes of a wavefront, then read by others: a
// wavefront-scope release / barrier / acquire
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// XCD-aware block index.  Workgroups are dispatched round-robin over the 8
// XCDs of an MI355X (workgroup w runs on XCD w % 8); this bijection on
// [0, gridDim.x) gives each XCD one contiguous range of logical blocks, so
// that neighbouring particles, whose gathers overlap, share one L2.
constexpr int kNumXcd = 8;
__device__ inline int xcd_block() {
    const int B = (int)gridDim.x, w = (int)blockIdx.x;
    const int q = B / kNumXcd, r = B % kNumXcd, x = w % kNumXcd, k = w / kNumXcd;
    return x * q + (x < r ? x : r) + k;
}

// abs() of the reference's slope / pairwise limiter (see mlh_limiter_abs).
template <typename T>
__device__ inline T lim_abs(T x, int mode) {
    if (mode == 1) return x < T(0) ? -x : x;
    double xd = (double)x;
    int32_t t;
    if (xd != xd || xd >= 2147483648.0 || xd <= -2147483649.0) t = INT32_MIN;
    else t = (int32_t)xd;
    if (t != INT32_MIN && t < 0) t = -t;
    return (T)t;
}

}  // namespace mlh
