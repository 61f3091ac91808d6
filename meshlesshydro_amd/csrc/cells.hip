// cells.hip -- GPU cell list and neighbour lists.
//
// Replaces Particles::assignParticlesAndCells (Particles.cpp:270-322),
// createGhostParticles (:2113-2191), gridNNS (:324-365) and ghostNNS
// (:2237-2260):
//   1. cell key per particle with the reference's floor((x-min)/cellSize)
//      (clamped when it equals the cell count), positions packed as double4;
//   2. stable LSD radix sort of (cell key, particle index), 8-bit digits:
//      LDS histogram per 2048-key tile, exclusive scan of the tile
//      histograms, then a scatter that ranks equal digits inside each
//      wavefront with eight __ballot masks (match + popcount below the
//      lane), and across the block's waves through LDS counts;
//   3. per-cell [start, end) ranges of the sorted order (ascending particle
//      index inside a cell, as the reference's push_back lists);
//   4. positions gathered into cell order as double4 (x, y, z, original
//      index); cell keys follow a Morton curve (device_common.h), so that a
//      block of the cell order is a compact brick of cells;
//   5. one thread per particle, in work order: the 3^D stencil of
//      non-periodic cells (x outer ... z inner, Domain.cpp:83-118) with the
//      strict test d^2 < h^2 in the reference's operation order, then the
//      periodic images: every candidate in the window of virtual cells
//      around the particle that wraps in at least one axis, kept when the
//      image exists under the reference's creation test and d^2 < h^2, and
//      inserted in ghost-index order (original parent index, then axis
//      subset).  A periodic image within h of x_i is less than one cell size
//      away per axis, so the window is complete: the result equals the
//      reference's exhaustive ghost loop.
// Neighbour entries hold work positions (cell-order positions, or original
// indices when the work order is the caller's); rows are stored transposed
// (entry n of work position t at nbr[n*N + t]) so that the writes here and
// the reads of the per-particle kernels are coalesced.  Inside one cell the
// cell order is ascending original index (stable sort of the identity), so
// the reference's neighbour order is kept.
#include <hip/hip_runtime.h>
#include <climits>
#include <type_traits>
#include "device_common.h"
#include "kernels.h"

#pragma clang fp contract(off)

namespace mlh {

__global__ void k_fill_i32(int *p, int v, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

template <int D>
__global__ void k_assign_cells(Geom g, const double *__restrict__ x0,
                               const double *__restrict__ x1,
                               const double *__restrict__ x2, int N,
                               uint32_t *__restrict__ key, uint32_t *__restrict__ idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double xs[3] = {x0[i], x1[i], D == 3 ? x2[i] : 0.0};
    int f[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) {
        f[a] = (int)floor((xs[a] - g.bmin[a]) / g.cs[a]);
        if (f[a] == g.nc[a]) f[a] -= 1;
        if (f[a] < 0) f[a] = 0;
        if (f[a] > g.nc[a] - 1) f[a] = g.nc[a] - 1;
    }
    key[i] = cell_key<D>(g, f[0], f[1], f[2]);
    idx[i] = (uint32_t)i;
}

// positions in cell order; w = original index (exact in a double); inv:
// original index -> cell-order position
__global__ void k_gather_pos(const double *__restrict__ x0, const double *__restrict__ x1,
                             const double *__restrict__ x2, int D, int N,
                             const uint32_t *__restrict__ sidx, double4 *__restrict__ pos,
                             uint32_t *__restrict__ inv) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    uint32_t i = sidx[t];
    pos[t] = make_double4(x0[i], x1[i], D == 3 ? x2[i] : 0.0, (double)i);
    inv[i] = (uint32_t)t;
}

// ---------------------------------------------------------------- radix sort
// LSD passes of up to kRsMaxBits bits: digit width ceil(bits / passes) with
// passes = ceil(bits / kRsMaxBits) (18-bit Morton keys: two 9-bit passes).
constexpr int kRsThreads = 256;
constexpr int kRsItems = 8;
constexpr int kRsTile = kRsThreads * kRsItems;
constexpr int kRsMaxBits = 9;
constexpr int kRsMaxBins = 1 << kRsMaxBits;

// per-tile digit counts, hist[digit * ntiles + tile]
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(
    const uint32_t *__restrict__ keys, int N, int shift, int dbits,
    uint32_t *__restrict__ hist, int ntiles) {
    __shared__ uint32_t h[kRsMaxBins];
    const int tid = threadIdx.x, nb = 1 << dbits;
    const uint32_t mask = (uint32_t)nb - 1u;
    for (int b = tid; b < nb; b += kRsThreads) h[b] = 0;
    __syncthreads();
    int base = blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; ++r) {
        int e = base + r * kRsThreads + tid;
        if (e < N) atomicAdd(&h[(keys[e] >> shift) & mask], 1u);
    }
    __syncthreads();
    for (int b = tid; b < nb; b += kRsThreads) hist[b * ntiles + blockIdx.x] = h[b];
}

// stable scatter of one tile: a key's rank among the equal digits of its
// wavefront from dbits ballots, the wavefronts before it from LDS counts
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, int N, int shift, int dbits,
    const uint32_t *__restrict__ offs, int ntiles) {
    constexpr int W = kRsThreads / 64;
    __shared__ uint32_t run[kRsMaxBins];
    __shared__ uint32_t wcnt[W][kRsMaxBins];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, nb = 1 << dbits;
    const uint32_t mask = (uint32_t)nb - 1u;
    for (int b = tid; b < nb; b += kRsThreads) {
        run[b] = offs[b * ntiles + blockIdx.x];
        for (int w = 0; w < W; ++w) wcnt[w][b] = 0;
    }
    __syncthreads();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int base = blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; ++r) {
        int e = base + r * kRsThreads + tid;
        bool valid = e < N;
        uint32_t k = valid ? kin[e] : 0u;
        uint32_t v = valid ? vin[e] : 0u;
        uint32_t d = (k >> shift) & mask;
        uint64_t m = __ballot(valid);
        for (int b = 0; b < dbits; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t bb = __ballot(valid && bit);
            m &= bit ? bb : ~bb;
        }
        int rank = __popcll(m & below);
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + (uint32_t)rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
        for (int b = tid; b < nb; b += kRsThreads) {
            uint32_t add = 0;
            for (int w = 0; w < W; ++w) {
                add += wcnt[w][b];
                wcnt[w][b] = 0;
            }
            run[b] += add;
        }
        __syncthreads();
    }
}

// exclusive scan of n uint32 in blocks of 1024, then block offsets
constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanBlock = kScanThreads * kScanItems;

__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *sh, uint32_t *total) {
    int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        uint32_t t = tid >= off ? sh[tid - off] : 0u;
        __syncthreads();
        sh[tid] += t;
        __syncthreads();
    }
    uint32_t incl = sh[tid];
    *total = sh[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_local(
    const uint32_t *__restrict__ in, uint32_t *__restrict__ out, int n,
    uint32_t *__restrict__ sums) {
    __shared__ uint32_t sh[kScanThreads];
    int base = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    uint32_t loc[kScanItems], s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        loc[k] = base + k < n ? in[base + k] : 0u;
        s += loc[k];
    }
    uint32_t tot;
    uint32_t pre = block_excl_scan(s, sh, &tot);
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = pre;
        pre += loc[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// single block: in-place exclusive scan of the block sums (any length)
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(uint32_t *__restrict__ sums, int n) {
    __shared__ uint32_t sh[kScanThreads];
    uint32_t carry = 0;
    for (int base = 0; base < n; base += kScanThreads) {
        int t = base + threadIdx.x;
        uint32_t v = t < n ? sums[t] : 0u;
        uint32_t tot;
        uint32_t pre = block_excl_scan(v, sh, &tot);
        if (t < n) sums[t] = pre + carry;
        carry += tot;
    }
}

__global__ void k_scan_add(uint32_t *__restrict__ out, int n, const uint32_t *__restrict__ sums) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] += sums[t / kScanBlock];
}

void scan_exclusive_u32(const uint32_t *in, uint32_t *out, int n, uint32_t *tmp,
                        hipStream_t s) {
    int nb = (n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(kScanThreads), 0, s, in, out, n, tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, tmp, nb);
    hipLaunchKernelGGL(k_scan_add, dim3((n + 255) / 256), dim3(256), 0, s, out, n, tmp);
}

size_t radix_sort_scratch_words(int N) {
    int ntiles = (N + kRsTile - 1) / kRsTile;
    size_t h = (size_t)kRsMaxBins * ntiles;
    return 2 * h + (h + kScanBlock - 1) / kScanBlock + 16;
}

// Stable sort of (key, val) by the low `bits` bits of key.  Result in
// (k0, v0); (k1, v1) is ping-pong space.
void radix_sort_pairs(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, int N,
                      int bits, uint32_t *scratch, hipStream_t s) {
    int ntiles = (N + kRsTile - 1) / kRsTile;
    const int passes = (bits + kRsMaxBits - 1) / kRsMaxBits;
    const int dbits = passes ? (bits + passes - 1) / passes : 0;
    const int h = (1 << dbits) * ntiles;
    uint32_t *hist = scratch, *offs = scratch + (size_t)kRsMaxBins * ntiles,
             *tmp = scratch + 2 * (size_t)kRsMaxBins * ntiles;
    for (int p = 0; p < passes; ++p) {
        int shift = dbits * p;
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(kRsThreads), 0, s, k0, N, shift, dbits,
                           hist, ntiles);
        scan_exclusive_u32(hist, offs, h, tmp, s);
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kRsThreads), 0, s, k0, v0, k1, v1, N,
                           shift, dbits, offs, ntiles);
        uint32_t *t = k0; k0 = k1; k1 = t;
        t = v0; v0 = v1; v1 = t;
    }
    if (passes & 1)  // result landed in the ping-pong buffers
    {
        hipMemcpyAsync(k1, k0, sizeof(uint32_t) * N, hipMemcpyDeviceToDevice, s);
        hipMemcpyAsync(v1, v0, sizeof(uint32_t) * N, hipMemcpyDeviceToDevice, s);
    }
}

__global__ void k_cell_ranges(const uint32_t *__restrict__ skeys, int N,
                              int *__restrict__ cstart, int *__restrict__ cend) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    uint32_t k = skeys[p];
    if (p == 0 || skeys[p - 1] != k) cstart[k] = p;
    if (p == N - 1 || skeys[p + 1] != k) cend[k] = p + 1;
}

// --------------------------------------------------------------- neighbours
template <int D>
__device__ inline double dist2(const double *a, const double *b) {
    double d = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]);
    if (D == 3) d += (a[2] - b[2]) * (a[2] - b[2]);
    return d;
}

// distance from x to the slab of cell k along one axis (0 inside); the first
// and last cells are open outward
__device__ inline double box_gap(double x, int k, double bmin, double cs, int nc) {
    double lo = bmin + (double)k * cs, hi = bmin + (double)(k + 1) * cs;
    double gap = 0.0;
    if (k > 0 && x < lo) gap = lo - x;
    if (k < nc - 1 && x > hi) gap = x - hi;
    return gap;
}

// One thread per work position t.  Candidates are scanned in cell order
// (pos_cell); an entry stores the candidate's work position: its cell-order
// position (kCaller false), or its original index when the work order is
// the caller's (kCaller true).
__device__ inline double wsqrt(double x) { return sqrt(x); }
__device__ inline float wsqrt(float x) { return __builtin_sqrtf(x); }

// |a - b|^2 with a * b + c fused, as the FP32 kernels (compiled with FMA
// contraction) evaluate the distance of their kernel weights
template <int D>
__device__ inline double dist2_fma(const double *a, const double *b) {
#pragma clang fp contract(on)
    double d = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]);
    if (D == 3) d += (a[2] - b[2]) * (a[2] - b[2]);
    return d;
}
// the distance the density of record type TW uses: the search's own for
// FP64 records, the fused form for FP32 records
template <int D, typename TW>
__device__ inline TW wdist(double d_search, const double *a, const double *b) {
    if constexpr (sizeof(TW) == 4) return wsqrt(TW(dist2_fma<D>(a, b)));
    else return wsqrt(TW(d_search));
}

// Candidate sources of the search.  range() gives the cell-order range
// [q0, q1) of cell c and a delta such that load(q + delta) is the position
// of candidate q; the loads of a group of four are issued together.
struct GlobalCand {
    const char *pcb;  // pos_cell as bytes
    const int *cstart, *cend;
    __device__ inline void range(int, int, int, uint32_t c, int &q0, int &q1, int &delta) const {
        q0 = cstart[c];
        q1 = cend[c];
        delta = 0;
    }
    // 32-bit byte offset from the array base (N < 2^27)
    __device__ inline double4 load(int q) const {
        return *(const double4 *)(pcb + ((uint32_t)q << 5));
    }
};

// whether the particle in cell cc has periodic images to search: its cell
// lies within the image window of a domain face
template <int D>
__device__ inline bool has_images(const NbrArgs &a, const int *cc) {
    if (!a.g.periodic) return false;
    const int W = a.img_window;
    bool near_edge = false;
    for (int ax = 0; ax < D; ++ax) near_edge |= cc[ax] - W < 0 || cc[ax] + W >= a.g.nc[ax];
    return near_edge;
}

// Entries a thread's image search sorts in its LDS stage before they go to
// the row (more images than this continue with the row's own insertion).
// 40 holds every image of the 3D lattice with h = 2.4 dx (5 to 40 per
// particle, mean 12) and keeps the image kernel at four wavefronts per SIMD
// (10 KB of LDS per wavefront).
constexpr int kImgStage = 40;
// parents of staged images lie below 2^kStageIdxBits
constexpr int kStageIdxBits = 26;

// Sort word of an image of parent jo (below 2^kStageIdxBits) with the shift
// s: jo << 6 | (7 - axis mask) << 3 | signs (bit a: +1 along a).  The words
// of one particle's images ascend in its ghost order (parent, then the
// descending mask) and never tie: the creation test fixes a parent's shift
// along an axis by the side of the domain it lies on (+1 wants it within h
// of the low face, -1 not there), so the parent and the mask fix the signs.
__device__ inline uint32_t stage_word(uint32_t jo, const int *s) {
    uint32_t m = 0, sg = 0;
    for (int ax = 0; ax < 3; ++ax) {
        m |= (uint32_t)(s[ax] != 0) << ax;
        sg |= (uint32_t)(s[ax] > 0) << ax;
    }
    return (jo << 6) | ((7u - m) << 3) | sg;
}

// Periodic images of work position t (position xi, cell cc) in the
// reference's ghost order, stored in its row from entry g0 on: only
// particles whose cell lies within the image window of a domain face
// search.  Returns the image count (an overflow sets the status bit and
// keeps the entries found so far).
//  * The first kImgStage hits are insertion-sorted in the thread's LDS
//    stage as sort words (parent << 6 | (7 - axis mask) << 3 | shift signs:
//    ascending words are the ghost order, see stage_word), each placed by a
//    binary search; they go to the row as work positions (a.inv) when the
//    search ends.  Parents from 2^26 on skip the stage.
//  * Further hits (more images than the stage holds) are insertion-sorted
//    in the row itself, each key read back through the entry's position.
template <int D, bool kCaller>
__device__ inline int image_search(const NbrArgs &a, int t, const double *xi, const int *cc,
                                   int g0, uint32_t *stage, int ss) {
    const Geom &g = a.g;
    const int N = a.N;
    int gcnt = 0, err = 0;
    const int W = a.img_window;
    bool near_edge = false;
    for (int ax = 0; ax < D; ++ax) near_edge |= cc[ax] - W < 0 || cc[ax] + W >= g.nc[ax];
    if (g.periodic && near_edge) {
        uint32_t *nb = a.nbr;
        // staged sort word -> row entry (index | code)
        auto to_row = [&](uint32_t w) {
            const uint32_t jo = w >> 6, m = 7u - ((w >> 3) & 7u);
            uint32_t code = 0, mul = 1;
            for (int ax = 0; ax < 3; ++ax, mul *= 3)
                code += mul * (((m >> ax) & 1u) ? (((w >> ax) & 1u) ? 2u : 0u) : 1u);
            return (kCaller ? jo : a.inv[jo]) | (code << kIdxBits);
        };
        // entries are in the row: the stage overflowed, or parents are too
        // large for sort words
        bool in_row = (uint32_t)N > (1u << kStageIdxBits);
        for (int vz = (D == 3 ? cc[2] - W : 0); vz <= (D == 3 ? cc[2] + W : 0) && !err; ++vz)
            for (int vy = cc[1] - W; vy <= cc[1] + W && !err; ++vy)
                for (int vx = cc[0] - W; vx <= cc[0] + W && !err; ++vx) {
                    int v[3] = {vx, vy, vz};
                    int s[3] = {0, 0, 0}, r[3] = {0, 0, 0};
                    bool wrapped = false;
                    for (int ax = 0; ax < D; ++ax) {
                        s[ax] = v[ax] < 0 ? -1 : (v[ax] >= g.nc[ax] ? 1 : 0);
                        r[ax] = v[ax] - s[ax] * g.nc[ax];
                        wrapped |= s[ax] != 0;
                    }
                    if (!wrapped) continue;
                    uint32_t c = cell_key<D>(g, r[0], r[1], r[2]);
                    uint32_t code = (uint32_t)((s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1));
                    const uint32_t cw = stage_word(0, s);
                    // one candidate: the reference's ghost creation test and
                    // image, the distance test, the entry in ghost order
                    auto visit = [&](int q, const double4 &pj) {
                        double xj[3] = {pj.x, pj.y, pj.z};
                        bool exists = true;
                        for (int ax = 0; ax < D; ++ax) {
                            if (s[ax] == 0) continue;
                            bool low = xj[ax] <= g.bmin[ax] + g.h;
                            bool high = g.bmax[ax] - g.h < xj[ax];
                            exists &= s[ax] > 0 ? low : (!low && high);
                            xj[ax] = image_coord(xj[ax], s[ax], g.bmin[ax], g.bmax[ax]);
                        }
                        if (!exists) return;
                        double d = dist2<D>(xj, xi);
                        if (!(d < g.hsq)) return;
                        if (gcnt >= a.maxg) {
                            err = 1;
                            return;
                        }
                        const uint32_t jo = (uint32_t)pj.w;
                        if (!in_row && gcnt < kImgStage) {
                            // after the last word not above this one, the
                            // words behind it moved up by one (four reads,
                            // then four writes)
                            const uint32_t w = (jo << 6) | cw;
                            int lo = 0, hi = gcnt;
                            while (lo < hi) {
                                const int mid = (lo + hi) >> 1;
                                if (stage[mid * ss] <= w) lo = mid + 1;
                                else hi = mid;
                            }
                            int p = gcnt;
                            for (; p - 4 >= lo; p -= 4) {
                                uint32_t b[4];
#pragma unroll
                                for (int u = 0; u < 4; ++u) b[u] = stage[(p - 1 - u) * ss];
#pragma unroll
                                for (int u = 0; u < 4; ++u) stage[(p - u) * ss] = b[u];
                            }
                            for (; p > lo; --p) stage[p * ss] = stage[(p - 1) * ss];
                            stage[p * ss] = w;
                            ++gcnt;
                            return;
                        }
                        const uint64_t ke = ghost_key(jo, code);
                        if (!in_row) {
                            for (int k = 0; k < gcnt; ++k) nb[nb_slot(N, t, g0 + k)] = to_row(stage[k * ss]);
                            in_row = true;
                        }
                        int p = gcnt;
                        while (p > 0) {
                            uint32_t prev = nb[nb_slot(N, t, g0 + p - 1)];
                            uint32_t pp = (uint32_t)a.pos_w[nb_index(prev)].w;
                            if (ghost_key(pp, nb_code(prev)) <= ke) break;
                            nb[nb_slot(N, t, g0 + p)] = prev;
                            --p;
                        }
                        nb[nb_slot(N, t, g0 + p)] = (kCaller ? jo : (uint32_t)q) | (code << kIdxBits);
                        ++gcnt;
                    };
                    // candidates four at a time (indices clamped to the cell,
                    // the loads in flight together), visited in cell order
                    const int q0 = a.cstart[c], q1 = a.cend[c];
                    for (int q = q0; q < q1 && !err; q += 4) {
                        double4 pb[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) pb[u] = a.pos_cell[q + u < q1 ? q + u : q1 - 1];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (q + u < q1 && !err) visit(q + u, pb[u]);
                    }
                }
        if (!in_row)
            for (int k = 0; k < gcnt; ++k) nb[nb_slot(N, t, g0 + k)] = to_row(stage[k * ss]);
        if (err) atomicOr(a.status, 1 << 3);
    }
    return gcnt;
}

// The search of work position t: real neighbours in the reference's stencil
// order (i itself excluded: in cell order its candidate position is t), and
// with kDens the step's density (whole, or up to the images' terms for a
// particle whose images k_images searches and k_images_finish adds).
//  * A stencil cell whose box is farther than h from x_i (with a relative
//    margin far above the rounding of positions and cell indices; edge
//    cells are open outward, where clamped particles sit) holds no
//    neighbour and is skipped.
//  * Candidates are read four at a time (indices clamped to the cell, no
//    branches around the loads); a hit goes to the thread's LDS stage, and
//    every eight hits are stored as one 32-byte sector of the chunked row.
//  * The capacity is checked per group of four: the list has four guard
//    entries past its last one for the overshoot.
// Returns the real count through cnt_out.
template <int D, bool kCaller, int kDens, typename Src>
__device__ inline void search_one(const NbrArgs &a, int t, const Src &src, uint32_t *stage,
                                  int &cnt_out) {
    using TW = typename std::conditional<kDens == 4, float, double>::type;
    const Geom &g = a.g;
    const int N = a.N;
    int cnt = 0;
    const TW wh2 = TW(g.h / 2.), wsig = TW(g.sigma);
    TW omg = TW(0);
    double4 pi = a.pos_w[t];
    double xi[3] = {pi.x, pi.y, pi.z};
    const uint32_t io = (uint32_t)pi.w;
    uint32_t ck = a.ck_w[t];
    int cc[3];
    cell_coords<D>(g, ck, cc);
    int err = 0;
    // hits go to the thread's LDS stage (two chunks of eight entries, entry
    // n at stage[n & 15]); each completed chunk is stored as one 32-byte
    // sector of the row, so every sector is written once and whole
    int flushed = 0;  // chunks stored
    auto flush = [&]() {
        const uint4 *sp = (const uint4 *)(stage + 8 * (flushed & 1));
        uint4 *dp = (uint4 *)(a.nbr + nb_slot(N, t, 8 * flushed));
        dp[0] = sp[0];
        dp[1] = sp[1];
        ++flushed;
    };
    const double hlim = g.hsq * (1.0 + 1e-9);
    for (int k = cc[0] - 1; k <= cc[0] + 1 && !err; ++k) {
        if (k < 0 || k >= g.nc[0]) continue;
        double bx = box_gap(xi[0], k, g.bmin[0], g.cs[0], g.nc[0]);
        for (int l = cc[1] - 1; l <= cc[1] + 1 && !err; ++l) {
            if (l < 0 || l >= g.nc[1]) continue;
            double by = box_gap(xi[1], l, g.bmin[1], g.cs[1], g.nc[1]);
            for (int m = (D == 3 ? cc[2] - 1 : 0); m <= (D == 3 ? cc[2] + 1 : 0) && !err; ++m) {
                if (D == 3 && (m < 0 || m >= g.nc[2])) continue;
                double bz = D == 3 ? box_gap(xi[2], m, g.bmin[2], g.cs[2], g.nc[2]) : 0.0;
                if (bx * bx + by * by + bz * bz > hlim) continue;
                uint32_t c = cell_key<D>(g, k, l, m);
                int q0, q1, delta;
                src.range(k, l, m, c, q0, q1, delta);
                for (int q = q0; q < q1; q += 4) {
                    double4 pj[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) pj[u] = src.load((q + u < q1 ? q + u : q1 - 1) + delta);
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int qq = q + u;
                        double xj[3] = {pj[u].x, pj[u].y, pj[u].z};
                        double d = dist2<D>(xj, xi);
                        bool other;
                        uint32_t val;
                        if constexpr (kCaller) {
                            const uint32_t jo = (uint32_t)pj[u].w;
                            other = jo != io;
                            val = jo;
                        } else {
                            other = qq != t;
                            val = (uint32_t)qq;
                        }
                        if (qq < q1 && other && d < g.hsq) {
                            stage[cnt & 15] = val | (kCodeReal << kIdxBits);
                            ++cnt;
                            // density: the hits in row order, with the
                            // distance of the test ((-a)^2 = a^2 exactly)
                            if constexpr (kDens != 0)
                                omg += kernel_w<TW>(wdist<D, TW>(d, xj, xi), wh2, wsig);
                        }
                    }
                    // at most one chunk completes per group of four
                    if (cnt >= 8 * (flushed + 1)) flush();
                    if (cnt > a.maxr) {
                        err = 1;
                        break;
                    }
                }
            }
        }
    }
    // the last, partial chunk (and, after an overflow, up to one more);
    // slots past the count are never read (images overwrite theirs)
    while (8 * flushed < cnt) flush();
    if (err) {  // rows stay valid: the first maxr entries, no images
        cnt = a.maxr;
        atomicOr(a.status, 1 << 1);
    }
    a.noi[t] = cnt;
    // a particle within the image window of a periodic domain face has its
    // images from k_images, moved into place by k_images_finish (nog = -1
    // marks it); the others have none
    const bool images = !err && has_images<D>(a, cc);
    a.nog[t] = images ? -1 : 0;
    if constexpr (kDens != 0) {
        // compDensity (Particles.cpp:1151-1184): the self term after the
        // real neighbours, then the images in ghost order (:2262-2290), which
        // k_images_finish adds to the partial sum left in the omega slot
        const TW om = omg + kernel_w<TW>(TW(0), wh2, wsig);
        TW *pr = (TW *)a.prim + (size_t)t * kPrimN;
        if (images)
            pr[kPrimOmega] = om;
        else
            write_prim<D, TW>(pr, om, a.m[io], a.u[io], a.v[0][io], a.v[1][io],
                              D == 3 ? a.v[2][io] : 0., a.gamma);
    }
    cnt_out = cnt;
}

// largest counts of the step: a wavefront maximum, one atomic per wavefront
// and count (every lane of the wavefront calls this); the maxima start at 0,
// so a zero maximum needs no atomic
__device__ inline void reduce_max_counts(const NbrArgs &a, int cnt, int gcnt) {
    for (int off = 32; off > 0; off >>= 1) {
        int o = __shfl_xor(cnt, off, 64);
        cnt = o > cnt ? o : cnt;
        o = __shfl_xor(gcnt, off, 64);
        gcnt = o > gcnt ? o : gcnt;
    }
    if (__lane_id() == 0) {
        if (cnt > 0) atomicMax(&a.max_counts[0], cnt);
        if (gcnt > 0) atomicMax(&a.max_counts[1], gcnt);
    }
}

// The periodic image search, on a stream of its own beside k_neighbours: a
// particle whose cell lies within the image window of a domain face stores
// its images from row entry img_base on (behind every sector the real
// search stores) and their count in img_cnt.  One thread per work position
// (cell order: a wavefront's particles lie near the same domain faces), one
// wavefront per block; a lane's kImgStage LDS slots are interleaved with
// its neighbours' (slot k of lane l at k * 64 + l: no bank conflicts).
constexpr int kImgThreads = 64;
template <int D, bool kCaller>
__global__ __launch_bounds__(kImgThreads) void k_images(NbrArgs a) {
    __shared__ uint32_t s_stage[kImgThreads * kImgStage];
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= a.N) return;
    int cc[3];
    cell_coords<D>(a.g, a.ck_w[t], cc);
    if (!has_images<D>(a, cc)) return;
    const double4 pi = a.pos_w[t];
    const double xi[3] = {pi.x, pi.y, pi.z};
    a.img_cnt[t] = image_search<D, kCaller>(a, t, xi, cc, a.img_base, s_stage + threadIdx.x,
                                            kImgThreads);
}

// After k_neighbours and k_images: for the particles k_neighbours left their
// images to (nog = -1; not after a real overflow), the images move behind
// the real entries (an ascending copy: the target starts below the source),
// and with kDens their density terms are added in row order to the partial
// sum in the record's omega slot, then the primitive record is written.
template <int D, bool kCaller, int kDens>
__global__ __launch_bounds__(256) void k_images_finish(NbrArgs a) {
    using TW = typename std::conditional<kDens == 4, float, double>::type;
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    int gcnt = 0;
    if (t < a.N && a.nog[t] < 0) {
        const Geom &g = a.g;
        const int N = a.N, cnt = a.noi[t], src = a.img_base;
        gcnt = a.img_cnt[t];
        a.nog[t] = gcnt;
        const double4 pi = a.pos_w[t];
        const double xi[3] = {pi.x, pi.y, pi.z};
        const TW wh2 = TW(g.h / 2.), wsig = TW(g.sigma);
        TW *pr = (TW *)a.prim + (size_t)t * kPrimN;
        TW om = TW(0);
        if constexpr (kDens != 0) om = pr[kPrimOmega];
        auto add = [&](uint32_t e, const double4 &pj) {
            double xj[3] = {pj.x, pj.y, pj.z};
            for (int ax = 0; ax < D; ++ax)
                xj[ax] = image_coord(xj[ax], code_shift(nb_code(e), ax), g.bmin[ax], g.bmax[ax]);
            om += kernel_w<TW>(wdist<D, TW>(dist2<D>(xi, xj), xi, xj), wh2, wsig);
        };
        const double4 *pw = kCaller ? a.pos_w : a.pos_cell;
        // four entries at a time: read, stored, and (with kDens) their
        // positions in flight together
        int p = 0;
        for (; p + 4 <= gcnt; p += 4) {
            uint32_t eb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) eb[u] = nb_get(a.nbr, N, t, src + p + u);
#pragma unroll
            for (int u = 0; u < 4; ++u) a.nbr[nb_slot(N, t, cnt + p + u)] = eb[u];
            if constexpr (kDens != 0) {
                double4 pb[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) pb[u] = pw[nb_index(eb[u])];
#pragma unroll
                for (int u = 0; u < 4; ++u) add(eb[u], pb[u]);
            }
        }
        for (; p < gcnt; ++p) {
            const uint32_t e = nb_get(a.nbr, N, t, src + p);
            a.nbr[nb_slot(N, t, cnt + p)] = e;
            if constexpr (kDens != 0) add(e, pw[nb_index(e)]);
        }
        if constexpr (kDens != 0) {
            const uint32_t io = (uint32_t)pi.w;
            write_prim<D, TW>(pr, om, a.m[io], a.u[io], a.v[0][io], a.v[1][io],
                              D == 3 ? a.v[2][io] : 0., a.gamma);
        }
    }
    reduce_max_counts(a, 0, gcnt);
}

// kDens: 0, or the bytes of the primitive records (4, 8) when the step's
// density is fused into the search (NbrArgs::dens).  One thread per work
// position, candidates from global memory: caller order and linear keys.
template <int D, bool kCaller, int kDens>
__global__ __launch_bounds__(256) void k_neighbours(NbrArgs a) {
    __shared__ __align__(16) uint32_t s_stage[256 * 16];
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    int cnt = 0;
    if (t < a.N) {
        const GlobalCand src{(const char *)a.pos_cell, a.cstart, a.cend};
        search_one<D, kCaller, kDens>(a, t, src, s_stage + 16 * threadIdx.x, cnt);
    }
    reduce_max_counts(a, cnt, 0);
}

template __global__ void k_assign_cells<2>(Geom, const double *, const double *, const double *,
                                           int, uint32_t *, uint32_t *);
template __global__ void k_assign_cells<3>(Geom, const double *, const double *, const double *,
                                           int, uint32_t *, uint32_t *);
// ------------------------------------------------------------ domain limits
// Particles::getDomainLimits (Particles.cpp:228-268) as the reference runs
// it, sequentially per axis: min starts at DBL_MAX, max at DBL_MIN (the
// smallest positive double), and the max test is an else-branch of the min
// test, so x_i only counts for the max when it is not a new running
// minimum, i.e. when x_i >= min(x_0 .. x_{i-1}).  In parallel: block
// minima, their exclusive prefix minima, then every element's exclusive
// prefix minimum inside its block, the candidate test and a max reduction.
// Blocks of 256 threads x 4 consecutive elements; gridDim.y = axis.
constexpr int kLimItems = 4, kLimBlock = 256 * kLimItems;

__device__ inline double block_min256(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_limits_blockmin(const double *__restrict__ x0,
                                                         const double *__restrict__ x1,
                                                         const double *__restrict__ x2, int N,
                                                         double *__restrict__ bmin) {
    __shared__ double sh[256];
    const double *x = blockIdx.y == 0 ? x0 : (blockIdx.y == 1 ? x1 : x2);
    double m = DBL_MAX;
    size_t base = (size_t)blockIdx.x * kLimBlock + threadIdx.x * kLimItems;
    for (int k = 0; k < kLimItems; ++k)
        if (base + k < (size_t)N) m = fmin(m, x[base + k]);
    m = block_min256(m, sh);
    if (threadIdx.x == 0) bmin[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = m;
}

// exclusive prefix minima of the block minima, one axis per block
__global__ __launch_bounds__(256) void k_limits_scan(const double *__restrict__ bmin, int nb,
                                                     double *__restrict__ bpre) {
    __shared__ double sh[256];
    const double *in = bmin + (size_t)blockIdx.x * nb;
    double *out = bpre + (size_t)blockIdx.x * nb;
    int per = (nb + 255) / 256;
    int a = threadIdx.x * per, b = min(a + per, nb);
    double m = DBL_MAX;
    for (int k = a; k < b; ++k) m = fmin(m, in[k]);
    sh[threadIdx.x] = m;
    __syncthreads();
    if (threadIdx.x == 0) {  // exclusive scan of 256 chunk minima
        double run = DBL_MAX;
        for (int k = 0; k < 256; ++k) {
            double v = sh[k];
            sh[k] = run;
            run = fmin(run, v);
        }
    }
    __syncthreads();
    double run = sh[threadIdx.x];
    for (int k = a; k < b; ++k) {
        out[k] = run;
        run = fmin(run, in[k]);
    }
}

__global__ __launch_bounds__(256) void k_limits_blockmax(const double *__restrict__ x0,
                                                         const double *__restrict__ x1,
                                                         const double *__restrict__ x2, int N,
                                                         const double *__restrict__ bpre,
                                                         double *__restrict__ bmax) {
    __shared__ double sh[256];
    const double *x = blockIdx.y == 0 ? x0 : (blockIdx.y == 1 ? x1 : x2);
    size_t base = (size_t)blockIdx.x * kLimBlock + threadIdx.x * kLimItems;
    double v[kLimItems];
    double tm = DBL_MAX;
    for (int k = 0; k < kLimItems; ++k) {
        v[k] = base + k < (size_t)N ? x[base + k] : DBL_MAX;
        tm = fmin(tm, v[k]);
    }
    // exclusive prefix minimum of the thread minima inside the block
    sh[threadIdx.x] = tm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = bpre[(size_t)blockIdx.y * gridDim.x + blockIdx.x];
        for (int k = 0; k < 256; ++k) {
            double t = sh[k];
            sh[k] = run;
            run = fmin(run, t);
        }
    }
    __syncthreads();
    double pm = sh[threadIdx.x];
    __syncthreads();
    double mx = DBL_MIN;
    for (int k = 0; k < kLimItems; ++k) {
        if (base + k >= (size_t)N) break;
        double xv = v[k];
        if (xv < pm) pm = xv;
        else if (xv > mx) mx = xv;
    }
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) bmax[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

// out[2a] = min, out[2a+1] = max of axis a (block = axis)
__global__ __launch_bounds__(256) void k_limits_final(const double *__restrict__ bmin,
                                                      const double *__restrict__ bmax, int nb,
                                                      double *__restrict__ out) {
    __shared__ double sh[256];
    double mn = DBL_MAX, mx = DBL_MIN;
    for (int k = threadIdx.x; k < nb; k += 256) {
        mn = fmin(mn, bmin[(size_t)blockIdx.x * nb + k]);
        mx = fmax(mx, bmax[(size_t)blockIdx.x * nb + k]);
    }
    mn = block_min256(mn, sh);
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = mn;
        out[2 * blockIdx.x + 1] = sh[0];
    }
}

int limits_blocks(int N) { return (N + kLimBlock - 1) / kLimBlock; }

void launch_domain_limits(const double *x0, const double *x1, const double *x2, int D, int N,
                          double *scratch, double *out, hipStream_t s) {
    int nb = limits_blocks(N);
    double *bmin = scratch, *bpre = scratch + (size_t)D * nb, *bmax = scratch + 2 * (size_t)D * nb;
    hipLaunchKernelGGL(k_limits_blockmin, dim3(nb, D), dim3(256), 0, s, x0, x1, x2, N, bmin);
    hipLaunchKernelGGL(k_limits_scan, dim3(D), dim3(256), 0, s, bmin, nb, bpre);
    hipLaunchKernelGGL(k_limits_blockmax, dim3(nb, D), dim3(256), 0, s, x0, x1, x2, N, bpre, bmax);
    hipLaunchKernelGGL(k_limits_final, dim3(D), dim3(256), 0, s, bmin, bmax, nb, out);
}

// positions in the caller's order (work order = caller order)
__global__ void k_pack_pos(const double *__restrict__ x0, const double *__restrict__ x1,
                           const double *__restrict__ x2, int D, int N,
                           double4 *__restrict__ pos) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    pos[i] = make_double4(x0[i], x1[i], D == 3 ? x2[i] : 0.0, (double)i);
}

void launch_cells(const Geom &g, int D, const double *x0, const double *x1,
                  const double *x2, int N, double4 *pos, uint32_t *key, uint32_t *idx,
                  uint32_t *skey, uint32_t *scratch, int *cstart, int *cend, uint32_t *inv,
                  hipStream_t s) {
    int nb = (N + 255) / 256;
    if (D == 2)
        hipLaunchKernelGGL(k_assign_cells<2>, dim3(nb), dim3(256), 0, s, g, x0, x1, x2, N, key, idx);
    else
        hipLaunchKernelGGL(k_assign_cells<3>, dim3(nb), dim3(256), 0, s, g, x0, x1, x2, N, key, idx);
    // sort a copy: key[i] stays the cell of particle i
    hipMemcpyAsync(skey, key, sizeof(uint32_t) * N, hipMemcpyDeviceToDevice, s);
    int bits = 1;
    while ((1 << bits) < g.nkeys) ++bits;
    uint32_t *k1 = scratch, *v1 = scratch + N, *rs = scratch + 2 * (size_t)N;
    radix_sort_pairs(skey, idx, k1, v1, N, bits, rs, s);
    int nc = g.nkeys;
    hipLaunchKernelGGL(k_fill_i32, dim3((nc + 255) / 256), dim3(256), 0, s, cstart, 0, nc);
    hipLaunchKernelGGL(k_fill_i32, dim3((nc + 255) / 256), dim3(256), 0, s, cend, 0, nc);
    hipLaunchKernelGGL(k_cell_ranges, dim3(nb), dim3(256), 0, s, skey, N, cstart, cend);
    hipLaunchKernelGGL(k_gather_pos, dim3(nb), dim3(256), 0, s, x0, x1, x2, D, N, idx, pos, inv);
}

void launch_pack_pos(const double *x0, const double *x1, const double *x2, int D, int N,
                     double4 *pos, hipStream_t s) {
    hipLaunchKernelGGL(k_pack_pos, dim3((N + 255) / 256), dim3(256), 0, s, x0, x1, x2, D, N, pos);
}

struct NbrStreams {
    hipStream_t s, s_img;
    hipEvent_t fork, join;
};

// periodic: the image search forks to s_img after the cell list and joins
// before k_images_finish
template <int D, bool kCaller, int kDens>
static void launch_nbr_dens(const NbrArgs &a, dim3 grid, const NbrStreams &st) {
    if (a.g.periodic) {
        hipEventRecord(st.fork, st.s);
        hipStreamWaitEvent(st.s_img, st.fork, 0);
        hipLaunchKernelGGL((k_images<D, kCaller>), dim3((a.N + kImgThreads - 1) / kImgThreads),
                           dim3(kImgThreads), 0, st.s_img, a);
        hipEventRecord(st.join, st.s_img);
    }
    hipLaunchKernelGGL((k_neighbours<D, kCaller, kDens>), grid, dim3(256), 0, st.s, a);
    if (a.g.periodic) {
        hipStreamWaitEvent(st.s, st.join, 0);
        hipLaunchKernelGGL((k_images_finish<D, kCaller, kDens>), grid, dim3(256), 0, st.s, a);
    }
}

template <int D, bool kCaller>
static void launch_nbr(const NbrArgs &a, dim3 grid, const NbrStreams &st) {
    if (a.dens == 4) launch_nbr_dens<D, kCaller, 4>(a, grid, st);
    else if (a.dens == 8) launch_nbr_dens<D, kCaller, 8>(a, grid, st);
    else launch_nbr_dens<D, kCaller, 0>(a, grid, st);
}

bool launch_neighbours(const NbrArgs &a, int D, hipStream_t s, hipStream_t s_img,
                       hipEvent_t fork, hipEvent_t join) {
    dim3 grid((a.N + 255) / 256);
    const NbrStreams st{s, s_img, fork, join};
    if (a.caller_order) {
        if (D == 2) launch_nbr<2, true>(a, grid, st);
        else launch_nbr<3, true>(a, grid, st);
    } else {
        if (D == 2) launch_nbr<2, false>(a, grid, st);
        else launch_nbr<3, false>(a, grid, st);
    }
    return a.dens != 0;
}

}  // namespace mlh
