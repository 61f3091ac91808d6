// kernels.h -- kernel argument structs and host-side launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "device_common.h"
#include "mlp_api.h"

namespace mlh {

struct NbrArgs {
    Geom g;
    int N, maxr, maxg;
    const double4 *pos_cell;     // cell order: x, y, z, original index
    const double4 *pos_w;        // the same in work order
    const uint32_t *ck_w;        // cell of the particle at each work position
    int caller_order;            // work order = caller order (else cell order)
    const int *cstart, *cend;    // cell -> [start, end) of the cell order
    const uint32_t *inv;         // original index -> cell-order position
    uint32_t *nbr;               // [maxr+maxg][N], work-order rows and entries
    int *noi, *nog;              // per work position
    int *status;                 // [0] bit 1: real overflow, bit 3: image overflow
    int *max_counts;             // [2]
    int img_window;              // virtual-cell reach of the image search
    // The image search runs beside the real search: a particle's images go
    // to its row from entry img_base on (behind every sector the real search
    // may store, guards included), their count to img_cnt; k_images_finish
    // moves them behind the real entries.
    int *img_cnt;
    int img_base;
    // The step's compDensity + compPressure fused into the search (dens =
    // 4: FP32 records, 8: FP64 records, 0: off): omega summed over the hits
    // in row order, the self term, the images in row order, then the
    // primitive record at the work position, bitwise as k_density.
    int dens;
    void *prim;
    const double *m, *u, *v[3];
    double gamma;
};

// Arguments of the physics kernels.  Storage type of derived per-particle
// data (prim, frec) is T; positions / state are FP64.  Threads, neighbour
// rows, pos, prim and frec are in work order (cell order or the caller's
// order); the state arrays, the next state and the flux sums are in the
// caller's particle order (sidx maps; null when the two orders coincide).
struct HydroArgs {
    Geom g;
    int N;
    const uint32_t *sidx;        // work position -> original index, or null
    double gamma, cfl, beta, psi1, psi2, riemann_tol;
    int riemann_max_iter, learned_iters;
    int limiter_abs, slope_limiting, pairwise_limiter, mfm, enforce_flux_sym;
    int move_particles, foqp, quirk3d, write_fluxes, learned;
    int debug;                   // DEBUG_LVL >= 1: singular-E / non-finite counters
    const double4 *pos;
    const uint32_t *nbr;
    const int *noi, *nog;
    // state (current) and next state
    const double *x[3], *v[3], *m, *u;
    double *xn[3], *vn[3], *mn, *un;
    void *prim;                  // T[N][8]
    void *frec;                  // T[N][FRec<D>::N]
    double *dt_block;            // per-block minimum of the CFL step
    double *dt;                  // [0] step used, device resident
    double *mF, *vF, *eF;        // optional flux sums (write_fluxes)
    unsigned long long *counters;  // mlh_diag counters
    // [0]: error bits of the neighbour search (kStatusErrMask), sticky over
    // the steps of one run; [1]: first step of the run that set one (-1:
    // none).  With an error bit set the face kernel leaves the state as it
    // is, so a run stops where the reference exits (Particles.cpp:351,2252).
    int *status;
    const struct MlpFrags *mlp_frags;  // learned seed network (device)
    double *collect;             // face collection [stripes][scap][6] (or null)
    unsigned long long *collect_n;  // [stripes] fill counters
    long long collect_scap;      // rows per stripe
    int collect_smask;           // stripes - 1 (a power of two)
    int collect_stride;
};

// Per-entry face arrays of the per-method stages (FP64, component-major:
// value c of row entry n of work position t at (c * rowlen + n) * N + t)
struct FaceArrays {
    double *A;        // [D][rowlen][N]   effective face, i's perspective
    double *WR, *WL;  // [D+2][rowlen][N] {rho, P, v...} of i (R) and j (L)
    double *vF;       // [D][rowlen][N]   frame velocity
    double *F;        // [D+2][rowlen][N] {mass, energy, momentum...}
    int rowlen;       // maxr + maxg
};

// the per-method stages (stages_impl.h)
enum StageOp {
    kOpLimit = 0,       // slopeLimiter on the stored gradients
    kOpGradField,       // gradient(f, grad)
    kOpFaceStates,      // compEffectiveFace (what & 1), compRiemannStatesLR (what & 2)
    kOpSolveOwn,        // solveRiemannProblems: the entries the particle solves
    kOpSolveMirror,     // solveRiemannProblems: F_ij = -F_ji for the others
    kOpCollect          // collectFluxes + the update
};

// error bits of HydroArgs::status[0] (the reference's exit codes 1 and 3)
constexpr int kStatusErrMask = (1 << 1) | (1 << 3);

enum CounterIdx {
    kCntNegReal = 0, kCntNegGhost, kCntVacuum, kCntNonconv, kCntMissing,
    kCntFallback, kCntIters, kCntFaces, kCntSeeded, kCntSingular, kCntNonfinite,
    kCntAcoustic,
    kNumCounters
};

void launch_cells(const Geom &g, int D, const double *x0, const double *x1,
                  const double *x2, int N, double4 *pos, uint32_t *key, uint32_t *idx,
                  uint32_t *skey, uint32_t *scratch, int *cstart, int *cend, uint32_t *inv,
                  hipStream_t s);
// returns whether the search computed the density (NbrArgs::dens): the
// cell-order search does not, and the caller then runs the density stage
// Row entry where the image search stores: past the sector holding the real
// search's last guard entry (maxr + 3).
inline int image_base(int maxr) { return (maxr + 4 + kNbChunk - 1) / kNbChunk * kNbChunk; }
// The search on stream s; for a periodic domain the image search runs on
// s_img between the events fork (recorded on s) and join (waited on by s).
bool launch_neighbours(const NbrArgs &a, int D, hipStream_t s, hipStream_t s_img,
                       hipEvent_t fork, hipEvent_t join);
// Particles::getDomainLimits with the reference's sequential semantics;
// scratch: 3 * D * limits_blocks(N) doubles; out: [min, max] per axis
int limits_blocks(int N);
void launch_domain_limits(const double *x0, const double *x1, const double *x2, int D, int N,
                          double *scratch, double *out, hipStream_t s);
void launch_pack_pos(const double *x0, const double *x1, const double *x2, int D, int N,
                     double4 *pos, hipStream_t s);

size_t radix_sort_scratch_words(int N);
// stable LSD sort of (k0, v0) by the low `bits` bits of k0, result in
// (k0, v0); (k1, v1) ping-pong space, scratch radix_sort_scratch_words(N)
void radix_sort_pairs(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, int N,
                      int bits, uint32_t *scratch, hipStream_t s);
void scan_exclusive_u32(const uint32_t *in, uint32_t *out, int n, uint32_t *tmp,
                        hipStream_t s);

// precision: 0 = FP64, 1 = FP32 pair arithmetic
void launch_density(const HydroArgs &a, int D, int precision, hipStream_t s);
void launch_gradients(const HydroArgs &a, int D, int precision, hipStream_t s);
int dt_blocks(int N);
void launch_dt_finalize(const HydroArgs &a, double dt_fixed, double t_now, double t_clip,
                        const double *dt_in, double *dt_hist, int step, hipStream_t s);
void launch_faces(const HydroArgs &a, int D, int precision, hipStream_t s);
// the FP32 instantiations (hydro_fp32.hip)
void launch_density_fp32(const HydroArgs &a, int D, hipStream_t s);
void launch_gradients_fp32(const HydroArgs &a, int D, hipStream_t s);
void launch_faces_fp32(const HydroArgs &a, int D, hipStream_t s);
// bytes of one deferred-particle record (FP32 learned seed)
void launch_sums(const HydroArgs &a, int D, int precision, double *out, hipStream_t s);
void launch_stage(const HydroArgs &a, int D, int precision, int op, int what,
                  const FaceArrays &fa, const double *f, double *grad, hipStream_t s);
void launch_stage_fp32(const HydroArgs &a, int D, int op, int what, const FaceArrays &fa,
                       const double *f, double *grad, hipStream_t s);
// checkFluxSymmetry on the stored F_ij: out[0] entries failing the test,
// out[1] entries without a mirror (added to)
void launch_flux_check(const HydroArgs &a, int D, const FaceArrays &fa, double tol,
                       unsigned long long *out, hipStream_t s);
// neighbour rows whose partner does not list the particle -> the
// missing_mirror_pairs counter
void launch_mirror_check(const HydroArgs &a, hipStream_t s);
void launch_riemann_flux(const HydroArgs &a, int D, long long n, const double *wl,
                         const double *wr, const double *vf, const double *aij, double *f,
                         int *flag, hipStream_t s);

}  // namespace mlh
