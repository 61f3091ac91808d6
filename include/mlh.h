/*
 * mlh.h -- C ABI of the MI355X meshless-finite-volume (MFV) time step.
 *
 * Plain C types only (no torch, no HIP types): a host program in C, C++,
 * Python/ctypes or any FFI binds these symbols from libmlh_hip.so.
 * Every entry point returns an mlh_status (0 = OK).  All calls on one
 * context are synchronous with respect to the caller and must come from one
 * host thread; device work of a context runs on its HIP stream (for a
 * periodic domain the image search also runs on a second stream of the
 * context, forked from and joined back into the first within each search).
 *
 * Reference interface replaced (jammartin/meshlessHydro, demonstrator/):
 *   Particles      include/Particles.h:28-240   (SoA state, per-stage methods)
 *   MeshlessScheme include/MeshlessScheme.h:15-44, src/MeshlessScheme.cpp:21-254
 *   Riemann        include/Riemann.h:15-43, src/Riemann.cpp:7-229
 *   RiemannSolver  call contract at src/Riemann.cpp:84,93-94 (un-vendored upstream)
 *   parameter.h    include/parameter.h:9-77 (compile-time flags -> mlh_params)
 *   config.info    demonstrator/config.info:2-33 (runtime values -> mlh_params)
 */
#ifndef MLH_H
#define MLH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MLH_ABI_VERSION 2

/* Status codes.  1, 3, 6, 9 are the reference's exit codes
 * (src/Particles.cpp:351, :2252, :1824-1827/:1878-1881,
 *  src/MeshlessScheme.cpp:95-97); the rest are new. */
typedef enum mlh_status {
    MLH_OK = 0,
    MLH_ERR_NEIGHBOUR_OVERFLOW = 1,   /* MAX_NUM_INTERACTIONS exceeded        */
    MLH_ERR_GHOST_OVERFLOW = 3,       /* MAX_NUM_GHOST_INTERACTIONS exceeded  */
    MLH_ERR_NEGATIVE_FACE_PRESSURE = 6,
    MLH_ERR_DUMP_OVERRUN = 9,
    MLH_ERR_INVALID_ARGUMENT = 20,
    MLH_ERR_DEVICE = 21,              /* HIP runtime error / no device        */
    MLH_ERR_GRID = 22,                /* periodic: < 3 search cells per axis; *
                                       * open: extent below one kernel size   */
    MLH_ERR_CALL_ORDER = 23,          /* stage called before its inputs exist */
    MLH_ERR_NOT_TRAINED = 24          /* learned seed requested, no network   */
} mlh_status;

typedef enum mlh_precision {
    MLH_FP64 = 0,   /* all arithmetic in FP64                                  */
    MLH_FP32 = 1    /* FP32 pair arithmetic; FP64 for E_i accumulation and     *
                     * inversion, position differences, the Riemann p* Newton  *
                     * residual step and the flux sums / state update          */
} mlh_precision;

typedef enum mlh_seed {
    MLH_SEED_STANDARD = 0,  /* Toro adaptive PVRS/TRRS/TSRS initial guess      */
    MLH_SEED_LEARNED = 1    /* MLP log p* seed (BF16 MFMA inference)           */
} mlh_seed;

typedef enum mlh_limiter_abs {
    /* The reference calls unqualified abs() on a double with only <cmath>
     * in scope, which binds to int abs(int) (src/Particles.cpp:1416-1417,
     * :1748-1749,:1754,:1759): the ratio is truncated toward zero, NaN and
     * |x| >= 2^31 become INT_MIN (x86 cvttsd2si) and abs(INT_MIN) stays
     * INT_MIN.  MLH_ABS_INT_X86 reproduces that; MLH_ABS_FABS is the
     * intended fabs(). */
    MLH_ABS_INT_X86 = 0,
    MLH_ABS_FABS = 1
} mlh_limiter_abs;

typedef enum mlh_work_order {
    /* Order in which the kernels visit particles.  Results do not depend on
     * it (sums run in the reference's neighbour order, faces are solved by
     * the lower original index); only memory locality does. */
    MLH_ORDER_AUTO = 0,    /* cell order (the faster one on every input     *
                            * measured)                                      */
    MLH_ORDER_CELL = 1,    /* particles sorted by search cell, cells along  *
                            * a Morton curve                                 */
    MLH_ORDER_CALLER = 2   /* the caller's particle order                    */
} mlh_work_order;

/* Compile-time flags of parameter.h plus the config.info values, as one
 * runtime struct.  Fill with mlh_params_default() and override. */
typedef struct mlh_params {
    int32_t dim;                    /* DIM (2 or 3)                           */
    int32_t periodic;               /* PERIODIC_BOUNDARIES                    */
    double box_min[3];              /* periodicBoxLimits lowerX..lowerZ       */
    double box_max[3];              /* periodicBoxLimits upperX..upperZ       */
    double kernel_size;             /* kernelSize (h; the cubic spline uses h/2) */
    double gamma;                   /* gamma                                  */
    double cfl;                     /* CFL                                    */
    int32_t max_interactions;       /* MAX_NUM_INTERACTIONS                   */
    int32_t max_ghost_interactions; /* MAX_NUM_GHOST_INTERACTIONS             */
    int32_t slope_limiting;         /* SLOPE_LIMITING                         */
    double beta;                    /* BETA                                   */
    int32_t pairwise_limiter;       /* PAIRWISE_LIMITER                       */
    double psi1, psi2;              /* PSI_1, PSI_2                           */
    int32_t mfm;                    /* MESHLESS_FINITE_MASS                   */
    int32_t enforce_flux_sym;       /* ENFORCE_FLUX_SYM                       */
    int32_t move_particles;         /* MOVE_PARTICLES                         */
    int32_t first_order_quad_point; /* FIRST_ORDER_QUAD_POINT                 */
    int32_t debug_level;            /* DEBUG_LVL (face-pressure aborts)       */
    int32_t limiter_abs;            /* mlh_limiter_abs                        */
    int32_t quirk_3d_predictor;     /* 1: j-side uses vz[i] (Particles.cpp:1717-1719) */
    int32_t precision;              /* mlh_precision                          */
    int32_t riemann_seed;           /* mlh_seed                               */
    double riemann_tol;             /* relative Newton step accepted as converged */
    int32_t riemann_max_iter;       /* FP64 Newton iteration cap              */
    int32_t learned_iters;          /* FP32 Newton iterations after the MLP seed */
    int32_t work_order;             /* mlh_work_order                         */
} mlh_params;

/* Counters of soft conditions the reference only logs (WARN lines). */
typedef struct mlh_diag {
    int64_t neg_face_pressure_real;   /* Particles.cpp:1813-1834              */
    int64_t neg_face_pressure_ghost;  /* Particles.cpp:1873-1882              */
    int64_t vacuum_faces;             /* Riemann.cpp:112-114,125-127          */
    int64_t nonconverged_faces;       /* Newton cap reached (bisection solved) */
    int64_t missing_mirror_pairs;     /* neighbour seen from one side only    */
    int64_t learned_fallbacks;        /* learned seed rejected by FP64 test   */
    int64_t newton_iterations;        /* total over all solved faces          */
    int64_t faces_solved;             /* Riemann solves performed             */
    int64_t learned_seeded;           /* faces whose Newton started from the  *
                                       * network's p* (strong jumps only)     */
    int32_t max_neighbours;           /* largest real neighbour count          */
    int32_t max_ghost_neighbours;     /* largest periodic-image neighbour count */
    /* counted at debug_level >= 1 only (results are the same either way): */
    int64_t singular_matrices;        /* E_i with a zero or non-finite pivot  *
                                       * (LAPACK INFO > 0, ignored by the     *
                                       * reference, Helper.cpp:13-14)         */
    int64_t nonfinite_particles;      /* particles whose new state is not     *
                                       * finite after an update               */
    int64_t acoustic_faces;           /* FP32 mode: near-uniform faces given  *
                                       * the acoustic (linearised) solution,  *
                                       * no Newton step (relative jumps below *
                                       * 1e-5, |u*| <= c/2; error ~4e-11)     */
} mlh_diag;

typedef struct mlh_ctx mlh_ctx;

/* ---- lifecycle ------------------------------------------------------- */
void mlh_params_default(mlh_params *p);
int mlh_version(void);
/* Particles(int N) (Particles.cpp:67-149) + MeshlessScheme ctor
 * (MeshlessScheme.cpp:8-19): allocates device arrays for n particles. */
int mlh_create(const mlh_params *p, int32_t n, mlh_ctx **out);
int mlh_destroy(mlh_ctx *ctx);
const char *mlh_last_error(void);

/* ---- state transfer (Particles public SoA fields, Particles.h:33-36,54) --
 * z / vz are ignored (may be NULL) when dim == 2. */
int mlh_upload_state(mlh_ctx *ctx, const double *x, const double *y,
                     const double *z, const double *vx, const double *vy,
                     const double *vz, const double *m, const double *u);
int mlh_download_state(mlh_ctx *ctx, double *x, double *y, double *z,
                       double *vx, double *vy, double *vz, double *m,
                       double *u);

/* ---- one full MFV step (MeshlessScheme.cpp:53-224 minus snapshot I/O) ----
 * dt_fixed > 0 : use that step (ADAPTIVE_TIMESTEP 0, or a replayed dt);
 * dt_fixed <= 0: CFL step from compGlobalTimestep (Particles.cpp:1446-1485);
 *               when t_clip > 0 and t_now + dt >= t_clip the step becomes
 *               t_clip - t_now (dump-time clipping, MeshlessScheme.cpp:98-101)
 *               and *clipped is set to 1.
 * *dt_used / *clipped receive the step taken (either may be NULL). */
int mlh_step(mlh_ctx *ctx, double dt_fixed, double t_now, double t_clip,
             double *dt_used, int32_t *clipped);

/* Run n_steps steps with the step size kept on the device (no host round
 * trip per step).  Step k uses dts_in[k] when dts_in is not NULL (a replayed
 * dt sequence), else dt_fixed as in mlh_step.  dts_out (may be NULL)
 * receives the n_steps step sizes after the last step. */
int mlh_run(mlh_ctx *ctx, int32_t n_steps, double dt_fixed,
            const double *dts_in, double *dts_out);

/* ---- stages, for parity tests and the host Particles class ------------ */
/* assignParticlesAndCells + createGhostParticles + gridNNS + ghostNNS
 * (Particles.cpp:270-365, 2113-2191, 2237-2260). */
int mlh_neighbours(mlh_ctx *ctx);
/* compDensity (+ghost part) + compPressure (Particles.cpp:1151-1184,
 * 2262-2290, 1272-1288). */
int mlh_density(mlh_ctx *ctx);
/* compGlobalTimestep (Particles.cpp:1446-1485). */
int mlh_timestep(mlh_ctx *ctx, double *dt);
/* compPsijTilde + gradient x(D+2) + slopeLimiter (Particles.cpp:1186-1270,
 * 1313-1444, 2292-2502). */
int mlh_gradients(mlh_ctx *ctx);
/* compPsijTilde + gradient x(D+2) without the limiter (Particles.h:137-138;
 * Particles.cpp:1186-1270, 2292-2502): for a caller that runs slopeLimiter
 * itself with mlh_limit.  Also the CFL step, as mlh_gradients. */
int mlh_psi_gradients(mlh_ctx *ctx);
/* slopeLimiter (Particles.h:139-140; Particles.cpp:1313-1444) on the
 * gradients of mlh_psi_gradients, whatever params.slope_limiting says; the
 * result is bitwise that of mlh_gradients. */
int mlh_limit(mlh_ctx *ctx);
/* gradient(f, grad) (Particles.h:138; Particles.cpp:1257-1270, periodic
 * :2457-2502 with the image's value f[parent]) of any field: f [N] and grad
 * [N][D] host arrays in the caller's particle order; psi-tilde from the last
 * gradient stage. */
int mlh_gradient_field(mlh_ctx *ctx, const double *f, double *grad);
/* compEffectiveFace + compRiemannStatesLR + solveRiemannProblems +
 * collectFluxes (Particles.cpp:1290-1311, 1488-2011, 2504-2683); also
 * prepares the updated state, committed by mlh_update. */
int mlh_fluxes(mlh_ctx *ctx, double dt);
/* updateStateAndPosition (Particles.cpp:2013-2110). */
int mlh_update(mlh_ctx *ctx);

/* ---- the face stages one method at a time ------------------------------
 * For callers that drive solveRiemannProblems' neighbours one by one, as the
 * reference's MeshlessScheme::run does (MeshlessScheme.cpp:120-160).  The
 * per-entry arrays live on the device (FP64, max_interactions +
 * max_ghost_interactions entries per particle), allocated on first use;
 * mlh_fluxes computes the same sums without them.  After
 * mlh_collect_fluxes, mlh_update commits the step, bitwise as after
 * mlh_fluxes. */
/* compEffectiveFace (Particles.h:142; Particles.cpp:1290-1311, 2504-2531):
 * A_ij of every row entry. */
int mlh_effective_faces(mlh_ctx *ctx);
/* compRiemannStatesLR(dt) (Particles.h:145; Particles.cpp:1488-1733,
 * 2533-2683): W_R (particle i), W_L (neighbour j) and vFrame of every row
 * entry, seen from the row's particle. */
int mlh_riemann_states(mlh_ctx *ctx, double dt);
/* solveRiemannProblems (Particles.h:147; Particles.cpp:1787-1911): F_ij of
 * every entry from the stored faces and states, solved once from the lower
 * index (standard initial guess) and negated for the other side. */
int mlh_solve_faces(mlh_ctx *ctx);
/* collectFluxes (Particles.h:149; Particles.cpp:1913-2011): the flux sums
 * of the stored F_ij in the reference's order. */
int mlh_collect_fluxes(mlh_ctx *ctx);
/* checkFluxSymmetry (Particles.cpp:2888-2976) on the stored F_ij with the
 * reference's one-sided test F_ij + F_ji > tol (FLUX_SYM_TOL = 1e-20; image
 * entries test components 0..3 as the reference's ghost branch does):
 * entries failing it, and entries whose partner does not list the particle. */
int mlh_check_flux_symmetry(mlh_ctx *ctx, double tol, int64_t *asymmetric,
                            int64_t *missing);

/* ---- readbacks (host buffers owned by the caller) ---------------------
 * Neighbour entries are packed int32: bits 0..26 particle index, bits
 * 27..31 periodic-image code s = (sx+1) + 3(sy+1) + 9(sz+1) (13 = the real
 * particle, no image).  Row i holds noi[i] real neighbours in the
 * reference's stencil order followed by noi_ghost[i] image neighbours in
 * ghost-index order; row stride = max_interactions + max_ghost_interactions. */
int mlh_get_neighbours(mlh_ctx *ctx, int32_t *noi, int32_t *noi_ghost,
                       int32_t *nbr);
int mlh_get_density(mlh_ctx *ctx, double *rho, double *P, double *omega);
/* grads: [(D+2) fields {rho, vx, vy, (vz), P}][N][D]; B: [N][D*D]. */
int mlh_get_gradients(mlh_ctx *ctx, double *grads, double *B);
/* mF [N], vF [N][D], eF [N] (collectFluxes, Particles.cpp:1913-2011). */
int mlh_get_fluxes(mlh_ctx *ctx, double *mF, double *vF, double *eF);
/* The per-entry arrays of the face stages, rows as in mlh_get_neighbours:
 * Aij [N][stride][D], WR / WL [N][stride][D+2] {rho, P, v...}, vFrame
 * [N][stride][D], Fij [N][stride][D+2] {mass, energy, momentum...}; any may
 * be NULL; zero past each row's entries. */
int mlh_get_face_arrays(mlh_ctx *ctx, double *Aij, double *WR, double *WL,
                        double *vFrame, double *Fij);
int mlh_get_diagnostics(mlh_ctx *ctx, mlh_diag *d);
/* sumVolume, sumMass, sumEnergy, sumMomentumX/Y/Z (Particles.cpp:2830-2886)
 * into out[0..5]. */
int mlh_conserved_sums(mlh_ctx *ctx, double *out);
/* Device milliseconds per step of the last mlh_step / mlh_run, averaged
 * over its steps, per kernel group: [0] cell list + neighbours,
 * [1] density, [2] gradients + time step, [3] faces + update. */
int mlh_get_timings(mlh_ctx *ctx, float *ms4);

/* ---- exact Riemann solver, batched (Riemann.cpp:83-141 contract) -------
 * Solves n 1D problems (rhoL,uL,PL | rhoR,uR,PR) with gamma, sampled at
 * x/t = 0.  Host arrays in, host arrays out.  flag: -1 left side sampled,
 * +1 right side, 0 vacuum.  iters: Newton iterations used (may be NULL).
 * precision/seed as in mlh_params; the learned seed needs a trained MLP in
 * ctx (ctx may be NULL for MLH_SEED_STANDARD). */
int mlh_riemann_exact_batch(mlh_ctx *ctx, int64_t n, double gamma,
                            const double *rhoL, const double *uL,
                            const double *PL, const double *rhoR,
                            const double *uR, const double *PR,
                            double *rho_s, double *u_s, double *P_s,
                            int32_t *flag, int32_t *iters, int32_t precision,
                            int32_t seed, double tol);

/* ---- device buffers ------------------------------------------------------
 * Memory of the library's own HIP runtime for the *_device entry points.
 * No reference counterpart: a caller that holds device data in another HIP
 * runtime (a second libamdhip64 in the process, e.g. a framework's bundled
 * one) must not pass those pointers here; it copies into these buffers.
 * kind: 1 host -> device, 2 device -> host, 3 device -> device; the copy is
 * complete on return. */
int mlh_device_alloc(int64_t bytes, void **ptr);
int mlh_device_free(void *ptr);
int mlh_device_copy(void *dst, const void *src, int64_t bytes, int32_t kind);

/* The same solve on device-resident arrays, for callers whose face states
 * already live on the GPU (no host copies; the Riemann microbenchmark
 * measures the kernel this way).  states: device [6][n] = rhoL, uL, PL,
 * rhoR, uR, PR; star: device [3][n] = rho*, u*, P* sampled at S = 0; flag,
 * iters: device [n] (may be NULL).  Runs on ctx's stream (default stream
 * when ctx is NULL) and returns when the kernel has finished.  fallbacks
 * (host, may be NULL): faces the FP64 solver finished; kernel_ms (host,
 * may be NULL): the kernel's duration from HIP events.  Same interface
 * reference as mlh_riemann_exact_batch (RiemannSolver::solve, call site
 * Riemann.cpp:84,93-94). */
int mlh_riemann_exact_batch_device(mlh_ctx *ctx, int64_t n, double gamma,
                                   const double *states, double *star,
                                   int32_t *flag, int32_t *iters, int32_t precision,
                                   int32_t seed, double tol, int64_t *fallbacks,
                                   float *kernel_ms);

/* Allocation-free variant of mlh_riemann_exact_batch_device.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as mlh_riemann_exact_batch_device
 * except that it reports no kernel time.  It calls no hipMalloc / hipFree:
 * its scratch (one fallback count per 256 faces and their total, 2 MiB) is
 * a static device array that each GPU holds until the process ends, so n
 * above MLH_NOALLOC_MAX_FACES is rejected with MLH_ERR_INVALID_ARGUMENT.
 * Every call on one GPU uses that GPU's one array: calls that overlap on one
 * GPU (from several host threads or streams) race. */
#define MLH_NOALLOC_MAX_FACES (1 << 26)
int mlh_riemann_exact_batch_device_noalloc(mlh_ctx *ctx, int64_t n, double gamma,
                                           const double *states, double *star,
                                           int32_t *flag, int32_t *iters,
                                           int32_t precision, int32_t seed,
                                           double tol, int64_t *fallbacks);

/* ---- one face flux, batched (Riemann surface) ---------------------------
 * Riemann::Riemann(WR, WL, vFrame, Aij, i) + Riemann::exact(Fij, gamma)
 * (Riemann.h:19,26; Riemann.cpp:7-229) for n given faces, FP64, with ctx's
 * dim, gamma, tolerance and MFM flag.  w_left / w_right: [n][dim+2]
 * {rho, P, v[dim]} of the solver's left state (particle i: the ctor's WL
 * member, Particles.cpp:1852) and right state (particle j), in the frame
 * moving with v_frame [n][dim]; a_ij [n][dim].  f_ij: [n][dim+2] {mass,
 * energy, momentum[dim]} (Riemann.cpp:164-181).  flag: sampled side (-1
 * left, +1 right, 0 vacuum, 2 = zero face, no flux). */
int mlh_riemann_flux_batch(mlh_ctx *ctx, int64_t n, const double *w_left,
                           const double *w_right, const double *v_frame,
                           const double *a_ij, double *f_ij, int32_t *flag);

/* Converged star pressure p* of the plain FP64 exact solver for n face
 * problems (0 where vacuum forms): the training target of the learned seed. */
int mlh_riemann_pstar_batch(int64_t n, double gamma, const double *rhoL,
                            const double *uL, const double *PL,
                            const double *rhoR, const double *uR,
                            const double *PR, double tol, double *p_star,
                            int32_t *iters);

/* ---- learned p* seed: MLP 10 -> 64 -> 64 -> 1 (tanh hidden) ------------ */
#define MLH_MLP_IN 10
#define MLH_MLP_HIDDEN 64
#define MLH_MLP_NPARAM (MLH_MLP_IN * MLH_MLP_HIDDEN + MLH_MLP_HIDDEN + \
                        MLH_MLP_HIDDEN * MLH_MLP_HIDDEN + MLH_MLP_HIDDEN + \
                        MLH_MLP_HIDDEN + 1)
/* Feature vector of one face (normalised states): feats[10]; writes the
 * regression target log(p_star / p_ref) normaliser p_ref to *p_ref. */
int mlh_mlp_features(int64_t n, double gamma, const double *rhoL,
                     const double *uL, const double *PL, const double *rhoR,
                     const double *uR, const double *PR, float *feats,
                     double *p_ref);
/* Set / get the network parameters: MLH_MLP_NPARAM FP32 values W1[64][10],
 * b1[64], W2[64][64], b2[64], W3[64], b3.  Setting them resets Adam's
 * moments. */
int mlh_mlp_set_params(mlh_ctx *ctx, const float *params);
int mlh_mlp_get_params(mlh_ctx *ctx, float *params);
/* Forward + backward on a batch (device kernels), loss = mean squared
 * error of log p*; returns the loss and the gradient w.r.t. all params. */
int mlh_mlp_loss_grad(mlh_ctx *ctx, int64_t n, const float *feats,
                      const float *target, float *loss, float *grad);
/* Train with Adam for n_epochs over (feats, target), copied to the device
 * once; each epoch visits the samples in an order fixed by (seed, epoch)
 * in batches of `batch` (a partial last batch is left out). */
int mlh_mlp_train(mlh_ctx *ctx, int64_t n, const float *feats,
                  const float *target, int32_t n_epochs, int32_t batch,
                  float lr, uint64_t seed, float *final_loss);

/* Training set and training of the seed network built from face problems
 * on the device.  No counterpart in the reference, which has no learned
 * seed: these replace the host-side set construction (exact FP64 p*,
 * features, targets, subsampling of trivial faces, shuffling) so that an
 * online retraining needs no copy of the record to the host.
 *   - faces with a non-positive or non-finite input state, or p* = 0, are
 *     dropped; the target is log(p* / p_ref) - f7 in FP32, p* from the
 *     plain FP64 exact solver at tol 1e-12 (mlh_riemann_pstar_batch);
 *   - faces with |target| < 1e-5 (the standard guess converged) are
 *     trivial; all non-trivial faces are kept, and each trivial one with
 *     probability keep / n_trivial, keep = min(n_trivial,
 *     max(keep_trivial * max(n_nontrivial, 1), 1024)), decided by a hash
 *     of the face's six values;
 *   - the kept faces are ordered by that hash, so the set does not depend
 *     on the order of the input (the record's order is not deterministic);
 *     the first max(n / 10, 64) are the validation faces;
 *   - training continues from the current parameters and Adam moments.
 * With fewer than 256 kept faces nothing is trained (train_faces = 0). */
typedef struct mlh_train_info {
    int64_t faces;          /* face problems in the input                    */
    int64_t valid;          /* not dropped                                  */
    int64_t nontrivial;     /* |target| >= 1e-5                             */
    int64_t trivial_kept;   /* trivial faces kept by the subsampling        */
    int64_t train_faces;    /* 0: too few faces, nothing trained            */
    int64_t val_faces;
    int32_t epochs, batch;
    float loss_init;        /* MSE over the set's first min(n, 65536) faces */
    float loss_final;       /* the last batch's loss                        */
    double val_mse_log_pstar;       /* BF16 network on the validation faces */
    double val_mse_standard_guess;  /* the standard guess there (target 0) */
} mlh_train_info;
/* On n faces states6[n][6] (rhoL, uL, PL, rhoR, uR, PR) copied from the host. */
int mlh_mlp_train_faces(mlh_ctx *ctx, int64_t n, const double *states6,
                        double gamma, int32_t n_epochs, int32_t batch,
                        float lr, uint64_t seed, double keep_trivial,
                        mlh_train_info *info);
/* On the faces this context recorded (mlh_set_face_collection), at the
 * context's gamma; restarts the record. */
int mlh_mlp_train_recorded(mlh_ctx *ctx, int32_t n_epochs, int32_t batch,
                           float lr, uint64_t seed, double keep_trivial,
                           mlh_train_info *info);
/* The training set alone, in its order: of states6[n][6], or of the record
 * when states6 is null (the record is not restarted).  Copies the first
 * min(set size, cap) entries to feats[.][10], target and (if not null)
 * rows, each kept face's input row (for the record: its row in the order
 * mlh_get_collected_faces copies it out), and the set size to *n_out. */
int mlh_mlp_dataset(mlh_ctx *ctx, int64_t n, const double *states6,
                    double gamma, double keep_trivial, int64_t cap,
                    float *feats, float *target, int64_t *rows,
                    int64_t *n_out);
/* Inference: FP32 reference path (bf16 = 0) or BF16 MFMA path (bf16 = 1). */
int mlh_mlp_infer(mlh_ctx *ctx, int64_t n, const float *feats, float *out,
                  int32_t bf16);
/* Record the 1D face problems (rhoL,uL,PL,rhoR,uR,PR along the face normal,
 * as handed to the exact solver) of every particle i with
 * i % particle_stride == 0 during the following steps, up to cap faces
 * (particle_stride <= 0 switches recording off).  Training data for the
 * learned seed, including online retraining during a run. */
int mlh_set_face_collection(mlh_ctx *ctx, int64_t cap, int32_t particle_stride);
/* Copy the recorded faces (n x 6 doubles; states6 holds at least cap rows)
 * out and restart the recording.  The record is filled in slices, one per
 * counter, to spread the atomics; faces come out slice by slice, not in
 * the order they were solved. */
int mlh_get_collected_faces(mlh_ctx *ctx, double *states6, int64_t *n_out);

#ifdef __cplusplus
}
#endif
#endif /* MLH_H */
