"""ctypes mirrors of the C ABI structs in include/mlh.h.

mlh_params carries the reference's compile-time flags (parameter.h,
demonstrator/include/parameter.h:9-77) and the config.info values
(demonstrator/config.info:2-33) as one runtime struct.
"""
import ctypes as C

MLH_OK = 0
MLH_ERR_NEIGHBOUR_OVERFLOW = 1
MLH_ERR_GHOST_OVERFLOW = 3
MLH_ERR_NEGATIVE_FACE_PRESSURE = 6
MLH_ERR_DUMP_OVERRUN = 9
MLH_ERR_INVALID_ARGUMENT = 20
MLH_ERR_DEVICE = 21
MLH_ERR_GRID = 22
MLH_ERR_CALL_ORDER = 23
MLH_ERR_NOT_TRAINED = 24

MLH_FP64 = 0
MLH_FP32 = 1
MLH_SEED_STANDARD = 0
MLH_SEED_LEARNED = 1
MLH_ABS_INT_X86 = 0
MLH_ABS_FABS = 1
MLH_ORDER_AUTO, MLH_ORDER_CELL, MLH_ORDER_CALLER = 0, 1, 2

MLP_IN = 10
MLP_HIDDEN = 64
MLP_NPARAM = MLP_IN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN + 1


class Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("periodic", C.c_int32),
        ("box_min", C.c_double * 3),
        ("box_max", C.c_double * 3),
        ("kernel_size", C.c_double),
        ("gamma", C.c_double),
        ("cfl", C.c_double),
        ("max_interactions", C.c_int32),
        ("max_ghost_interactions", C.c_int32),
        ("slope_limiting", C.c_int32),
        ("beta", C.c_double),
        ("pairwise_limiter", C.c_int32),
        ("psi1", C.c_double),
        ("psi2", C.c_double),
        ("mfm", C.c_int32),
        ("enforce_flux_sym", C.c_int32),
        ("move_particles", C.c_int32),
        ("first_order_quad_point", C.c_int32),
        ("debug_level", C.c_int32),
        ("limiter_abs", C.c_int32),
        ("quirk_3d_predictor", C.c_int32),
        ("precision", C.c_int32),
        ("riemann_seed", C.c_int32),
        ("riemann_tol", C.c_double),
        ("riemann_max_iter", C.c_int32),
        ("learned_iters", C.c_int32),
        ("work_order", C.c_int32),
    ]


class TrainInfo(C.Structure):
    """mlh_train_info: one training of the seed network on a device-built set."""
    _fields_ = [
        ("faces", C.c_int64),
        ("valid", C.c_int64),
        ("nontrivial", C.c_int64),
        ("trivial_kept", C.c_int64),
        ("train_faces", C.c_int64),
        ("val_faces", C.c_int64),
        ("epochs", C.c_int32),
        ("batch", C.c_int32),
        ("loss_init", C.c_float),
        ("loss_final", C.c_float),
        ("val_mse_log_pstar", C.c_double),
        ("val_mse_standard_guess", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Diag(C.Structure):
    _fields_ = [
        ("neg_face_pressure_real", C.c_int64),
        ("neg_face_pressure_ghost", C.c_int64),
        ("vacuum_faces", C.c_int64),
        ("nonconverged_faces", C.c_int64),
        ("missing_mirror_pairs", C.c_int64),
        ("learned_fallbacks", C.c_int64),
        ("newton_iterations", C.c_int64),
        ("faces_solved", C.c_int64),
        ("learned_seeded", C.c_int64),
        ("max_neighbours", C.c_int32),
        ("max_ghost_neighbours", C.c_int32),
        ("singular_matrices", C.c_int64),
        ("nonfinite_particles", C.c_int64),
        ("acoustic_faces", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def default_params(**kw):
    """Flags of testcases/kelvin-helmholtz/parameter_long_run.h and the values
    of config_long_run.info (box [0,1]^2, kernelSize .04, gamma 5/3), the only
    MFV set the reference compiles.  Same values as mlh_params_default()."""
    p = Params()
    p.dim = 2
    p.periodic = 1
    for a in range(3):
        p.box_min[a] = 0.0
        p.box_max[a] = 1.0
    p.kernel_size = 0.04
    p.gamma = 5.0 / 3.0
    p.cfl = 0.4
    p.max_interactions = 400
    p.max_ghost_interactions = 300
    p.slope_limiting = 1
    p.beta = 4.0
    p.pairwise_limiter = 0
    p.psi1 = 0.5
    p.psi2 = 0.25
    p.mfm = 0
    p.enforce_flux_sym = 1
    p.move_particles = 1
    p.first_order_quad_point = 1
    p.debug_level = 1
    p.limiter_abs = MLH_ABS_INT_X86
    p.quirk_3d_predictor = 0
    p.precision = MLH_FP64
    p.riemann_seed = MLH_SEED_STANDARD
    p.riemann_tol = 1e-6
    p.riemann_max_iter = 50
    p.learned_iters = 3
    p.work_order = MLH_ORDER_AUTO
    for k, v in kw.items():
        if k in ("box_min", "box_max"):
            for a, x in enumerate(v):
                getattr(p, k)[a] = x
        else:
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
    return p
