"""ctypes binding of libmlh_hip.so (the C ABI of include/mlh.h).

The library is the product: this module only marshals numpy arrays.  It
raises if the HIP library is missing; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from ._abi import Diag, Params, TrainInfo, MLP_IN, MLP_NPARAM, MLH_FP64, MLH_SEED_STANDARD

_HERE = os.path.dirname(os.path.abspath(__file__))
# MLH_LIBRARY selects another build of the same library (e.g. a variant
# compiled with different launch attributes); default: the in-tree build
LIB_PATH = os.environ.get("MLH_LIBRARY") or os.path.join(_HERE, "libmlh_hip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p

# symbol -> (restype, argtypes); the complete exported C ABI
SIGNATURES = {
    "mlh_params_default": (None, [C.POINTER(Params)]),
    "mlh_version": (C.c_int, []),
    "mlh_create": (C.c_int, [C.POINTER(Params), C.c_int32, C.POINTER(_vp)]),
    "mlh_destroy": (C.c_int, [_vp]),
    "mlh_last_error": (C.c_char_p, []),
    "mlh_upload_state": (C.c_int, [_vp] + [_dp] * 8),
    "mlh_download_state": (C.c_int, [_vp] + [_dp] * 8),
    "mlh_step": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, _dp, _ip]),
    "mlh_run": (C.c_int, [_vp, C.c_int32, C.c_double, _dp, _dp]),
    "mlh_neighbours": (C.c_int, [_vp]),
    "mlh_density": (C.c_int, [_vp]),
    "mlh_timestep": (C.c_int, [_vp, _dp]),
    "mlh_gradients": (C.c_int, [_vp]),
    "mlh_fluxes": (C.c_int, [_vp, C.c_double]),
    "mlh_update": (C.c_int, [_vp]),
    "mlh_psi_gradients": (C.c_int, [_vp]),
    "mlh_limit": (C.c_int, [_vp]),
    "mlh_gradient_field": (C.c_int, [_vp, _dp, _dp]),
    "mlh_effective_faces": (C.c_int, [_vp]),
    "mlh_riemann_states": (C.c_int, [_vp, C.c_double]),
    "mlh_solve_faces": (C.c_int, [_vp]),
    "mlh_collect_fluxes": (C.c_int, [_vp]),
    "mlh_check_flux_symmetry": (C.c_int, [_vp, C.c_double, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mlh_get_face_arrays": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "mlh_get_neighbours": (C.c_int, [_vp, _ip, _ip, _ip]),
    "mlh_get_density": (C.c_int, [_vp, _dp, _dp, _dp]),
    "mlh_get_gradients": (C.c_int, [_vp, _dp, _dp]),
    "mlh_get_fluxes": (C.c_int, [_vp, _dp, _dp, _dp]),
    "mlh_get_diagnostics": (C.c_int, [_vp, C.POINTER(Diag)]),
    "mlh_conserved_sums": (C.c_int, [_vp, _dp]),
    "mlh_get_timings": (C.c_int, [_vp, _fp]),
    "mlh_riemann_exact_batch": (C.c_int, [_vp, C.c_int64, C.c_double] + [_dp] * 9 + [_ip, _ip, C.c_int32, C.c_int32, C.c_double]),
    "mlh_riemann_flux_batch": (C.c_int, [_vp, C.c_int64] + [_dp] * 5 + [_ip]),
    "mlh_device_alloc": (C.c_int, [C.c_int64, C.POINTER(C.c_void_p)]),
    "mlh_device_free": (C.c_int, [C.c_void_p]),
    "mlh_device_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32]),
    "mlh_riemann_exact_batch_device": (C.c_int, [_vp, C.c_int64, C.c_double, C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                                 C.c_double, C.POINTER(C.c_int64), _fp]),
    "mlh_riemann_exact_batch_device_noalloc": (C.c_int, [_vp, C.c_int64, C.c_double, C.c_void_p,
                                                         C.c_void_p, C.c_void_p, C.c_void_p,
                                                         C.c_int32, C.c_int32, C.c_double,
                                                         C.POINTER(C.c_int64)]),
    "mlh_mlp_features": (C.c_int, [C.c_int64, C.c_double] + [_dp] * 6 + [_fp, _dp]),
    "mlh_mlp_set_params": (C.c_int, [_vp, _fp]),
    "mlh_mlp_get_params": (C.c_int, [_vp, _fp]),
    "mlh_mlp_loss_grad": (C.c_int, [_vp, C.c_int64, _fp, _fp, _fp, _fp]),
    "mlh_mlp_train": (C.c_int, [_vp, C.c_int64, _fp, _fp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _fp]),
    "mlh_mlp_infer": (C.c_int, [_vp, C.c_int64, _fp, _fp, C.c_int32]),
    "mlh_mlp_train_faces": (C.c_int, [_vp, C.c_int64, _dp, C.c_double, C.c_int32, C.c_int32, C.c_float,
                                      C.c_uint64, C.c_double, C.POINTER(TrainInfo)]),
    "mlh_mlp_train_recorded": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, C.c_double,
                                         C.POINTER(TrainInfo)]),
    "mlh_mlp_dataset": (C.c_int, [_vp, C.c_int64, _dp, C.c_double, C.c_double, C.c_int64, _fp, _fp,
                                  C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mlh_riemann_pstar_batch": (C.c_int, [C.c_int64, C.c_double] + [_dp] * 6 + [C.c_double, _dp, _ip]),
    "mlh_set_face_collection": (C.c_int, [_vp, C.c_int64, C.c_int32]),
    "mlh_get_collected_faces": (C.c_int, [_vp, _dp, C.POINTER(C.c_int64)]),
}

_lib = None


class MLHError(RuntimeError):
    def __init__(self, status, what, msg=""):
        super().__init__(f"{what}: status {status} {msg}")
        self.status = status


def lib():
    """Load libmlh_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run python -m meshlesshydro_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f(a):
    return a.ctypes.data_as(_fp)


def _chk(st, what):
    if st != 0:
        raise MLHError(st, what, lib().mlh_last_error().decode())


class Simulation:
    """One MFV context on the GPU: the Particles state plus the step."""

    def __init__(self, params, state):
        self.L = lib()
        self.p = params
        self.D = params.dim
        self.N = len(state["x"])
        h = _vp()
        _chk(self.L.mlh_create(C.byref(params), self.N, C.byref(h)), "mlh_create")
        self.h = h
        self.upload(state)

    def close(self):
        if getattr(self, "h", None):
            self.L.mlh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def upload(self, s):
        g = lambda k: np.ascontiguousarray(s[k], dtype=np.float64) if k in s else np.zeros(self.N)
        keep = [g(k) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "u")]
        _chk(self.L.mlh_upload_state(self.h, *[_d(a) for a in keep]), "mlh_upload_state")

    def state(self):
        out = {k: np.zeros(self.N) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "u")}
        _chk(self.L.mlh_download_state(self.h, *[_d(out[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "u")]), "download")
        if self.D == 2:
            out.pop("z"), out.pop("vz")
        return out

    def step(self, dt_fixed=-1.0, t_now=-1.0, t_clip=-1.0):
        dt = C.c_double()
        cl = C.c_int32()
        _chk(self.L.mlh_step(self.h, dt_fixed, t_now, t_clip, C.byref(dt), C.byref(cl)), "mlh_step")
        return dt.value

    def run(self, n_steps, dt_fixed=-1.0, dts_in=None):
        dts = np.zeros(max(n_steps, 1))
        din = None
        if dts_in is not None:
            din = np.ascontiguousarray(dts_in[:n_steps], dtype=np.float64)
            assert len(din) == n_steps
        _chk(self.L.mlh_run(self.h, n_steps, dt_fixed, _d(din) if din is not None else None, _d(dts)), "mlh_run")
        return dts[:n_steps]

    # stages
    def neighbours(self):
        _chk(self.L.mlh_neighbours(self.h), "mlh_neighbours")

    def density(self):
        _chk(self.L.mlh_density(self.h), "mlh_density")

    def timestep(self):
        dt = C.c_double()
        _chk(self.L.mlh_timestep(self.h, C.byref(dt)), "mlh_timestep")
        return dt.value

    def gradients(self):
        _chk(self.L.mlh_gradients(self.h), "mlh_gradients")

    def fluxes(self, dt):
        _chk(self.L.mlh_fluxes(self.h, dt), "mlh_fluxes")

    def update(self):
        _chk(self.L.mlh_update(self.h), "mlh_update")

    # the reference's methods one at a time
    def psi_gradients(self):
        """compPsijTilde + gradient x (D+2), unlimited."""
        _chk(self.L.mlh_psi_gradients(self.h), "mlh_psi_gradients")

    def limit(self):
        """slopeLimiter on the gradients of psi_gradients()."""
        _chk(self.L.mlh_limit(self.h), "mlh_limit")

    def gradient_field(self, f):
        """gradient(f, grad) of any field f [N] (caller's order) -> [N][D]."""
        f = np.ascontiguousarray(f, np.float64)
        assert f.shape == (self.N,)
        g = np.zeros(self.N * self.D)
        _chk(self.L.mlh_gradient_field(self.h, _d(f), _d(g)), "mlh_gradient_field")
        return g.reshape(self.N, self.D)

    def effective_faces(self):
        _chk(self.L.mlh_effective_faces(self.h), "mlh_effective_faces")

    def riemann_states(self, dt):
        _chk(self.L.mlh_riemann_states(self.h, dt), "mlh_riemann_states")

    def solve_faces(self):
        _chk(self.L.mlh_solve_faces(self.h), "mlh_solve_faces")

    def collect_fluxes(self):
        _chk(self.L.mlh_collect_fluxes(self.h), "mlh_collect_fluxes")

    def check_flux_symmetry(self, tol=1e-20):
        asym, miss = C.c_int64(), C.c_int64()
        _chk(self.L.mlh_check_flux_symmetry(self.h, tol, C.byref(asym), C.byref(miss)),
             "mlh_check_flux_symmetry")
        return asym.value, miss.value

    def get_face_arrays(self, which=("A", "WR", "WL", "vF", "F")):
        """Per-entry arrays of the face stages, [N][stride][components]."""
        D = self.D
        S = self.p.max_interactions + self.p.max_ghost_interactions
        nc = {"A": D, "WR": D + 2, "WL": D + 2, "vF": D, "F": D + 2}
        out = {k: np.zeros((self.N, S, nc[k])) for k in which}
        args = [_d(out[k]) if k in out else None for k in ("A", "WR", "WL", "vF", "F")]
        _chk(self.L.mlh_get_face_arrays(self.h, *args), "mlh_get_face_arrays")
        return out

    def get_neighbours(self):
        stride = self.p.max_interactions + self.p.max_ghost_interactions
        noi = np.zeros(self.N, np.int32)
        nog = np.zeros(self.N, np.int32)
        nbr = np.zeros(self.N * stride, np.int32)
        _chk(self.L.mlh_get_neighbours(self.h, _i(noi), _i(nog), _i(nbr)), "get_neighbours")
        return noi, nog, nbr.reshape(self.N, stride)

    def get_density(self):
        r, P, w = np.zeros(self.N), np.zeros(self.N), np.zeros(self.N)
        _chk(self.L.mlh_get_density(self.h, _d(r), _d(P), _d(w)), "get_density")
        return r, P, w

    def get_gradients(self):
        D = self.D
        g = np.zeros((D + 2) * self.N * D)
        B = np.zeros(self.N * D * D)
        _chk(self.L.mlh_get_gradients(self.h, _d(g), _d(B)), "get_gradients")
        return g.reshape(D + 2, self.N, D), B.reshape(self.N, D, D)

    def get_fluxes(self):
        mF, vF, eF = np.zeros(self.N), np.zeros(self.N * self.D), np.zeros(self.N)
        _chk(self.L.mlh_get_fluxes(self.h, _d(mF), _d(vF), _d(eF)), "get_fluxes")
        return mF, vF.reshape(self.N, self.D), eF

    def conserved(self):
        o = np.zeros(6)
        _chk(self.L.mlh_conserved_sums(self.h, _d(o)), "conserved")
        return o

    def diag(self):
        d = Diag()
        _chk(self.L.mlh_get_diagnostics(self.h, C.byref(d)), "diag")
        return d.as_dict()

    def timings(self):
        t = np.zeros(4, np.float32)
        _chk(self.L.mlh_get_timings(self.h, _f(t)), "timings")
        return t

    # learned seed
    def mlp_set_params(self, params):
        params = np.ascontiguousarray(params, np.float32)
        assert params.size == MLP_NPARAM
        _chk(self.L.mlh_mlp_set_params(self.h, _f(params)), "mlp_set_params")

    def mlp_get_params(self):
        out = np.zeros(MLP_NPARAM, np.float32)
        _chk(self.L.mlh_mlp_get_params(self.h, _f(out)), "mlp_get_params")
        return out

    def mlp_loss_grad(self, feats, target):
        feats = np.ascontiguousarray(feats, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        _chk(self.L.mlh_mlp_loss_grad(self.h, len(target), _f(feats), _f(target), _f(loss), _f(g)), "mlp_loss_grad")
        return float(loss[0]), g

    def mlp_train(self, feats, target, epochs=1, batch=4096, lr=1e-3, seed=1):
        feats = np.ascontiguousarray(feats, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        loss = np.zeros(1, np.float32)
        _chk(self.L.mlh_mlp_train(self.h, len(target), _f(feats), _f(target), epochs, batch, lr, seed, _f(loss)), "mlp_train")
        return float(loss[0])

    def mlp_infer(self, feats, bf16=True):
        feats = np.ascontiguousarray(feats, np.float32)
        out = np.zeros(len(feats), np.float32)
        _chk(self.L.mlh_mlp_infer(self.h, len(feats), _f(feats), _f(out), int(bf16)), "mlp_infer")
        return out

    def mlp_train_faces(self, faces, gamma, epochs=4, batch=4096, lr=2e-3, seed=1, keep_trivial=1.0):
        """Build the training set of faces [n, 6] on the device and train
        the network on it (mlh_mlp_train_faces); returns the info dict."""
        faces = np.ascontiguousarray(faces, np.float64).reshape(-1, 6)
        info = TrainInfo()
        _chk(self.L.mlh_mlp_train_faces(self.h, len(faces), _d(faces), gamma, epochs, batch, lr, seed,
                                        keep_trivial, C.byref(info)), "mlp_train_faces")
        return info.as_dict()

    def mlp_train_recorded(self, epochs=2, batch=4096, lr=2e-3, seed=1, keep_trivial=1.0):
        """Train on the faces this context recorded and restart the record
        (mlh_mlp_train_recorded); returns the info dict."""
        info = TrainInfo()
        _chk(self.L.mlh_mlp_train_recorded(self.h, epochs, batch, lr, seed, keep_trivial, C.byref(info)),
             "mlp_train_recorded")
        return info.as_dict()

    def mlp_dataset(self, faces=None, gamma=None, keep_trivial=1.0, rows=False):
        """The device-built training set (features [n, 10], targets [n]) of
        faces [n, 6], or of the record when faces is None; with rows=True
        also each kept face's input row."""
        gamma = self.p.gamma if gamma is None else gamma
        n = C.c_int64()
        if faces is None:
            cap = max(getattr(self, "_collect_cap", 0), 0)
            src, nf = None, 0
        else:
            src = np.ascontiguousarray(faces, np.float64).reshape(-1, 6)
            cap, nf = len(src), len(src)
        feats = np.zeros((max(cap, 1), MLP_IN), np.float32)
        target = np.zeros(max(cap, 1), np.float32)
        r = np.zeros(max(cap, 1), np.int64)
        _chk(self.L.mlh_mlp_dataset(self.h, nf, None if src is None else _d(src), gamma, keep_trivial, cap,
                                    _f(feats), _f(target), r.ctypes.data_as(C.POINTER(C.c_int64)),
                                    C.byref(n)), "mlp_dataset")
        k = n.value
        out = (feats[:k].copy(), target[:k].copy())
        return out + (r[:k].copy(),) if rows else out

    def set_face_collection(self, cap, particle_stride):
        self._collect_cap = int(cap)
        _chk(self.L.mlh_set_face_collection(self.h, int(cap), int(particle_stride)), "set_face_collection")

    def collected_faces(self):
        cap = getattr(self, "_collect_cap", 0)
        out = np.zeros((max(cap, 1), 6))
        n = C.c_int64()
        _chk(self.L.mlh_get_collected_faces(self.h, _d(out), C.byref(n)), "get_collected_faces")
        return out[: n.value].copy()

    def riemann_flux(self, w_left, w_right, v_frame, a_ij):
        """Riemann::Riemann + exact on given faces (device, FP64): w_* [n][D+2]
        {rho, P, v...}, v_frame / a_ij [n][D] -> F [n][D+2], flags [n]."""
        D = self.p.dim
        wl = np.ascontiguousarray(w_left, np.float64).reshape(-1, D + 2)
        wr = np.ascontiguousarray(w_right, np.float64).reshape(-1, D + 2)
        vf = np.ascontiguousarray(v_frame, np.float64).reshape(-1, D)
        A = np.ascontiguousarray(a_ij, np.float64).reshape(-1, D)
        n = len(wl)
        F = np.zeros((n, D + 2))
        fl = np.zeros(n, np.int32)
        _chk(lib().mlh_riemann_flux_batch(self.h, n, _d(wl), _d(wr), _d(vf), _d(A), _d(F), _i(fl)),
             "mlh_riemann_flux_batch")
        return F, fl

    def riemann(self, rl, ul, pl, rr, ur, pr, precision=MLH_FP64, seed=MLH_SEED_STANDARD, tol=1e-6):
        return riemann_batch(self.p.gamma, rl, ul, pl, rr, ur, pr, precision, seed, tol, ctx=self.h)


def riemann_batch(gamma, rl, ul, pl, rr, ur, pr, precision=MLH_FP64, seed=MLH_SEED_STANDARD, tol=1e-6, ctx=None):
    L = lib()
    n = len(rl)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rl, ul, pl, rr, ur, pr)]
    rs, us, ps = np.zeros(n), np.zeros(n), np.zeros(n)
    flag, it = np.zeros(n, np.int32), np.zeros(n, np.int32)
    _chk(L.mlh_riemann_exact_batch(ctx, n, gamma, *[_d(a) for a in arrs], _d(rs), _d(us), _d(ps), _i(flag), _i(it), precision, seed, tol), "mlh_riemann_exact_batch")
    return rs, us, ps, flag, it


class DeviceBuffer:
    """A device allocation of the library's HIP runtime (mlh_device_alloc)
    holding an array of `shape` and numpy `dtype`."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(x) for x in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _chk(lib().mlh_device_alloc(self.nbytes, C.byref(p)), "mlh_device_alloc")
        self.ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.shape, a.dtype)
        _chk(lib().mlh_device_copy(b.ptr, a.ctypes.data_as(C.c_void_p), b.nbytes, 1), "mlh_device_copy")
        return b

    def copy_from(self, a, offset_bytes=0):
        a = np.ascontiguousarray(a)
        if offset_bytes < 0 or offset_bytes + a.nbytes > self.nbytes:
            raise ValueError("copy past the buffer")
        dst = C.c_void_p(self.ptr.value + offset_bytes)
        _chk(lib().mlh_device_copy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, 1), "mlh_device_copy")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _chk(lib().mlh_device_copy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 2), "mlh_device_copy")
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().mlh_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def riemann_batch_device(gamma, states, star, flag=None, iters=None, precision=MLH_FP64,
                         seed=MLH_SEED_STANDARD, tol=1e-6, ctx=None):
    """mlh_riemann_exact_batch_device on DeviceBuffers: float64 states
    [6, n] and star [3, n], int32 flag / iters [n] or None.  Returns
    (fallbacks, kernel milliseconds)."""
    n = states.shape[1]
    if states.dtype != np.float64 or star.dtype != np.float64 or states.shape != (6, n) \
            or star.shape != (3, n):
        raise ValueError("states [6, n] and star [3, n] must be float64 device buffers")
    for b in (flag, iters):
        if b is not None and (b.dtype != np.int32 or b.shape != (n,)):
            raise ValueError("flag / iters must be int32 device buffers of n entries")
    fb = C.c_int64()
    ms = (C.c_float * 1)()
    ptr = lambda b: b.ptr if b is not None else None
    _chk(lib().mlh_riemann_exact_batch_device(ctx, n, gamma, states.ptr, star.ptr, ptr(flag), ptr(iters),
                                              precision, seed, tol, C.byref(fb), ms),
         "mlh_riemann_exact_batch_device")
    return fb.value, float(ms[0])


def riemann_batch_device_noalloc(gamma, states, star, flag=None, iters=None, precision=MLH_FP64,
                                 seed=MLH_SEED_STANDARD, tol=1e-6, ctx=None, n=None):
    """mlh_riemann_exact_batch_device_noalloc on DeviceBuffers laid out as
    for riemann_batch_device ([6, n] states, [3, n] star); n defaults to the
    buffers' length.  Returns the fallback count."""
    if n is None:
        n = states.shape[1]
    fb = C.c_int64()
    ptr = lambda b: b.ptr if b is not None else None
    _chk(lib().mlh_riemann_exact_batch_device_noalloc(ctx, n, gamma, ptr(states), ptr(star), ptr(flag),
                                                      ptr(iters), precision, seed, tol, C.byref(fb)),
         "mlh_riemann_exact_batch_device_noalloc")
    return fb.value


def riemann_pstar(gamma, rl, ul, pl, rr, ur, pr, tol=1e-6):
    L = lib()
    n = len(rl)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rl, ul, pl, rr, ur, pr)]
    ps = np.zeros(n)
    it = np.zeros(n, np.int32)
    _chk(L.mlh_riemann_pstar_batch(n, gamma, *[_d(a) for a in arrs], tol, _d(ps), _i(it)), "mlh_riemann_pstar_batch")
    return ps, it


def mlp_features(gamma, rl, ul, pl, rr, ur, pr):
    n = len(rl)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rl, ul, pl, rr, ur, pr)]
    feats = np.zeros((n, MLP_IN), np.float32)
    pref = np.zeros(n)
    _chk(lib().mlh_mlp_features(n, gamma, *[_d(a) for a in arrs], _f(feats), _d(pref)), "mlh_mlp_features")
    return feats, pref


def default_params_c():
    p = Params()
    lib().mlh_params_default(C.byref(p))
    return p
