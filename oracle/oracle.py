"""ctypes binding of the CPU oracle (oracle/_build/libmlh_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from meshlesshydro_amd._abi import Diag, Params, MLP_IN, MLP_NPARAM

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "libmlh_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Params), C.c_int32]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_state.argtypes = [C.c_void_p] + [_dp] * 8
        L.orc_get_state.argtypes = [C.c_void_p] + [_dp] * 8
        for f in ("orc_neighbours", "orc_density", "orc_gradients"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.orc_timestep.argtypes = [C.c_void_p, _dp]
        L.orc_fluxes.argtypes = [C.c_void_p, C.c_double]
        L.orc_update.argtypes = [C.c_void_p, C.c_double]
        L.orc_step.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp]
        L.orc_get_neighbours.argtypes = [C.c_void_p, _ip, _ip, _ip]
        L.orc_get_density.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_get_gradients.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_get_fluxes.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_conserved_sums.argtypes = [C.c_void_p, _dp]
        L.orc_get_diag.argtypes = [C.c_void_p, C.POINTER(Diag)]
        L.orc_set_bruteforce_ghosts.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_ghosts.argtypes = [C.c_void_p, C.c_int32, _ip, _ip, _dp, _dp, _dp, _ip]
        L.orc_psi_gradients.argtypes = [C.c_void_p]
        L.orc_limit.argtypes = [C.c_void_p]
        L.orc_gradient_field.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_get_face_arrays.argtypes = [C.c_void_p, C.c_double] + [_dp] * 5
        L.orc_check_flux_symmetry.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_int64),
                                              C.POINTER(C.c_int64)]
        L.orc_riemann_flux_batch.argtypes = [C.c_void_p, C.c_int64] + [_dp] * 5 + [_ip]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_riemann_solve.restype = C.c_int
        L.orc_riemann_solve.argtypes = [C.c_double] * 7 + [C.c_double, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.orc_riemann_batch.argtypes = [C.c_int64, C.c_double] + [_dp] * 6 + [C.c_double, C.c_int] + [_dp] * 3 + [_ip, _ip]
        L.orc_riemann_guess.restype = C.c_double
        L.orc_riemann_pstar.argtypes = [C.c_double] * 8 + [C.c_int, C.POINTER(C.c_int)]
        L.orc_riemann_pstar.restype = C.c_double
        L.orc_riemann_guess.argtypes = [C.c_double] * 7
        L.orc_riemann_f.argtypes = [C.c_double] * 8 + [_dp, _dp]
        L.orc_test_face_states.restype = C.c_int
        L.orc_test_face_states.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_int, C.c_double,
                                           _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_test_pairwise_limiter.restype = C.c_double
        L.orc_test_pairwise_limiter.argtypes = [C.POINTER(Params)] + [C.c_double] * 5
        L.orc_kernel.restype = C.c_double
        L.orc_kernel.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_mlp_loss_grad.argtypes = [C.c_int64, _fp, _fp, _fp, _dp]
        L.orc_mlp_forward.argtypes = [C.c_int64, _fp, _fp, _dp]
        L.orc_mlp_features.argtypes = [C.c_int64, C.c_double] + [_dp] * 6 + [_fp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f(a):
    return a.ctypes.data_as(_fp)


class Oracle:
    """The reference's MFV step on the CPU, in FP64."""

    def __init__(self, params, state, threads=None):
        self.L = lib()
        if threads:
            self.L.orc_set_threads(int(threads))
        self.p = params
        self.D = params.dim
        self.N = len(state["x"])
        self.h = self.L.orc_create(C.byref(params), self.N)
        if not self.h:
            raise RuntimeError("orc_create failed")
        self.set_state(state)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_state(self, s):
        g = lambda k: np.ascontiguousarray(s[k], dtype=np.float64) if k in s else np.zeros(self.N)
        self._keep = [g(k) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "u")]
        self.L.orc_set_state(self.h, *[_d(a) for a in self._keep])

    def state(self):
        out = {k: np.zeros(self.N) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "u")}
        self.L.orc_get_state(self.h, *[_d(out[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "u")])
        if self.D == 2:
            out.pop("z"), out.pop("vz")
        return out

    def _chk(self, st, what):
        if st != 0:
            raise OracleError(st, what)

    def neighbours(self):
        self._chk(self.L.orc_neighbours(self.h), "neighbours")

    def density(self):
        self._chk(self.L.orc_density(self.h), "density")

    def timestep(self):
        dt = C.c_double()
        self._chk(self.L.orc_timestep(self.h, C.byref(dt)), "timestep")
        return dt.value

    def gradients(self):
        self._chk(self.L.orc_gradients(self.h), "gradients")

    def fluxes(self, dt):
        st = self.L.orc_fluxes(self.h, dt)
        self._chk(st, "fluxes")

    def update(self, dt):
        self._chk(self.L.orc_update(self.h, dt), "update")

    def step(self, dt_fixed=-1.0, dt_max=-1.0):
        dt = C.c_double()
        self._chk(self.L.orc_step(self.h, dt_fixed, dt_max, C.byref(dt)), "step")
        return dt.value

    def riemann_flux(self, w_left, w_right, v_frame, a_ij):
        """Riemann::Riemann + exact on given faces: w_* = [n][D+2] {rho, P,
        v...} (left = the solver's particle i), v_frame / a_ij = [n][D].
        Returns F [n][D+2] {mass, energy, momentum...} and the sample flags."""
        D = self.D
        wl = np.ascontiguousarray(w_left, np.float64).reshape(-1, D + 2)
        wr = np.ascontiguousarray(w_right, np.float64).reshape(-1, D + 2)
        vf = np.ascontiguousarray(v_frame, np.float64).reshape(-1, D)
        A = np.ascontiguousarray(a_ij, np.float64).reshape(-1, D)
        n = len(wl)
        F = np.zeros((n, D + 2))
        fl = np.zeros(n, np.int32)
        self.L.orc_riemann_flux_batch(self.h, n, _d(wl), _d(wr), _d(vf), _d(A), _d(F), _i(fl))
        return F, fl

    def get_neighbours(self):
        stride = self.p.max_interactions + self.p.max_ghost_interactions
        noi = np.zeros(self.N, np.int32)
        nog = np.zeros(self.N, np.int32)
        nbr = np.zeros(self.N * stride, np.int32)
        self._chk(self.L.orc_get_neighbours(self.h, _i(noi), _i(nog), _i(nbr)), "get_neighbours")
        return noi, nog, nbr.reshape(self.N, stride)

    def get_density(self):
        r, P, w = np.zeros(self.N), np.zeros(self.N), np.zeros(self.N)
        self._chk(self.L.orc_get_density(self.h, _d(r), _d(P), _d(w)), "get_density")
        return r, P, w

    def get_gradients(self):
        D = self.D
        g = np.zeros((D + 2) * self.N * D)
        B = np.zeros(self.N * D * D)
        self._chk(self.L.orc_get_gradients(self.h, _d(g), _d(B)), "get_gradients")
        return g.reshape(D + 2, self.N, D), B.reshape(self.N, D, D)

    def get_fluxes(self):
        mF, vF, eF = np.zeros(self.N), np.zeros(self.N * self.D), np.zeros(self.N)
        self._chk(self.L.orc_get_fluxes(self.h, _d(mF), _d(vF), _d(eF)), "get_fluxes")
        return mF, vF.reshape(self.N, self.D), eF

    def conserved(self):
        o = np.zeros(6)
        self.L.orc_conserved_sums(self.h, _d(o))
        return o

    def diag(self):
        d = Diag()
        self.L.orc_get_diag(self.h, C.byref(d))
        return d.as_dict()

    def face_states(self, i, j, xj, dt, rho, P, grads, ghost=False):
        """compRiemannStatesLR for the pair (i, j) with the given rho, P [N]
        and gradients [D+2][N][D] in place of the stages' results: returns
        WR (i's state), WL (j's state), vFrame."""
        D = self.D
        xj = np.ascontiguousarray(xj, np.float64)
        rho = np.ascontiguousarray(rho, np.float64)
        P = np.ascontiguousarray(P, np.float64)
        grads = np.ascontiguousarray(grads, np.float64)
        assert grads.shape == (D + 2, self.N, D)
        WR, WL, vF = np.zeros(D + 2), np.zeros(D + 2), np.zeros(D)
        self._chk(self.L.orc_test_face_states(self.h, i, j, _d(xj), int(ghost), dt, _d(rho), _d(P),
                                              _d(grads), _d(WR), _d(WL), _d(vF)), "face_states")
        return WR, WL, vF

    def set_bruteforce_ghosts(self, on=True):
        self.L.orc_set_bruteforce_ghosts(self.h, int(on))

    def ghosts(self):
        """(parent, code, coordinates [n][D]) of the ghost particles of the
        last neighbour search."""
        cap = 7 * self.N
        par, code = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        g = [np.zeros(cap) for _ in range(3)]
        n = C.c_int32()
        self._chk(self.L.orc_get_ghosts(self.h, cap, _i(par), _i(code), _d(g[0]), _d(g[1]), _d(g[2]),
                                        C.byref(n)), "get_ghosts")
        k = n.value
        return par[:k].copy(), code[:k].copy(), np.stack([a[:k] for a in g[: self.D]], 1)

    def psi_gradients(self):
        self._chk(self.L.orc_psi_gradients(self.h), "psi_gradients")

    def limit(self):
        self._chk(self.L.orc_limit(self.h), "limit")

    def gradient_field(self, f):
        f = np.ascontiguousarray(f, np.float64)
        g = np.zeros(self.N * self.D)
        self._chk(self.L.orc_gradient_field(self.h, _d(f), _d(g)), "gradient_field")
        return g.reshape(self.N, self.D)

    def get_face_arrays(self, dt, which=("A", "WR", "WL", "vF", "F")):
        """Per-pair arrays of the face stages, [N][stride][components] (after
        fluxes(dt))."""
        D = self.D
        S = self.p.max_interactions + self.p.max_ghost_interactions
        nc = {"A": D, "WR": D + 2, "WL": D + 2, "vF": D, "F": D + 2}
        out = {k: np.zeros((self.N, S, nc[k])) for k in which}
        args = [_d(out[k]) if k in out else None for k in ("A", "WR", "WL", "vF", "F")]
        self._chk(self.L.orc_get_face_arrays(self.h, dt, *args), "get_face_arrays")
        return out

    def check_flux_symmetry(self, tol=1e-20):
        a, m = C.c_int64(), C.c_int64()
        self._chk(self.L.orc_check_flux_symmetry(self.h, tol, C.byref(a), C.byref(m)),
                  "check_flux_symmetry")
        return a.value, m.value


class OracleError(RuntimeError):
    def __init__(self, status, what):
        super().__init__(f"oracle {what} returned status {status}")
        self.status = status


def riemann_batch(gamma, rl, ul, pl, rr, ur, pr, tol=1e-6, max_iter=50):
    L = lib()
    n = len(rl)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rl, ul, pl, rr, ur, pr)]
    rs, us, ps = np.zeros(n), np.zeros(n), np.zeros(n)
    flag, it = np.zeros(n, np.int32), np.zeros(n, np.int32)
    L.orc_riemann_batch(n, gamma, *[_d(a) for a in arrs], tol, max_iter, _d(rs), _d(us), _d(ps), _i(flag), _i(it))
    return rs, us, ps, flag, it


def riemann_guess(gamma, rl, ul, pl, rr, ur, pr):
    return lib().orc_riemann_guess(gamma, rl, ul, pl, rr, ur, pr)


def riemann_f(gamma, rl, ul, pl, rr, ur, pr, p):
    f, df = C.c_double(), C.c_double()
    lib().orc_riemann_f(gamma, rl, ul, pl, rr, ur, pr, p, C.byref(f), C.byref(df))
    return f.value, df.value


def pairwise_limiter(params, phi0, phi_i, phi_j, xijxi_abs, xjxi_abs):
    """pairwiseLimiter (Particles.cpp:1735-1785) with params' PSI_1, PSI_2
    and abs() mode."""
    return lib().orc_test_pairwise_limiter(C.byref(params), phi0, phi_i, phi_j, xijxi_abs, xjxi_abs)


def kernel(dim, r, h):
    return lib().orc_kernel(dim, r, h)


def mlp_features(gamma, rl, ul, pl, rr, ur, pr):
    n = len(rl)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rl, ul, pl, rr, ur, pr)]
    feats = np.zeros((n, MLP_IN), np.float32)
    pref = np.zeros(n)
    lib().orc_mlp_features(n, gamma, *[_d(a) for a in arrs], _f(feats), _d(pref))
    return feats, pref


def mlp_loss_grad(params, feats, target):
    params = np.ascontiguousarray(params, np.float32)
    feats = np.ascontiguousarray(feats, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    g = np.zeros(MLP_NPARAM)
    loss = lib().orc_mlp_loss_grad(len(target), _f(params), _f(feats), _f(target), _d(g))
    return loss, g


def mlp_forward(params, feats):
    params = np.ascontiguousarray(params, np.float32)
    feats = np.ascontiguousarray(feats, np.float32)
    out = np.zeros(len(feats))
    lib().orc_mlp_forward(len(feats), _f(params), _f(feats), _d(out))
    return out


def riemann_pstar(gamma, dl, ul, pl, dr, ur, pr, tol=1e-6, max_iter=50):
    """Star pressure of the oracle's exact solver (0 for vacuum) and the
    iterations it took (Newton, then bisection beyond max_iter)."""
    L = lib()
    n = len(dl)
    out, its = np.zeros(n), np.zeros(n, np.int32)
    it = C.c_int()
    for k in range(n):
        out[k] = L.orc_riemann_pstar(gamma, dl[k], ul[k], pl[k], dr[k], ur[k], pr[k], tol,
                                     max_iter, C.byref(it))
        its[k] = it.value
    return out, its
