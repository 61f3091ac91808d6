"""ctypes binding of the host C++ layer (meshlesshydro_amd/host/libmlh_host.so):
parameter.h / config.info readers, HDF5 I/O in the reference's layout and
the whole `mlh` driver.  The `mlh` executable next to the library is the
command-line form of `run`."""
import ctypes as C
import os

import numpy as np

from . import lib as _lib
from ._abi import Params

_HERE = os.path.dirname(os.path.abspath(__file__))
HOST_DIR = os.path.join(_HERE, "host")
HOST_LIB = os.path.join(HOST_DIR, "libmlh_host.so")
MLH_EXE = os.path.join(HOST_DIR, "mlh")

_L = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def hostlib():
    global _L
    if _L is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is not built: run `make -C {HOST_DIR}` "
                               "(after python -m meshlesshydro_amd.build)")
        _lib.lib()  # libmlh_hip.so first (and the system C++ runtime with it)
        L = C.CDLL(HOST_LIB)
        L.mlh_host_error.restype = C.c_char_p
        L.mlh_host_last_error.restype = C.c_char_p
        L.mlh_host_read_parameters.argtypes = [C.c_char_p, C.POINTER(Params), _ip]
        L.mlh_host_config_value.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32]
        L.mlh_host_run.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, C.POINTER(C.c_int64), _dp, _ip]
        L.mlh_host_run_methods.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, C.c_int32,
                                           C.POINTER(C.c_int64), _dp, _ip]
        L.mlh_host_ghost_images.argtypes = [C.POINTER(Params), C.c_int32, _dp, _dp, _dp, C.c_int32,
                                            _ip, _dp, _dp, _dp, _ip]
        L.mlh_h5_write_ic.argtypes = [C.c_char_p, C.c_int32, C.c_int32] + [_dp] * 4 + [_ip]
        L.mlh_h5_write_snapshot.argtypes = ([C.c_char_p, C.c_int32, C.c_int32] + [C.c_double] * 3 +
                                            [_dp] * 8 + [_ip])
        L.mlh_h5_dataset_shape.argtypes = [C.c_char_p, C.c_char_p, _ip, C.POINTER(C.c_int64)]
        L.mlh_h5_read_f64.argtypes = [C.c_char_p, C.c_char_p, _dp, C.c_int64]
        L.mlh_h5_read_i32.argtypes = [C.c_char_p, C.c_char_p, _ip, C.c_int64]
        _L = L
    return _L


class HostError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"status {status}: {msg}")
        self.status = status


def _b(s):
    return s.encode() if isinstance(s, str) else s


def read_parameters(path):
    """parameter.h -> (mlh_params, adaptive_timestep)"""
    p = Params()
    a = C.c_int32()
    if hostlib().mlh_host_read_parameters(_b(path), C.byref(p), C.byref(a)) != 0:
        raise HostError(-1, hostlib().mlh_host_error().decode())
    return p, bool(a.value)


def config_value(path, key):
    buf = C.create_string_buffer(4096)
    if hostlib().mlh_host_config_value(_b(path), _b(key), buf, len(buf)) != 0:
        raise HostError(-1, hostlib().mlh_host_error().decode())
    return buf.value.decode()


def write_ic(path, state, dim):
    """Initial distribution file (/m, /x, /v, /u, /materialId) from a state dict."""
    n = len(state["x"])
    ax = ("x", "y", "z")[:dim]
    vx = ("vx", "vy", "vz")[:dim]
    X = np.ascontiguousarray(np.stack([state[k] for k in ax], 1), np.float64)
    V = np.ascontiguousarray(np.stack([state[k] for k in vx], 1), np.float64)
    m = np.ascontiguousarray(state["m"], np.float64)
    u = np.ascontiguousarray(state["u"], np.float64)
    mat = np.zeros(n, np.int32)
    d = lambda a: a.ctypes.data_as(_dp)
    if hostlib().mlh_h5_write_ic(_b(path), n, dim, d(X), d(V), d(m), d(u), mat.ctypes.data_as(_ip)) != 0:
        raise HostError(-1, hostlib().mlh_host_last_error().decode())


def write_snapshot(path, snap, dim):
    """Snapshot file in the dump2file layout from a dict with `time`,
    `totalMass`, `energy`, `momentum` (dim values) and per-particle `rho`,
    `m`, `u`, `P`, `noi`, `x`, `v`, `rhoGrad` ([N][dim])."""
    f = lambda k: np.ascontiguousarray(snap[k], np.float64)
    n = len(snap["m"])
    arrs = {k: f(k) for k in ("momentum", "rho", "m", "u", "x", "v", "rhoGrad", "P")}
    noi = np.ascontiguousarray(snap["noi"], np.int32)
    d = lambda k: arrs[k].ctypes.data_as(_dp)
    if hostlib().mlh_h5_write_snapshot(_b(path), n, dim, float(snap["time"]), float(snap["totalMass"]),
                                       float(snap["energy"]), d("momentum"), d("rho"), d("m"), d("u"),
                                       d("x"), d("v"), d("rhoGrad"), d("P"),
                                       noi.ctypes.data_as(_ip)) != 0:
        raise HostError(-1, hostlib().mlh_host_last_error().decode())


def read_dataset(path, name):
    L = hostlib()
    nd = C.c_int32()
    dims = (C.c_int64 * 2)()
    if L.mlh_h5_dataset_shape(_b(path), _b(name), C.byref(nd), dims) != 0:
        raise HostError(-1, L.mlh_host_last_error().decode())
    shape = tuple(dims[k] for k in range(nd.value))
    count = int(np.prod(shape)) if shape else 1
    out = np.zeros(count, np.float64)
    if L.mlh_h5_read_f64(_b(path), _b(name), out.ctypes.data_as(_dp), count) != 0:
        raise HostError(-1, L.mlh_host_last_error().decode())
    return out.reshape(shape)


def read_dataset_int(path, name):
    L = hostlib()
    v = read_dataset(path, name)  # shape query
    out = np.zeros(v.size, np.int32)
    if L.mlh_h5_read_i32(_b(path), _b(name), out.ctypes.data_as(_ip), v.size) != 0:
        raise HostError(-1, L.mlh_host_last_error().decode())
    return out.reshape(v.shape)


def run(config, parameter_h=None, log_level=2, per_method=False):
    """The mlh driver in-process: returns (steps, t_end, dumps); raises
    HostError with the reference's exit code (1, 3, 6, 9) or the status.
    per_method: MeshlessScheme calls the Particles methods one at a time in
    the reference's order (the driver's --per-method)."""
    steps, t, dumps = C.c_int64(), C.c_double(), C.c_int32()
    st = hostlib().mlh_host_run_methods(_b(config), _b(parameter_h) if parameter_h else None,
                                        log_level, int(per_method), C.byref(steps), C.byref(t),
                                        C.byref(dumps))
    if st != 0:
        raise HostError(st, hostlib().mlh_host_error().decode())
    return steps.value, t.value, dumps.value


def ghost_images(params, state):
    """createGhostParticles on the host (Particles::createGhostParticles of
    the C++ mirror): (parent [n], coordinates [n][dim]) of the periodic
    images of the particles in `state`, in the reference's ghost order."""
    N = len(state["x"])
    g = lambda k: np.ascontiguousarray(state[k], np.float64) if k in state else np.zeros(N)
    xs = [g(k) for k in ("x", "y", "z")]
    cap = 7 * N
    par = np.zeros(cap, np.int32)
    out = [np.zeros(cap) for _ in range(3)]
    n = C.c_int32()
    st = hostlib().mlh_host_ghost_images(C.byref(params), N, *[a.ctypes.data_as(_dp) for a in xs], cap,
                                         par.ctypes.data_as(_ip), *[a.ctypes.data_as(_dp) for a in out],
                                         C.byref(n))
    if st != 0:
        raise HostError(st, hostlib().mlh_host_error().decode())
    k = n.value
    return par[:k].copy(), np.stack([a[:k] for a in out[: params.dim]], 1)
