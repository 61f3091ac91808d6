"""Decaying turbulent box: a particle glass from seeded jitter, a Gaussian
random solenoidal velocity field, and the run with online retraining of the
learned p* seed.

Workload: N = n^3 particles in the periodic unit cube, rho = 1, P = 1,
gamma = 5/3; the lattice is jittered by U(-0.3, 0.3) dx per axis (seed
4242) and relaxed towards a glass; the velocity is a Gaussian random
solenoidal field with power spectrum |v(k)|^2 ~ k^-4 on integer wave
numbers 2 <= |k| <= 8, normalised to rms Mach 1 (c_s = sqrt(gamma P /
rho)); h = 2.48 dx gives about 64 neighbours.

The glass relaxation and the run use the HIP library (its density and
gradient stages, its step); nothing here runs on the CPU oracle.
"""
import time

import numpy as np

from . import learned as L
from . import lib
from ._abi import MLH_FP64, default_params

GAMMA = 5.0 / 3.0


def kernel_size(n, neighbours=64.0):
    """h for a mean of `neighbours` particles within h at number density n^3:
    (4 pi / 3) h^3 n^3 = neighbours."""
    return (3.0 * neighbours / (4.0 * np.pi)) ** (1.0 / 3.0) / n


def jittered_lattice(n, amplitude=0.3, seed=4242):
    """Cell-centred n^3 lattice in [0, 1)^3 (x outer, z inner), each
    coordinate displaced by U(-amplitude, amplitude) dx, wrapped."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / n
    c = (np.arange(n) + 0.5) * dx
    X = np.repeat(c, n * n)
    Y = np.tile(np.repeat(c, n), n)
    Z = np.tile(c, n * n)
    J = rng.uniform(-amplitude, amplitude, size=(3, n ** 3)) * dx
    return (X + J[0]) % 1.0, (Y + J[1]) % 1.0, (Z + J[2]) % 1.0


def solenoidal_field(ng, kmin=2.0, kmax=8.0, slope=-4.0, seed=4243):
    """Gaussian random divergence-free vector field on an ng^3 periodic grid
    (points i / ng): white noise filtered to |v(k)|^2 ~ |k|^slope on
    kmin <= |k| <= kmax (integer wave numbers, k = 2 pi m), with the
    longitudinal part k (k . v) / |k|^2 removed.  Returns (3, ng, ng, ng),
    normalised to rms |v| = 1."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((3, ng, ng, ng))
    W = np.fft.rfftn(w, axes=(1, 2, 3))
    m = np.fft.fftfreq(ng) * ng
    mz = np.fft.rfftfreq(ng) * ng
    KX, KY, KZ = np.meshgrid(m, m, mz, indexing="ij")
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    K = np.sqrt(K2)
    band = (K >= kmin) & (K <= kmax)
    amp = np.where(band, np.where(K > 0, K, 1.0) ** (0.5 * slope), 0.0)
    W *= amp
    kdotw = KX * W[0] + KY * W[1] + KZ * W[2]
    inv = np.where(K2 > 0, 1.0 / np.where(K2 > 0, K2, 1.0), 0.0)
    W[0] -= KX * kdotw * inv
    W[1] -= KY * kdotw * inv
    W[2] -= KZ * kdotw * inv
    v = np.fft.irfftn(W, s=(ng, ng, ng), axes=(1, 2, 3))
    v /= np.sqrt(np.mean(np.sum(v * v, axis=0)))
    return v


def sample_periodic(field, x, y, z):
    """Trilinear interpolation of a periodic grid field (C, ng, ng, ng),
    grid points at i / ng, at positions in [0, 1)."""
    ng = field.shape[1]
    out = np.empty((field.shape[0], len(x)))
    gx, gy, gz = x * ng, y * ng, z * ng
    i0, j0, k0 = np.floor(gx).astype(np.int64), np.floor(gy).astype(np.int64), np.floor(gz).astype(np.int64)
    fx, fy, fz = gx - i0, gy - j0, gz - k0
    i0 %= ng; j0 %= ng; k0 %= ng
    i1, j1, k1 = (i0 + 1) % ng, (j0 + 1) % ng, (k0 + 1) % ng
    for c in range(field.shape[0]):
        f = field[c]
        out[c] = ((1 - fx) * ((1 - fy) * ((1 - fz) * f[i0, j0, k0] + fz * f[i0, j0, k1]) +
                              fy * ((1 - fz) * f[i0, j1, k0] + fz * f[i0, j1, k1])) +
                  fx * ((1 - fy) * ((1 - fz) * f[i1, j0, k0] + fz * f[i1, j0, k1]) +
                        fy * ((1 - fz) * f[i1, j1, k0] + fz * f[i1, j1, k1])))
    return out


def relax_glass(x, y, z, h, iterations=8, step=0.25, max_frac=0.2, max_interactions=160):
    """Relax particle positions towards uniform density on the GPU: each
    iteration computes rho and its (unlimited) gradient with the library's
    FP64 stages and moves every particle down the density gradient,
    dx = -step h^2 grad(rho) / rho, capped at max_frac times the mean
    spacing.  Returns the positions and the relative density spread
    std(rho)/mean(rho) before each iteration and after the last."""
    n = len(x)
    dx_mean = n ** (-1.0 / 3.0)
    p = default_params(dim=3, kernel_size=h, max_interactions=max_interactions,
                       max_ghost_interactions=max_interactions, debug_level=0,
                       slope_limiting=0, precision=MLH_FP64)
    z0 = np.zeros(n)
    state = dict(x=x.copy(), y=y.copy(), z=z.copy(), vx=z0, vy=z0, vz=z0,
                 m=np.full(n, 1.0 / n), u=np.full(n, 1.5))
    sim = lib.Simulation(p, state)
    spread = []
    try:
        for it in range(iterations + 1):
            sim.upload(state)
            sim.neighbours()
            sim.density()
            rho, _, _ = sim.get_density()
            spread.append(float(np.std(rho) / np.mean(rho)))
            if it == iterations:
                break
            sim.gradients()
            G, _ = sim.get_gradients()
            d = -step * h * h * G[0] / rho[:, None]
            norm = np.sqrt(np.sum(d * d, axis=1))
            cap = max_frac * dx_mean
            d *= np.minimum(1.0, cap / np.maximum(norm, 1e-300))[:, None]
            for a, k in enumerate(("x", "y", "z")):
                state[k] = (state[k] + d[:, a]) % 1.0
    finally:
        sim.close()
    return state["x"], state["y"], state["z"], spread


def turbulent_box(n, mach=1.0, relax_iterations=8, ng=None, seed=4242, rho0=1.0, P0=1.0,
                  gamma=GAMMA):
    """The decaying-turbulence initial state: glass positions, solenoidal
    velocity of rms Mach `mach`, rho0, P0.  Returns (state, info)."""
    x, y, z = jittered_lattice(n, seed=seed)
    h = kernel_size(n)
    spread = []
    if relax_iterations > 0:
        x, y, z, spread = relax_glass(x, y, z, h, iterations=relax_iterations)
    ng = ng or max(32, min(128, 2 * n))
    v = sample_periodic(solenoidal_field(ng, seed=seed + 1), x, y, z)
    cs = np.sqrt(gamma * P0 / rho0)
    v *= mach * cs / np.sqrt(np.mean(np.sum(v * v, axis=0)))
    N = n ** 3
    state = dict(x=x, y=y, z=z, vx=v[0], vy=v[1], vz=v[2], m=np.full(N, rho0 / N),
                 u=np.full(N, P0 / ((gamma - 1.0) * rho0)))
    return state, {"kernel_size": h, "density_spread": spread, "velocity_grid": ng}


def run_online(sim, steps, retrain_every=50, stride=8, epochs=2, dts=None, log=None):
    """Run `steps` steps of Simulation `sim` (learned seed) and retrain its
    network every `retrain_every` steps on the faces the run recorded since
    the last retraining (every `stride`-th particle's faces that needed
    Newton iterations), starting from the current weights.  Returns the dt
    sequence and one info dict per retraining (with its wall seconds)."""
    infos = []
    out = []
    done = 0
    sim.set_face_collection(1 << 21, stride)
    while done < steps:
        k = min(retrain_every, steps - done)
        din = None if dts is None else dts[done:done + k]
        out.append(sim.run(k, dts_in=din))
        done += k
        if done < steps:
            t0 = time.perf_counter()
            try:
                _, info = L.train_seed_from_record(sim, n_epochs=epochs, log=log)
                info["step"] = done
            except ValueError as e:  # too few recorded faces: keep the weights
                info = {"step": done, "skipped": str(e)}
            info["seconds"] = time.perf_counter() - t0
            infos.append(info)
    sim.set_face_collection(0, 0)
    return np.concatenate(out) if out else np.zeros(0), infos
