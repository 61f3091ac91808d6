"""Learned p* seed: training data from the exact solver's own faces, GPU
training, and the measured gain over the standard guess.

The network (10-64-64-1, tanh) predicts log(p*/p_guess), the correction of
Toro's adaptive PVRS/TRRS/TSRS guess, from the face features of
csrc/mlp_device.h.  With a zero output layer it reproduces the standard
guess exactly, so training starts from the method without it.  Targets are
the converged p* of the plain FP64 exact solver on the GPU; the set is
built and the network trained with Adam on the device
(mlh_mlp_train_faces, mlh_mlp_train_recorded), with dataset() as the
host construction of the same set.
"""
import numpy as np

from . import lib
from ._abi import MLP_HIDDEN, MLP_IN, MLP_NPARAM, MLH_FP32, MLH_FP64, MLH_SEED_LEARNED, MLH_SEED_STANDARD

H = MLP_HIDDEN


def init_params(seed=1):
    """Glorot-uniform hidden layers, zero output layer (= the standard guess)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_NPARAM, np.float32)
    o = 0
    lim1 = np.sqrt(6.0 / (MLP_IN + H))
    p[o:o + H * MLP_IN] = rng.uniform(-lim1, lim1, H * MLP_IN); o += H * MLP_IN
    o += H  # b1
    lim2 = np.sqrt(6.0 / (H + H))
    p[o:o + H * H] = rng.uniform(-lim2, lim2, H * H); o += H * H
    o += H  # b2
    # W3, b3 zero
    return p


_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
SELECT_SALT = np.uint64(0x5851F42D4C957F2D)


def mix64(z):
    """splitmix64's output function on uint64 arrays (wrapping arithmetic)."""
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def face_hash(faces):
    """Content hash of face problems [n, 6]: the doubles' bit patterns
    chained through mix64 (the key that orders the training set)."""
    bits = np.ascontiguousarray(faces, np.float64).view(np.uint64).reshape(-1, 6)
    h = np.zeros(len(bits), np.uint64)
    for q in range(6):
        h = mix64(h ^ bits[:, q])
    return h


def dataset(faces, gamma, keep_trivial=1.0, rows=False):
    """(features, target) of recorded faces [n, 6], built on the host with
    the rules of the device construction (mlh_mlp_dataset): target =
    log(p*/p_guess); faces whose standard guess is already converged
    (|target| < 1e-5) are kept with probability keep / n_trivial, keep =
    min(n_trivial, max(keep_trivial * max(n_nontrivial, 1), 1024)), by a
    hash of the face; the kept faces are ordered by their content hash.
    With rows=True also each kept face's row of `faces`."""
    if len(faces) == 0:
        out = np.zeros((0, MLP_IN), np.float32), np.zeros(0, np.float32)
        return out + (np.zeros(0, np.int64),) if rows else out
    faces = np.ascontiguousarray(faces, np.float64).reshape(-1, 6)
    h = face_hash(faces)
    rl, ul, pl, rr, ur, pr = (faces[:, k] for k in range(6))
    ok = (rl > 0) & (rr > 0) & (pl > 0) & (pr > 0) & np.isfinite(faces).all(axis=1)
    idx = np.flatnonzero(ok)
    ps, _ = lib.riemann_pstar(gamma, rl[idx], ul[idx], pl[idx], rr[idx], ur[idx], pr[idx], tol=1e-12)
    feats, pref = lib.mlp_features(gamma, rl[idx], ul[idx], pl[idx], rr[idx], ur[idx], pr[idx])
    good = ps > 0
    idx, feats, pref, ps = idx[good], feats[good], pref[good], ps[good]
    target = (np.log(ps / pref) - feats[:, 7]).astype(np.float32)
    trivial = np.abs(target) < np.float32(1e-5)
    ntr, nt = int(trivial.sum()), int((~trivial).sum())
    keep = min(ntr, max(int(keep_trivial * max(nt, 1)), 1024))
    thr53 = 0
    if ntr > 0:
        thr53 = 1 << 53 if keep >= ntr else int(keep / ntr * 9007199254740992.0)
    hk = h[idx]
    u53 = mix64(hk ^ SELECT_SALT) >> np.uint64(11)
    sel = ~trivial | (trivial & (u53 < np.uint64(thr53)))
    hk, feats, target, idx = hk[sel], feats[sel], target[sel], idx[sel]
    order = np.lexsort((hk & np.uint64(0xFFFFFFFF), hk >> np.uint64(33)))
    out = feats[order], target[order]
    return out + (idx[order],) if rows else out


def newton_stats(ctx_sim, gamma, faces, precision=MLH_FP32):
    """Mean Newton iterations (and fallback share) on faces, standard vs learned seed."""
    rl, ul, pl, rr, ur, pr = (faces[:, k] for k in range(6))
    out = {}
    for name, seed in (("standard", MLH_SEED_STANDARD), ("learned", MLH_SEED_LEARNED)):
        rs, us, ps, fl, it = ctx_sim.riemann(rl, ul, pl, rr, ur, pr, precision=precision, seed=seed)
        out[name] = float(np.mean(it))
    return out


def train_seed_from_faces(sim, faces, gamma, n_epochs=4, batch=4096, lr=2e-3, seed=1, log=None,
                          init=None):
    """Train the seed network of Simulation `sim` on faces [n, 6], from
    `init` (weights to continue from) or from init_params(seed): the set is
    built and trained on the device (mlh_mlp_train_faces).  Returns
    (params, info); ValueError when fewer than 256 faces are kept."""
    sim.mlp_set_params(init_params(seed) if init is None else np.asarray(init, np.float32))
    info = sim.mlp_train_faces(faces, gamma, epochs=n_epochs, batch=batch, lr=lr, seed=seed)
    return _finish(sim, info, log)


def train_seed_from_record(sim, n_epochs=2, batch=4096, lr=2e-3, seed=1, log=None):
    """Retrain the seed network of `sim` on the faces it recorded since the
    record last restarted, continuing from its current weights with fresh
    Adam moments, without copying the record to the host
    (mlh_mlp_train_recorded).  Returns (params, info)."""
    sim.mlp_set_params(sim.mlp_get_params())
    info = sim.mlp_train_recorded(epochs=n_epochs, batch=batch, lr=lr, seed=seed)
    return _finish(sim, info, log)


def _finish(sim, info, log):
    if info["train_faces"] == 0:
        raise ValueError(f"too few training faces ({info['valid']} valid, "
                         f"{info['nontrivial'] + info['trivial_kept']} kept)")
    if log:
        log(f"[learned] {info['faces']} faces, {info['train_faces']} trained, val MSE(log p*) "
            f"{info['val_mse_log_pstar']:.3e} vs standard guess {info['val_mse_standard_guess']:.3e}")
    return sim.mlp_get_params(), info


def collect_faces(params, state, steps=1, stride=8, cap=1 << 21, dts=None):
    """Run `steps` FP64 steps from `state` recording the exact solver's face
    problems of every `stride`-th particle."""
    sim = lib.Simulation(params, state)
    sim.set_face_collection(cap, stride)
    faces = []
    for k in range(steps):
        sim.step(dt_fixed=-1.0 if dts is None else float(dts[k]))
        faces.append(sim.collected_faces())
    sim.close()
    return np.concatenate(faces) if faces else np.zeros((0, 6))


def train_seed_from_state(params, state, n_epochs=4, steps=1, stride=8, log=None):
    faces = collect_faces(params, state, steps=steps, stride=stride)
    sim = lib.Simulation(params, state)
    p, info = train_seed_from_faces(sim, faces, params.gamma, n_epochs=n_epochs, log=log)
    sim.close()
    return p, info
