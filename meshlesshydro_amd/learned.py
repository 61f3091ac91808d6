"""Learned p* seed: training data from the exact solver's own faces, GPU
training, and the measured gain over the standard guess.

The network (10-64-64-1, tanh) predicts log(p*/p_guess), the correction of
Toro's adaptive PVRS/TRRS/TSRS guess, from the face features of
csrc/mlp_device.h.  With a zero output layer it reproduces the standard
guess exactly, so training starts from the method without it.  Targets are
the converged p* of the plain FP64 exact solver on the GPU
(mlh_riemann_pstar_batch); training runs the library's HIP forward /
backward kernels with Adam (mlh_mlp_train).
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


def dataset(faces, gamma, keep_trivial=1.0, seed=3):
    """(features, target) from recorded faces [n, 6]: target = log(p*/p_guess).

    Faces whose standard guess is already converged (|target| < 1e-5) are
    subsampled to keep_trivial times the number of non-trivial faces."""
    if len(faces) == 0:
        return np.zeros((0, MLP_IN), np.float32), np.zeros(0, np.float32)
    rl, ul, pl, rr, ur, pr = (faces[:, k] for k in range(6))
    ok = (rl > 0) & (rr > 0) & (pl > 0) & (pr > 0) & np.isfinite(faces).all(axis=1)
    rl, ul, pl, rr, ur, pr = (a[ok] for a in (rl, ul, pl, rr, ur, pr))
    ps, _ = lib.riemann_pstar(gamma, rl, ul, pl, rr, ur, pr, tol=1e-12)
    feats, pref = lib.mlp_features(gamma, rl, ul, pl, rr, ur, pr)
    good = ps > 0
    feats, pref, ps = feats[good], pref[good], ps[good]
    target = (np.log(ps / pref) - feats[:, 7]).astype(np.float32)
    trivial = np.abs(target) < 1e-5
    nt = int((~trivial).sum())
    rng = np.random.default_rng(seed)
    tidx = np.flatnonzero(trivial)
    keep = min(len(tidx), max(int(keep_trivial * max(nt, 1)), 1024))
    sel = np.concatenate([np.flatnonzero(~trivial), rng.choice(tidx, keep, replace=False) if keep else []]).astype(np.int64)
    rng.shuffle(sel)
    return feats[sel], target[sel]


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
    """Train the seed network of Simulation `sim` on recorded faces, from
    `init` (weights to continue from) or from init_params(seed); returns
    (params, info)."""
    feats, target = dataset(faces, gamma)
    n = len(target)
    if n < 256:
        raise ValueError(f"too few training faces ({n})")
    nval = max(n // 10, 64)
    sim.mlp_set_params(init_params(seed) if init is None else np.asarray(init, np.float32))
    loss0, _ = sim.mlp_loss_grad(feats[:min(n, 65536)], target[:min(n, 65536)])
    loss = sim.mlp_train(feats[nval:], target[nval:], epochs=n_epochs, batch=min(batch, n - nval), lr=lr, seed=seed)
    val = sim.mlp_infer(feats[:nval], bf16=True)
    val_mse = float(np.mean((val - target[:nval]) ** 2))
    base_mse = float(np.mean(target[:nval] ** 2))
    info = {"train_faces": int(n - nval), "val_faces": int(nval), "epochs": n_epochs,
            "loss_init": loss0, "loss_final": loss, "val_mse_log_pstar": val_mse,
            "val_mse_standard_guess": base_mse}
    if log:
        log(f"[learned] {n} faces, val MSE(log p*) {val_mse:.3e} vs standard guess {base_mse:.3e}")
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
