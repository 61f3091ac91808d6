"""The seed network's training set built and trained on the device
(mlh_mlp_dataset, mlh_mlp_train_faces, mlh_mlp_train_recorded) against
the host construction of the same set (learned.dataset: numpy, the exact
FP64 p* of mlh_riemann_pstar_batch and the host features of
mlh_mlp_features), and the determinism the set's hash order gives: the set
and the trained weights do not depend on the order of the faces, and a
retraining on the run's record equals one on the record copied out."""
import numpy as np
import pytest

from meshlesshydro_amd import ics
from meshlesshydro_amd import learned as L

pytestmark = pytest.mark.gpu

GAMMA = 5 / 3


def _faces(n, seed):
    """Random face problems plus converged (trivial) and invalid ones."""
    rng = np.random.default_rng(seed)
    f = np.stack(ics.riemann_states(n, seed=seed), axis=1)
    k = n // 4
    f[:k, 3:] = f[:k, :3]              # left = right: p* = p exactly
    f[k, 0] = -1.0                     # negative density
    f[k + 1, 5] = 0.0                  # zero pressure
    f[k + 2, 1] = np.nan
    f[k + 3, 4] = np.inf
    return f[rng.permutation(n)]


@pytest.fixture(scope="module")
def ctx(hip):
    from cases import kh_lattice
    p, s = kh_lattice(16)
    sim = hip.Simulation(p, s)
    yield sim
    sim.close()


@pytest.mark.parametrize("kt", [0.0, 0.25, 1.0])
def test_device_set_rules(ctx, hip, oracle_mod, kt):
    """The device set keeps exactly the faces its rules name: every valid
    face (positive, finite states and p* > 0, as on the host), all
    non-trivial ones, the trivial ones whose selection hash is below the
    threshold the reported counts give, in ascending order of the content
    hash; its features and targets meet the bars the host features meet
    against the FP64 oracle (test_gpu_mlp.test_features_vs_oracle)."""
    f = _faces(40_000, 11)               # 10,000 trivial faces
    xd, td, rd = ctx.mlp_dataset(f, GAMMA, keep_trivial=kt, rows=True)
    ctx.mlp_set_params(L.init_params(1))
    info = ctx.mlp_train_faces(f, GAMMA, epochs=0, keep_trivial=kt)
    assert info["faces"] == len(f) and len(td) == info["nontrivial"] + info["trivial_kept"]
    ok = (f[:, 0] > 0) & (f[:, 3] > 0) & (f[:, 2] > 0) & (f[:, 5] > 0) & np.isfinite(f).all(axis=1)
    g = np.where(ok[:, None], f, 1.0)
    ps, _ = hip.riemann_pstar(GAMMA, *(g[:, k] for k in range(6)), tol=1e-12)
    valid = ok & (ps > 0)
    assert info["valid"] == int(valid.sum()) and int(valid.sum()) <= len(f) - 4
    assert np.all(valid[rd]) and len(np.unique(rd)) == len(rd)
    # the order: ascending (h >> 33, then h mod 2^32)
    h = L.face_hash(f[rd])
    k1, k0 = h >> np.uint64(33), h & np.uint64(0xFFFFFFFF)
    step = (k1[1:] > k1[:-1]) | ((k1[1:] == k1[:-1]) & (k0[1:] >= k0[:-1]))
    assert np.all(step)
    # all non-trivial faces kept; kept trivial faces pass the selection hash
    triv = np.abs(td) < 1e-5
    assert int((~triv).sum()) == info["nontrivial"]
    ntr = info["valid"] - info["nontrivial"]
    keep = min(ntr, max(int(kt * max(info["nontrivial"], 1)), 1024))
    thr = 1 << 53 if keep >= ntr else int(keep / ntr * 9007199254740992.0)
    u53 = L.mix64(h[triv] ^ L.SELECT_SALT) >> np.uint64(11)
    assert np.all(u53 < np.uint64(thr))
    assert abs(int(triv.sum()) - keep) <= 5 * np.sqrt(keep) + 1, (int(triv.sum()), keep)
    if keep == ntr:
        assert np.array_equal(np.sort(rd), np.flatnonzero(valid))
        # the host construction keeps the same faces in the same order
        assert np.array_equal(L.dataset(f, GAMMA, keep_trivial=kt, rows=True)[2], rd)
    # values against the oracle's FP64 features and the FP64 p*: FP32
    # cannot resolve the PVRS / TRRS pressures where they cancel.  The
    # features meet the host features' bars; the target carries the error
    # of f7 alone, and is held to the host evaluation's error on the same
    # faces
    fo, pro = oracle_mod.mlp_features(GAMMA, *(f[rd, k] for k in range(6)))
    to = np.log(ps[rd] / pro) - fo[:, 7]
    d = np.abs(xd - fo).ravel()
    assert np.median(d) < 1e-6
    assert np.quantile(d, 0.99) < 1e-4, np.quantile(d, 0.99)
    assert np.mean(d > 1e-2) < 2e-3, np.mean(d > 1e-2)
    xh, prh = hip.mlp_features(GAMMA, *(f[rd, k] for k in range(6)))
    dh = np.abs((np.log(ps[rd] / prh) - xh[:, 7]).astype(np.float32) - to)
    dt = np.abs(td - to)
    assert np.median(dt) < 1e-6
    for q in (0.9, 0.99, 0.999):
        assert np.quantile(dt, q) <= 2 * np.quantile(dh, q) + 1e-6, (q, np.quantile(dt, q), np.quantile(dh, q))
    assert np.mean(dt > 1e-2) < 2e-3, np.mean(dt > 1e-2)


def test_set_and_weights_independent_of_face_order(ctx):
    f = _faces(20_000, 13)
    perm = np.random.default_rng(5).permutation(len(f))
    xa, ta = ctx.mlp_dataset(f, GAMMA)
    xb, tb = ctx.mlp_dataset(f[perm], GAMMA)
    assert np.array_equal(xa, xb) and np.array_equal(ta, tb)
    w = []
    for faces in (f, f[perm]):
        ctx.mlp_set_params(L.init_params(1))
        ctx.mlp_train_faces(faces, GAMMA, epochs=2, batch=1024, seed=7)
        w.append(ctx.mlp_get_params())
    assert np.array_equal(w[0], w[1])
    assert not np.array_equal(w[0], L.init_params(1))


def test_train_faces_equals_train_on_its_set(ctx):
    """mlh_mlp_train_faces = mlh_mlp_loss_grad on the set's first faces,
    mlh_mlp_train on its training faces (the same epoch order), and the
    BF16 network's MSE on its validation faces: bitwise for the weights and
    the losses."""
    f = _faces(24_000, 14)
    x, t = ctx.mlp_dataset(f, GAMMA)
    n = len(t)
    nval = max(n // 10, 64)
    init = L.init_params(3)
    ctx.mlp_set_params(init)
    info = ctx.mlp_train_faces(f, GAMMA, epochs=3, batch=2048, lr=3e-3, seed=9)
    wa = ctx.mlp_get_params()
    assert info["train_faces"] == n - nval and info["val_faces"] == nval and info["batch"] == 2048
    ctx.mlp_set_params(init)
    loss0, _ = ctx.mlp_loss_grad(x[:min(n, 65536)], t[:min(n, 65536)])
    last = ctx.mlp_train(x[nval:], t[nval:], epochs=3, batch=2048, lr=3e-3, seed=9)
    wb = ctx.mlp_get_params()
    assert np.array_equal(wa, wb)
    assert info["loss_init"] == loss0 and info["loss_final"] == last
    y = ctx.mlp_infer(x[:nval], bf16=True).astype(np.float64)
    tv = t[:nval].astype(np.float64)
    assert info["val_mse_log_pstar"] == pytest.approx(np.mean((y - tv) ** 2), rel=1e-12)
    assert info["val_mse_standard_guess"] == pytest.approx(np.mean(tv ** 2), rel=1e-12)
    assert info["val_mse_log_pstar"] < info["val_mse_standard_guess"], info


def test_too_few_faces_trains_nothing(ctx):
    f = _faces(600, 15)[:200]
    ctx.mlp_set_params(L.init_params(1))
    info = ctx.mlp_train_faces(f, GAMMA, epochs=2)
    assert info["train_faces"] == 0
    assert np.array_equal(ctx.mlp_get_params(), L.init_params(1))
    with pytest.raises(ValueError):
        L.train_seed_from_faces(ctx, f, GAMMA)


def test_train_recorded_equals_train_on_copied_record(hip):
    """A retraining on the run's record (no copy to the host) gives the
    weights of a training on the same record copied out, and restarts the
    record; the record's set equals the set of the copied faces."""
    from cases import jitter3d
    p, s = jitter3d(12)

    def run():
        sim = hip.Simulation(p, s)
        sim.set_face_collection(1 << 18, 1)
        sim.step()
        sim.step()
        sim.mlp_set_params(L.init_params(2))
        return sim

    a = run()
    xr, tr = a.mlp_dataset()          # the record, not restarted
    ia = a.mlp_train_recorded(epochs=2, batch=1024, seed=4)
    assert len(a.collected_faces()) == 0
    wa = a.mlp_get_params()
    a.close()
    b = run()
    faces = b.collected_faces()
    assert ia["faces"] == len(faces) > 1000
    xf, tf = b.mlp_dataset(faces)
    assert np.array_equal(xr, xf) and np.array_equal(tr, tf)
    ib = b.mlp_train_faces(faces, p.gamma, epochs=2, batch=1024, seed=4)
    assert np.array_equal(wa, b.mlp_get_params())
    assert {k: ia[k] for k in ia} == {k: ib[k] for k in ib}
    b.close()
