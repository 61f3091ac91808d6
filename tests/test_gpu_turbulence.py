"""Decaying turbulent box on the GPU: glass relaxation, Mach-1 solenoidal
velocity, and FP32 with the learned p* seed retrained online against the
FP64 run.

* The glass relaxation reduces the relative density spread of the jittered
  lattice.
* FP64 plain runs 20 steps with its own CFL steps; its first steps record
  the faces the exact solver iterated on, and the seed is trained on them.
* FP32 with the learned seed replays the FP64 dt sequence and retrains the
  network once, after 10 steps, on the faces its own run recorded
  (continuing from the current weights).
* After 20 steps the FP32 density, pressure and speed are within 1e-4
  relative L1 of the FP64 run (rho recomputed by one FP64 density pass for
  both), mass is conserved to 1e-13 and energy to 1e-12.

This test covers the face recording, the online retraining and the learned
run's bookkeeping.  On this flow the standard guess passes the FP32 probe
or comes close enough that no face reaches the network (the deferral gate),
so it does not exercise the seeded solve; tests/test_gpu_learned.py does,
against the oracle, with seeding asserted.
"""
import numpy as np
import pytest

from meshlesshydro_amd import learned as L
from meshlesshydro_amd import turbulence as T
from meshlesshydro_amd._abi import MLH_FP32, MLH_FP64, MLH_SEED_LEARNED, default_params

pytestmark = pytest.mark.gpu


def _fields(hip, base, st):
    s = hip.Simulation(default_params(precision=MLH_FP64, **base), st)
    s.neighbours()
    s.density()
    rho, P, _ = s.get_density()
    s.close()
    return rho, P, np.sqrt(st["vx"] ** 2 + st["vy"] ** 2 + st["vz"] ** 2)


def test_turbulent_box_online_retraining(hip):
    n = 24
    st, info = T.turbulent_box(n, relax_iterations=6)
    spread = info["density_spread"]
    print(f"\n[turbulence {n}^3] density spread {spread}", flush=True)
    assert spread[-1] < 0.7 * spread[0]
    vrms = np.sqrt(np.mean(st["vx"] ** 2 + st["vy"] ** 2 + st["vz"] ** 2))
    assert abs(vrms - np.sqrt(5.0 / 3.0)) < 1e-12
    base = dict(dim=3, kernel_size=info["kernel_size"], max_interactions=160,
                max_ghost_interactions=160, debug_level=0)
    p64 = default_params(precision=MLH_FP64, **base)
    sim = hip.Simulation(p64, st)
    sim.set_face_collection(1 << 20, 4)
    dts_a = sim.run(5)
    faces = sim.collected_faces()
    sim.set_face_collection(0, 0)
    dts_b = sim.run(15)
    ref = sim.state()
    weights, _ = L.train_seed_from_faces(sim, faces, p64.gamma, n_epochs=2)
    sim.close()
    dts = np.concatenate([dts_a, dts_b])

    p32 = default_params(precision=MLH_FP32, **base)
    p32.riemann_seed = MLH_SEED_LEARNED
    s32 = hip.Simulation(p32, st)
    s32.mlp_set_params(weights)
    c0 = s32.conserved()
    _, infos = T.run_online(s32, 20, retrain_every=10, stride=4, epochs=2, dts=dts)
    print(f"[turbulence] retraining {infos}", flush=True)
    assert len(infos) == 1 and "skipped" not in infos[0] and infos[0]["train_faces"] > 0
    d = s32.diag()
    print(f"[turbulence] seeded {d['learned_seeded']} of {d['faces_solved']} faces", flush=True)
    out = s32.state()
    c1 = s32.conserved()
    s32.close()
    assert abs(c1[1] - c0[1]) <= 1e-13 * abs(c0[1])
    assert abs(c1[2] - c0[2]) <= 1e-12 * abs(c0[2])
    fr, fo = _fields(hip, base, ref), _fields(hip, base, out)
    for name, a, b in zip(("rho", "P", "|v|"), fo, fr):
        l1 = np.sum(np.abs(a - b)) / np.sum(np.abs(b))
        print(f"[turbulence] L1_rel({name}) = {l1:.3e}", flush=True)
        assert l1 < 1e-4, (name, l1)
