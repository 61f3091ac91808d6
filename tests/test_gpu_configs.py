"""The bench workloads against the oracle: the 3D blast at its full 128^3,
the shear layer at its full 1000^2, and the decaying turbulent box at 20^3.

* 128^3 blast (2,097,152 particles, 96 + 96 list limits): one FP64 step
  stage by stage (neighbour lists bit-exact, gridNNS Particles.cpp:324-365
  and the image search; the other stages to the FP64 bars), then, on the
  first state of the run whose faces the network seeds, one FP32 step's
  flux sums with the standard and with the learned seed within the FP32
  bars (the seed trained on the blast's own FP64 faces, seeding asserted).
* 1000^2 shear layer (10^6 particles, the reference's list limits): lists
  bit-exact, one FP32 step's density, dt and flux sums within the FP32 bars.
* turbulent box (glass, Mach-1 solenoidal velocity, the reference's list
  limits 400 / 300) at 20^3: lists bit-exact, three FP64 steps with their
  own CFL steps within 1e-9, three FP32 steps with the learned seed within
  the FP32 step bar, mass conserved to round-off.
"""
import numpy as np
import pytest

from cases import assert_fp32_fluxes, first_seeded_state, interior, relerr
from meshlesshydro_amd import ics
from meshlesshydro_amd import learned as L
from meshlesshydro_amd import turbulence as T
from meshlesshydro_amd._abi import MLH_FP32, MLH_FP64, MLH_SEED_LEARNED, default_params

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def _lists_equal(sim, orc, chunk=1 << 17):
    """Neighbour counts and rows bit-exact (same entries, same order),
    compared in chunks of rows."""
    gn, gg, gl = sim.get_neighbours()
    on, og, ol = orc.get_neighbours()
    assert np.array_equal(gn, on) and np.array_equal(gg, og)
    cnt = gn + gg
    col = np.arange(gl.shape[1])[None, :]
    for a in range(0, len(cnt), chunk):
        b = a + chunk
        m = col < cnt[a:b, None]
        assert np.array_equal(np.where(m, gl[a:b], 0), np.where(m, ol[a:b], 0)), a
    return int(cnt.max())


def _fp32_sim(p64_kw, state, weights=None):
    from meshlesshydro_amd import lib
    p = default_params(precision=MLH_FP32, **p64_kw)
    if weights is not None:
        p.riemann_seed = MLH_SEED_LEARNED
    sim = lib.Simulation(p, state)
    if weights is not None:
        sim.mlp_set_params(weights)
    return sim


def _stages_to_fluxes(sim, dt):
    sim.neighbours()
    sim.density()
    sim.gradients()
    sim.fluxes(dt)


def test_blast_128_full_size(hip, oracle_mod):
    n = 128
    kw = dict(dim=3, kernel_size=ics.blast_kernel_size(n), max_interactions=96,
              max_ghost_interactions=96, debug_level=0)
    p = default_params(precision=MLH_FP64, **kw)
    s = ics.blast_3d(n)
    # the seed network, from the faces of the first two FP64 steps
    faces = L.collect_faces(p, s, steps=2, stride=8)
    sim = hip.Simulation(p, s)
    weights, _ = L.train_seed_from_faces(sim, faces, p.gamma, n_epochs=4)
    print(f"\n[blast128] FP64 stages vs the oracle ({len(faces)} training faces)", flush=True)
    orc = oracle_mod.Oracle(p, s)
    sim.neighbours(); orc.neighbours()
    kmax = _lists_equal(sim, orc)
    print(f"[blast128] lists bit-exact, longest row {kmax}", flush=True)
    sim.density(); orc.density()
    gr, gP, gw = sim.get_density()
    orr, oP, ow = orc.get_density()
    assert relerr(gw, ow) < 1e-13 and relerr(gr, orr) < 1e-13 and relerr(gP, oP) < 1e-13
    dt = orc.timestep()
    assert abs(sim.timestep() - dt) <= 1e-13 * dt
    sim.gradients(); orc.gradients()
    gG, gB = sim.get_gradients()
    oG, oB = orc.get_gradients()
    assert relerr(gB, oB) < FP64_TOL
    for f in range(5):
        assert relerr(gG[f], oG[f]) < FP64_TOL, f
    del gG, gB, oG, oB
    sim.fluxes(dt); orc.fluxes(dt)
    for a, b in zip(sim.get_fluxes(), orc.get_fluxes()):
        assert relerr(a, b) < FP64_TOL
    sim.update(); orc.update(dt)
    gs, os_ = sim.state(), orc.state()
    for k in gs:
        assert relerr(gs[k], os_[k]) < 1e-12, k
    sim.close()
    del orc
    # FP32 with the standard and the learned seed on the first state of the
    # run whose faces the network seeds (none at the first step: every
    # standard guess passes the probe there), the oracle's dt from that state
    p32 = default_params(precision=MLH_FP32, riemann_seed=MLH_SEED_LEARNED, **kw)
    st, seeded, k = first_seeded_state(hip, p, p32, s, weights)
    print(f"[blast128] FP32 fluxes after {k} FP64 steps ({seeded} faces seeded)", flush=True)
    assert seeded > 0
    orc = oracle_mod.Oracle(p, st)
    orc.neighbours()
    orc.density()
    dt = orc.timestep()
    orc.gradients()
    orc.fluxes(dt)
    inner = interior(p, st)
    for name, w in (("standard", None), ("learned", weights)):
        print(f"[blast128] FP32 {name} fluxes", flush=True)
        s32 = _fp32_sim(kw, st, w)
        _stages_to_fluxes(s32, dt)
        assert_fp32_fluxes(p, orc, s32, inner, "blast128 " + name)
        if w is not None:
            d = s32.diag()
            print(f"[blast128] seeded {d['learned_seeded']} of {d['faces_solved']} faces", flush=True)
            assert d["learned_seeded"] > 0
        s32.close()


def test_shear_layer_1000_full_size(hip, oracle_mod):
    n = 1000
    kw = dict(kernel_size=4.0 / n, debug_level=0)
    p = default_params(precision=MLH_FP64, **kw)
    s = ics.kelvin_helmholtz(n)
    print("\n[shear1000] oracle stages", flush=True)
    orc = oracle_mod.Oracle(p, s)
    orc.neighbours()
    orc.density()
    dt = orc.timestep()
    orc.gradients()
    orc.fluxes(dt)
    print("[shear1000] FP32 stages", flush=True)
    sim = _fp32_sim(kw, s)
    sim.neighbours()
    kmax = _lists_equal(sim, orc)
    print(f"[shear1000] lists bit-exact, longest row {kmax}", flush=True)
    sim.density()
    gr, gP, _ = sim.get_density()
    orr, oP, _ = orc.get_density()
    assert relerr(gr, orr) < 1e-6 and relerr(gP, oP) < 1e-6
    assert abs(sim.timestep() - dt) <= 1e-6 * dt
    sim.gradients()
    sim.fluxes(dt)
    assert_fp32_fluxes(p, orc, sim, interior(p, s), "shear1000")
    sim.close()


def test_turbulent_box_vs_oracle(hip, oracle_mod):
    n, steps = 20, 3
    st, info = T.turbulent_box(n, relax_iterations=6)
    kw = dict(dim=3, kernel_size=info["kernel_size"], max_interactions=400,
              max_ghost_interactions=300, debug_level=0)
    p = default_params(precision=MLH_FP64, **kw)
    print(f"\n[turbulence {n}^3] FP64 vs the oracle", flush=True)
    sim = hip.Simulation(p, st)
    orc = oracle_mod.Oracle(p, st)
    sim.neighbours(); orc.neighbours()
    kmax = _lists_equal(sim, orc)
    print(f"[turbulence] lists bit-exact, longest row {kmax}", flush=True)
    sim.upload(st)
    orc.set_state(st)
    c0 = sim.conserved()
    dts = []
    for k in range(steps):
        dt_o = orc.step()
        dt_g = sim.step()
        assert abs(dt_g - dt_o) <= 1e-11 * dt_o
        dts.append(dt_o)
    gs, os_ = sim.state(), orc.state()
    for k in gs:
        assert relerr(gs[k], os_[k]) < 1e-9, k
    c1 = sim.conserved()
    assert abs(c1[1] - c0[1]) <= 1e-13 * abs(c0[1])
    sim.close()
    faces = L.collect_faces(p, st, steps=2, stride=1)
    tr = hip.Simulation(p, st)
    weights, _ = L.train_seed_from_faces(tr, faces, p.gamma, n_epochs=2)
    tr.close()
    print(f"[turbulence] FP32 learned, {len(faces)} training faces", flush=True)
    s32 = _fp32_sim(kw, st, weights)
    c0 = s32.conserved()
    s32.run(steps, dts_in=np.array(dts))
    c1 = s32.conserved()
    assert abs(c1[1] - c0[1]) <= 1e-13 * abs(c0[1])
    out = s32.state()
    d = s32.diag()
    print(f"[turbulence] seeded {d['learned_seeded']} of {d['faces_solved']} faces", flush=True)
    for k in ("m", "u", "vx", "vy", "vz"):
        l1 = np.sum(np.abs(out[k] - os_[k])) / np.sum(np.abs(os_[k]))
        assert l1 < 1e-5, (k, l1)
    s32.close()
