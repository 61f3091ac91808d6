"""The reference's Particles methods one at a time through the C ABI
(include/mlh.h), against the oracle and against the fused stages:

* compPsijTilde + gradient (mlh_psi_gradients) and slopeLimiter (mlh_limit,
  Particles.cpp:1313-1444): the unlimited gradients within the FP64 bar of
  the oracle's, and limit(psi_gradients) bitwise equal to the fused
  mlh_gradients, in both precisions;
* gradient(f) of any field (mlh_gradient_field, Particles.cpp:1257-1270,
  2457-2502): a smooth field against the oracle, and gradient(rho) bitwise
  the unlimited density gradient;
* compEffectiveFace, compRiemannStatesLR, solveRiemannProblems
  (mlh_effective_faces, mlh_riemann_states, mlh_solve_faces,
  Particles.cpp:1290-1911): every row entry's A_ij, W_R, W_L, vFrame and
  F_ij against the oracle's per-pair arrays;
* collectFluxes (mlh_collect_fluxes, :1913-2011) then the update: bitwise
  the fused mlh_fluxes and its update, in both precisions;
* checkFluxSymmetry (mlh_check_flux_symmetry, :2888-2976): the oracle's
  counts with ENFORCE_FLUX_SYM, and a positive count without it;
* the mirror check of the fused step at DEBUG_LVL >= 1 and the counters of
  singular E_i and non-finite states (mlh_diag), which leave the results
  unchanged.
"""
import numpy as np
import pytest

from cases import blast3d, jitter3d, kh_lattice, kh_random, relerr
from meshlesshydro_amd._abi import MLH_FP32, MLH_FP64, default_params

pytestmark = pytest.mark.gpu

CASES = {
    "kh100": lambda **kw: kh_lattice(100, **kw),
    "khrand48": lambda **kw: kh_random(48, **kw),
    "blast16": lambda **kw: blast3d(16, **kw),
    "jitter12": lambda **kw: jitter3d(12, **kw),
}


def _entry_mask(sim):
    noi, nog, _ = sim.get_neighbours()
    S = sim.p.max_interactions + sim.p.max_ghost_interactions
    return np.arange(S)[None, :] < (noi + nog)[:, None]


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec", [MLH_FP64, MLH_FP32])
def test_split_limiter_equals_fused(hip, oracle_mod, case, prec):
    p, s = CASES[case](precision=prec)
    print(f"\n[stages] limiter {case} precision {prec}", flush=True)
    fused = hip.Simulation(p, s)
    fused.neighbours(); fused.density(); fused.gradients()
    want, wB = fused.get_gradients()
    fused.close()
    sim = hip.Simulation(p, s)
    sim.neighbours(); sim.density(); sim.psi_gradients()
    unl, _ = sim.get_gradients()
    sim.limit()
    got, gB = sim.get_gradients()
    assert np.array_equal(got, want) and np.array_equal(gB, wB)
    with pytest.raises(hip.MLHError):
        sim.limit()  # already limited
    sim.close()
    if prec == MLH_FP64:
        p64, _ = CASES[case]()
        orc = oracle_mod.Oracle(p64, s)
        orc.neighbours(); orc.density(); orc.psi_gradients()
        ou, _ = orc.get_gradients()
        for f in range(p.dim + 2):
            assert relerr(unl[f], ou[f]) < 1e-10, f


@pytest.mark.parametrize("case", list(CASES))
def test_gradient_field(hip, oracle_mod, case):
    p, s = CASES[case]()
    sim = hip.Simulation(p, s)
    orc = oracle_mod.Oracle(p, s)
    for o in (sim, orc):
        o.neighbours(); o.density(); o.psi_gradients()
    X = [s[k] for k in ("x", "y", "z")[: p.dim]]
    f = np.sin(2 * np.pi * X[0]) * np.cos(2 * np.pi * X[1]) + 0.5 * X[-1] ** 2
    g, go = sim.gradient_field(f), orc.gradient_field(f)
    assert relerr(g, go) < 1e-10
    rho, _, _ = sim.get_density()
    G, _ = sim.get_gradients()
    assert np.array_equal(sim.gradient_field(rho), G[0])
    sim.close()


@pytest.mark.parametrize("case", list(CASES))
def test_face_arrays_vs_oracle(hip, oracle_mod, case):
    p, s = CASES[case]()
    print(f"\n[stages] face arrays {case}", flush=True)
    orc = oracle_mod.Oracle(p, s)
    orc.neighbours(); orc.density()
    dt = orc.timestep()
    orc.gradients(); orc.fluxes(dt)
    want = orc.get_face_arrays(dt)
    sim = hip.Simulation(p, s)
    sim.neighbours(); sim.density(); sim.psi_gradients(); sim.limit()
    sim.effective_faces()
    sim.riemann_states(dt)
    sim.solve_faces()
    got = sim.get_face_arrays()
    live = _entry_mask(sim)
    for k, tol in (("A", 1e-12), ("WR", 1e-12), ("WL", 1e-12), ("vF", 1e-12), ("F", 1e-10)):
        a, b = got[k][live], want[k][live]
        for c in range(a.shape[1]):
            assert relerr(a[:, c], b[:, c]) < tol, (k, c, relerr(a[:, c], b[:, c]))
        assert not np.any(got[k][~live])
    sim.close()


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec", [MLH_FP64, MLH_FP32])
def test_split_faces_equal_fused(hip, case, prec):
    p, s = CASES[case](precision=prec)
    fused = hip.Simulation(p, s)
    fused.neighbours(); fused.density()
    dt = fused.timestep()
    fused.gradients(); fused.fluxes(dt)
    want = fused.get_fluxes()
    fused.update()
    want_state = fused.state()
    fused.close()
    sim = hip.Simulation(p, s)
    sim.neighbours(); sim.density(); sim.psi_gradients(); sim.limit()
    sim.effective_faces()
    sim.riemann_states(dt)
    sim.solve_faces()
    sim.collect_fluxes()
    for a, b in zip(sim.get_fluxes(), want):
        assert np.array_equal(a, b)
    sim.update()
    got_state = sim.state()
    for k in want_state:
        assert np.array_equal(got_state[k], want_state[k]), k
    sim.close()


@pytest.mark.parametrize("case", ["khrand48", "jitter12"])
def test_flux_symmetry_check(hip, oracle_mod, case):
    for sym in (1, 0):
        p, s = CASES[case](enforce_flux_sym=sym)
        orc = oracle_mod.Oracle(p, s)
        orc.neighbours(); orc.density()
        dt = orc.timestep()
        orc.gradients(); orc.fluxes(dt)
        sim = hip.Simulation(p, s)
        sim.neighbours(); sim.density(); sim.gradients()
        sim.effective_faces(); sim.riemann_states(dt); sim.solve_faces()
        got = sim.check_flux_symmetry(1e-20)
        want = orc.check_flux_symmetry(1e-20)
        print(f"\n[stages] flux symmetry {case} ENFORCE_FLUX_SYM {sym}: {got} (oracle {want})",
              flush=True)
        if sym:
            assert got == want == (0, 0)
        else:
            assert got[1] == want[1] == 0 and got[0] > 0 and want[0] > 0
        sim.close()


def test_flux_symmetry_fp32_acoustic(hip):
    """FP32 mode on the blast, where most faces take the acoustic solution
    and the centre's faces take Newton: every face's flux is the exact
    negative of its mirror's (F_ji = -F_ij bitwise), none missing."""
    p, s = CASES["blast16"](precision=MLH_FP32)
    sim = hip.Simulation(p, s)
    sim.neighbours(); sim.density(); dt = sim.timestep(); sim.gradients()
    sim.effective_faces(); sim.riemann_states(dt); sim.solve_faces()
    got = sim.check_flux_symmetry(0.0)
    d = sim.diag()
    sim.close()
    print(f"\n[stages] FP32 flux symmetry: {got}, acoustic {d['acoustic_faces']} of "
          f"{d['faces_solved']} faces, {d['newton_iterations']} Newton iterations", flush=True)
    assert got == (0, 0)
    assert 0 < d["acoustic_faces"] < d["faces_solved"] and d["newton_iterations"] > 0


def test_mirror_check_and_debug_counters(hip):
    """At DEBUG_LVL 1 the fused step checks that every entry's partner lists
    the particle (none missing here, as in the oracle) and counts singular
    E_i (the three isolated particles of a sparse box) and non-finite states;
    the state is bitwise that of DEBUG_LVL 0."""
    rng = np.random.default_rng(5)
    iso = np.array([[0.1, 0.1], [0.5, 0.15], [0.15, 0.6]])
    g = (np.arange(4) - 1.5) * 0.03
    clu = np.array([0.7, 0.7]) + np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    clu += rng.uniform(-0.003, 0.003, clu.shape)
    X = np.concatenate([iso, clu])
    n = len(X)
    s = dict(x=X[:, 0].copy(), y=X[:, 1].copy(), vx=rng.uniform(-0.01, 0.01, n),
             vy=rng.uniform(-0.01, 0.01, n), m=np.full(n, 1.0 / n), u=np.full(n, 1.5))
    out = {}
    for dbg in (0, 1):
        p = default_params(kernel_size=0.1, debug_level=dbg)
        sim = hip.Simulation(p, s)
        sim.run(2)
        out[dbg] = (sim.state(), sim.diag())
        sim.close()
    d0, d1 = out[0][1], out[1][1]
    for k in out[0][0]:
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert d0["singular_matrices"] == 0 and d0["missing_mirror_pairs"] == 0
    assert d1["singular_matrices"] == 2 * 3  # three isolated particles, two steps
    assert d1["missing_mirror_pairs"] == 0 and d1["nonfinite_particles"] == 0
    # a particle with negative internal energy: its sound speed and its
    # faces' states are not finite, so are the states it reaches
    p, st = kh_lattice(32, debug_level=1)
    st = {k: v.copy() for k, v in st.items()}
    st["u"][500] = -1.0
    sim = hip.Simulation(p, st)
    sim.step(dt_fixed=1e-4)
    d = sim.diag()
    bad = ~np.all(np.isfinite(np.stack([sim.state()[k] for k in ("x", "y", "vx", "vy", "m", "u")])), 0)
    print(f"\n[stages] non-finite particles {d['nonfinite_particles']} (state: {bad.sum()})", flush=True)
    assert d["nonfinite_particles"] == bad.sum() > 0
    sim.close()


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("prec", [MLH_FP64, MLH_FP32])
@pytest.mark.parametrize("order", [0, 2])
def test_fused_step_equals_stages(hip, case, prec, order):
    """mlh_step runs compDensity + compPressure inside the neighbour search;
    mlh_neighbours + mlh_density run them as two kernels.  The same dt and
    the same state, bitwise, in both precisions and both work orders."""
    p, s = CASES[case](precision=prec, work_order=order)
    a = hip.Simulation(p, s)
    dt_a = a.step()
    sa = a.state()
    a.close()
    b = hip.Simulation(p, s)
    b.neighbours(); b.density()
    dt_b = b.timestep()
    b.gradients(); b.fluxes(dt_b); b.update()
    sb = b.state()
    b.close()
    assert dt_a == dt_b
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
