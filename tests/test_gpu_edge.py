"""Edge cases through the C ABI.

* An empty state is rejected with MLH_ERR_INVALID_ARGUMENT (nothing is
  launched).
* A sparse periodic box: three isolated particles (no neighbour within h,
  so E_i is singular and they have no faces) beside a 4 x 4 patch.
  Against the oracle: bit-exact neighbour lists, the same dt, and the same
  state after three steps. The isolated particles only drift: m is
  unchanged, u and v to a few ulps (the update divides m v by m), and x
  moves by v times the elapsed time. FP32 mode leaves them the same way.
* Error paths: negative face pressures abort at the reference's debug
  levels with status 6 (against the oracle); stages called out of order,
  a learned seed without a network and a too-coarse periodic grid report
  their own status codes.
"""
import numpy as np
import pytest

from cases import relerr
from meshlesshydro_amd._abi import MLH_FP32, default_params

pytestmark = pytest.mark.gpu


def sparse_case(**kw):
    rng = np.random.default_rng(5)
    iso = np.array([[0.1, 0.1], [0.5, 0.15], [0.15, 0.6]])
    # a jittered 4 x 4 patch, spacing 0.03: 12-15 neighbours each
    g = (np.arange(4) - 1.5) * 0.03
    clu = np.array([0.7, 0.7]) + np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    clu += rng.uniform(-0.003, 0.003, clu.shape)
    X = np.concatenate([iso, clu])
    n = len(X)
    s = dict(x=X[:, 0].copy(), y=X[:, 1].copy(), vx=rng.uniform(-0.01, 0.01, n),
             vy=rng.uniform(-0.01, 0.01, n), m=np.full(n, 1.0 / n), u=np.full(n, 1.5))
    return default_params(kernel_size=0.1, debug_level=0, **kw), s


def test_empty_state_rejected(hip):
    p = default_params(kernel_size=0.1)
    empty = {k: np.zeros(0) for k in ("x", "y", "vx", "vy", "m", "u")}
    with pytest.raises(hip.MLHError) as e:
        hip.Simulation(p, empty)
    assert "n <= 0" in str(e.value)


def test_isolated_particles(hip, oracle_mod):
    p, s = sparse_case()
    sim = hip.Simulation(p, s)
    orc = oracle_mod.Oracle(p, s)
    sim.neighbours()
    orc.neighbours()
    gn, gg, gl = sim.get_neighbours()
    on, og, ol = orc.get_neighbours()
    assert np.array_equal(gn, on) and np.array_equal(gg, og)
    assert list(gn[:3]) == [0, 0, 0] and list(gg[:3]) == [0, 0, 0]
    cnt = gn + gg
    mask = np.arange(gl.shape[1])[None, :] < cnt[:, None]
    assert np.array_equal(np.where(mask, gl, 0), np.where(mask, ol, 0))
    t = 0.0
    for _ in range(3):
        dt_o = orc.step()
        dt_g = sim.step()
        assert abs(dt_g - dt_o) <= 1e-13 * dt_o
        t += dt_g
    gs, os_ = sim.state(), orc.state()
    for k in gs:
        assert relerr(gs[k], os_[k]) < 1e-10, k
    # v = (m v - dt F_v) / m and u = Q_E / m - v^2 / 2 with zero fluxes:
    # unchanged to a few ulps (as in the reference), m exactly
    assert np.array_equal(gs["m"][:3], s["m"][:3])
    for k in ("vx", "vy", "u"):
        assert np.allclose(gs[k][:3], s[k][:3], rtol=1e-14, atol=0), k
    for k, v in (("x", "vx"), ("y", "vy")):
        assert np.allclose(gs[k][:3], (s[k][:3] + t * s[v][:3]) % 1.0, rtol=0, atol=1e-14), k
    sim.close()

    p32, _ = sparse_case(precision=MLH_FP32)
    s32 = hip.Simulation(p32, s)
    s32.run(3)
    o32 = s32.state()
    s32.close()
    assert all(np.all(np.isfinite(o32[k])) for k in o32)
    assert np.array_equal(o32["m"][:3], s["m"][:3])
    for k in ("vx", "vy", "u"):
        assert np.allclose(o32[k][:3], s[k][:3], rtol=1e-14, atol=0), k


def test_negative_face_pressure_status(hip, oracle_mod):
    """Slope limiting off on the 16^3 blast: linear reconstruction next to
    the energy spike gives negative face pressures on real faces in the
    first step (none across images). The reference aborts with exit code 6
    on real faces only above DEBUG_LVL 1 (Particles.cpp:1824-1827, ghost
    faces :1878-1881 from level 1): level 2 fails on the first step like the
    oracle; level 1 runs and counts the same faces as the oracle."""
    from cases import CASES
    from meshlesshydro_amd._abi import MLH_ERR_NEGATIVE_FACE_PRESSURE
    p, s = CASES["blast16"](debug_level=2, slope_limiting=0)
    sim = hip.Simulation(p, s)
    with pytest.raises(hip.MLHError) as e:
        sim.step()
    assert e.value.status == MLH_ERR_NEGATIVE_FACE_PRESSURE
    sim.close()
    orc = oracle_mod.Oracle(p, s)
    with pytest.raises(oracle_mod.OracleError) as eo:
        orc.step()
    assert eo.value.status == MLH_ERR_NEGATIVE_FACE_PRESSURE
    p1, _ = CASES["blast16"](debug_level=1, slope_limiting=0)
    sim = hip.Simulation(p1, s)
    orc = oracle_mod.Oracle(p1, s)
    sim.step()
    orc.step()
    dg, do = sim.diag(), orc.diag()
    sim.close()
    assert dg["neg_face_pressure_real"] == do["neg_face_pressure_real"] > 0
    assert dg["neg_face_pressure_ghost"] == do["neg_face_pressure_ghost"]


def test_call_order_and_configuration_errors(hip):
    """Stages called before their inputs exist, a learned seed without a
    network, and a periodic box with fewer than 3 search cells per axis are
    reported with their own status codes; nothing runs."""
    from cases import CASES
    from meshlesshydro_amd import ics
    from meshlesshydro_amd._abi import (MLH_ERR_CALL_ORDER, MLH_ERR_GRID, MLH_ERR_NOT_TRAINED,
                                        MLH_SEED_LEARNED)
    p, s = CASES["jitter12"]()
    sim = hip.Simulation(p, s)
    for call in (sim.density, sim.gradients, lambda: sim.fluxes(1e-3), sim.get_density):
        with pytest.raises(hip.MLHError) as e:
            call()
        assert e.value.status == MLH_ERR_CALL_ORDER
    sim.neighbours()
    for call in (sim.gradients, lambda: sim.fluxes(1e-3), sim.get_gradients):
        with pytest.raises(hip.MLHError) as e:
            call()
        assert e.value.status == MLH_ERR_CALL_ORDER
    sim.close()
    p32, _ = CASES["jitter12"](precision=MLH_FP32)
    p32.riemann_seed = MLH_SEED_LEARNED
    sim = hip.Simulation(p32, s)
    with pytest.raises(hip.MLHError) as e:
        sim.step()
    assert e.value.status == MLH_ERR_NOT_TRAINED
    sim.close()
    # kernel size 0.4 in the unit box: floor(1 / 0.4) = 2 cells per axis
    with pytest.raises(hip.MLHError) as e:
        hip.Simulation(default_params(kernel_size=0.4), ics.kelvin_helmholtz(16))
    assert e.value.status == MLH_ERR_GRID
