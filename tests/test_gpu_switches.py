"""GPU parity of the parameter.h switches against the CPU oracle.

Each variant changes one compile-time flag of the reference
(demonstrator/include/parameter.h) from the Kelvin-Helmholtz long-run set:
PAIRWISE_LIMITER (Particles.cpp:1735-1785), MESHLESS_FINITE_MASS
(Riemann.cpp:171-175, Particles.cpp:2041-2043), FIRST_ORDER_QUAD_POINT
(:1500-1530), MOVE_PARTICLES (:2082-2087), SLOPE_LIMITING (:1313-1444),
ENFORCE_FLUX_SYM (:1838-1858), the intended fabs() limiter and the 3D
predictor quirk (:1717,1719); plus the flag set of the reference's own blast
case (testcases/sedov/parameter.h: CFL .25, BETA 1, PAIRWISE_LIMITER 1).

FP64: three full steps, dt within 1e-11 and the state within 1e-9 of the
oracle.  FP32: one step's fluxes against the FP64 oracle with the bars of
the FP32 stage test (face-flux scale, 1e-5 for all but 1% of particles).
"""
import numpy as np
import pytest

from cases import CASES, relerr
from meshlesshydro_amd._abi import MLH_ABS_FABS, MLH_FP32

pytestmark = pytest.mark.gpu

VARIANTS = {
    "pairwise": dict(pairwise_limiter=1),
    "blast_set": dict(pairwise_limiter=1, beta=1.0, cfl=0.25),
    "mfm": dict(mfm=1),
    "quad_point": dict(first_order_quad_point=0),
    "fixed_particles": dict(move_particles=0),
    "unlimited": dict(slope_limiting=0),
    "no_flux_sym": dict(enforce_flux_sym=0),
    "fabs": dict(limiter_abs=MLH_ABS_FABS),
    "quirk_3d_predictor": dict(quirk_3d_predictor=1),
}
# FIRST_ORDER_QUAD_POINT 0 puts the face state of j at x_i + h/4 (x_j - x_i),
# almost a full pair distance from x_j: on the random 2D set the reference
# algorithm (oracle) extrapolates two particles to non-finite states in the
# first step, so that variant's 2D case is the lattice
PARAMS = [("kh100" if (c, v) == ("khrand48", "quad_point") else c, v)
          for c in ("khrand48", "jitter12") for v in VARIANTS
          if not (v == "quirk_3d_predictor" and c == "khrand48")]


@pytest.mark.parametrize("case,variant", PARAMS)
def test_switch_fp64_steps(case, variant, hip, oracle_mod):
    p, s = CASES[case](debug_level=0, **VARIANTS[variant])
    print(f"\n[switch {case}/{variant}] N={len(s['x'])}", flush=True)
    sim = hip.Simulation(p, s)
    orc = oracle_mod.Oracle(p, s)
    for k in range(3):
        dt_o = orc.step()
        dt_g = sim.step()
        assert abs(dt_g - dt_o) <= 1e-11 * dt_o, (k, dt_g, dt_o)
    gs, os_ = sim.state(), orc.state()
    for k in gs:
        assert relerr(gs[k], os_[k]) < 1e-9, k
    sim.close()


@pytest.mark.parametrize("case,variant", PARAMS)
def test_switch_fp32_fluxes(case, variant, hip, oracle_mod):
    p, s = CASES[case](debug_level=0, **VARIANTS[variant])
    p32, _ = CASES[case](debug_level=0, precision=MLH_FP32, **VARIANTS[variant])
    sim = hip.Simulation(p32, s)
    orc = oracle_mod.Oracle(p, s)
    sim.neighbours(); orc.neighbours()
    sim.density(); orc.density()
    dt = orc.timestep()
    sim.gradients(); orc.gradients()
    sim.fluxes(dt); orc.fluxes(dt)
    gm, gv, ge = sim.get_fluxes()
    om, ov, oe = orc.get_fluxes()
    orr, oP, _ = orc.get_density()
    st = orc.state()
    vmax = max(np.max(np.abs(st[k])) for k in ("vx", "vy", "vz")[: p.dim])
    cmax = np.max(np.sqrt(p.gamma * oP / orr))
    area = p.kernel_size ** (p.dim - 1)
    fm = np.max(orr) * (vmax + cmax) * area
    scales = (fm, fm * (vmax + cmax) + np.max(oP) * area,
              fm * (vmax + cmax) ** 2 + np.max(oP) * (vmax + cmax) * area)
    for a, b, scale in zip((gm, gv, ge), (om, ov, oe), scales):
        err = np.abs(a - b).reshape(len(gm), -1).max(axis=1) / scale
        assert np.mean(err > 1e-5) < 0.01, (case, variant, np.mean(err > 1e-5))
        assert np.max(err) < 1.0, (case, variant, np.max(err))
    sim.close()


def test_blast_reference_flag_set_steps(hip, oracle_mod):
    """The 3D point blast (16^3) with the reference's blast flag set: strong
    shocks through the pairwise limiter over five steps."""
    p, s = CASES["blast16"](debug_level=0, **VARIANTS["blast_set"])
    sim = hip.Simulation(p, s)
    orc = oracle_mod.Oracle(p, s)
    c0 = sim.conserved()
    for k in range(5):
        dt_o = orc.step()
        dt_g = sim.step()
        assert abs(dt_g - dt_o) <= 1e-11 * dt_o, (k, dt_g, dt_o)
    gs, os_ = sim.state(), orc.state()
    for k in gs:
        assert relerr(gs[k], os_[k]) < 1e-9, k
    c1 = sim.conserved()  # [volume, mass, energy, momentum...]
    assert abs(c1[1] - c0[1]) <= 1e-13 * abs(c0[1])
    assert abs(c1[2] - c0[2]) <= 1e-12 * abs(c0[2])
    sim.close()
