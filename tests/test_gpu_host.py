"""The host driver (Particles / MeshlessScheme over the C ABI, HDF5 snapshots)
on the GPU against the library's own step, and the reference's dump schedule."""
import os

import numpy as np
import pytest

from meshlesshydro_amd import ics

pytestmark = pytest.mark.gpu
host = pytest.importorskip("meshlesshydro_amd.host")

PARAMS = """#define DIM 2
#define PERIODIC_BOUNDARIES 1
#define ADAPTIVE_TIMESTEP {adaptive}
#define CFL .4
#define MAX_NUM_INTERACTIONS 400
#define MAX_NUM_GHOST_INTERACTIONS 300
#define SLOPE_LIMITING 1
#define BETA 4.
#define PAIRWISE_LIMITER 0
#define PSI_1 .5
#define PSI_2 .25
#define MESHLESS_FINITE_MASS 0
#define ENFORCE_FLUX_SYM 1
#define MOVE_PARTICLES 1
#define DEBUG_LVL 1
#define FIRST_ORDER_QUAD_POINT 1
"""

CONFIG = """initFile ic.h5
outDir out
timeStep {dt}
timeEnd {tend}
h5DumpInterval {every}
periodicBoxLimits {{
    lowerX 0.
    lowerY 0.
    upperX 1.
    upperY 1.
}}
kernelSize {h}
gamma 1.6666666666666667
"""


def _setup(tmp_path, adaptive, dt, tend, every, n=32):
    s = ics.kelvin_helmholtz(n)
    h = 4.0 / n
    (tmp_path / "parameter.h").write_text(PARAMS.format(adaptive=adaptive))
    (tmp_path / "config.info").write_text(CONFIG.format(dt=dt, tend=tend, every=every, h=h))
    os.makedirs(tmp_path / "out", exist_ok=True)
    host.write_ic(str(tmp_path / "ic.h5"), s, 2)
    return s, h


def test_fixed_step_driver_matches_library(tmp_path, hip):
    """ADAPTIVE_TIMESTEP 0: snapshots every 3 steps hold the state the
    library's own stepping reaches with the same dt (bitwise), plus rho / P /
    noi of that step and the conserved sums."""
    s, h = _setup(tmp_path, 0, 0.002, 0.012, 3)
    steps, t, dumps = host.run(str(tmp_path / "config.info"), str(tmp_path / "parameter.h"))
    files = sorted(os.listdir(tmp_path / "out"))
    assert files == ["000000.h5", "000003.h5", "000006.h5"], files
    p = host.read_parameters(str(tmp_path / "parameter.h"))[0]
    p.kernel_size = h
    sim = hip.Simulation(p, s)
    for k, name in enumerate(files):
        f = str(tmp_path / "out" / name)
        X = host.read_dataset(f, "/x")
        V = host.read_dataset(f, "/v")
        st = sim.state()
        assert np.array_equal(X[:, 0], st["x"]) and np.array_equal(X[:, 1], st["y"]), name
        assert np.array_equal(V[:, 0], st["vx"]) and np.array_equal(host.read_dataset(f, "/u"), st["u"])
        assert host.read_dataset(f, "/time")[0] == pytest.approx(0.006 * k, abs=1e-15)
        sim.neighbours()
        sim.density()
        rho, P, _ = sim.get_density()
        assert np.array_equal(host.read_dataset(f, "/rho"), rho)
        assert np.array_equal(host.read_dataset(f, "/P"), P)
        noi = host.read_dataset_int(f, "/noi")
        assert np.array_equal(noi, sim.get_neighbours()[0])
        assert host.read_dataset(f, "/totalMass")[0] == pytest.approx(np.sum(st["m"]), rel=1e-13)
        assert host.read_dataset(f, "/rhoGrad").shape == (len(rho), 2)
        sim.run(3, dt_fixed=0.002)
    sim.close()


def _reference_loop(sim, time_step, time_end, every):
    """MeshlessScheme::run (MeshlessScheme.cpp:21-254) with ADAPTIVE_TIMESTEP,
    restated over the library's stages: (dump number, time, state) of every
    dump.  The read past dumpTimes' end on the last dump is taken as no clip,
    as in the C++ driver."""
    num = int(time_end / time_step) // every + 1
    dump_times = [k * time_step * every for k in range(num)]
    t, dump_step, dump, dump_next = 0.0, 0, True, False
    out = []
    while True:
        sim.neighbours()
        sim.density()
        dt = sim.timestep()
        assert dump_step < num
        tclip = dump_times[dump_step + 1] if dump_step + 1 < num else np.inf
        if t + dt >= tclip:
            dump_next = True
            dt = tclip - t
        if dump:
            out.append((dump_step, t, sim.state()))
            dump_step += 1
            dump = False
        if t >= time_end:
            break
        sim.fluxes(dt)
        sim.update()
        t += dt
        if dump_next:
            dump, dump_next = True, False
        if not t < time_end + time_step:
            break
    return out


def test_adaptive_dump_schedule(tmp_path, hip):
    """ADAPTIVE_TIMESTEP 1: the driver's dumps (numbers, times and states)
    equal those of the reference's loop restated above, CFL steps shorter
    than the dump interval included (the reference's first clip target is
    dumpTimes[1] only until the first dump has incremented dumpStep)."""
    ts, te = 2.0 ** -6, 3 * 2.0 ** -6
    s, h = _setup(tmp_path, 1, ts, te, 1)
    steps, t, dumps = host.run(str(tmp_path / "config.info"), str(tmp_path / "parameter.h"))
    files = sorted(os.listdir(tmp_path / "out"))
    p = host.read_parameters(str(tmp_path / "parameter.h"))[0]
    p.kernel_size = h
    sim = hip.Simulation(p, s)
    ref = _reference_loop(sim, ts, te, 1)
    sim.close()
    assert files == ["%06d.h5" % k for k, _, _ in ref], (files, [r[:2] for r in ref])
    assert dumps == len(ref)
    for (k, tk, st), f in zip(ref, files):
        f = str(tmp_path / "out" / f)
        assert host.read_dataset(f, "/time")[0] == tk
        X = host.read_dataset(f, "/x")
        assert np.array_equal(X[:, 0], st["x"]) and np.array_equal(host.read_dataset(f, "/u"), st["u"])


@pytest.mark.parametrize("adaptive", [0, 1])
def test_per_method_driver_equals_default(tmp_path, hip, adaptive):
    """MeshlessScheme with --per-method calls the Particles methods one at a
    time in the reference's order (compPsijTilde, slopeLimiter,
    compEffectiveFace, compRiemannStatesLR, solveRiemannProblems,
    checkFluxSymmetry, collectFluxes, ...): every snapshot equals the
    default run's, dataset for dataset, bitwise."""
    if adaptive:
        ts, te, every = 2.0 ** -6, 3 * 2.0 ** -6, 1
    else:
        ts, te, every = 0.002, 0.012, 3
    _setup(tmp_path, adaptive, ts, te, every)
    out = {}
    for per_method in (False, True):
        d = tmp_path / ("out_m" if per_method else "out")
        os.makedirs(d, exist_ok=True)
        cfg = (tmp_path / "config.info").read_text().replace("outDir out", f"outDir {d.name}")
        (tmp_path / f"config_{int(per_method)}.info").write_text(cfg)
        res = host.run(str(tmp_path / f"config_{int(per_method)}.info"), str(tmp_path / "parameter.h"),
                       per_method=per_method)
        out[per_method] = (res, sorted(os.listdir(d)), d)
    assert out[False][0] == out[True][0]
    assert out[False][1] == out[True][1] and len(out[False][1]) >= 3
    for name in out[False][1]:
        a, b = str(out[False][2] / name), str(out[True][2] / name)
        for ds in ("/time", "/x", "/v", "/m", "/u", "/rho", "/P", "/rhoGrad", "/totalMass", "/energy"):
            assert np.array_equal(host.read_dataset(a, ds), host.read_dataset(b, ds)), (name, ds)
        assert np.array_equal(host.read_dataset_int(a, "/noi"), host.read_dataset_int(b, "/noi"))
