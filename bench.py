#!/usr/bin/env python3
"""Benchmark of the MFV time step on one MI355X.

Workload (default): the 3D point-energy blast, 128^3 particles on a
cell-centred lattice in the periodic unit cube, rho = 1, P = 1e-6, E = 1 on
the 32 central particles, gamma = 5/3, h = 2.4 dx (56 neighbours).

Metric: MFV particle-steps/s (N x steps / wall time of the step loop, from
the cell list through the state update; no snapshot I/O), FP32 pair
arithmetic + learned p* seed against plain FP64, at fixed L1 error:
  1. FP64 plain runs W + K steps with its own CFL steps (dt recorded);
  2. the FP32 run replays that dt sequence from the same initial state and
     its density after W + K steps is compared with the FP64 one
     (L1_rel(rho) = sum|rho - rho64| / sum|rho64|), so the error measures
     arithmetic, not time-step drift;
  3. the timed region is the last K steps of each run, inputs resident in
     HBM, bracketed by device synchronisation.
`value` is the FP32 (learned seed when trained, else standard seed) rate.

    python bench.py [--steps K] [--warmup W] [--n 128] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue: one wave64 instruction per CU per clock (4 SIMDs x 16 lanes),
# 256 CUs at 2.4 GHz.  Full-rate FP32 and FP64 ops issue at this rate;
# transcendentals and FP64 divides take several issue slots.
VALU_ISSUE_PEAK = 256 * 2.4e9


def face_kernel_bytes(D, tsz, k_mean):
    """Algorithmic HBM bytes per particle of the faces+update kernel: own
    position (double4), primitive record (8 values), face record (grads +
    B), FP64 state read (x, v, m, u), neighbour counts and row, next state
    written.  Neighbour records are gathered through the caches and not
    counted."""
    nf = D + 2
    frn = nf * D + D * D
    own = 32 + 8 * tsz + frn * tsz + 8 * (2 * D + 2) + 8
    row = 4 * k_mean
    write = 8 * (2 * D + 2)
    return own + row + write


def profiled(key, kernel):
    """A per-launch figure of `kernel` from the committed rocprofv3 PMC
    summary profiles/pmc_traffic_blast128.md: the line `<key> <kernel> =
    <value>`, or None when absent.  Keys: traffic_bytes_per_launch (HBM
    bytes, FETCH_SIZE + WRITE_SIZE each scaled by the counters of a 1 GiB
    copy) and valu_instr_per_launch (SQ_INSTS_VALU, wave instructions)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "pmc_traffic_blast128.md")
    try:
        with open(path) as fh:
            for line in fh:
                line = line.strip().strip("`")
                if line.startswith(key + " ") and "=" in line:
                    name, val = line[len(key) + 1:].rsplit("=", 1)
                    if name.strip() == kernel:
                        return float(val)
    except OSError:
        pass
    return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary."""
    return profiled("traffic_bytes_per_launch", kernel)


def valu_roof(kernel, seconds):
    """The face kernel's VALU issue rate against the chip's: wave-level
    VALU instructions per launch (committed PMC summary) over the launch
    time measured in this run."""
    instr = profiled("valu_instr_per_launch", kernel)
    if instr is None or seconds <= 0:
        return None
    rate = instr / seconds
    return {"instr_per_launch": instr, "achieved_per_s": rate, "peak_per_s": VALU_ISSUE_PEAK,
            "frac": rate / VALU_ISSUE_PEAK}


def roofline(kernel, seconds, bytes_per_launch, hbm_gbs):
    """The dominant kernel's roofline.  The face kernel is bound by VALU
    issue (its HBM traffic is a few % of peak): `achieved` / `peak` are VALU
    wave instructions per second (instructions per launch from the
    committed PMC summary, launch time from this run's HIP events), and the
    HBM figures (algorithmic bytes per launch over the same time, against
    8 TB/s) follow as a secondary roof."""
    v = valu_roof(kernel, seconds)
    hbm = {"achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_gbs / HBM_PEAK_GBS,
           "algorithmic_bytes_per_launch": bytes_per_launch}
    base = {"kernel": kernel, "traffic": measured_traffic(kernel), "launch_ms": 1e3 * seconds,
            "hbm": hbm}
    if v is None:  # no VALU count in the committed profile: HBM only
        return dict(base, bound="hbm", achieved=hbm_gbs, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=hbm_gbs / HBM_PEAK_GBS)
    return dict(base, bound="valu", achieved=v["achieved_per_s"], peak=v["peak_per_s"],
                unit="VALU wave-instructions/s", frac=v["frac"],
                valu_instr_per_launch=v["instr_per_launch"])


def sedov_run(args, state, base, log):
    """The blast run to the time t_end at which the Sedov-Taylor shock
    radius is args.sedov_radius, with each mode's own CFL steps (clipped to
    land on t_end), then compared with the analytic solution: shell-mean
    density and pressure profiles, shock position, conservation drift,
    steps and wall time.  Prints one JSON line."""
    from meshlesshydro_amd import analysis, lib
    from meshlesshydro_amd import learned as L
    from meshlesshydro_amd._abi import default_params, MLH_FP32, MLH_FP64, MLH_SEED_LEARNED

    N = len(state["x"])
    # the flag set of the reference's own blast case (testcases/sedov/
    # parameter.h: CFL .25, BETA 1, PAIRWISE_LIMITER 1); under the
    # Kelvin-Helmholtz long-run set (BETA 4, no pairwise limiter) this blast
    # goes non-finite within a few dozen steps, the CPU oracle's too
    base = dict(base, cfl=0.25, beta=1.0, pairwise_limiter=1)
    sol = analysis.SedovTaylor(E=1.0, rho0=1.0, gamma=5.0 / 3.0, nu=3)
    t_end = sol.time_for_radius(args.sedov_radius)
    log(f"[sedov] t_end = {t_end:.6g} (R_s = {args.sedov_radius}), N = {N}")

    def fields(st):
        s = lib.Simulation(default_params(precision=MLH_FP64, **base), st)
        s.neighbours()
        s.density()
        r, P, _ = s.get_density()
        s.close()
        return r, P

    def advance(s, n_max):
        """step with CFL dt clipped to land on t_end; stop before the first
        step whose state has non-finite conserved sums.  Returns (steps, t,
        failed): the reference algorithm ignores a singular E_i (its LAPACK
        call's INFO is unchecked, Helper.cpp:7-18), and the blast's emptied
        centre leaves particles with too few neighbours for a fixed kernel
        size, so a long run can end in NaN as the reference's would."""
        t, steps = 0.0, 0
        while t_end - t > 1e-12 * t_end and steps < n_max:
            dt = s.step(t_now=t, t_clip=t_end)
            if not (np.all(np.isfinite(s.conserved())) and np.isfinite(dt)):
                return steps, t, True
            t += dt
            steps += 1
        return steps, t, False

    def run(prec, seed_params=None, collect=False):
        p = default_params(precision=prec, **base)

        def fresh():
            s = lib.Simulation(p, state)
            if seed_params is not None:
                s.mlp_set_params(seed_params)
            return s

        if seed_params is not None:
            p.riemann_seed = MLH_SEED_LEARNED
        s = fresh()
        faces = None
        if collect:
            s.set_face_collection(1 << 21, 8)
            s.run(5)
            faces = s.collected_faces()
            s.close()
            s = fresh()
        c0 = s.conserved()
        t0 = time.perf_counter()
        steps, t, failed = advance(s, args.sedov_max_steps)
        wall = time.perf_counter() - t0
        if failed:
            # the run is deterministic: replay to the last finite state
            s.close()
            s = fresh()
            advance(s, steps)
        c1 = s.conserved()
        d = s.diag()
        st = s.state()
        s.close()
        rho, P = fields(st)
        x = np.stack([st["x"], st["y"], st["z"]], axis=1)
        cmp = analysis.sedov_compare(x, rho, P, t, solution=sol, nbins=64)
        m0, e0 = c0[1], c0[2]  # conserved sums: volume, mass, energy, momentum
        out = dict(steps=steps, reached_t_end=not failed, nonfinite_after_step=steps if failed else None,
                   wall_s=wall, particle_steps_per_s=N * steps / wall,
                   mass_drift=abs(c1[1] - m0) / abs(m0), energy_drift=abs(c1[2] - e0) / abs(e0),
                   newton_iters_per_face=d["newton_iterations"] / max(1, d["faces_solved"]), **cmp)
        return out, faces

    res = {}
    log("[sedov] FP64 plain")
    res["fp64"], faces = run(MLH_FP64, collect=True)
    log("[sedov] FP32 standard seed")
    res["fp32_standard"], _ = run(MLH_FP32)
    if faces is not None and len(faces):
        p64 = default_params(precision=MLH_FP64, **base)
        s = lib.Simulation(p64, state)
        mlp_params, _ = L.train_seed_from_faces(s, faces, p64.gamma, n_epochs=4, log=log)
        s.close()
        log("[sedov] FP32 learned seed")
        res["fp32_learned"], _ = run(MLH_FP32, seed_params=mlp_params)
    print(json.dumps({"workload": f"3D point blast {round(N ** (1 / 3))}^3 periodic to the Sedov-Taylor "
                                  f"shock radius {args.sedov_radius}",
                      "flags": "CFL .25, BETA 1, PAIRWISE_LIMITER 1 (testcases/sedov/parameter.h)",
                      "t_end": t_end, "xi0": sol.xi0, "modes": res}))


def c3_run(args, log):
    """Riemann microbenchmark: n = 2^c3_log2n random face states
    (ics.riemann_states, seed 20261202) resident in HBM; the batched exact
    solver in FP64 and FP32 with the standard guess; the seed network trained
    on the first n/4 states (log p*/p_guess targets from the FP64 star
    pressure, the library's GPU training); FP32 and FP64 with the network's
    seed.  Reports
    faces/s from the kernel's HIP-event time, Newton iterations per face
    (all states and the held-out 3n/4), fallback rates, and the kernel's
    algorithmic bytes (6 doubles in, 3 doubles + flag + iterations out)
    against HBM.  Prints one JSON line."""
    from meshlesshydro_amd import ics, lib
    from meshlesshydro_amd import learned as L
    from meshlesshydro_amd._abi import (default_params, MLH_FP32, MLH_FP64, MLH_SEED_LEARNED,
                                        MLH_SEED_STANDARD)
    n = 1 << args.c3_log2n
    ntrain = n // 4
    gamma = 5.0 / 3.0
    t0 = time.perf_counter()
    st = ics.riemann_states(n)
    log(f"[c3] {n} states in {time.perf_counter() - t0:.1f} s")
    # device-resident inputs and outputs (the library's own allocations)
    states = lib.DeviceBuffer((6, n), np.float64)
    for k in range(6):
        states.copy_from(st[k], offset_bytes=k * n * 8)
    star = lib.DeviceBuffer((3, n), np.float64)
    flag = lib.DeviceBuffer(n, np.int32)
    iters = lib.DeviceBuffer(n, np.int32)
    ctx = lib.Simulation(default_params(kernel_size=0.1), {k: np.array([0.5]) for k in ("x", "y", "m", "u")} |
                         {"vx": np.zeros(1), "vy": np.zeros(1)})
    bytes_face = 6 * 8 + 3 * 8 + 4 + 4
    res = {}

    def run(name, prec, seed):
        lib.riemann_batch_device(gamma, states, star, flag, iters, prec, seed, ctx=ctx.h)  # warm-up
        fb, ms = lib.riemann_batch_device(gamma, states, star, flag, iters, prec, seed, ctx=ctx.h)
        it = iters.numpy()
        r = {"faces_per_s": n / (ms * 1e-3), "kernel_ms": ms,
             "newton_iters_per_face": float(it.mean()),
             "newton_iters_per_face_heldout": float(it[ntrain:].mean()),
             "fallback_rate": fb / n,
             "hbm_gbs_algorithmic": n * bytes_face / (ms * 1e-3) / 1e9}
        log(f"[c3] {name}: {r}")
        res[name] = r

    run("fp64_standard", MLH_FP64, MLH_SEED_STANDARD)
    run("fp32_standard", MLH_FP32, MLH_SEED_STANDARD)
    # training set: the first n/4 states, targets log(p*/p_guess) from the
    # FP64 star-region pressure, built and trained on the device
    # (trivially converged faces subsampled, a tenth held out)
    t0 = time.perf_counter()
    ctx.mlp_set_params(L.init_params(1))
    info = ctx.mlp_train_faces(np.stack([a[:ntrain] for a in st], axis=1), gamma, epochs=4,
                               batch=4096, lr=2e-3, seed=1)
    train = {"samples": int(info["train_faces"]), "seconds": time.perf_counter() - t0,
             "loss_init": info["loss_init"], "loss_final": info["loss_final"],
             "val_mse_log_pstar": info["val_mse_log_pstar"],
             "val_mse_standard_guess": info["val_mse_standard_guess"]}
    log(f"[c3] training: {train}")
    run("fp32_learned", MLH_FP32, MLH_SEED_LEARNED)
    run("fp64_learned", MLH_FP64, MLH_SEED_LEARNED)
    ctx.close()
    for b in (states, star, flag, iters):
        b.free()
    print(json.dumps({
        "workload": f"Riemann microbenchmark: 2^{args.c3_log2n} random face states (C3), gamma 5/3",
        "faces": n, "training": train, "modes": res,
        "iterations_saved_fp32": res["fp32_standard"]["newton_iters_per_face_heldout"]
        - res["fp32_learned"]["newton_iters_per_face_heldout"],
        "iterations_saved_fp64": res["fp64_standard"]["newton_iters_per_face_heldout"]
        - res["fp64_learned"]["newton_iters_per_face_heldout"],
        "speedup_fp32_learned_vs_fp64": res["fp32_learned"]["faces_per_s"] / res["fp64_standard"]["faces_per_s"],
        "speedup_fp64_learned_vs_fp64": res["fp64_learned"]["faces_per_s"] / res["fp64_standard"]["faces_per_s"]}))


def shear_run(args, log):
    """The 2D shear layer (Kelvin-Helmholtz, reference flags).

    --c1: 100x100 lattice, h = 4 dx, 100 FP64 steps on the GPU and on the CPU
    oracle (both with their own CFL steps): largest state difference
    relative to each field's scale, dt sequence difference, and wall times.
    --c2: 1000x1000, h = 4 dx, 200 steps: FP64 plain against FP32 with the
    standard seed replaying the FP64 dt sequence; rates, L1/Linf errors of
    rho, P, |v| and the conservation drift.  Prints one JSON line."""
    from meshlesshydro_amd import ics, lib
    from meshlesshydro_amd._abi import default_params, MLH_FP32, MLH_FP64

    n, K = (100, 100) if args.c1 else (1000, 200)
    st = ics.kelvin_helmholtz(n)
    p = default_params(kernel_size=4.0 / n, debug_level=0)
    N = n * n

    def fields(state):
        s = lib.Simulation(default_params(kernel_size=4.0 / n, debug_level=0), state)
        s.neighbours()
        s.density()
        r, P, _ = s.get_density()
        s.close()
        return {"rho": r, "P": P, "v": np.sqrt(state["vx"] ** 2 + state["vy"] ** 2)}

    sim = lib.Simulation(p, st)
    c0 = sim.conserved()
    sim.run(1)  # first launch of each kernel outside the timed region
    sim.close()
    sim = lib.Simulation(p, st)
    t0 = time.perf_counter()
    dts = sim.run(K)
    t64 = time.perf_counter() - t0
    c1 = sim.conserved()
    g64 = sim.state()
    sim.close()
    drift = lambda a, b: {"mass": abs(b[1] - a[1]) / abs(a[1]), "energy": abs(b[2] - a[2]) / abs(a[2]),
                          "momentum": float(np.hypot(b[3] - a[3], b[4] - a[4]))}
    if args.c1:
        from oracle import oracle as O
        o = O.Oracle(p, st, threads=args.cpu_threads)
        t0 = time.perf_counter()
        odt = np.array([o.step() for _ in range(K)])
        tc = time.perf_counter() - t0
        os_ = o.state()
        diff = {k: float(np.max(np.abs(g64[k] - os_[k])) / max(np.max(np.abs(os_[k])), 1e-300)) for k in g64}
        print(json.dumps({
            "workload": "shear layer 100x100 (C1), reference flags, 100 FP64 steps",
            "gpu_particle_steps_per_s": N * K / t64, "cpu_oracle_particle_steps_per_s": N * K / tc,
            "cpu_threads": args.cpu_threads, "gpu_over_cpu": tc / t64,
            "max_rel_state_difference": diff,
            "max_rel_dt_difference": float(np.max(np.abs(dts - odt) / odt)),
            "t_end": float(np.sum(dts)), "drift": drift(c0, c1)}))
        return
    ref = fields(g64)
    p32 = default_params(kernel_size=4.0 / n, debug_level=0, precision=MLH_FP32)
    s32 = lib.Simulation(p32, st)
    s32.run(1, dts_in=dts[:1])
    s32.close()
    s32 = lib.Simulation(p32, st)
    e0 = s32.conserved()
    t0 = time.perf_counter()
    s32.run(K, dts_in=dts)
    t32 = time.perf_counter() - t0
    e1 = s32.conserved()
    tim32 = s32.timings()
    out = fields(s32.state())
    s32.close()
    err = {}
    for k, b in ref.items():
        d = np.abs(out[k] - b)
        err["L1_rel_" + k] = float(np.sum(d) / np.sum(np.abs(b)))
        err["Linf_rel_" + k] = float(np.max(d) / np.max(np.abs(b)))
    print(json.dumps({
        "workload": "shear layer 1000x1000 (C2), reference flags, 200 steps",
        "fp64_plain": {"particle_steps_per_s": N * K / t64, "drift": drift(c0, c1)},
        "fp32_standard": {"particle_steps_per_s": N * K / t32, "kernel_ms": [round(float(x), 4) for x in tim32],
                          "drift": drift(e0, e1)},
        "speedup_vs_fp64": t64 / t32, "errors_vs_fp64": err}))


def c5_run(args, log):
    """The decaying turbulent box: glass relaxation and Mach-1 solenoidal
    velocity (meshlesshydro_amd.turbulence), FP64 plain with its own CFL
    steps (an untimed run of its first 5 steps records the faces for the
    initial training), then
    FP32 with the learned seed replaying that dt sequence and retraining
    online every args.retrain_every steps.  Rates count the step loop;
    `with_retraining` also counts the retraining.  Prints one JSON line."""
    from meshlesshydro_amd import learned as L
    from meshlesshydro_amd import lib, turbulence
    from meshlesshydro_amd._abi import default_params, MLH_FP32, MLH_FP64, MLH_SEED_LEARNED

    n, K = args.c5_n, args.steps
    N = n ** 3
    t0 = time.perf_counter()
    st, info = turbulence.turbulent_box(n)
    t_ic = time.perf_counter() - t0
    log(f"[c5] {n}^3 initial state in {t_ic:.1f} s, density spread {info['density_spread']}")
    # the reference's list limits (parameter_long_run.h: 400 / 300): Mach-1
    # compressions raise the 64 mean neighbours several-fold
    base = dict(dim=3, kernel_size=info["kernel_size"], max_interactions=400,
                max_ghost_interactions=300, debug_level=0)

    def fields(state):
        s = lib.Simulation(default_params(precision=MLH_FP64, **base), state)
        s.neighbours()
        s.density()
        r, P, _ = s.get_density()
        s.close()
        return {"rho": r, "P": P, "v": np.sqrt(state["vx"] ** 2 + state["vy"] ** 2 + state["vz"] ** 2)}

    p64 = default_params(precision=MLH_FP64, **base)
    # training faces: the FP64 solver's faces over the first 5 steps, recorded
    # in an untimed run (the timed FP64 run below repeats those steps
    # bitwise, without recording)
    pre = lib.Simulation(p64, st)
    pre.set_face_collection(1 << 21, 8)
    pre.run(5)
    faces = pre.collected_faces()
    pre.close()
    sim = lib.Simulation(p64, st)
    c0 = sim.conserved()
    ta = time.perf_counter()
    dts = sim.run(K)
    t64 = time.perf_counter() - ta
    tim64 = sim.timings()
    c1 = sim.conserved()
    ref = fields(sim.state())
    weights, train0 = L.train_seed_from_faces(sim, faces, p64.gamma, n_epochs=4, log=log)
    sim.close()

    # FP32, standard seed, same dt sequence
    p32s = default_params(precision=MLH_FP32, **base)
    s32s = lib.Simulation(p32s, st)
    e0 = s32s.conserved()
    tc = time.perf_counter()
    s32s.run(K, dts_in=dts)
    t32s = time.perf_counter() - tc
    std = {"particle_steps_per_s": N * K / t32s,
           "kernel_ms": [round(float(x), 4) for x in s32s.timings()]}
    dstd = s32s.diag()
    std["newton_iters_per_face"] = dstd["newton_iterations"] / max(1, dstd["faces_solved"])
    std["acoustic_rate"] = dstd["acoustic_faces"] / max(1, dstd["faces_solved"])
    std["drift"] = {"mass": abs(s32s.conserved()[1] - e0[1]) / abs(e0[1]),
                    "energy": abs(s32s.conserved()[2] - e0[2]) / abs(e0[2])}
    outs = fields(s32s.state())
    s32s.close()
    std["errors_vs_fp64"] = {"L1_rel_" + k: float(np.sum(np.abs(outs[k] - b)) / np.sum(np.abs(b)))
                             for k, b in ref.items()}

    p32 = default_params(precision=MLH_FP32, **base)
    p32.riemann_seed = MLH_SEED_LEARNED
    s32 = lib.Simulation(p32, st)
    s32.mlp_set_params(weights)
    d0 = s32.conserved()
    tb = time.perf_counter()
    _, infos = turbulence.run_online(s32, K, retrain_every=args.retrain_every, stride=8, epochs=2,
                                     dts=dts, log=log)
    t32 = time.perf_counter() - tb
    t_train = sum(i.get("seconds", 0.0) for i in infos)
    d1 = s32.conserved()
    diag = s32.diag()
    tim32 = s32.timings()
    out = fields(s32.state())
    s32.close()
    err = {}
    for k, b in ref.items():
        d = np.abs(out[k] - b)
        err["L1_rel_" + k] = float(np.sum(d) / np.sum(np.abs(b)))
        err["Linf_rel_" + k] = float(np.max(d) / np.max(np.abs(b)))
    drift = lambda a, b: {"mass": abs(b[1] - a[1]) / abs(a[1]), "energy": abs(b[2] - a[2]) / abs(a[2])}
    print(json.dumps({
        "workload": f"decaying turbulent box {n}^3 (C5): glass, Mach 1 solenoidal P(k)~k^-4, k in [2,8]",
        "steps": K, "particles": N, "retrain_every": args.retrain_every,
        "fp32_learned_online": {"particle_steps_per_s": N * K / (t32 - t_train),
                                "with_retraining": N * K / t32, "retraining_s": t_train,
                                "kernel_ms_last_segment": [round(float(x), 4) for x in tim32],
                                "newton_iters_per_face": diag["newton_iterations"] / max(1, diag["faces_solved"]),
                                "seeded_rate": diag["learned_seeded"] / max(1, diag["faces_solved"]),
                                "acoustic_rate": diag["acoustic_faces"] / max(1, diag["faces_solved"]),
                                "drift": drift(d0, d1)},
        "fp64_plain": {"particle_steps_per_s": N * K / t64, "kernel_ms": [round(float(x), 4) for x in tim64],
                       "drift": drift(c0, c1)},
        "fp32_standard": std,
        "speedup_vs_fp64": t64 / (t32 - t_train),
        "errors_vs_fp64": err,
        "initial_training": train0,
        "retraining": infos,
        "initial_state_s": t_ic, "density_spread": info["density_spread"],
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128, help="particles per side")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=0, help="side of the CPU sample (default: --n)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="profile helper: one 1 GiB device-to-device copy kernel (torch) "
                         "to scale FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--only", choices=["fp64", "fp32", "learned"], default=None,
                    help="profile helper: run one mode for --steps steps and print its timings")
    ap.add_argument("--sedov", action="store_true",
                    help="run the blast to the time its Sedov-Taylor shock reaches --sedov-radius "
                         "(own CFL steps, FP64, FP32 and FP32 + learned seed) and compare the "
                         "final density and pressure profiles with the analytic solution")
    ap.add_argument("--sedov-radius", type=float, default=0.07)
    ap.add_argument("--sedov-max-steps", type=int, default=20000)
    ap.add_argument("--c5", action="store_true",
                    help="the decaying turbulent box (glass, Mach-1 solenoidal velocity): FP64 plain "
                         "against FP32 + learned seed retrained online every --retrain-every steps, "
                         "--steps steps each, at --c5-n^3 particles")
    ap.add_argument("--c5-n", type=int, default=256)
    ap.add_argument("--c3", action="store_true",
                    help="Riemann microbenchmark: 2^26 random face states resident on the GPU, FP64 and "
                         "FP32 with the standard guess and with the network trained on 2^24 solver outputs")
    ap.add_argument("--c3-log2n", type=int, default=26)
    ap.add_argument("--c1", action="store_true",
                    help="the reference's shear layer (100x100, FP64, 100 steps) on the GPU and on "
                         "the CPU oracle: state difference and both times")
    ap.add_argument("--c2", action="store_true",
                    help="the shear layer at 1000x1000: FP64 against FP32 (standard seed), 200 steps")
    ap.add_argument("--retrain-every", type=int, default=50)
    ap.add_argument("--order", choices=["auto", "cell", "caller"], default="auto",
                    help="work order of the kernels (results are identical)")
    ap.add_argument("--shuffle", action="store_true",
                    help="randomly permute the particle order of the initial state")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this MFV step has no multi-GPU path "
              "(it would shard by space, one slab of cells per GPU); use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.calib_copy:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4}))
        return

    from meshlesshydro_amd import ics, lib
    from meshlesshydro_amd import learned as L
    from meshlesshydro_amd._abi import default_params, MLH_FP32, MLH_FP64, MLH_SEED_LEARNED

    if args.c5:
        c5_run(args, lambda *a: print(*a, file=sys.stderr, flush=True))
        return
    if args.c3:
        c3_run(args, lambda *a: print(*a, file=sys.stderr, flush=True))
        return
    if args.c1 or args.c2:
        shear_run(args, lambda *a: print(*a, file=sys.stderr, flush=True))
        return

    n = args.n
    N = n ** 3
    W, K = args.warmup, args.steps
    state = ics.blast_3d(n)
    if args.shuffle:
        perm = np.random.default_rng(77).permutation(N)
        state = {k: v[perm] for k, v in state.items()}
    order = {"auto": 0, "cell": 1, "caller": 2}[args.order]
    base = dict(dim=3, kernel_size=ics.blast_kernel_size(n), max_interactions=96,
                max_ghost_interactions=96, debug_level=0, work_order=order)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)

    if args.only:
        # one mode only, for rocprofv3 runs (no error measurement)
        prec = MLH_FP64 if args.only == "fp64" else MLH_FP32
        p = default_params(precision=prec, **base)
        s = lib.Simulation(p, state)
        if args.only == "learned":
            faces = L.collect_faces(default_params(precision=MLH_FP64, **base), state,
                                    steps=max(W, 1), stride=8)
            params_mlp, _ = L.train_seed_from_faces(s, faces, p.gamma, n_epochs=2, log=log)
            s.close()
            p.riemann_seed = MLH_SEED_LEARNED
            s = lib.Simulation(p, state)
            s.mlp_set_params(params_mlp)
        s.run(W)
        t0 = time.perf_counter()
        s.run(K)
        t = time.perf_counter() - t0
        print(json.dumps({"mode": args.only, "particle_steps_per_s": N * K / t,
                          "kernel_ms": [float(x) for x in s.timings()], "diag": s.diag()}))
        return

    if args.sedov:
        sedov_run(args, state, base, log)
        return

    # 1. FP64 plain, own CFL steps; the warmup steps also record the exact
    #    solver's faces (every 8th particle) as training data for the seed
    p64 = default_params(precision=MLH_FP64, **base)
    sim = lib.Simulation(p64, state)
    log(f"[bench] FP64 plain: {W} warmup + {K} timed steps, N={N}")
    if not args.no_learned:
        sim.set_face_collection(1 << 21, 8)
    c64_start = sim.conserved()
    dts_w = sim.run(W)
    faces = sim.collected_faces() if not args.no_learned else None
    sim.set_face_collection(0, 0)
    t0 = time.perf_counter()
    dts_k = sim.run(K)
    t64 = time.perf_counter() - t0
    tim64 = sim.timings()
    d64 = sim.diag()
    cons64 = (c64_start, sim.conserved())
    ref = sim.state()
    dts = np.concatenate([dts_w, dts_k])
    rate64 = N * K / t64
    mlp_params, train_info = None, None
    if not args.no_learned:
        # held-out faces: two further FP64 steps after the timed ones (the
        # seed trains on the warmup steps' faces only)
        sim.set_face_collection(1 << 21, 8)
        sim.run(2)
        held = sim.collected_faces()
        sim.set_face_collection(0, 0)
        log(f"[bench] training the p* seed on {len(faces)} FP64 faces of the warmup steps")
        mlp_params, train_info = L.train_seed_from_faces(sim, faces, p64.gamma, n_epochs=4, log=log)
        sub = held[np.random.default_rng(9).choice(len(held), min(len(held), 200_000), replace=False)]
        train_info["heldout_faces"] = int(len(sub))
        train_info["heldout_source"] = f"2 FP64 steps after the {W}+{K} steps (not trained on)"
        train_info["newton_iters_heldout_fp32"] = L.newton_stats(sim, p64.gamma, sub)
    sim.close()

    # 2. FP32 runs replaying the FP64 dt sequence
    def fp32_run(seed_params=None):
        p = default_params(precision=MLH_FP32, **base)
        if seed_params is not None:
            p.riemann_seed = MLH_SEED_LEARNED
        s = lib.Simulation(p, state)
        if seed_params is not None:
            s.mlp_set_params(seed_params)
        c_start = s.conserved()
        s.run(W, dts_in=dts[:W])
        t0 = time.perf_counter()
        s.run(K, dts_in=dts[W:])
        t = time.perf_counter() - t0
        out = (N * K / t, s.timings(), s.diag(), s.state(), (c_start, s.conserved()))
        s.close()
        return out

    log("[bench] FP32 standard seed, replayed dt")
    rate32, tim32, d32, st32, cons32 = fp32_run()
    learned = None
    if mlp_params is not None:
        log("[bench] FP32 learned seed, replayed dt")
        learned = fp32_run(mlp_params) + (train_info,)

    # density error: rho = m * omega, omega from the FP64 neighbour sets of
    # each final state (one FP64 density pass)
    def final_fields(st):
        """rho, P (one FP64 density pass over the state) and |v|"""
        p = default_params(precision=MLH_FP64, **base)
        s = lib.Simulation(p, st)
        s.neighbours()
        s.density()
        r, P, _ = s.get_density()
        s.close()
        v = np.sqrt(st["vx"] ** 2 + st["vy"] ** 2 + st["vz"] ** 2)
        return {"rho": r, "P": P, "v": v}

    ref_f = final_fields(ref)

    def field_err(st):
        """relative L1 and Linf errors of rho, P, |v| against the FP64 run"""
        f = final_fields(st)
        out = {}
        for k, b in ref_f.items():
            d = np.abs(f[k] - b)
            out["L1_rel_" + k] = float(np.sum(d) / np.sum(np.abs(b)))
            out["Linf_rel_" + k] = float(np.max(d) / np.max(np.abs(b)))
        return out

    def drift(cons):
        """relative change of total mass and energy, momentum change over
        the total |momentum| scale sum m |v| of the end state"""
        (c0, c1) = cons
        return {"mass": abs(c1[1] - c0[1]) / abs(c0[1]), "energy": abs(c1[2] - c0[2]) / abs(c0[2]),
                "momentum": float(np.sqrt(sum((c1[k] - c0[k]) ** 2 for k in (3, 4, 5))))}

    e32 = field_err(st32)
    head = dict(rate=rate32, tim=tim32, diag=d32, err=e32, drift=drift(cons32), seed="standard")
    if learned is not None:
        head = dict(rate=learned[0], tim=learned[1], diag=learned[2], err=field_err(learned[3]),
                    drift=drift(learned[4]), seed="learned", train=learned[5])

    face_kernel = "k_faces<3, float, %s>" % ("true" if head["seed"] == "learned" else "false")
    k_mean = (d64["faces_solved"] / max(1, (W + K) * N))
    tface = float(head["tim"][3]) * 1e-3
    bytes_face = face_kernel_bytes(3, 4, k_mean) * N
    achieved = bytes_face / tface / 1e9

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        cn = args.cpu_n or n
        cst = ics.blast_3d(cn)
        cp = default_params(dim=3, kernel_size=ics.blast_kernel_size(cn), max_interactions=96,
                            max_ghost_interactions=96, debug_level=0)
        log(f"[bench] CPU oracle sample: 1 step of the {cn}^3 blast, {args.cpu_threads} threads")
        o = O.Oracle(cp, cst, threads=args.cpu_threads)
        t0 = time.perf_counter()
        o.step()
        tc = time.perf_counter() - t0
        cpu = {"value": cn ** 3 / tc, "unit": "particle-steps/s", "cores": args.cpu_threads,
               "kind": "port",
               "sample": f"1 FP64 step of the {cn}^3 blast (same workload) by oracle/ (C, OpenMP), {tc:.1f} s"}
        del o

    out = {
        "metric": "MFV particle-steps/s, FP32+learned p* vs FP64 plain, at fixed L1 error",
        "value": head["rate"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * N / head["rate"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (FP64 E_i inverse, Newton residual, flux sums, state)",
        "data": "synthetic",
        "config": {"workload": f"3D point blast {n}^3 periodic (C4)", "particles": N,
                   "neighbours_mean": round(k_mean, 2), "seed": head["seed"],
                   "parallelism": "single GPU"},
        "fp64_plain": {"value": rate64, "ms_per_step": 1e3 * t64 / K,
                       "kernel_ms": [round(float(x), 4) for x in tim64],
                       "drift": drift(cons64)},
        "speedup_vs_fp64": head["rate"] / rate64,
        "fp32_standard": dict(value=rate32, kernel_ms=[round(float(x), 4) for x in tim32],
                              newton_iters_per_face=d32["newton_iterations"] / max(1, d32["faces_solved"]),
                              acoustic_rate=d32["acoustic_faces"] / max(1, d32["faces_solved"]),
                              drift=drift(cons32), **e32),
        "L1_rel_rho": head["err"]["L1_rel_rho"],
        "Linf_rel_rho": head["err"]["Linf_rel_rho"],
        "errors_vs_fp64": head["err"],
        "conservation_drift": head["drift"],
        "kernel_ms": [round(float(x), 4) for x in head["tim"]],
        "newton_iters_per_face": head["diag"]["newton_iterations"] / max(1, head["diag"]["faces_solved"]),
        # FP32 faces given the acoustic solution (relative jumps below 1e-5)
        "acoustic_rate": head["diag"]["acoustic_faces"] / max(1, head["diag"]["faces_solved"]),
        "fallback_rate": head["diag"]["learned_fallbacks"] / max(1, head["diag"]["faces_solved"]),
        "seeded_rate": head["diag"]["learned_seeded"] / max(1, head["diag"]["faces_solved"]),
        "nonconverged_rate": head["diag"]["nonconverged_faces"] / max(1, head["diag"]["faces_solved"]),
        "roofline": roofline(face_kernel, tface, bytes_face, achieved),
        "cpu_baseline": cpu,
    }
    if "train" in head:
        out["learned_seed"] = head["train"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
