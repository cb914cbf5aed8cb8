#!/usr/bin/env python3
"""Throughput of the finite-volume time step on one MI355X.

Default workload: 3D radial Sod, 256^3 uniform cells on [-1,1]^3, gamma 1.4,
inside r<0.4 rho=1 p=1, outside rho=0.125 p=0.1, at rest, transmissive
(free-flow) borders, first-order Rusanov + SSP-RK3 at CFL 0.5 towards t=0.25,
FP64.  --workload sphere192 runs the immersed-sphere case instead (192^3 on
[-2,2]^3, sphere r=0.25, Mach 0.3 free stream, Dirichlet inflow at x-,
free-flow elsewhere); --workload batch512 runs 512 independent seeded 2D
cases together (persistent stage kernel, graph-replayed step); --workload
learned128 trains the learned face reconstruction on the device and compares
its L1 density error and wall time with the plain limiter; --workload
octree128 runs radial Sod on a 2:1 refined octree mesh through the explicit
interface-list kernels.

A step is one SSP-RK3 time step (three fused flux+update stage kernels and
the device-side dt/t update) over the whole grid, with the state resident
in HBM.  value = solved cells x 3 stages x steps / wall time of the timed
steps (cell-updates per second per RK stage).

Prints ONE JSON line on stdout.  The roofline entry is for the stage kernel:
achieved = algorithmic bytes per launch (80/120/120 B per cell for stages
1/2/3, mean 106.67 B; +1 B per cell for the body mask) / average launch
duration: one HIP event pair recorded on the solver's stream around the
timed steps (replayed from the step graphs, as in any run) divided by their
stage launches, so the short gaps between launches count as kernel time.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_STAGE = (80.0 + 120.0 + 120.0) / 3.0
GAMMA = 1.4


def log(msg):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


# ---- workloads ----------------------------------------------------------------

def radsod_setup(mmf, n, order):
    pb = mmf.problem("radsod", 3, sod_r0=0.4, domain_origin=[-1.0, -1.0, -1.0], domain_length=2.0)
    s = mmf.Solver.for_problem(pb, n, order=order, side_bc=[mmf.BC_FREE_FLOW] * 6)
    s.set_initial_state(pb)
    # a fixed number of steps from t = 0 (the problem's end time 0.25 comes after ~140 steps;
    # longer runs go on past it, so any --steps is measured)
    s.set_time(0.0, 1e30, 0.5)
    cfg = {"domain": "[-1,1]^3", "initial": "r<0.4: rho=1 p=1; else rho=0.125 p=0.1; u=0",
           "boundaries": "transmissive (free flow)",
           "time": "SSP-RK3, CFL 0.5, fixed step count from t = 0 (t_end 0.25 after ~140 steps)"}
    return s, 0.0, cfg


def sphere_setup(mmf, n, order):
    fs = [1.0 / GAMMA, 0.3, 0.0, 0.0, 1.0 / GAMMA]  # rho = 1, p = 1/gamma, a = 1, Mach 0.3
    pb = mmf.problem("freestream", 3, fs_prim=fs, domain_origin=[-2.0, -2.0, -2.0], domain_length=4.0)
    side_bc = [mmf.BC_DIRICHLET] + [mmf.BC_FREE_FLOW] * 5
    s = mmf.Solver(3, [n] * 3, [-2.0, -2.0, -2.0], 4.0 / n, side_bc, [fs] * 6, order)
    s.set_bodies(spheres=[[0.0, 0.0, 0.0, 0.25]])
    s.set_initial_state(pb)
    s.set_time(0.0, 1e30, 0.5)  # the case runs a fixed number of steps
    cfg = {"domain": "[-2,2]^3", "body": "sphere r=0.25 at the origin (level set at centroids)",
           "initial": "free stream rho=1, p=1/gamma, u=0.3 (Mach 0.3)",
           "boundaries": "Dirichlet inflow at x-, free flow elsewhere", "time": "SSP-RK3, CFL 0.5, 2000 steps"}
    return s, 1.0, cfg


PREWARM_S = 0.1        # GPU time that brings the clocks to their steady state
PREWARM_MAX_STEPS = 60


def prewarm(advance, sync):
    """Run steps of a second instance of the workload for about PREWARM_S of
    wall time right before the timed instance starts (whose state stays at
    t = 0): the first ~20 ms of work after an idle GPU run up to 1.4x slower,
    and steps after a pause of a few hundred ms up to 13% slower."""
    n = 0
    t0 = time.perf_counter()
    while n < PREWARM_MAX_STEPS and time.perf_counter() - t0 < PREWARM_S:
        advance(5)
        sync()
        n += 5
    return n


WORKLOADS = {"radsod256": (radsod_setup, 256, "radial_sod_3d"), "sphere192": (sphere_setup, 192, "sphere_3d"),
             "batch512": (None, 512, "batch_2d_mixed"), "learned128": (None, 128, "learned_2d"),
             "octree128": (None, 128, "octree_sod_3d")}
MLP_FLOPS_PER_CELL = 2 * (100 * 64 + 64 * 64 + 64 * 16)


def cpu_threads():
    """Host threads for the all-cores CPU baseline: OMP_NUM_THREADS when set
    (the share a GPU box grants), else the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _host():
    return "%s, %d CPUs visible" % (cpu_model(), os.cpu_count() or 0)


def _both(single, multi, threads, what):
    """cpu_baseline entry: the all-cores rate as the value, the single-thread
    rate beside it (the reference runs serially per rank)."""
    return {
        "value": multi[0],
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": "%s; all cores: %s, oracle/mmf_oracle.c -O3 -fopenmp (threaded face loop and RK update, bit for "
                  "bit the serial oracle) with %d threads on %s" % (what, multi[1], threads, _host()),
        "single_thread": {"value": single[0], "cores": 1,
                          "sample": "%s, oracle/mmf_oracle.c -O3, one thread" % single[1]},
    }


def cpu_baseline(n, steps, budget_mt=6.0):
    """The oracle on a bounded sample of the radial-Sod workload: 'steps' RK3
    steps on one thread, then steps on all host cores for about budget_mt s."""
    from oracle import oracle as O
    pb = O.problem(O.RADSOD, 3, sod_r0=0.4)
    s = O.Solver(pb, 1, [n] * 3, [-1.0, -1.0, -1.0], 2.0 / n, side_bc=[O.BC_FREE_FLOW] * 6)
    threads = cpu_threads()
    res = []
    for nt in (1, threads):
        U = s.initial_state()
        W = U.copy()
        t = 0.0
        k = 0
        t0 = time.perf_counter()
        while (nt == 1 and k < steps) or (nt > 1 and (k < 2 or time.perf_counter() - t0 < budget_mt)):
            dt, _ = s.rk3_step(U, W, 0.5, t, 0.25, threads=nt)
            t += dt
            k += 1
        el = time.perf_counter() - t0
        res.append((n ** 3 * 3 * k / el, "%d SSP-RK3 step(s) (%d stage updates of %d^3 cells) in %.1f s" % (k, 3 * k, n, el)))
    return _both(res[0], res[1], threads, "the same radial-Sod workload from its initial condition")


def cpu_baseline_sphere(gpu_solver, n, budget_s=10.0, budget_mt=5.0):
    """The oracle stepping the immersed-sphere workload (the GPU solver's own
    fluid flags, Dirichlet inflow) from its initial condition: about budget_s
    on one thread, then about budget_mt on all host cores."""
    import numpy as np
    from oracle import oracle as O
    fs = [1.0 / GAMMA, 0.3, 0.0, 0.0, 1.0 / GAMMA]
    fluid = gpu_solver.cell_flags() & 1
    po = O.problem(O.FFSTEP + 1, 3, fs_prim=fs)  # the uniform free stream
    s = O.Solver(po, 1, [n] * 3, [-2.0] * 3, 4.0 / n, side_bc=[O.BC_DIRICHLET] + [O.BC_FREE_FLOW] * 5,
                 side_info=np.array([fs] * 6), fluid=fluid)
    solved = int(fluid.sum())
    threads = cpu_threads()
    res = []
    for nt, budget in ((1, budget_s), (threads, budget_mt)):
        U = s.initial_state()
        W = U.copy()
        t = 0.0
        k = 0
        t0 = time.perf_counter()
        while k < 1 or time.perf_counter() - t0 < budget:
            dt, _ = s.rk3_step(U, W, 0.5, t, 1e30, threads=nt)
            t += dt
            k += 1
        el = time.perf_counter() - t0
        res.append((solved * 3 * k / el, "%d SSP-RK3 step(s) (%d stage updates of the %d solved cells of %d^3) in "
                    "%.1f s" % (k, 3 * k, solved, n, el)))
    return _both(res[0], res[1], threads, "the same sphere workload from its initial condition")


def cpu_baseline_batch(descs, problems, budget_s=10.0, budget_mt=5.0):
    """The oracle stepping the cases of the batch from their initial
    conditions, one SSP-RK3 step per case per pass in case order: on one
    thread for about budget_s, then the cases spread over all host cores
    (one case per thread at a time) for about budget_mt."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    keys = {0: O.VORTEX_XY, 3: O.RADSOD}

    def make_cases():
        cases = []
        for d, pm in zip(descs, problems):
            po = O.problem(keys[pm.type], 2, vortex_beta=pm.vortex_beta, vortex_center=list(pm.vortex_center),
                           vortex_periodic=pm.vortex_periodic, domain_origin=list(pm.domain_origin),
                           domain_length=pm.domain_length, sod_r0=pm.sod_r0, sod_p_in=pm.sod_p_in,
                           sod_t_in=pm.sod_t_in, sod_p_out=pm.sod_p_out, sod_t_out=pm.sod_t_out)
            n = [int(d.n[0]), int(d.n[1])]
            sv = O.Solver(po, 1, n, list(d.origin), d.h, side_bc=[d.side_bc[k] for k in range(6)])
            U = sv.initial_state()
            cases.append([sv, U, U.copy(), 0.0, n[0] * n[1]])
        return cases

    def step(c):
        dt, _ = c[0].rk3_step(c[1], c[2], 0.5, c[3], 1e30)
        c[3] += dt
        return c[4]

    threads = cpu_threads()
    res = []
    cases = make_cases()
    cells, updates, el, passes = 0, 0, 0.0, 0
    while el < budget_s:
        for c in cases:
            t0 = time.perf_counter()
            cells += step(c)
            el += time.perf_counter() - t0
            updates += 1
            if el >= budget_s:
                break
        passes += 1
    res.append((cells * 3 / el, "%d SSP-RK3 case-steps (%d pass(es) in case order, %d cells stepped) in %.1f s"
                % (updates, passes, cells, el)))
    cases = make_cases()
    cells, updates, passes = 0, 0, 0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:  # the oracle releases the GIL inside each step
        while passes < 1 or time.perf_counter() - t0 < budget_mt:
            cells += sum(ex.map(step, cases))
            updates += len(cases)
            passes += 1
    el = time.perf_counter() - t0
    res.append((cells * 3 / el, "%d SSP-RK3 case-steps (%d whole pass(es), %d cells stepped) in %.1f s"
                % (updates, passes, cells, el)))
    return _both(res[0], res[1], threads, "the %d cases of the same 512-case set" % len(descs))


def run_batch(args, mmf):
    """512 independent 2D cases advanced together: persistent stage kernel,
    graph-replayed step, launch shape from the autotuner."""
    from minimmerflow_amd import workloads as WL
    t_setup = time.perf_counter()
    descs, problems, params = WL.mixed_2d_cases(mmf, args.ncases, WL.DEFAULT_SEED)
    b = mmf.Batch(descs)
    b.set_initial_states(problems)
    b.set_time(0.0, 1e30, 0.5)  # a fixed number of steps: every case stays active
    w = mmf.Batch(descs)  # the pre-warm instance
    w.set_initial_states(problems)
    w.set_time(0.0, 1e30, 0.5)
    N = b.total_cells
    log("setup %.1fs: %d cases, %d cells, %d work items" % (time.perf_counter() - t_setup, b.ncases, N, b.nitems))
    if args.chunk or args.blocks:
        b.set_tail(args.tail)
        b.set_shape(args.chunk, args.blocks)
        tuned = None
    else:
        chunk, blocks, ms = b.autotune(steps_per_trial=3)
        tuned = {"chunk": chunk, "tail": b.tail(), "blocks": blocks, "ms_per_step": ms}
        log("autotune: %d rows per item, tail items of %d rows, %d persistent blocks (%.4f ms/step)"
            % (chunk, b.tail(), blocks, ms))
    chunk, blocks = b.shape()
    w.set_tail(b.tail())
    w.set_shape(chunk, blocks)
    nprewarm = prewarm(w.advance, w.times)
    b.advance(args.warmup)
    b.times()
    b.enable_timing(True)
    t0 = time.perf_counter()
    b.advance(args.steps)
    t, steps = b.times()
    el = time.perf_counter() - t0
    step_ms, nsteps = b.read_timing()
    b.enable_timing(False)
    if not (steps == args.warmup + args.steps).all():
        raise SystemExit("bench: some cases did not take every step")
    value = N * 3 * args.steps / el
    avg_step_ms = step_ms / max(nsteps, 1)
    achieved = BYTES_PER_CELL_STAGE * 3 * N / (avg_step_ms * 1e-3) / 1e9
    workload = "batch_2d_mixed_%d_order1" % b.ncases
    log("timed %d steps in %.3f s: %.3f ms/step (graph replay avg %.4f ms), %.0f GB/s"
        % (args.steps, el, el / args.steps * 1e3, avg_step_ms, achieved))
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline: oracle over a sample of the cases ...")
        cpu = cpu_baseline_batch(descs, problems)
        log("cpu baseline: %.3e cell-updates/s" % cpu["value"])
    config = {"workload": workload, "cases": b.ncases, "cells": N, "work_items": b.nitems,
              "case_set": "SplitMix64 seed 0x6D6D66: side ~ {32,48,...,256}; cases 0-255 isentropic vortex "
                          "(beta~U[2,6], centre~U[-2,2]^2, [-5,5]^2, periodic), cases 256-511 radial Sod "
                          "(rho and p ratios ~U[4,16], [-1,1]^2, r0=0.4, transmissive)",
              "sides": WL.side_histogram(params), "time": "SSP-RK3, CFL 0.5, fixed step count, per-case dt",
              "shape": {"rows_per_item": chunk, "tail_rows": b.tail(), "persistent_blocks": blocks,
                        "autotuned": tuned},
              "prewarm_steps": nprewarm,
              "reconstruction": "order 1", "flux": "Rusanov", "parallelism": "single GPU"}
    out = {
        "metric": "cell-updates/s per RK stage",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded initial conditions sampled at cell centroids)",
        "config": config,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic(workload),
            "kernel": "k_batch2d (persistent, fused flux + RK stage), whole graph-replayed step "
                      "(three stage launches + per-case time update)",
            "algorithmic_bytes_per_step": BYTES_PER_CELL_STAGE * 3 * N,
            "avg_step_ms": avg_step_ms,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline_order2(n, budget_s=8.0, budget_mt=4.0):
    """The oracle: order-2 periodic vortex at n^2, RK3 steps from the initial
    condition, about budget_s on one thread and budget_mt on all host cores."""
    from oracle import oracle as O
    pb = O.problem(O.VORTEX_XY, 2, vortex_periodic=1)
    s = O.Solver(pb, 2, [n, n], list(pb.domain_origin), pb.domain_length / n, side_bc=[O.BC_PERIODIC] * 4 + [0, 0])
    threads = cpu_threads()
    res = []
    for nt, budget in ((1, budget_s), (threads, budget_mt)):
        U = s.initial_state()
        W = U.copy()
        t, k = 0.0, 0
        t0 = time.perf_counter()
        while k < 1 or time.perf_counter() - t0 < budget:
            dt, _ = s.rk3_step(U, W, 0.5, t, 100.0, threads=nt)
            t += dt
            k += 1
        el = time.perf_counter() - t0
        res.append((n * n * 3 * k / el, "%d SSP-RK3 steps in %.1f s" % (k, el)))
    return _both(res[0], res[1], threads, "the plain order-2 periodic vortex at %d^2 (limited linear, Rusanov)" % n)


def run_learned(args, mmf):
    """Learned face reconstruction against the plain limiter: train on the
    device from fine plain runs, then run held-out cases at n^2 (learned and
    plain) and at 2n^2, 4n^2 (plain) to the same end time."""
    import numpy as np
    from minimmerflow_amd import learned as LR
    n = args.n or 128
    theta = LR.THETA if args.theta is None else args.theta
    t0 = time.perf_counter()
    train_cases = LR.case_params(args.train_cases, seed=11)
    per_case = args.snapshots * 64 * 8
    net = mmf.Learner(seed=5, capacity=len(train_cases) * per_case)
    train_theta = LR.THETA_TRAIN if args.train_theta is None else args.train_theta
    LR.build_training_set(net, train_cases, 4 * n, 4, args.snapshots, 64, seed=1, theta=train_theta)
    t_data = time.perf_counter() - t0
    t1 = time.perf_counter()
    loss = None
    for r in range(3):
        loss = net.train(3000, batch=8192, lr0=2e-3 if r == 0 else 5e-4, lr1=2e-5, seed=r + 1)
    t_train = time.perf_counter() - t1
    log("training set %d rows from %d fine %d^2 runs x %d snapshots (%.1fs); 9000 Adam steps, batch 8192 (%.1fs), "
        "loss %.3e" % (net.num_samples(), len(train_cases), 4 * n, args.snapshots, t_data, t_train, loss))
    cases = []
    ratios, speed256, speed512 = [], [], []
    for c in LR.case_params(args.eval_cases, seed=99):
        runs = {}
        for m in (n, 2 * n, 4 * n):
            runs["plain%d" % m] = LR.run_case(c, m)
        runs["learned%d" % n] = LR.run_case(c, n, net=net, theta=theta)
        ref = runs["plain%d" % (4 * n)][0]
        out = {"case": c}
        for key, (sv, pb, steps, wall) in runs.items():
            m = sv.n[0]
            out[key] = {"l1": LR.l1_error(c, sv, pb, ref, 4 * n // m), "steps": steps, "wall_s": wall}
        lc, fc = runs["learned%d" % n][0].learned_counts()
        out["network_fraction"] = lc / max(1, lc + fc)
        L, P1 = out["learned%d" % n], out["plain%d" % n]
        ratios.append(P1["l1"] / L["l1"])
        speed256.append(out["plain%d" % (2 * n)]["wall_s"] / L["wall_s"])
        speed512.append(out["plain%d" % (4 * n)]["wall_s"] / L["wall_s"])
        log("%s: L1 plain %d^2 %.5f, %d^2 %.5f, %d^2 %.5f | learned %d^2 %.5f; wall learned %.4fs vs plain %d^2 %.4fs"
            % (c["kind"], n, P1["l1"], 2 * n, out["plain%d" % (2 * n)]["l1"], 4 * n, out["plain%d" % (4 * n)]["l1"],
               n, L["l1"], L["wall_s"], 2 * n, out["plain%d" % (2 * n)]["wall_s"]))
        cases.append(out)
        for sv_, _, _, _ in runs.values():
            sv_.close()
    # stage timing of the learned tile kernel (HIP events)
    c = LR.case_params(1, seed=99)[0]
    sv, _ = LR.solver(c, n)
    sv.set_learned(net, theta)
    sv.set_time(0.0, 1e30, LR.CFL)
    sv.advance(3)
    sv.enable_timing(True)
    sv.advance(20)
    stage_ms, launches = sv.read_timing()
    sv.enable_timing(False)
    sv.close()
    avg = stage_ms / max(1, launches)
    flops = MLP_FLOPS_PER_CELL * n * n
    achieved = flops / (avg * 1e-3) / 1e12
    cpu = None if args.no_cpu_baseline else cpu_baseline_order2(n)
    gm = float(np.exp(np.mean(np.log(ratios))))
    out = {
        "metric": "L1 rho error at equal time, learned vs plain (plain / learned at %d^2, geometric mean)" % n,
        "value": gm,
        "unit": "x (error reduction)",
        "n_gpus": 1,
        "steps": None,
        "warmup": None,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (network bf16 in, fp32 accumulate)",
        "data": "synthetic (seeded initial conditions); training rows built on the device from fine plain runs",
        "config": {"workload": "learned_2d_%d" % n, "train_cases": len(train_cases), "fine": 4 * n,
                   "snapshots": len(train_cases) * args.snapshots, "sensor_threshold": theta,
                   "training_sensor_threshold": train_theta if train_theta != float("inf") else None,
                   "training_rows": net.num_samples(), "training_s": t_train, "data_s": t_data, "final_loss": loss,
                   "cases": cases,
                   "wall_speedup_vs_plain_2n": float(np.exp(np.mean(np.log(speed256)))),
                   "wall_speedup_vs_plain_4n": float(np.exp(np.mean(np.log(speed512)))),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": achieved / 2500.0, "traffic": None,
                     "kernel": "k_stage_learned (tile kernel with the network: reconstruction, flux and RK "
                               "update of one stage), network FLOPs only",
                     "algorithmic_flops_per_launch": flops, "avg_launch_ms": avg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def mesh_bytes(fm):
    """Algorithmic HBM bytes of one stage of the face-list kernels (axis-aligned
    faces, a handful of areas: 12 B of geometry per face).  Face kernel: owner,
    neighbour, kind (12 B) and A*F written (40 B) per face, the stage state read
    once per cell (40 B).  Cell kernel: offset (4 B) and volume (8 B) per cell,
    one 4 B incidence per face side, A*F read once per face (40 B), the RK
    update's state traffic per cell (80 / 120 / 120 B for stages 1 / 2 / 3)."""
    import numpy as np
    F, N = fm.nfaces, fm.ncells
    sides = int(2 * F - (fm.neigh < 0).sum())
    face = 52.0 * F + 40.0 * N
    cell = 12.0 * N + 4.0 * sides + 40.0 * F + BYTES_PER_CELL_STAGE * N
    return face, cell


def cpu_baseline_mesh(fm, pb_o, budget_s=10.0, budget_mt=5.0):
    """The oracle's face-list loop (orc_solver_init_faces) stepping the same
    mesh from its initial state: about budget_s on one thread, then about
    budget_mt on all host cores (faces, then per-cell sums in face order)."""
    from oracle import oracle as O
    s = O.Solver.from_faces(pb_o, fm.dim, fm.volume, fm.owner, fm.neigh, fm.area, fm.normal, fm.min_cell_size,
                            border_bc=fm.face_bc, border_info=fm.face_info, centroids=fm.centroid)
    threads = cpu_threads()
    res = []
    for nt, budget in ((1, budget_s), (threads, budget_mt)):
        U = s.initial_state()
        W = U.copy()
        t, k = 0.0, 0
        t0 = time.perf_counter()
        while k < 1 or time.perf_counter() - t0 < budget:
            dt, _ = s.rk3_step(U, W, 0.5, t, 1e30, threads=nt)
            t += dt
            k += 1
        el = time.perf_counter() - t0
        res.append((fm.ncells * 3 * k / el, "%d SSP-RK3 step(s) of the %d cells in %.1f s" % (k, fm.ncells, el)))
    return _both(res[0], res[1], threads, "the same octree mesh from its initial state")


def run_mesh(args, mmf):
    """The explicit-interface-list path on a 2:1 refined octree mesh: radial Sod
    on [-1,1]^3, a 128^3 base grid whose cells inside |x|_inf < 0.5 are split
    in eight (hanging faces on the refinement boundary), free-flow borders."""
    from minimmerflow_amd import meshes
    from oracle import oracle as O
    n = args.n or 128
    t0 = time.perf_counter()
    fm = meshes.refined(3, [n] * 3, [-1.0] * 3, 2.0 / n, ([-0.5] * 3, [0.5] * 3), [0] * 6)
    pb = mmf.problem("radsod", 3, sod_r0=0.4, domain_origin=[-1.0] * 3, domain_length=2.0)
    U0 = mmf.exact_conservatives(pb, fm.centroid, 0.0)

    def make():
        g = mmf.Mesh(fm)
        g.set_state(U0)
        g.set_time(0.0, 1e30, 0.5)
        return g

    g = make()
    log("setup %.1fs: octree mesh %d cells (%d fine), %d faces" % (time.perf_counter() - t0, fm.ncells,
                                                                     int((fm.size < 2.0 / n).sum()), fm.nfaces))
    w = make()
    nprewarm = prewarm(w.advance, lambda: w.time())
    w.close()
    g.advance(args.warmup)
    g.time()
    g.enable_timing(True)
    t1 = time.perf_counter()
    g.advance(args.steps)
    t, dt, step, _ = g.time()
    el = time.perf_counter() - t1
    face_ms, cell_ms, stages = g.read_timing()
    g.enable_timing(False)
    if step != args.warmup + args.steps:
        raise SystemExit("bench: only %d of %d steps were active" % (step, args.warmup + args.steps))
    npc = g.counters()
    if npc[0] or npc[1] or npc[2]:
        raise SystemExit("bench: negative or non-finite states encountered (%d, %d, %d)" % npc)
    N = fm.ncells
    value = N * 3 * args.steps / el
    fb, cb = mesh_bytes(fm)
    fa, ca = face_ms / stages, cell_ms / stages
    dom = "face" if fa >= ca else "cell"
    achieved = (fb / (fa * 1e-3) if dom == "face" else cb / (ca * 1e-3)) / 1e9
    workload = "octree_sod_3d_%d_order1" % n
    log("timed %d steps in %.3f s: %.3f ms/step; face kernel %.4f ms (%.0f GB/s), cell kernel %.4f ms (%.0f GB/s)"
        % (args.steps, el, el / args.steps * 1e3, fa, fb / fa / 1e6, ca, cb / ca / 1e6))
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline: oracle face-list loop ...")
        cpu = cpu_baseline_mesh(fm, O.problem(O.RADSOD, 3, sod_r0=0.4))
        log("cpu baseline: %.3e cell-updates/s (single thread %.3e)" % (cpu["value"], cpu["single_thread"]["value"]))
    out = {
        "metric": "cell-updates/s per RK stage",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (initial condition sampled at cell centroids)",
        "config": {"workload": workload, "cells": N, "faces": fm.nfaces,
                   "mesh": "%d^3 base cells on [-1,1]^3, cells inside |x|_inf < 0.5 split in eight (2:1, hanging "
                           "faces), explicit interface list" % n,
                   "initial": "r<0.4: rho=1 p=1; else rho=0.125 p=0.1; u=0", "boundaries": "free flow",
                   "time": "SSP-RK3, CFL 0.5 (min cell size), fixed step count", "prewarm_steps": nprewarm,
                   "reconstruction": "order 1", "flux": "Rusanov", "t_after_timed_steps": t,
                   "kernels_ms_per_stage": {"faces": fa, "cells": ca},
                   "algorithmic_bytes_per_stage": {"faces": fb, "cells": cb},
                   "parallelism": "single GPU"},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic(workload),
            "kernel": "k_mesh_%s (%s)" % (dom + "s", "per-interface flux, A*F per face" if dom == "face"
                                          else "segmented sum in face order + RK update"),
            "algorithmic_bytes_per_launch": fb if dom == "face" else cb,
            "avg_launch_ms": fa if dom == "face" else ca,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def load_traffic(workload):
    p = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if os.path.exists(p):
        try:
            return json.load(open(p)).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            return None
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="radsod256")
    ap.add_argument("--n", type=int, default=0, help="cells per axis (default: the workload's)")
    ap.add_argument("--order", type=int, default=1, choices=(1, 2))
    ap.add_argument("--kernel", choices=("auto", "gather", "march", "tile"), default="auto")
    ap.add_argument("--rows", type=int, default=0, help="marching kernel: tile rows (4, 8 or 16)")
    ap.add_argument("--cols", type=int, default=0, help="marching kernel: tile columns (64 or 32)")
    ap.add_argument("--chunk", type=int, default=0, help="marching kernel: planes per block")
    ap.add_argument("--blocks", type=int, default=0, help="batch: persistent blocks (default: autotuned)")
    ap.add_argument("--tail", type=int, default=0, help="batch: rows of the small items ending each stage "
                                                         "(with --chunk/--blocks; 0: none)")
    ap.add_argument("--ncases", type=int, default=512, help="batch: number of cases")
    ap.add_argument("--train-cases", type=int, default=128, help="learned: fine training runs")
    ap.add_argument("--eval-cases", type=int, default=4, help="learned: held-out cases")
    ap.add_argument("--snapshots", type=int, default=80, help="learned: snapshots per fine training run")
    ap.add_argument("--theta", type=float, default=None, help="learned: shock-sensor threshold")
    ap.add_argument("--train-theta", type=float, default=None,
                    help="learned: sensor threshold that zeroes training rows (default: none)")
    ap.add_argument("--cpu-steps", type=int, default=3, help="radial Sod: oracle steps of the CPU sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only; --gpus %d is not supported "
              "(the multi-GPU sharding is a design, not a measured path)" % args.gpus)
        sys.exit(2)

    import minimmerflow_amd as mmf

    if args.copy_calibration:
        ms = mmf.calibration_copy(args.copy_calibration, reps=5)
        print(json.dumps({"copy_bytes": args.copy_calibration, "ms_per_copy": ms,
                          "GBs_read_plus_write": 2 * args.copy_calibration / (ms * 1e-3) / 1e9}), flush=True)
        return

    if args.workload == "batch512":
        run_batch(args, mmf)
        return
    if args.workload == "learned128":
        run_learned(args, mmf)
        return
    if args.workload == "octree128":
        run_mesh(args, mmf)
        return

    setup, n_default, wname = WORKLOADS[args.workload]
    n = args.n or n_default
    N = n ** 3
    t_setup = time.perf_counter()
    s, mask_bytes, cfg = setup(mmf, n, args.order)
    kind = {"auto": mmf.KERNEL_AUTO, "gather": mmf.KERNEL_GATHER, "march": mmf.KERNEL_MARCH,
            "tile": mmf.KERNEL_TILE}[args.kernel]
    s.set_kernel(kind, args.rows, args.chunk)
    if args.cols:
        s.set_march_shape(args.cols, args.rows, args.chunk)
    kinfo = s.kernel()
    solved = N if mask_bytes == 0 else int(((s.cell_flags() & mmf.CELL_SOLVED) != 0).sum())
    log("setup %.1fs: %d^3 cells (%d solved), %d interfaces, order %d"
        % (time.perf_counter() - t_setup, n, solved, s.ninterfaces, args.order))

    w, _, _ = setup(mmf, n, args.order)  # the pre-warm instance
    w.set_kernel(kind, args.rows, args.chunk)
    if args.cols:
        w.set_march_shape(args.cols, args.rows, args.chunk)
    cfg["prewarm_steps"] = prewarm(w.advance, w.synchronize)
    s.advance(args.warmup)
    s.synchronize()
    s.enable_timing(2)  # one event pair around the timed, graph-replayed steps
    t0 = time.perf_counter()
    s.advance(args.steps)
    s.synchronize()
    el = time.perf_counter() - t0
    stage_ms, launches = s.read_timing()
    s.enable_timing(False)
    t, dt, step, eig = s.time()
    if step != args.warmup + args.steps:
        raise SystemExit("bench: only %d of %d steps were active (t=%.4f reached t_end)" % (step, args.warmup + args.steps, t))
    np_, nr_ = s.positivity_counters()
    if np_ or nr_:
        raise SystemExit("bench: negative pressure/density encountered (%d, %d)" % (np_, nr_))

    value = solved * 3 * args.steps / el
    avg_ms = stage_ms / max(launches, 1)
    bytes_per_cell = BYTES_PER_CELL_STAGE + mask_bytes
    achieved = bytes_per_cell * N / (avg_ms * 1e-3) / 1e9
    workload = "%s_%d_order%d" % (wname, n, args.order)
    log("timed %d steps in %.3f s: %.3f ms/step, stage kernel avg %.4f ms (%d launches), %.0f GB/s, t=%.5f dt=%.3e"
        % (args.steps, el, el / args.steps * 1e3, avg_ms, launches, achieved, t, dt))

    cpu = None
    if not args.no_cpu_baseline and args.workload == "radsod256":
        log("cpu baseline: oracle, %d step(s) at %d^3 ..." % (args.cpu_steps, n))
        cpu = cpu_baseline(n, args.cpu_steps)
        log("cpu baseline: %.3e cell-updates/s" % cpu["value"])
    elif not args.no_cpu_baseline and args.workload == "sphere192" and args.order == 1:
        log("cpu baseline: oracle, sphere at %d^3 for about 12 s ..." % n)
        cpu = cpu_baseline_sphere(s, n)
        log("cpu baseline: %.3e cell-updates/s" % cpu["value"])

    config = {"workload": workload, "cells": N, "solved_cells": solved, "interfaces": s.ninterfaces}
    config.update(cfg)
    config.update({"reconstruction": "order %d" % args.order, "flux": "Rusanov", "t_after_timed_steps": t,
                   "parallelism": "single GPU"})
    out = {
        "metric": "cell-updates/s per RK stage",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (initial condition sampled at cell centroids)",
        "config": config,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic(workload),
            "kernel": "%s (fused flux + RK stage), mean of stages 1-3" % (
                ("k_march_o2 (LDS-staged limited reconstruction) tile 64x8, %d planes/block" % kinfo[2]
                 if args.order == 2 else
                 ("k_march tile %dx%d, %d planes/block" % (args.cols or 64, kinfo[1], kinfo[2]) if kinfo[2]
                  else "k_march tile %dx%d, the planes split over one block per CU in shares of equal cost" % (args.cols or 64,
                                                                                               kinfo[1])))
                if kinfo[0] == mmf.KERNEL_MARCH
                else "k_stage (per-cell gather)"),
            "algorithmic_bytes_per_launch": bytes_per_cell * N,
            "avg_launch_ms": avg_ms,
            "timing": "one HIP event pair on the solver's stream around the timed steps (graph-replayed), "
                      "divided by their %d stage launches; the gaps between launches are included" % launches,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
