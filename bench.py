"""Benchmark of the per-step fluid pipeline on one MI355X.

Workload (the configuration the headline metric is quoted on): a 1024 x 1024
FP32 MAC grid, fluid filling the lower 55 % of the domain, 16 rigid bodies
(4 circles, 4 rectangles, 4 capsules, 4 triangles) at seeded positions,
densities 0.5-2.0 rho, two-way area-fraction coupling (the bodies are
integrated on the device from the fluid forces and gravity), dt = 1/60, and a pressure solve iterated to
||r||2/||b||2 <= 1e-5 (CG preconditioned by one symmetric RBGS V-cycle of the
reference's multigrid; the reference's single fixed V-cycle cannot reach that
target).  One step = one efl_step call: the whole render-node pipeline
(solids, area fractions, advection, forces, divergence, projection, velocity
update, extrapolation, level-set advection, sources, redistancing, gradient,
fluid-to-solid forces), inputs resident in HBM.

Prints one JSON line.  `value` is steps/s over the timed window (simulation
steps warmup+1 .. warmup+steps, three later with --autotune, named in
`timed_steps`); `pressure_solve_ms`
and `max_norm_divergence` (dt * max|div u| after the projection) come with it.
`roofline` prices the kernel that dominates the timed step: k_mg_down<2,5,true>,
the level-0 launch of every MGPCG iteration, which runs the iteration itself
(p = z + beta p, q = A p, x += alpha p, r -= alpha q, ||r||^2 partials), the
two red-black pre-smoothing sweeps of the preconditioner's V-cycle, the
residual and its restriction.  Its algorithmic bytes follow the per-unit
model of the reference dataflow (each kernel reads each input once and
writes each output once, 4 B per word, no credit for fusing): the search
direction 9 words per cell (z, p, air level set and 4 area fractions read;
p, q written), the update 6 (x, p, r, q read; x, r written), 2 sweeps x 8
(x, b, air level set and 4 area fractions read, x written), the residual pass
8 (the same 7 inputs, r written) and the restriction of the residual 1.25 (r
read, one coarse b per 4 cells written) = 40.25 words per cell; it is timed with hipEvents around back-
to-back launches of exactly that launch on the library's stream
(efl_time_stage CG_SMOOTH); `traffic` is its measured HBM bytes per launch from
the committed FETCH_SIZE / WRITE_SIZE profile.  `step_frac` prices the whole
step: (176 + 69 k) words per cell for a step with k V-cycles (k = mean CG
iterations, one V-cycle each) over ms_per_step.  `cpu_baseline` times the CPU
oracle (the reference pipeline restated in C++) on a bounded sample of the same
workload after 20 warm-up steps, on all host cores and on one core, and the
reference's own projection (one 2-3-2 V-cycle per step) on all cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sim steps/s + pressure-solve ms at 1024² to ‖∇·u‖∞<1e-4; % HBM roofline"
HBM_PEAK_GBS = 8000.0
SMOOTH_SWEEPS = 2  # red-black sweeps of one MG_SMOOTH launch
# words per cell of the level-0 CG launch: search direction 9 + update 6 + 2 sweeps x 8 + residual 8
# + restriction 1.25
CG_SMOOTH_WORDS = 9 + 6 + SMOOTH_SWEEPS * 8 + 8 + 1.25
CG_SMOOTH_KERNEL = ("k_mg_down<2,5,true>: level-0 launch of each MGPCG iteration (p = z + beta p, q = A p, "
                    "x/r update, 2 red-black sweeps, residual, restriction)")
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_mg_down_cg_level0_1024.json")
DT = 1.0 / 60.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_scene(size, rel_tol):
    from bevy_eulerian_fluid_amd import scenes
    proj = scenes.projection("mgpcg", max_cycles=200, rel_tol=rel_tol, check_every=1)
    return scenes.coupled_tank(size, proj=proj)


def roofline_only(args):
    """Initialise the workload, run two steps, then only the timed launches of
    the roofline kernel (k_mg_down<2,5,true> as an MGPCG iteration issues it,
    or with --roofline-kernel mg_smooth the standalone two-sweep smoother): a
    profile of this run holds that kernel alone at level 0."""
    from bevy_eulerian_fluid_amd import native, scenes
    sc = build_scene(args.size, args.rel_tol)
    ctx = native.Context(0)
    inst = ctx.instance(sc["settings"])
    inst.set_sources(sc["init_sources"], True)
    ctx.set_obstacles(sc["obstacles"])
    inst.step(DT)
    inst.step(DT)
    ctx.sync()
    n = args.size * args.size
    if args.roofline_kernel == "mg_smooth":
        ms = inst.time_stage("MG_SMOOTH", DT, args.kernel_reps)
        name, nbytes = "k_mg_down<2,5> level 0 (MG_SMOOTH)", SMOOTH_SWEEPS * 32 * n
    else:
        ms = inst.time_stage("CG_SMOOTH", DT, args.kernel_reps)
        name, nbytes = CG_SMOOTH_KERNEL, int(CG_SMOOTH_WORDS * 4 * n)
    print(json.dumps({"kernel": name, "grid": [args.size, args.size], "reps": args.kernel_reps, "ms": ms,
                      "algorithmic_bytes": nbytes, "achieved_gbs": nbytes / (ms * 1e-3) / 1e9}), flush=True)
    inst.close()
    ctx.close()


def copy_probe(args):
    """A plain device-to-device copy of known size (256 MiB, 10 times): the
    reference point that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 64 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": 4 * n, "reps": 10}), flush=True)


def gpu_run(args):
    from bevy_eulerian_fluid_amd import native, scenes
    sc = build_scene(args.size, args.rel_tol)
    ctx = native.Context(0)
    if args.graphs:
        ctx.set_graphs(True)
    inst = ctx.instance(sc["settings"])
    inst.set_sources(sc["init_sources"], True)
    ctx.set_obstacles(sc["obstacles"])
    ctx.set_bodies(scenes.bodies_for(sc["obstacles"], sc["densities"]))
    log(f"[bench] {args.size}^2, {len(sc['obstacles'])} bodies: initialise + {args.warmup} warm-up steps")
    inst.step(DT)  # initialisation dispatch
    for _ in range(args.warmup):
        inst.step(DT)
    tuning = None
    if args.autotune:
        tuning = inst.autotune(DT, apply=True)
        log(f"[bench] autotuned: {tuning}")
        for _ in range(3):
            inst.step(DT)
    ctx.sync()
    log(f"[bench] timing {args.steps} steps" + (f" (asynchronous, margin {args.async_margin})" if args.async_margin else ""))
    cycles = []
    missed0 = inst.stats().unconverged_solves
    if args.async_margin:
        ctx.set_async(args.async_margin)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        inst.step(DT)
        if not args.async_margin:
            cycles.append(inst.stats().cycles)
    ctx.sync()
    t1 = time.perf_counter()
    if args.async_margin:
        cycles.append(inst.stats().cycles)  # the last step's (statistics resolve lazily)
        ctx.set_async(0)
    unconverged = inst.stats().unconverged_solves - missed0
    wall = t1 - t0
    steps_per_s = args.steps / wall

    # pressure-solve time and divergence on extra, separately timed steps
    ctx.set_timing(True)
    inst.set_diagnostics(True)
    proj_ms, step_ms, divs, rels = [], [], [], []
    for _ in range(args.diag_steps):
        inst.step(DT)
        st = inst.stats()
        proj_ms.append(st.projection_ms)
        step_ms.append(st.step_ms)
        divs.append(st.max_norm_divergence)
        rels.append(st.rel_residual)
    ctx.set_timing(False)
    inst.set_diagnostics(False)

    # the dominant kernel of the timed step: the level-0 launch of an MGPCG
    # iteration, timed with hipEvents around back-to-back launches of it on the
    # library's stream; beside it the standalone two-sweep smoother
    n = args.size * args.size
    cg_ms = inst.time_stage("CG_SMOOTH", DT, args.kernel_reps)
    cg_bytes = int(CG_SMOOTH_WORDS * 4 * n)
    sweep_ms = inst.time_stage("MG_SMOOTH", DT, args.kernel_reps)
    sweep_bytes = SMOOTH_SWEEPS * 8 * 4 * n
    traffic = None  # HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE profile
    if args.size == 1024 and os.path.exists(TRAFFIC_PROFILE):
        with open(TRAFFIC_PROFILE) as f:
            traffic = json.load(f).get("traffic_bytes")
    inst.close()
    ctx.close()
    return dict(steps_per_s=steps_per_s, ms_per_step=1e3 * wall / args.steps,
                cycles_mean=sum(cycles) / len(cycles), cycles_max=max(cycles), unconverged=unconverged,
                pressure_solve_ms=sum(proj_ms) / len(proj_ms), device_step_ms=sum(step_ms) / len(step_ms),
                max_norm_divergence=max(divs), rel_residual=max(rels),
                cg_ms=cg_ms, cg_bytes=cg_bytes, sweep_ms=sweep_ms, sweep_bytes=sweep_bytes,
                traffic=traffic, tuning=tuning, first_timed=args.warmup + 1 + (3 if tuning else 0))


def run_config(args):
    """The other workloads of the benchmark suite, one JSON line each:
      1  dam break 256^2, Gauss-Seidel 100 sweeps (weight 1.9), no bodies
      3  64 independent instances, sides log-uniform in [128, 1024], 0-8
         two-way coupled bodies each, MG-preconditioned CG to --rel-tol
      5  4096^2 pool + 200 droplets (an initial level set), 24 bodies, 64
         local forces resampled every 30 steps, MG-preconditioned CG to
         1e-5 (with --learned: the learned preconditioner, shipped weights,
         MGPCG fallback), stepped together with 8 instances of 512^2 (MGPCG)"""
    from bevy_eulerian_fluid_amd import native, scenes
    if args.config == 4:
        return run_learned(args)
    ctx = native.Context(0)
    if args.graphs:
        ctx.set_graphs(True)
    insts, forces, desc = [], None, {}
    if args.config == 1:
        sc = scenes.dam_break(256)
        i = ctx.instance(sc["settings"])
        i.set_sources(sc["init_sources"], True)
        insts = [i]
        desc = {"workload": "dam break 256x256, GS 100 sweeps w=1.9, FIM(10)", "grid": [256, 256], "instances": 1}
    elif args.config == 3:
        for sc in scenes.batch(64, proj=scenes.projection("mgpcg", max_cycles=200, rel_tol=args.rel_tol)):
            i = ctx.instance(sc["settings"])
            i.set_sources(sc["init_sources"], True)
            i.set_obstacles(sc["obstacles"], scenes.bodies_for(sc["obstacles"], sc["densities"]))
            insts.append(i)
        cells = sum(i.W * i.H for i in insts)
        desc = {"workload": "64 instances, sides log-uniform [128, 1024], 0-8 coupled bodies each, "
                            f"MG-preconditioned CG to {args.rel_tol}", "instances": 64, "cells": cells}
    elif args.config == 5:
        if args.learned:
            from bevy_eulerian_fluid_amd import learned
            w = learned.load_weights()
            if w is None:
                raise SystemExit("--learned: no shipped weights (bevy_eulerian_fluid_amd/data)")
            sc = scenes.combined(4096, proj=scenes.projection("learned", max_cycles=200, rel_tol=1e-5))
        else:
            sc = scenes.combined(4096)
        big = ctx.instance(sc["settings"])
        if args.learned:
            net = native.Network(ctx)
            net.set_params(w)
            big.set_learned(net)
        big.init_levelset(sc["levelset"])
        ctx.set_obstacles(sc["obstacles"])
        ctx.set_bodies(scenes.bodies_for(sc["obstacles"], sc["densities"]))
        insts = [big]
        for k in range(8):
            s2 = scenes.coupled_tank(512, seed=scenes.SEED_BASE + 50 + k, n_each=1,
                                     proj=scenes.projection("mgpcg", max_cycles=200, rel_tol=1e-5))
            i = ctx.instance(s2["settings"])
            i.set_sources(s2["init_sources"], True)
            i.set_obstacles(s2["obstacles"], scenes.bodies_for(s2["obstacles"], s2["densities"]))
            insts.append(i)
        forces = (sc["forces"], sc["rng"])
        desc = {"workload": "4096x4096 pool + 200 droplets, 24 coupled bodies, 64 local forces, "
                            + ("learned preconditioner (flexible CG, MGPCG fallback)" if args.learned
                               else "MG-preconditioned CG") + " to 1e-5, + 8 instances of 512x512 (MGPCG)",
                "instances": 9, "grid": [4096, 4096]}
    else:
        raise SystemExit(f"unknown --config {args.config}")
    cells = sum(i.W * i.H for i in insts)
    frc = forces[0] if forces else None

    def step(k):
        nonlocal frc
        if forces:
            if k % 30 == 0 and k > 0:
                frc = scenes.random_forces(forces[1], 4096, 4096, 64)
            insts[0].set_forces(frc)
        ctx.step(insts, DT)

    log(f"[bench] config {args.config}: {len(insts)} instance(s), {cells} cells; init + {args.warmup} warm-up steps")
    step(0)
    for k in range(args.warmup):
        step(k + 1)
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(args.warmup + 1 + k)
    ctx.sync()
    wall = time.perf_counter() - t0
    stats = [i.stats() for i in insts]
    out = {
        "metric": f"config {args.config} steps/s", "value": round(args.steps / wall, 3), "unit": "steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * wall / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": desc, "cell_steps_per_s": round(cells * args.steps / wall, 1),
        "instance_steps_per_s": round(len(insts) * args.steps / wall, 3),
        "cycles_last_step": [int(s.cycles) for s in stats[:9]],
        "roofline": None, "cpu_baseline": None,
    }
    for i in insts:
        i.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def run_learned(args):
    """Config 4: the learned preconditioner (shipped weights, trained at
    512^2) at args.size^2 on the coupled tank: steps/s with
    EFL_PROJECTION_LEARNED (flexible CG with the network, MGPCG fallback)
    beside MGPCG's, the per-step solver comparison on identical systems, and
    one network application timed with hipEvents against the dense BF16 MFMA
    peak."""
    from bevy_eulerian_fluid_amd import learned, native, scenes
    w = learned.load_weights()
    if w is None:
        raise SystemExit("no shipped weights: bevy_eulerian_fluid_amd/data/learned_preconditioner.f32")
    size = args.size
    ctx = native.Context(0)
    log(f"[bench] config 4: learned vs MGPCG on identical systems, {size}^2")
    cmp = learned.evaluate(ctx, w, size=size, steps=10)
    rates = {}
    for kind in ("learned", "mgpcg"):
        sc = scenes.coupled_tank(size, proj=scenes.projection(kind, max_cycles=200, rel_tol=args.rel_tol))
        i = ctx.instance(sc["settings"])
        i.set_sources(sc["init_sources"], True)
        i.set_obstacles(sc["obstacles"], scenes.bodies_for(sc["obstacles"], sc["densities"]))
        net = None
        if kind == "learned":
            net = native.Network(ctx)
            net.set_params(w)
            i.set_learned(net)
        for _ in range(6):
            i.step(DT)
        ctx.sync()
        n = 20
        t0 = time.perf_counter()
        for _ in range(n):
            i.step(DT)
        ctx.sync()
        rates[kind] = n / (time.perf_counter() - t0)
        if kind == "learned":
            reps = max(20, args.kernel_reps // 4)
            ms = i.time_stage("NN_PRECONDITION", DT, reps)
            wp, hp = (size + 7) // 8 * 8, (size + 7) // 8 * 8
            flops = native.Network.flops_per_pixel() * wp * hp
            tflops = flops / (ms * 1e-3) / 1e12
            nn_timing = dict(ms=round(ms, 4), gflop=round(flops / 1e9, 2), tflops=round(tflops, 2))
            net.close()
        i.close()
    ctx.close()
    peak = 2500.0
    out = {
        "metric": "config 4 steps/s (learned preconditioner)", "value": round(rates["learned"], 3), "unit": "steps/s",
        "n_gpus": 1, "steps": 20, "warmup": 6, "ms_per_step": round(1e3 / rates["learned"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 network, f32 solve",
        "data": "synthetic", "config": {"workload": f"coupled tank {size}x{size}, 16 two-way coupled bodies, "
                                                    f"pressure to ||r||/||b||<={args.rel_tol}, learned U-Net "
                                                    "preconditioner (4 levels, 272k params, trained at 512x512)",
                                        "grid": [size, size], "projection": "learned"},
        "mgpcg_steps_per_s": round(rates["mgpcg"], 3),
        "solver": {k: v for k, v in cmp.items() if k != "rows"},
        "network_application": nn_timing,
        "roofline": {"bound": "mfma", "kernel": "U-Net forward (15 k_conv3x3 + 3 k_pool2 + features/output)",
                     "achieved": nn_timing["tflops"], "peak": peak, "unit": "TFLOP/s",
                     "frac": round(nn_timing["tflops"] / peak, 4), "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(args):
    """The CPU oracle on the same scene: init, 20 warm-up steps on all cores,
    then bounded samples of the same workload on all cores and on one core, and
    of the reference's own projection (one 2-3-2 V-cycle per step)."""
    from bevy_eulerian_fluid_amd import scenes
    from oracle.oracle import OracleInstance, set_threads
    cores = min(os.cpu_count() or 1, args.cpu_threads)
    set_threads(cores)
    sc = build_scene(args.size, args.rel_tol)
    o = OracleInstance(sc["settings"])
    o.set_sources(sc["init_sources"], True)
    o.set_obstacles(sc["obstacles"])
    o.set_bodies(scenes.bodies_for(sc["obstacles"], sc["densities"]))
    o.step(DT)
    for _ in range(args.cpu_warmup):
        o.step(DT)

    def sample(budget):
        t0 = time.perf_counter()
        k = 0
        while k < args.cpu_steps:
            o.step(DT)
            k += 1
            if time.perf_counter() - t0 > budget:
                break
        return k, time.perf_counter() - t0

    k, el = sample(args.cpu_budget_s)
    set_threads(1)
    k1, el1 = sample(args.cpu_budget_s)
    set_threads(cores)
    o.set_projection(scenes.projection("multigrid", max_cycles=1))
    kv, elv = sample(args.cpu_budget_s / 2)
    o.close()
    what = (f"{args.size}^2 scene after init + {args.cpu_warmup} warm-up steps, oracle (reference pipeline "
            "restated in C++, OpenMP over rows)")
    return dict(value=k / el, unit="steps/s", cores=cores, kind="port",
                sample=f"{k} steps of the same {what}, {el:.1f} s",
                one_core={"value": k1 / el1, "unit": "steps/s", "cores": 1, "sample": f"the next {k1} steps, {el1:.1f} s"},
                reference_vcycle={"value": kv / elv, "unit": "steps/s", "cores": cores,
                                  "sample": f"the next {kv} steps with the reference's projection (one 2-3-2 "
                                            f"V-cycle per step, projection.rs:146-162), {elv:.1f} s"})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--rel-tol", type=float, default=1e-5)
    ap.add_argument("--diag-steps", type=int, default=10)
    ap.add_argument("--kernel-reps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=64)
    ap.add_argument("--cpu-warmup", type=int, default=20)
    ap.add_argument("--cpu-budget-s", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graphs", action="store_true", help="replay steps as HIP graphs instead of launching every kernel")
    ap.add_argument("--async-margin", type=int, default=0,
                    help="time the steps on an asynchronous context (efl_set_async): MGPCG budgets of the last "
                         "iteration count + this margin, no read-back inside efl_step")
    ap.add_argument("--autotune", action="store_true",
                    help="pick sweep counts, the bottom-solve size and the smoother tile height "
                         "with efl_autotune before timing")
    ap.add_argument("--config", type=int, default=2,
                    help="workload: 2 = coupled tank 1024^2 (the headline line); 1, 3, 4 (learned), 5 = the other suite configs")
    ap.add_argument("--learned", action="store_true",
                    help="config 5: the 4096^2 instance solves with the learned preconditioner")
    ap.add_argument("--roofline-only", action="store_true",
                    help="time only the roofline kernel (for profiling it alone)")
    ap.add_argument("--roofline-kernel", choices=["cg_smooth", "mg_smooth"], default="cg_smooth")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 256 MiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU layout is "
              "a design note, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args)
        return
    if args.roofline_only:
        roofline_only(args)
        return
    if args.config != 2:
        run_config(args)
        return
    g = gpu_run(args)
    cpu = None if args.no_cpu else cpu_baseline(args)
    n = args.size * args.size
    cg_gbs = g["cg_bytes"] / (g["cg_ms"] * 1e-3) / 1e9
    step_bytes = (176 + 69 * g["cycles_mean"]) * 4 * n  # step with k = mean CG iterations V-cycles
    step_gbs = step_bytes / (g["ms_per_step"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(g["steps_per_s"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "timed_steps": [g["first_timed"], g["first_timed"] + args.steps - 1],
        "ms_per_step": round(g["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"coupled tank {args.size}x{args.size}, lower 55% fluid, 16 two-way coupled bodies, "
                               f"dt=1/60, pressure to ||r||/||b||<={args.rel_tol:g} (MG-preconditioned CG)",
                   "async_margin": args.async_margin,
                   "grid": [args.size, args.size], "bodies": 16, "projection": "mgpcg",
                   "rel_tol": args.rel_tol, "reinit": "fast_iterative(10)",
                   "autotuned": None if g["tuning"] is None else
                   {k: g["tuning"][k] for k in ("pre", "post", "coarsest", "bottom_cells", "tile_h")}},
        "pressure_solve_ms": round(g["pressure_solve_ms"], 4),
        "device_step_ms": round(g["device_step_ms"], 4),
        "cg_iterations_mean": round(g["cycles_mean"], 2),
        "unconverged_steps": g["unconverged"],
        "max_norm_divergence": g["max_norm_divergence"],
        "rel_residual": g["rel_residual"],
        "roofline": {"bound": "hbm", "kernel": CG_SMOOTH_KERNEL,
                     "achieved": round(cg_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(cg_gbs / HBM_PEAK_GBS, 4), "traffic": g["traffic"],
                     "traffic_source": "profiles/k_mg_down_cg_level0_1024.json (rocprofv3 FETCH_SIZE, WRITE_SIZE)",
                     "algorithmic_bytes": g["cg_bytes"], "words_per_cell": CG_SMOOTH_WORDS, "ms": round(g["cg_ms"], 5),
                     "frac_sweeps_only": round(g["sweep_bytes"] / (g["cg_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "step_frac": round(step_gbs / HBM_PEAK_GBS, 4), "step_achieved": round(step_gbs, 1),
                     "step_bytes": int(step_bytes),
                     "smoother_alone": {"kernel": "k_mg_down<2,5> level 0 (MG_SMOOTH)", "ms": round(g["sweep_ms"], 5),
                                        "algorithmic_bytes": g["sweep_bytes"],
                                        "frac": round(g["sweep_bytes"] / (g["sweep_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                                      4)}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
