#!/usr/bin/env python3
"""The liquid step's benchmark on one MI355X.

A step is one liquid step (encodeFluid) of the plain FP32 world at scale:
a 1024 x 1024 MAC grid with ~10% of its cells solid as 200 random discs
(radius U[4, 24], seeded), water in the non-solid cells of the lowest 40% of
the rows at 4 particles a cell (~1.47 M particles), each cell's velocity
N(0, 0.5^2) cells/s, dt 1/120, gravity 9.81 m/s^2 at 1 m a cell, MG-PCG with
7 levels to a relative residual of 1e-6.  The world's buffers are resident
in HBM before the timed region starts.

Prints one JSON line: particle-steps/s over the timed steps (and the solves/s
and p50/p99 step latency beside it), the HBM roofline of the kernel that
takes the most time (its algorithmic bytes per launch over its mean launch
time, from HIP events on the world's stream over the timed region), and the
oracle (the CPU restatement of the reference's liquid step) timed on a
bounded sample of the same workload on the host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s & MG-PCG solves/s at 1e-6 resid; p50/p99 step latency"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per unit of the kernels the roofline may name (the
# per-step byte model's terms: P2G reads a particle's 48 bytes; G2P reads 48
# and writes 32; the sort counts (reads 8) and scatters (reads and writes 48
# each): 224 a particle in all).  Units: particles in the step.
KERNEL_BYTES_PER_PARTICLE = {"gather": 48, "advect": 80, "sortCount": 8, "scatter": 96 - 8,
                             "syncParticles": 96}


def build_world(size, seed=2, workload="single"):
    from worldinagrain_amd import scenes
    from worldinagrain_amd import elements as E
    if workload == "dam":
        # the small plain configuration: a 1-cell solid border, water on
        # cells [1, size/2) x [1, size/2) at rest
        cells = scenes.dam_break(size, size, block=(size // 2, size // 2), wall=1)
        return cells, (size // 2 - 1) ** 2, E.kinds()
    cells, fluid_cells = scenes.discs_dam(size, size, seed=seed)
    return cells, fluid_cells, E.kinds()


def config_simulation(world, size, tol, storage=0):
    import numpy as np
    from worldinagrain_amd.host import Simulation, SolverSettings
    g = float(np.float32(9.81 * (1.0 / 120.0) ** 2 / 1.0))
    return Simulation(world, size, size, dx=1.0, earth_g=g, dt=1.0 / 120.0,
                      solver=SolverSettings.relative(tol, storage=storage))


# The V-cycle's storage (--storage), and the byte model's bytes a cell a PCG
# iteration for it: 9 + 24 + 16 + (4/3)(7.5 s_v + 3), s_v the bytes a stored
# V-cycle value (4 float, 2 half or bfloat16)
STORAGES = {"fp32": 0, "fp16": 1, "bf16": 2}
B_ITER = {0: 93.0, 1: 73.0, 2: 73.0}


def step_bytes(totals_delta, b_iter=93.0):
    """The per-step algorithmic byte model summed over the timed steps:
    N*224 + C_a*65 + C_a*16 + K*C_a*B_iter (93 FP32, 73 16-bit V-cycle storage)."""
    st, cyc, parts, ca = totals_delta
    if st == 0:
        return 0.0
    ca_mean = ca / st
    return parts * 224.0 + ca * (65.0 + 16.0) + cyc * ca_mean * b_iter


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cells, kinds, size, tol, steps, solver=None, runs=3, one_thread_runs=3):
    """The oracle (the CPU restatement of the reference's liquid step) on the
    host, the same world with the same solver settings: its first (seeding)
    step untimed, then `runs` runs of `steps` steps each, the median run
    reported.  On all the threads OpenMP is given (OMP_NUM_THREADS, else the
    host's cores; the OpenMP build, whose results are the serial build's),
    and on one thread (the serial build)."""
    import statistics
    from oracle.oracle import OracleWorld, lib, OMP_PATH
    from worldinagrain_amd import abi

    def measure(L, runs):
        w = OracleWorld(size, size, kinds=kinds, dot_mode=1, L=L)
        w.upload(abi.PART_CELLS, cells)
        sim = config_simulation(w, size, tol)
        if solver is not None:
            sim.solver = solver
        sim.step()
        n = int(sim.last_stats[abi.STAT_PARTICLES])
        rates, secs = [], []
        for _ in range(runs):
            t0 = time.perf_counter()
            for _ in range(steps):
                sim.step()
            dt = time.perf_counter() - t0
            rates.append(n * steps / dt)
            secs.append(dt)
        w.close()
        return statistics.median(rates), n, secs

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    all_rate, n, all_secs = measure(lib(OMP_PATH), runs)
    one_rate, _, one_secs = measure(lib(), one_thread_runs)
    return {"value": all_rate, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "one_thread_value": one_rate, "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "sample": f"{steps} steps of the same {size}x{size} world ({n} particles) after its seeding step, median of "
                      f"{runs} runs: the C restatement of fluid.wgsl (oracle/wig_oracle.c) built -O3 with OpenMP on "
                      f"{threads} threads (runs {', '.join(f'{t:.1f}' for t in all_secs)} s), and the serial -O2 build "
                      f"on one thread ({len(one_secs)} run{'s' if len(one_secs) > 1 else ''}: "
                      f"{', '.join(f'{t:.1f}' for t in one_secs)} s)"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=("single", "dam", "batch", "serve", "precision", "snapshot", "air"),
                    default="single",
                    help="single: the 1024x1024 world (the metric's configuration); dam: the 256x256 dam break "
                         "(5 levels, 2+2 red-black pairs, 1e-5, 200 steps); batch: many worlds in one batch; "
                         "serve: the serving loop over many sessions")
    ap.add_argument("--worlds", type=int, default=512, help="batch: how many worlds")
    ap.add_argument("--sessions", type=int, default=2000, help="serve: sessions")
    ap.add_argument("--rate", type=float, default=20000.0, help="serve: Poisson arrivals a second")
    ap.add_argument("--duration", type=float, default=0.25, help="serve: seconds of arrivals")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: HBM budget for the worlds")
    ap.add_argument("--round-cells", type=int, default=1 << 22,
                    help="serve: most cells a round steps (4 M keeps warm rounds with 512x512 worlds under 16 ms at p99)")
    ap.add_argument("--seeds", type=int, default=2, help="precision: worlds (seeds) per variant")
    ap.add_argument("--air-tile", type=int, default=0,
                    help="air: every level's LDS tile side (8, 16 or 32); 0: 32 on level 0, 16 on level 1, 8 below")
    ap.add_argument("--air-tiles", default="",
                    help="air: tile sides by level, comma-separated (as --air-tile, level by level)")
    ap.add_argument("--air-coarse", type=int, default=0,
                    help="air: 1, the levels that fit 64x64 in one workgroup's kernel; 0 (default), a launch each")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (single: 100, batch: 50)")
    ap.add_argument("--storage", choices=tuple(STORAGES), default=None,
                    help="V-cycle storage (single: fp32, the plain configuration; batch and serve: fp16, mixed precision)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=None, help="grid side (single: 1024, dam: 256)")
    ap.add_argument("--tol", type=float, default=None, help="relative residual (single: 1e-6, dam: 1e-5)")
    ap.add_argument("--kernel", default="solve",
                    help="what the roofline prices: 'solve' (the MG-PCG solve, the byte model's unit of work) "
                         "or a particle kernel (gather, advect, sortCount, scatter, syncParticles)")
    ap.add_argument("--kernel-table", action="store_true", help="also time every kernel over one step each")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="oracle steps in each of 3 timed runs on the host (single: 4, about 1.4 s a step on one "
                         "thread at 1024x1024; dam: 200, the configuration's)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="single: leave out the batched and served forms' short runs beside the line (configs 4 and 5)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default coarse-level tile sides")
    ap.add_argument("--settings", default="",
                    help="the autotuner's choices as a printed line gives them, in place of tuning: JSON with "
                         "'tiles', 'sweeps', 'lds_level' and 'coarse_wave_cells' (a profile's run takes the "
                         "bench line's own)")
    ap.add_argument("--graphs", type=int, default=1, help="1: replay the step's phases as hipGraphs")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (the HBM counters' calibration)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"batch": 50, "air": 30, "dam": 200}.get(args.workload, 100)
    if args.size is None:
        args.size = 256 if args.workload == "dam" else 1024
    if args.cpu_steps is None:
        args.cpu_steps = 200 if args.workload == "dam" else 4
    if args.tol is None:
        args.tol = 1e-5 if args.workload == "dam" else 1e-6
    if args.storage is None:
        args.storage = "fp16" if args.workload in ("batch", "serve") else "fp32"
    world_size = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world_size != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(independent worlds would shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from worldinagrain_amd import abi
    from worldinagrain_amd.gpu import GpuWorld, kernel_names

    if args.workload == "batch":
        print(json.dumps(run_batch(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(run_serve(args)), flush=True)
        return
    if args.workload == "precision":
        return run_precision(args)
    if args.workload == "snapshot":
        return run_snapshot(args)

    if args.copy_mib:
        src = torch.empty(args.copy_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload == "air":
        return run_air(args)

    size = args.size
    cells, fluid_cells, kinds = build_world(size, workload=args.workload)
    world = GpuWorld(size, size, kinds=kinds)
    world.upload(abi.PART_CELLS, cells)
    stream = torch.cuda.Stream() if args.graphs else torch.cuda.current_stream()
    world.set_stream(stream.cuda_stream)
    world.set_graphs(bool(args.graphs))
    storage = STORAGES[args.storage]
    sim = config_simulation(world, size, args.tol, storage)
    for _ in range(max(args.warmup, 1)):
        sim.step()
    # the coarse-level tile sides, by the autotuner on the world's own steps
    # (they do not change results)
    tiles, sweeps, lds_level, coarse_wave = {}, None, None, None
    if args.settings:
        import dataclasses
        pin = json.loads(args.settings)
        sweeps = pin.get("sweeps")
        if sweeps:
            sim.solver = dataclasses.replace(sim.solver, sweep_pairs=sweeps[0], level_pairs=sweeps[1],
                                             coarse_sweeps=sweeps[2])
        lds_level, coarse_wave = pin.get("lds_level"), pin.get("coarse_wave_cells")
        if lds_level is not None:
            world.set_lds_level(int(lds_level))
        if coarse_wave is not None:
            world.set_coarse_wave(int(coarse_wave))
        tiles = {int(k): int(v) for k, v in pin.get("tiles", {}).items()}
        for level, side in tiles.items():
            world.set_tile(level, side)
        for _ in range(2):
            sim.step()
    elif not args.no_tune:
        from worldinagrain_amd.tune import tune_tiles, tune_sweeps, tune_lds_level
        lds_level, coarse_wave = tune_lds_level(world, sim)
        tiles = tune_tiles(world, sim)
        # and the sweep counts, on copies of the world (they change the
        # iteration, so each is scored by its solve time at the tolerance);
        # then the speed settings again, for those counts.  The dam keeps
        # its configuration's 2 + 2 red-black pairs.
        if args.workload != "dam":
            solver, _ = tune_sweeps(world, sim)
            sim.solver = solver
            lds_level, coarse_wave = tune_lds_level(world, sim)
            tiles = tune_tiles(world, sim)
        solver = sim.solver
        sweeps = [solver.sweep_pairs, solver.level_pairs, solver.coarse_sweeps]
        for _ in range(2):
            sim.step()
    torch.cuda.synchronize()

    # each kernel's time over one step (untimed steps), for the record
    per_kernel = {}
    if args.kernel_table:
        for name in kernel_names():
            world.time_kernel(name)
            sim.step(read_report=False)
            ms, n = world.kernel_time()
            per_kernel[name] = (ms, n)
        world.time_kernel(None)
    torch.cuda.synchronize()

    # the timed steps, the solve (cgStart .. cgFinish: the unit of work the
    # byte model prices) bracketed by HIP events on the world's stream
    K = args.steps
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    t0 = world.run_totals()
    world.time_kernel(args.kernel)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for k in range(K):
        starts[k].record(stream)
        sim.step(read_report=False)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    kms, klaunches = world.kernel_time()
    world.time_kernel(None)
    t1 = world.run_totals()
    lat = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])
    total_ms = starts[0].elapsed_time(ends[-1])
    d = (t1.steps - t0.steps, t1.cycles - t0.cycles, t1.particles - t0.particles, t1.active_cells - t0.active_cells)
    assert d[0] == K, d
    particles_per_step = d[2] / K
    value = d[2] / (total_ms / 1e3)
    solves = K / (total_ms / 1e3)
    bytes_total = step_bytes(d, B_ITER[storage])
    ca_mean = d[3] / K

    roof = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None, "kernel": args.kernel,
            "launches": klaunches, "mean_launch_ms": kms / max(klaunches, 1)}
    if args.kernel == "solve":
        # B_solve = C_a*16 + K*C_a*B_iter, per solve
        per_launch = (d[3] * 16.0 + d[1] * ca_mean * B_ITER[storage]) / K
        roof["unit_of_work"] = (f"one MG-PCG solve (cgStart..cgFinish), B_solve = C_a*16 + K*C_a*{B_ITER[storage]:g} bytes "
                                f"({args.storage} V-cycle storage)")
        roof["C_a"] = ca_mean
        roof["pcg_iterations"] = d[1] / K
    elif args.kernel in KERNEL_BYTES_PER_PARTICLE:
        per_launch = KERNEL_BYTES_PER_PARTICLE[args.kernel] * particles_per_step * (K / max(klaunches, 1))
    else:
        per_launch = None
    roof["algorithmic_bytes_per_launch"] = per_launch
    roof["achieved"] = per_launch / (kms / max(klaunches, 1) / 1e3) / 1e9 if per_launch and klaunches else None
    roof["frac"] = roof["achieved"] / HBM_PEAK_GBS if roof["achieved"] else None
    step_roof = bytes_total / (total_ms / 1e3) / 1e9

    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": total_ms / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"plain small: {size}x{size} dam break, 1-cell solid border, water on cells "
                                f"[1,{size // 2})^2 at 4 particles/cell, g 9.81, dt 1/120, MG-PCG "
                                f"{1 + world.layout.levels} levels with 2+2 red-black pairs to {args.tol:g} relative "
                                f"residual, {K} steps") if args.workload == "dam" else
                               (f"{'plain FP32' if storage == 0 else args.storage + ' V-cycle storage'} at scale: "
                                f"{size}x{size} MAC grid, ~10% solid discs, water in the lowest "
                                f"40% of rows at 4 particles/cell, dt 1/120, MG-PCG {1 + world.layout.levels} levels "
                                f"to {args.tol:g} relative residual"),
                   "grid": [size, size], "particles": particles_per_step, "levels": 1 + world.layout.levels,
                   "tol": args.tol, "vcycle_storage": args.storage, "parallelism": "single world, one GPU",
                   "graphs": bool(args.graphs),
                   "tiles": {str(k): v for k, v in tiles.items()},
                   "sweeps": sweeps, "lds_level": lds_level, "coarse_wave_cells": coarse_wave},
        "solves_per_s": solves, "pcg_iterations_per_solve": d[1] / K,
        "p50_step_ms": float(np.percentile(lat, 50)), "p99_step_ms": float(np.percentile(lat, 99)),
        "step_roofline_frac": step_roof / HBM_PEAK_GBS, "step_achieved_GBs": step_roof,
        "wall_s": wall,
        "roofline": roof,
    }
    out["solve_ms_mean"] = kms / max(klaunches, 1) if args.kernel == "solve" else None
    if per_kernel:
        out["kernel_ms_one_step"] = {k: round(v[0], 4) for k, v in sorted(per_kernel.items(), key=lambda kv: -kv[1][0])[:12]}
    if not args.no_cpu:
        # (the dam's one-thread figure is one run of its 200 steps, about 20 s)
        out["cpu_baseline"] = cpu_baseline(cells, kinds, size, args.tol, args.cpu_steps, solver=sim.solver,
                                           one_thread_runs=1 if args.workload == "dam" else 3)
    if args.workload == "single" and not args.no_forms:
        del sim
        world.close()
        torch.cuda.synchronize()
        out["forms"] = run_forms(args)
    print(json.dumps(out), flush=True)


def run_forms(args):
    """The batched and the served forms (configs 4 and 5: the 512-world
    batch, and the serving loop over 2000 sessions) run after the
    single-world line in the same process, each with its workload's
    defaults, and summed up beside it; a form that fails is reported, not
    raised."""
    import copy
    forms = {}
    for wl in ("batch", "serve"):
        a = copy.copy(args)
        a.workload, a.storage, a.tol, a.steps, a.warmup, a.no_tune = wl, "fp16", 1e-6, 50, 5, False
        a.kernel, a.kernel_table, a.settings = "solve", False, ""
        try:
            r = run_batch(a) if wl == "batch" else run_serve(a)
        except Exception as e:       # (the single-world line stands on its own)
            forms[wl] = {"error": repr(e)}
            continue
        if wl == "batch":
            forms[wl] = {k: r[k] for k in ("metric", "value", "unit", "ms_per_step", "p50_step_ms", "p99_step_ms",
                                           "particle_steps_per_s", "pcg_iterations_per_solve", "speedup_vs_sequential")}
            forms[wl]["roofline"] = {k: r["roofline"][k] for k in ("kernel", "achieved", "peak", "frac", "unit")}
            forms[wl]["config"] = {k: r["config"][k] for k in ("workload", "worlds", "cells", "vcycle_storage", "graphs")}
        else:
            forms[wl] = {k: r[k] for k in ("metric", "value", "unit", "p50_request_ms", "p99_request_ms", "p50_step_ms_512",
                                           "p99_step_ms_512", "p50_step_ms_512_warm", "p99_step_ms_512_warm", "rounds",
                                           "evictions", "admissions", "elapsed_s", "speedup_vs_single_world_loop")}
            forms[wl]["config"] = r["config"]
    return forms


def run_batch(args):
    """Many worlds stepped together: sides drawn from {64, 128, 256, 512}
    with probabilities {.4, .3, .2, .1}; each a dam block over 20-50% of its
    cells and 0-15% solid discs (scenes.batch_world), log2(side) - 2 MG
    levels, relative residual `tol`, the V-cycle in `storage` (16-bit by
    default: the mixed-precision V-cycle).  Tile sides and sweep counts per side
    class come from the autotuner (on copies of a world of the class).  The
    same worlds stepped one at a time (a sample of each class) give the
    sequential rate beside it."""
    import math
    import numpy as np
    import torch
    from worldinagrain_amd import abi, scenes, elements as E
    from worldinagrain_amd.gpu import GpuWorld, GpuBatch
    from worldinagrain_amd.host import step_batch
    from worldinagrain_amd.tune import tune_sweeps, tune_tiles_batch, tune_coarse_wave_batch

    kinds = E.kinds()
    sides = scenes.batch_sides(args.worlds, seed=7)
    worlds, sims = [], []
    for i, side in enumerate(sides):
        cells, _ = scenes.batch_world(side, seed=1000 + i)
        w = GpuWorld(side, side, coarse_levels=int(math.log2(side)) - 3, kinds=kinds)
        w.upload(abi.PART_CELLS, cells)
        worlds.append(w)
        sims.append(config_simulation(w, side, args.tol, STORAGES[args.storage]))
    stream = torch.cuda.Stream() if args.graphs else torch.cuda.current_stream()
    batch = GpuBatch(worlds)
    batch.set_stream(stream.cuda_stream)
    batch.set_graphs(bool(args.graphs))
    for w in worlds:
        w.set_stream(stream.cuda_stream)
    for _ in range(max(args.warmup, 1)):
        step_batch(batch, sims)
    # the autotuner: sweep counts per side class on copies of the class's
    # first world; then tile sides per class in the batch itself
    tuned = {}
    if not args.no_tune:
        import copy
        for side in sorted(set(sides)):
            i = sides.index(side)
            clone = worlds[i].clone()
            clone.set_stream(stream.cuda_stream)
            sim = copy.copy(sims[i])
            sim.world = clone
            solver, scores = tune_sweeps(clone, sim)
            clone.close()
            tuned[side] = {"sweeps": [solver.sweep_pairs, solver.level_pairs, solver.coarse_sweeps]}
            for j, sj in enumerate(sides):
                if sj == side:
                    sims[j].solver = solver
        groups = {side: [i for i, s in enumerate(sides) if s == side] for side in sorted(set(sides))}
        for side, tiles in tune_tiles_batch(batch, sims, groups).items():
            tuned[side]["tiles"] = tiles
        for side, cells in tune_coarse_wave_batch(batch, sims, groups).items():
            tuned[side]["coarse_wave_cells"] = cells
    from worldinagrain_amd.host import BatchHost
    host = BatchHost(sims)          # the per-step host words for all worlds at once
    host.step(batch)
    per_kernel = {}
    if args.kernel_table:
        from worldinagrain_amd.gpu import kernel_names
        for name in kernel_names():
            batch.time_kernel(name)
            host.step(batch, read_report=False)
            per_kernel[name] = batch.kernel_time()
        batch.time_kernel(None)
    torch.cuda.synchronize()
    K = args.steps
    t0 = [w.run_totals() for w in worlds]
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    lat = []
    ev0.record(stream)
    for _ in range(K):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        host.step(batch, read_report=False)
        b.record(stream)
        lat.append((a, b))
    ev1.record(stream)
    torch.cuda.synchronize()
    total_ms = ev0.elapsed_time(ev1)
    t1 = [w.run_totals() for w in worlds]
    # the solves' roofline over the batch: SURVEY's B_solve summed over the
    # worlds (each world's C_a and PCG iterations from its run totals) over
    # the batch's solve time (HIP events around cgStart..cgFinish of a
    # separate set of steps: event timing is off in the timed ones)
    batch.time_kernel("solve")
    t2 = [w.run_totals() for w in worlds]
    for _ in range(max(2, K // 4)):
        host.step(batch, read_report=False)
    solve_ms, solve_n = batch.kernel_time()
    batch.time_kernel(None)
    t3 = [w.run_totals() for w in worlds]
    bsolve = 0.0
    for a, b in zip(t2, t3):
        st = b.steps - a.steps
        if st:
            ca = (b.active_cells - a.active_cells) / st
            bsolve += (b.active_cells - a.active_cells) * 16.0 + (b.cycles - a.cycles) * ca * B_ITER[STORAGES[args.storage]]
    solve_achieved = bsolve / solve_n / (solve_ms / solve_n / 1e3) / 1e9 if solve_n else None
    parts = sum(b.particles - a.particles for a, b in zip(t0, t1))
    cycles = sum(b.cycles - a.cycles for a, b in zip(t0, t1))
    steps = sum(b.steps - a.steps for a, b in zip(t0, t1))
    lat_ms = np.array([a.elapsed_time(b) for a, b in lat])
    # the same worlds one at a time: a sample of each class, stepped alone
    seq_ms_per_round = 0.0
    seq = {}
    for side in sorted(set(sides)):
        idx = [i for i, s in enumerate(sides) if s == side][:4]
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for i in idx:
            for _ in range(2):
                sims[i].step(read_report=False)
        e1.record(stream)
        torch.cuda.synchronize()
        per = e0.elapsed_time(e1) / (2 * len(idx))
        seq[side] = per
        seq_ms_per_round += per * sides.count(side)
    out = {
        "metric": "solves/s and particle-steps/s of a batch of worlds", "workload": "batch",
        "value": steps / (total_ms / 1e3), "unit": "solves/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": total_ms / K, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "particle_steps_per_s": parts / (total_ms / 1e3), "pcg_iterations_per_solve": cycles / max(steps, 1),
        "p50_step_ms": float(np.percentile(lat_ms, 50)), "p99_step_ms": float(np.percentile(lat_ms, 99)),
        "config": {"workload": f"{len(sides)} worlds, sides {{64,128,256,512}} w.p. {{.4,.3,.2,.1}}, dam blocks over "
                               f"20-50% of cells, 0-15% solid discs, log2(side)-2 levels, {args.tol:g} relative residual",
                   "worlds": {str(s): sides.count(s) for s in sorted(set(sides))}, "cells": int(sum(s * s for s in sides)),
                   "vcycle_storage": args.storage, "graphs": bool(args.graphs),
                   "tuned": {str(k): v for k, v in tuned.items()}},
        "sequential": {"ms_per_round": seq_ms_per_round, "solves_per_s": len(sides) / (seq_ms_per_round / 1e3),
                       "ms_per_world_step": {str(k): v for k, v in seq.items()},
                       "note": "the same worlds stepped one at a time (4 of each side, 2 steps each), scaled to the batch"},
    }
    out["speedup_vs_sequential"] = out["value"] / out["sequential"]["solves_per_s"]
    out["roofline"] = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None, "kernel": "solve",
                       "unit_of_work": "one batch step's MG-PCG solves (cgStart..cgFinish over all worlds), "
                                       f"sum of each world's C_a*16 + K*C_a*{B_ITER[STORAGES[args.storage]]:g} bytes "
                                       f"({args.storage} V-cycle storage)",
                       "launches": solve_n, "mean_launch_ms": solve_ms / max(solve_n, 1),
                       "algorithmic_bytes_per_launch": bsolve / max(solve_n, 1),
                       "achieved": solve_achieved,
                       "frac": solve_achieved / HBM_PEAK_GBS if solve_achieved else None}
    if per_kernel:
        out["kernel_ms_one_step"] = {k: [round(v[0], 4), v[1]] for k, v in sorted(per_kernel.items(), key=lambda kv: -kv[1][0])[:16]}
    batch.close()
    for w in worlds:
        w.close()
    return out


def run_precision(args):
    """Isolated pressure solves (the first step's solve of a world) on a
    `size` x `size` grid: per seed 70% water with velocities i.i.d. N(0, 1)
    per component, 20% air in random rectangles (p = 0 there), 10% solid
    discs, no gravity (scenes.precision_world), so the rhs is the
    divergence of the random velocities (regions closed off from the air
    are pinned, as the reference does).  For each tolerance (1e-6, 1e-8 on
    ||r|| / ||b||) and V-cycle storage (float, half, bfloat16), a copy of the
    same world solves it; its iterations, solve time, and its pressure's
    error and true residual against FP64 CG on the same matrix (wig_cg64)."""
    import numpy as np
    import torch
    from worldinagrain_amd import abi, scenes, elements as E
    from worldinagrain_amd.gpu import GpuWorld
    from worldinagrain_amd.host import Simulation, SolverSettings

    kinds = E.kinds()
    size = args.size if args.size != 1024 else 2048
    names = {abi.WIG_STORE_F32: "fp32", abi.WIG_STORE_F16: "fp16", abi.WIG_STORE_BF16: "bf16"}
    results = []
    for seed in range(1, args.seeds + 1):
        cells, nf = scenes.precision_world(size, seed)
        base = GpuWorld(size, size, kinds=kinds)
        base.upload(abi.PART_CELLS, cells)
        for tol in (1e-6, 1e-8):
            for st in (abi.WIG_STORE_F32, abi.WIG_STORE_F16, abi.WIG_STORE_BF16):
                w = base.clone()
                w.capture(True)
                solver = SolverSettings.relative(tol, max_cycles=200, storage=st)
                sim = Simulation(w, size, size, dx=1.0, earth_g=0.0, dt=1.0 / 120.0, solver=solver)
                w.time_kernel("solve")
                sim.step()
                ms, _ = w.kernel_time()
                w.time_kernel(None)
                ss = w.solve_stats()
                chk, _ = w.cg64(tol=1e-12)
                results.append({"seed": seed, "tol": tol, "vcycle": names[st], "iterations": int(ss.cycles),
                                "solve_ms": ms, "err_vs_fp64": chk.err, "true_relres": chk.relres,
                                "recursive_relres": float(np.sqrt(ss.rr / ss.bb)) if ss.bb > 0 else 0.0,
                                "fp64_iterations": int(chk.iterations), "fp64_relres": chk.relres64})
                print(f"[precision] seed {seed} tol {tol:g} {names[st]}: {ss.cycles} iterations, {ms:.2f} ms, "
                      f"err {chk.err:.2e}, true relres {chk.relres:.2e}", file=sys.stderr, flush=True)
                w.close()
        base.close()
    out = {"metric": "iterations, solve time and error vs FP64 CG of isolated pressure solves", "workload": "precision",
           "config": {"workload": f"{size}x{size}: 70% water with N(0,1) velocities, 20% air rectangles, 10% solid discs; "
                                  f"MG-PCG, FP32 outer PCG with FP64 sums, V-cycle storage float/half/bfloat16",
                      "seeds": args.seeds, "fluid_cells": nf},
           "results": results}
    print(json.dumps(out), flush=True)


def run_snapshot(args):
    """A world's snapshot (worldinagrain_amd.snapshot, the reference's
    SANDSNAP container) taken and loaded back: the 1024x1024 discs world
    after 3 steps.  Times, each the best of 3: take (every part copied to
    pageable host memory, synchronous), the container packed, gzip at
    levels 1 and 6, and back: gunzip, unpack, put (the parts uploaded).
    The copies' rate is against the host link; the rest is host CPU work."""
    import numpy as np
    from worldinagrain_amd import abi, snapshot
    from worldinagrain_amd.gpu import GpuWorld

    size = args.size
    cells, _, kinds = build_world(size)
    w = GpuWorld(size, size, kinds=kinds)
    w.upload(abi.PART_CELLS, cells)
    sim = config_simulation(w, size, args.tol)
    for _ in range(3):
        sim.step()
    w.sync()

    def best(f, n=3):
        t = []
        for _ in range(n):
            t0 = time.perf_counter()
            r = f()
            t.append(time.perf_counter() - t0)
        return min(t), r

    t_take, (header, parts) = best(lambda: snapshot.take(sim))
    nbytes = sum(p["bytes"] for p in header["parts"])
    t_pack, payload = best(lambda: snapshot.pack_payload(header, parts))
    t_gz1, gz1 = best(lambda: snapshot.pack(header, parts, level=1), 1)
    t_gz6, gz6 = best(lambda: snapshot.pack(header, parts, level=6), 1)
    t_unpack, (h2, p2) = best(lambda: snapshot.unpack(gz1), 1)
    w2 = GpuWorld(size, size, kinds=kinds)
    sim2 = config_simulation(w2, size, args.tol)
    t_put, _ = best(lambda: (snapshot.put(sim2, h2, p2), w2.sync()))
    same = all(np.array_equal(np.asarray(w.download(p)).view(np.uint8), np.asarray(w2.download(p)).view(np.uint8))
               for p in range(abi.PART_COUNT))
    mb = nbytes / 1e6
    out = {"metric": "snapshot save and load rates (MB of parts a second)", "workload": "snapshot", "n_gpus": 1,
           "config": {"workload": f"{size}x{size} discs dam after 3 steps, {len(header['parts'])} parts, {mb:.1f} MB"},
           "part_mb": mb, "take_s": t_take, "take_GBs": nbytes / t_take / 1e9,
           "pack_s": t_pack, "gzip1_s": t_gz1, "gzip1_MBs": mb / t_gz1, "gzip1_ratio": len(gz1) / len(payload),
           "gzip6_s": t_gz6, "gzip6_MBs": mb / t_gz6, "gzip6_ratio": len(gz6) / len(payload),
           "gunzip_unpack_s": t_unpack, "put_s": t_put, "put_GBs": nbytes / t_put / 1e9,
           "loaded_equal": bool(same)}
    print(json.dumps(out), flush=True)
    if not same:
        sys.exit(1)


def air_bytes(dims):
    """The air projection's algorithmic bytes: (per solve start, per
    iteration).  cgStart reads level 0's two float4s, the potential and the
    zones' mean and writes the CG's float4 and the potential's and rhs's
    words: 64 B a cell.  An iteration: each level pass pair but the
    coarsest's (down: 32 B read, the smoothed correction's 4 B and a quarter
    of the coarse cell's 32 B written; up: 32 B, the correction's 4 B and a
    quarter of the coarse potential's 4 B read, 4 B written) 85 B a cell of
    its level; the CG passes 88 B a level-0 cell (cgRho 8, cgDirection 16,
    cgApply 24, cgUpdate 40)."""
    c0 = dims[0][0] * dims[0][1]
    levels = sum(w * h for w, h in dims[:-1])
    return 64.0 * c0, 85.0 * levels + 88.0 * c0


def run_air(args):
    """The air's projection (worldinagrain_amd.air; shaders/air.wgsl's
    MG-PCG): a size x size air grid of 60 walled rooms (sealed air inside,
    some with a door), the lower 15 % liquid, two black holes, the divergence
    of N(0, 0.05) face velocities (scenes.air_rooms).  Each timed solve
    starts from the same level 0 and warm potentials, restored by an upload
    before it (outside the timed region), and runs up to AIR_CYCLES (64)
    iterations, those after the most |r| falls under 1e-4 skipping; HIP
    events around each solve on its stream."""
    import numpy as np
    from worldinagrain_amd import scenes
    from worldinagrain_amd.air import GpuAir, PART_MG, PART_CG, PART_STATUS, AIR_CYCLES

    size = args.size
    mg0, x0 = scenes.air_rooms(size, size, seed=21, walls=60)
    g = GpuAir(size, size)
    if args.air_tile:
        g.set_tile(args.air_tile)
    for l, side in enumerate(int(v) for v in args.air_tiles.split(",") if v):
        if l < g.levels:
            g.set_tile(side, l)
    g.set_coarse(args.air_coarse)
    g.load(mg0, x0)
    mg_init, cg_init = g.download(PART_MG), g.download(PART_CG)

    def restore():
        g.upload(PART_MG, mg_init)
        g.upload(PART_CG, cg_init)

    for _ in range(args.warmup):
        restore()
        g.project(AIR_CYCLES)
        g.sync()
    its = int(g.download(PART_STATUS)[4])
    g.time(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        restore()
        g.project(AIR_CYCLES)
    wall = time.perf_counter() - t0
    ms, n = g.elapsed()
    its_last = int(g.download(PART_STATUS)[4])
    r = g.download(PART_CG)[:4 * size * size].reshape(size, size, 4)[..., 1]
    ms_solve = ms / n
    start_b, iter_b = air_bytes(g.dims)
    solve_b = start_b + its * iter_b
    out = {
        "metric": "air projection solves/s (MG-PCG, most |r| under 1e-4)", "workload": "air",
        "value": 1000.0 / ms_solve, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_solve, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{size}x{size} air grid, 60 walled rooms, lower 15% liquid, 2 black holes, "
                               f"divergence of N(0,0.05) face velocities, {len(g.dims)} levels, LDS tiles "
                               f"{args.air_tiles or args.air_tile or '32 on level 0, 16 on level 1, 8 below'}", "iterations": its,
                   "tile": args.air_tiles or args.air_tile or "32/16/8", "coarse_fused": bool(args.air_coarse)},
        "iterations": its, "ms_per_iteration": ms_solve / its if its else None,
        "most_r": float(np.abs(r).max()), "wall_s_with_uploads": wall,
        "roofline": {"bound": "hbm", "achieved": solve_b / (ms_solve * 1e-3) / 1e9, "peak": 8000.0,
                     "unit": "GB/s", "frac": solve_b / (ms_solve * 1e-3) / 8e12, "traffic": None,
                     "unit_of_work": "one solve: cgStart and its iterations",
                     "bytes_per_solve": solve_b, "bytes_per_iteration": iter_b},
    }
    assert its_last == its, (its, its_last)
    # (a grid that needs more than AIR_CYCLES runs them all, as the
    # reference's host gives no more; the line says so)
    out["converged"] = bool(its < AIR_CYCLES)
    if not args.no_cpu:
        from oracle.air import OracleAir
        o = OracleAir(size, size)
        o.load(mg0, x0)
        t0 = time.perf_counter()
        o.project(AIR_CYCLES)
        dt = time.perf_counter() - t0
        assert int(o.status[4]) == its
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": f"one solve of the same {size}x{size} grid ({its} iterations), "
                                         f"single-threaded C restatement of air.wgsl's projection "
                                         f"(oracle/air_oracle.c, -O2), {dt:.1f} s"}
    g.close()
    print(json.dumps(out), flush=True)


def run_serve(args):
    """The serving loop: `sessions` sessions with sides log-uniform on
    [64, 1024] rounded to a multiple of 32, each a dam block with solid discs
    (scenes.batch_world; the reference seeds 4 particles a water cell);
    requests of 1-8 steps of a uniformly chosen session, Poisson arrivals at
    `rate` for `duration` seconds, the V-cycle in `storage` (16-bit by
    default); continuous batching under an HBM budget of
    `budget_gb` with LRU eviction of idle worlds to host memory
    (worldinagrain_amd.serve).  Beside it, the plain single-world loop: the
    served world-steps timed one world at a time on sample worlds of each
    size."""
    import math
    import numpy as np
    import torch
    from worldinagrain_amd import abi, scenes, elements as E
    from worldinagrain_amd.host import SolverSettings
    from worldinagrain_amd.serve import Server, Session

    kinds = E.kinds()
    rng = np.random.default_rng(11)
    sides = [int(min(1024, max(64, round(math.exp(rng.uniform(math.log(64), math.log(1024))) / 32) * 32)))
             for _ in range(args.sessions)]
    scene = {}
    for side in sorted(set(sides)):
        scene[side] = scenes.batch_world(side, seed=side)[0]
    g = float(np.float32(9.81 * (1.0 / 120.0) ** 2))
    kw = dict(dx=1.0, earth_g=g, dt=1.0 / 120.0)
    tol = args.tol
    st = STORAGES[args.storage]
    sessions = [Session(i, s, s, scene[s], lambda: SolverSettings.relative(tol, storage=st), kinds, sim_kw=kw)
                for i, s in enumerate(sides)]
    t, arrivals = 0.0, []
    while True:
        t += rng.exponential(1.0 / args.rate)
        if t > args.duration:
            break
        arrivals.append((t, int(rng.integers(args.sessions)), int(rng.integers(1, 9))))
    stream = torch.cuda.Stream()
    copy = torch.cuda.Stream()      # the evictions' (a thread of their own)
    server = Server(sessions, int(args.budget_gb * (1 << 30)), max_round_cells=args.round_cells, stream=stream.cuda_stream,
                    copy_stream=copy.cuda_stream)
    last = [time.perf_counter()]

    def progress(sv):
        if time.perf_counter() - last[0] > 20.0:
            last[0] = time.perf_counter()
            print(f"[serve] round {sv.round_no}, {sv.evictions} evictions", file=sys.stderr, flush=True)
    reqs = server.run(arrivals, on_round=progress)
    lat = np.array([r.done_at - r.arrival for r in reqs]) * 1e3
    world_steps = sum(s.steps_run for s in sessions)
    elapsed = server.elapsed
    rounds = np.array([r[2] for r in server.rounds])
    round_ms = np.array([(r[1] - r[0]) * 1e3 for r in server.rounds])
    # a step's latency for the worlds of 512 x 512 cells and more: the
    # rounds that stepped one, from the round's start to its end
    big_ms = np.array([(r[1] - r[0]) * 1e3 for r in server.rounds if r[4] >= 512 * 512] or [0.0])
    # and those whose worlds were all resident and past their first step
    warm_ms = np.array([(r[1] - r[0]) * 1e3 for r in server.rounds if r[4] >= 512 * 512 and not r[5]] or [0.0])
    # the plain single-world loop: the served steps one world at a time
    from worldinagrain_amd.gpu import GpuWorld
    from worldinagrain_amd.host import Simulation
    probe = sorted(set([64, 128, 256, 512, 1024]) & set(sides) or set(sides))
    solo_ms = {}
    for side in probe:
        w = GpuWorld(side, side, kinds=kinds)
        w.set_stream(stream.cuda_stream)
        w.upload(abi.PART_CELLS, scene[side])
        sim = Simulation(w, side, side, solver=SolverSettings.relative(tol), **kw)
        for _ in range(2):
            sim.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            sim.step()
        torch.cuda.synchronize()
        solo_ms[side] = (time.perf_counter() - t0) / 3 * 1e3
        w.close()
    cells = np.array(sorted(solo_ms)) ** 2.0
    ms = np.array([solo_ms[k] for k in sorted(solo_ms)])
    seq_s = sum(s.steps_run * float(np.interp(s.width * s.height, cells, ms)) for s in sessions) / 1e3
    out = {
        "metric": "request latency p50/p99 and particle-steps/s, served", "workload": "serve",
        "value": server.particle_steps / elapsed, "unit": "particle-steps/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "p50_request_ms": float(np.percentile(lat, 50)), "p99_request_ms": float(np.percentile(lat, 99)),
        "requests": len(reqs), "world_steps": world_steps, "world_steps_per_s": world_steps / elapsed,
        "elapsed_s": elapsed, "rounds": server.round_no, "worlds_per_round_mean": float(rounds.mean()),
        "p50_round_ms": float(np.percentile(round_ms, 50)), "p99_round_ms": float(np.percentile(round_ms, 99)),
        "p50_step_ms_512": float(np.percentile(big_ms, 50)), "p99_step_ms_512": float(np.percentile(big_ms, 99)),
        "p50_step_ms_512_warm": float(np.percentile(warm_ms, 50)), "p99_step_ms_512_warm": float(np.percentile(warm_ms, 99)),
        "warm_rounds_512": int(sum(1 for r in server.rounds if r[4] >= 512 * 512 and not r[5])),
        "evictions": server.evictions, "admissions": server.admissions,
        "config": {"workload": f"{args.sessions} sessions, sides log-uniform on [64,1024] (multiples of 32), requests of "
                               f"1-8 steps, Poisson arrivals at {args.rate:g}/s for {args.duration:g} s, HBM budget "
                               f"{args.budget_gb:g} GB, LRU eviction to host, {args.tol:g} relative residual, {args.storage} V-cycle storage",
                   "round_cells": args.round_cells},
        "single_world_loop": {"solo_ms_per_step": {str(k): v for k, v in solo_ms.items()},
                              "particle_steps_per_s": server.particle_steps / seq_s,
                              "note": "the same world-steps run one world at a time, from per-size step times"},
    }
    out["speedup_vs_single_world_loop"] = out["value"] / out["single_world_loop"]["particle_steps_per_s"]
    server.close()
    return out


if __name__ == "__main__":
    main()
