"""Benchmark of the per-step liquid solver on one MI355X.

Workload (default): the sloshing tank at 256^3 in 8^3 bricks with the 16M-particle fill
y < 0.125 + 0.025 sin(2 pi x) (15,999,968 particles, 8 per cell jittered, emit seed 1), FP32 particles and grid,
tank-frame gravity (-2 sin(2 pi t), -9.8, 0) m/s^2, dt = 1/240, pressure by MGPCG to ||r||/||b|| <= 1e-6.
A step is one full substep of the solver: sort, brick pool, P2G, normalise, divergence, pressure solve,
projection, extrapolation and G2P, with inputs resident in HBM.

Prints ONE JSON line with steps/s, the roofline of the dominant kernel (HIP events on the library's stream over
further steps, against the algorithmic bytes of that kernel; its measured HBM traffic and VALU issue share from the
committed rocprofv3 summaries), the whole step against the per-step byte model, and a CPU baseline: the oracle (a C
restatement of the reference kernels, OpenMP) on all host threads over the workload's first steps, and on one.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "spheres512_fp16": "512^3, 64 fluid spheres r ~ U(0.03, 0.08) over a 0.15-deep pool (seed 2), 8 particles per "
                       "filled cell, dt 1/240, MGPCG with FP16 multigrid levels",
    "tanks64": "64 independent tanks, each axis in {32..128} cells (seed 4), 2-6 static spheres each, half-filled "
               "below a random tilted plane, dt 1/240, stepped as one batch",
    "slosh256_16m": "sloshing tank 256^3, 8^3 bricks, 16.0M particles, dt 1/240, tank acceleration 2 m/s^2 at 1 Hz",
    "slosh256_literal": "sloshing tank 256^3, 8^3 bricks, fill y < 0.5 + 0.1 sin(2 pi x) (65.5M particles), dt 1/240",
    "dambreak64": "dam break 64^3, block x < 0.4, y < 0.6 (471,200 particles), dt 1/120",
}


def build_scene(workload):
    from slosh_amd import scenes

    if workload == "slosh256_16m":
        grid, h, reqs = scenes.sloshing_tank(256, seed=1)
        return grid, h, reqs, 1.0 / 240.0, scenes.tank_gravity
    if workload == "slosh256_literal":
        grid, h, reqs = scenes.sloshing_tank(256, seed=1, literal=True)
        return grid, h, reqs, 1.0 / 240.0, scenes.tank_gravity
    grid, h, reqs = scenes.dam_break(64, seed=0)
    return grid, h, reqs, 1.0 / 120.0, lambda t: (0.0, -9.8, 0.0)


def algorithmic_bytes(kernel, np_, nc, nf, nb, ns=None):
    """Bytes a kernel must move per launch: every array element it reads or writes, once (neighbour reuse
    perfect). np_ particles, nc allocated cells, nf fluid cells, nb allocated bricks, ns cells of the bricks that
    hold fluid (the level-0 solver passes visit only those bricks; defaults to nc)."""
    if ns is None:
        ns = nc
    table = {
        # G2P first pass: pos+age 16 B read, new brick key 4 B written; staged face velocities once per cell
        "k_g2p_keys": 20 * np_ + 12 * nc,
        # G2P second pass: pos+age, key and in-tile rank read (24 B), the 64 B record written at its sorted place
        "k_g2p_move": 88 * np_ + 12 * nc,
        # 64-byte record read and written, 4 B in-tile rank read
        "k_sort_scatter": 132 * np_,
        # the first G2P pass's key read (4 B), 4 B rank written (a step after emission reads pos+age instead)
        "k_sort_tiles": 8 * np_,
        # pos, vel, C read (64 B); brick tile of 6000 ints written; flags written; FP64 pressure read+written
        "k_p2g": 64 * np_ + 24000 * nb + 4 * nc + 16 * nc,
        # tiles read; flags read; velocity, weight (FP32 x3) and face flags (1 B x3) written
        "k_normalize": 24000 * nb + 4 * nc + 27 * nc,
        "k_divergence": 24 * nc + 4 * nc + 8 * nc,
        "k_project": 12 * nc + 12 * nc + 4 * nc + 8 * nc + 3 * nc + 15 * nc,
        "k_extrapolate": 12 * nc + 3 * nc + 12 * nc + 3 * nc,
        # level-0 MGPCG passes: vectors on fluid cells, the one-byte stencil mask on every cell of a brick that
        # holds fluid; p (8+8), d, r_in, q, diag read; r_out, z written
        "k_update_black": 40 * nf + ns,
        # z1, d_in, diag, r read; d_out, q written
        "k_direction": 24 * nf + ns,
        # diag, z, r read; level-1 b and x written (one level-1 cell per 8 cells)
        "k_restrict0": 12 * nf + 2 * ns,
        # diag, r, z read, z1 written; level-1 stencil word and x1
        "k_prolong_black": 16 * nf + ns + 6 * ns // 8,
        # diag, r, z1 read; z1 written on red cells
        "k_red_dot": 14 * nf + ns,
        "k_pcg_init": 8 * nf + 16 * nf + 12 * nc + 4 * nc,
        # a red and a black sweep: p, diag, div read once, p written on every fluid cell
        "k_sor_pair": 16 * nf + 8 * nf + nc,
    }
    return table.get(kernel)


def step_bytes(np_, nc, nf, iters):
    """Bytes one plain-MGPCG step must move by the reference pass structure's byte model: 328 per particle (sort
    pass 1 12, sort scatter 128, P2G particle side 60, G2P particle side 128), 258 per allocated cell (P2G grid
    side 28, classify 13, normalise 52, divergence 33, project 47, two extrapolations 60, G2P grid side 15, solver
    set-up 10), 108 per fluid cell of solver set-up, and per iteration 136 per fluid cell plus 8 per allocated
    cell."""
    return 328.0 * np_ + 258.0 * nc + 108.0 * nf + iters * (136.0 * nf + 8.0 * nc)


def fluid_brick_cells(sim):
    """512 x the number of 8^3 bricks holding at least one fluid cell (read back after the timed region)."""
    from slosh_amd import abi
    t = sim.read_grid(abi.FIELD_CELL_FLAGS) & 3
    nz, ny, nx = t.shape
    pz, py, px = (-nz) % 8, (-ny) % 8, (-nx) % 8
    f = np.pad(t == abi.CELL_FLUID, ((0, pz), (0, py), (0, px)))
    b = f.reshape((nz + pz) // 8, 8, (ny + py) // 8, 8, (nx + px) // 8, 8).any(axis=(1, 3, 5))
    return int(b.sum()) * 512


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 7 and not line.startswith("#") and f[0] == workload and f[1] == kernel:
            return int(float(f[6]))
    return None


def measured_valu(workload, kernel):
    """Share of the kernel's time the GPU's SIMDs spend issuing VALU instructions, from the committed rocprofv3 PMC
    summary (4 cycles per wave64 instruction per SIMD over the dispatch's elapsed cycles), or None."""
    path = os.path.join(ROOT, "profiles", "valu_issue.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 9 and not line.startswith("#") and f[0] == workload and f[1] == kernel:
            return float(f[8])
    return None


def copy_calibration(mib=512, reps=3):
    """Device-to-device copies of a known size, so a PMC pass over this run can scale FETCH_SIZE/WRITE_SIZE."""
    import torch

    src = torch.empty(mib << 20, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()


def host_cpu():
    """Logical CPUs this process may run on, and the CPU model."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count()
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, model


def _oracle_steps(O, abi, grid, h, reqs, dt, grav, steps, threads):
    """Seconds per step of the C oracle over the first `steps` steps after emission, on `threads` threads."""
    n = sum(int(r.cell_dims[0]) * int(r.cell_dims[1]) * int(r.cell_dims[2]) * 8 for r in reqs)
    O.set_threads(threads)
    o = O.OracleSim(abi.make_settings(grid, h, n))
    for r in reqs:
        o.emit(r)
    t0 = time.perf_counter()
    for k in range(steps):
        o.step(dt, grav(k * dt))
    el = time.perf_counter() - t0
    o.close()
    return el / steps


def cpu_baseline(workload):
    """The C oracle (the reference's kernels restated; OpenMP) on the same workload over its first steps after
    emission, on all host threads (`value`) and on one, plus the 64^3 dam break over 10 steps on both."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    from slosh_amd import abi, scenes

    nthreads = O.threads()
    ncpu, model = host_cpu()
    grid, h, reqs, dt, grav = build_scene(workload)
    all_s = _oracle_steps(O, abi, grid, h, reqs, dt, grav, 3, nthreads)
    one_s = _oracle_steps(O, abi, grid, h, reqs, dt, grav, 1, 1)
    dgrid, dh, dreqs = scenes.dam_break(64, seed=0)
    dgrav = lambda t: (0.0, -9.8, 0.0)  # noqa: E731
    dam_all = _oracle_steps(O, abi, dgrid, dh, dreqs, 1.0 / 120.0, dgrav, 10, nthreads)
    dam_one = _oracle_steps(O, abi, dgrid, dh, dreqs, 1.0 / 120.0, dgrav, 10, 1)
    O.set_threads(nthreads)
    return {"value": round(1.0 / all_s, 4), "unit": "steps/s", "cores": nthreads, "kind": "port",
            "sample": f"steps 1-3 after emission of the {workload} workload on the C oracle (OpenMP, "
                      f"{nthreads} threads, {all_s:.2f} s per step)",
            "one_thread": {"value": round(1.0 / one_s, 4), "unit": "steps/s", "sample": "step 1, 1 thread",
                           "s_per_step": round(one_s, 2)},
            "dam_break_64": {"sample": "64^3 dam break, 471,200 particles, MGPCG to 1e-6, steps 1-10",
                             "all_threads_steps_per_s": round(1.0 / dam_all, 3),
                             "one_thread_steps_per_s": round(1.0 / dam_one, 3)},
            "host": {"logical_cpus": ncpu, "model": model}}


def bench_other(args):
    """The mixed-precision (512^3, FP16 levels) and batched (64 tanks) workloads: steps/s over the timed steps."""
    from slosh_amd import Batch, Sim, abi, scenes

    g = (0.0, -9.8, 0.0)
    dt = 1.0 / 240.0
    base = {"unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic"}
    if args.workload == "spheres512_fp16":
        grid, h, reqs = scenes.spheres_pool(512, seed=2)
        n = sum(int(r.cell_dims[0]) * int(r.cell_dims[1]) * int(r.cell_dims[2]) * 8 for r in reqs)
        sim = Sim(abi.make_settings(grid, h, n, solver=abi.SOLVER_MGPCG, tolerance=1e-6, budget=0.5,
                                    flags=abi.FLAG_MG_FP16 | abi.FLAG_MG_FLUID_COARSENING,
                                    mg_coarse_sweeps=args.mg_coarse_sweeps))
        for r in reqs:
            sim.emit(r)
        for _ in range(args.warmup):
            sim.step(dt, g)
        per = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sim.step(dt, g)
            r = sim.results()
            per.append((r.solver_iterations, r.residual_rel, r.true_residual_rel, r.converged))
        el = time.perf_counter() - t0
        res = sim.results()
        sim.close()
        return dict(base, metric="steps/s @512^3 spheres + pool, FP16 multigrid levels, MGPCG at 1e-6 residual",
                    value=round(args.steps / el, 3), ms_per_step=round(1e3 * el / args.steps, 3),
                    config={"workload": WORKLOADS[args.workload], "particles": res.live_particles, "grid": list(grid),
                            "fluid_cells": res.fluid_cells, "allocated_bricks": res.active_bricks,
                            "mean_solver_iterations": round(sum(p[0] for p in per) / len(per), 2),
                            "max_rel_residual": max(p[1] for p in per),
                            "max_true_rel_residual_fp64": max(p[2] for p in per),
                            "converged_steps": sum(p[3] for p in per),
                            "warm_start": "reference",
                            "mg_preconditioner": {"fluid_first_coarse_types": True,
                                                  "coarse_sweeps": max(args.mg_coarse_sweeps, 1)},
                            "particles_per_s": round(args.steps / el * res.live_particles, 1)})
    specs = scenes.tank_batch(64, seed=4)
    sims = []
    for grid, h, reqs, cols in specs:
        n = sum(int(r.cell_dims[0]) * int(r.cell_dims[1]) * int(r.cell_dims[2]) * 8 for r in reqs) + 1024
        sim = Sim(abi.make_settings(grid, h, n, solver=abi.SOLVER_MGPCG, tolerance=1e-6,
                                    flags=abi.FLAG_MG_FLUID_COARSENING | abi.FLAG_WARM_START_EXTRAPOLATE,
                                    mg_coarse_sweeps=args.mg_coarse_sweeps))
        sim.set_colliders(cols)
        for r in reqs:
            sim.emit(r)
        sims.append(sim)
    b = Batch(sims)
    grav = np.tile(np.array(g, np.float32), (len(sims), 1))
    if args.no_autotune:
        # lockstep MGPCG iterations, 4 host threads for the passes around the solve, chunks of 4, graphs, one brick
        # per workgroup of the level-0 passes
        b.set_config(4, 4, 1, 1, 1)
        chosen = b.config()
    else:
        chosen = b.autotune(dt, grav, steps_per_candidate=1)
    for _ in range(args.warmup):
        b.step(dt, grav)
    t0 = time.perf_counter()
    its, conv = 0, 0
    for _ in range(args.steps):
        b.step(dt, grav)
        for s in sims:
            r = s.results()
            its += r.solver_iterations
            conv += r.converged
    el = time.perf_counter() - t0
    parts = sum(s.results().live_particles for s in sims)
    for s in sims:
        s.close()
    b.close()
    return dict(base, metric="batch steps/s, 64 tanks (all tanks advance one step), MGPCG at 1e-6 residual",
                value=round(args.steps / el, 3), ms_per_step=round(1e3 * el / args.steps, 3),
                config={"workload": WORKLOADS[args.workload], "tanks": len(specs), "particles": parts,
                        "autotuned": chosen, "mean_solver_iterations": round(its / (len(specs) * args.steps), 2),
                        "warm_start": "extrapolated",
                        "mg_preconditioner": {"fluid_first_coarse_types": True,
                                              "coarse_sweeps": max(args.mg_coarse_sweeps, 1)},
                        "converged_tank_steps": conv, "tank_steps": len(specs) * args.steps,
                        "particles_per_s": round(args.steps / el * parts, 1)})


def learned_on_dam_break_256(weights, mg, warmup=5, steps=20):
    """The dam break of the 64^3 configuration at 256^3 (31.7M particles; dt 1/480 keeps its CFL number), MGPCG and
    the learned flexible CG with the same V-cycle options: steps/s over steps warmup+1..warmup+steps, iterations,
    fallbacks."""
    from slosh_amd import Sim, abi, scenes

    grid, h, reqs = scenes.dam_break(256, seed=0)
    n = scenes.particle_count(reqs)
    out = {"workload": "dam break 256^3, block x < 0.4, y < 0.6, 8 jittered particles per cell (seed 0), dt 1/480",
           "particles": n, "timed_window": f"steps {warmup + 1}-{warmup + steps}"}
    for name, solver in (("mgpcg", abi.SOLVER_MGPCG), ("learned_fcg", abi.SOLVER_FCG_UNET)):
        sim = Sim(abi.make_settings(grid, h, n + 16, solver=solver, tolerance=1e-6, flags=mg["flags"],
                                    mg_coarse_sweeps=mg["mg_coarse_sweeps"]))
        for r in reqs:
            sim.emit(r)
        if solver == abi.SOLVER_FCG_UNET:
            sim.unet_set_weights(weights)
        its, fb, conv = [], 0, 0
        for k in range(warmup + steps):
            if k == warmup:
                t0 = time.perf_counter()
            sim.step(1.0 / 480.0, (0.0, -9.8, 0.0))
            if k >= warmup:
                r = sim.results()
                its.append(r.solver_iterations)
                fb += r.fallback_used == 1
                conv += r.converged
        el = time.perf_counter() - t0
        sim.close()
        out[name] = {"steps_per_s": round(steps / el, 3), "mean_iterations": round(sum(its) / len(its), 2),
                     "fallback_steps": fb, "converged_steps": conv}
    out["speedup_vs_mgpcg"] = round(out["learned_fcg"]["steps_per_s"] / out["mgpcg"]["steps_per_s"], 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # defaults: the sloshing tank's 200-step run, its first 5 steps untimed
    ap.add_argument("--steps", type=int, default=195)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="slosh256_16m", choices=sorted(WORKLOADS))
    ap.add_argument("--solver", default="mgpcg", choices=["mgpcg", "rbsor"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-report", action="store_true", help="print per-kernel times to stderr")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-FCG comparison run")
    ap.add_argument("--mg-reference-cycle", action="store_true",
                    help="time the reference's V-cycle (air-first coarse types, one red-black pair per level) and "
                         "warm start for `value` instead of the tuned options")
    ap.add_argument("--mg-coarse-sweeps", type=int, default=3,
                    help="tuned preconditioner: red-black pairs per visit of levels 4 and coarser")
    ap.add_argument("--no-reference-cycle-run", action="store_true",
                    help="skip the comparison run with the reference's V-cycle")
    ap.add_argument("--roofline-steps", type=int, default=10,
                    help="further steps over which the dominant kernel is timed with HIP events")
    ap.add_argument("--no-kernel-events", action="store_true",
                    help="no HIP events around the dominant kernel in the timed steps (no roofline figure)")
    ap.add_argument("--fcg-steps", type=int, default=None,
                    help="timed steps of the learned-FCG comparison run (default: the same window as `value`)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="tanks64: the lockstep batch configuration instead of the autotuner's choice")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run 512 MiB device copies first (scale reference for rocprofv3 --pmc FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solver is a single-GPU library (one context per device); "
              "multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    from slosh_amd import Sim, abi, unet

    if args.copy_calibration:
        copy_calibration()
    if args.workload in ("spheres512_fp16", "tanks64"):
        print(json.dumps(bench_other(args)), flush=True)
        return
    grid, h, reqs, dt, grav = build_scene(args.workload)
    n = sum(int(r.cell_dims[0]) * int(r.cell_dims[1]) * int(r.cell_dims[2]) * 8 for r in reqs)

    # MGPCG options: tuned (fluid-first coarse cell types, extra red-black pairs on levels 4 and coarser, the
    # warm start extrapolated from the last two solutions) or the reference's cycle and warm start; both solve to the
    # same tolerance
    tuned = {"flags": abi.FLAG_MG_FLUID_COARSENING | abi.FLAG_WARM_START_EXTRAPOLATE,
             "mg_coarse_sweeps": args.mg_coarse_sweeps}
    reference_cycle = {"flags": 0, "mg_coarse_sweeps": 0}
    mg = reference_cycle if args.mg_reference_cycle else tuned

    def run(solver, weights, steps, warmup, mg_opts, among=None):
        """Times `steps` steps after `warmup`; returns (steps/s, per-step data, dominant kernel, its timing, sim).
        The dominant kernel is the one with the largest total time among `among` (default: the kernels with a byte
        model)."""
        s = abi.make_settings(grid, h, n, solver=solver, tolerance=1e-6,
                              max_iterations=20000 if solver == abi.SOLVER_RBSOR else 200, sor_weight=1.8,
                              flags=mg_opts["flags"], mg_coarse_sweeps=mg_opts["mg_coarse_sweeps"])
        sim = Sim(s)
        for r in reqs:
            sim.emit(r)
        if weights is not None:
            sim.unet_set_weights(weights)
        t = 0.0
        for _ in range(warmup):
            sim.step(dt, grav(t))
            t += dt
        # the dominant kernel from a fully instrumented step: the largest total time among the kernels that move
        # data by a byte model (the one-workgroup reductions and set-up passes are latency, not traffic)
        sim.set_kernel_timing(~0)
        sim.kernel_times(reset=True)
        sim.step(dt, grav(t))
        t += dt
        prof = sim.kernel_times(reset=True)
        if among is None:
            modelled = {k: v for k, v in prof.items() if algorithmic_bytes(k, 1, 1, 1, 1) is not None}
        else:
            modelled = {k: v for k, v in prof.items() if k in among}
        dominant = max((modelled or prof).items(), key=lambda kv: kv[1][1])[0]
        sim.set_kernel_timing(0)
        # the timed steps: no instrumentation (slosh_step returns with the library's stream idle)
        per = []
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.step(dt, grav(t))
            t += dt
            r = sim.results()
            per.append((r.solver_iterations, r.residual_rel, r.converged, r.fallback_used))
        el = time.perf_counter() - t0
        # the dominant kernel's launches timed with HIP events on the library's stream over further steps of the
        # same workload (event timing turns the replay of solver chunks as HIP graphs off, so the timed steps
        # above run without it)
        kt = (0, 0.0)
        if not args.no_kernel_events and args.roofline_steps > 0:
            sim.set_kernel_timing(1 << sim.kernel_index(dominant))
            sim.kernel_times(reset=True)
            for _ in range(args.roofline_steps):
                sim.step(dt, grav(t))
                t += dt
            kt = sim.kernel_times(reset=True)[dominant]
            sim.set_kernel_timing(0)
        return steps / el, el, per, dominant, kt, prof, sim

    solver = {"mgpcg": abi.SOLVER_MGPCG, "rbsor": abi.SOLVER_RBSOR}[args.solver]
    sps, el, per, dominant, (launches, total_ms), prof, sim = run(solver, None, args.steps, args.warmup, mg)
    res = sim.results()
    avg_s = total_ms / 1e3 / max(launches, 1)
    nb = res.active_bricks
    nc = nb * 512
    nf = res.fluid_cells
    ns = fluid_brick_cells(sim)
    abytes = algorithmic_bytes(dominant, res.live_particles, nc, nf, nb, ns)
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": measured_traffic(args.workload, dominant),
                "valu_issue_frac": measured_valu(args.workload, dominant), "launches": launches,
                "avg_us": avg_s * 1e6,
                "timing": f"a HIP event pair carried by each launch's dispatch on the library's stream (the kernel's "
                          f"own begin and end), {args.roofline_steps} steps after the timed ones"}
    if abytes and avg_s > 0:
        ach = abytes / avg_s / 1e9
        roofline.update({"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes": int(abytes)})
    sim.close()
    iters = [p[0] for p in per]
    rels = [p[1] for p in per]
    conv = sum(p[2] for p in per)

    # learned preconditioner: flexible CG with the trained network (MGPCG fallback), same workload and tolerance
    learned = None
    wpath = unet.default_weights_path()
    if args.solver == "mgpcg" and not args.no_learned and os.path.exists(wpath):
        fcg_steps = args.fcg_steps if args.fcg_steps is not None else args.steps
        fcg_warmup = args.warmup if args.fcg_steps is None else 2
        # its roofline: the network's dominant convolution (MFMA-bound)
        fsps, fel, fper, fdom, (flaunch, fms), fprof, fsim = run(abi.SOLVER_FCG_UNET, unet.load(wpath),
                                                                 fcg_steps, fcg_warmup, mg,
                                                                 among=("k_unet_conv_l0", "k_unet_conv_l1",
                                                                        "k_unet_conv_l2"))
        fsim.close()
        favg = fms / 1e3 / max(flaunch, 1) if fms > 0 else float("nan")
        # level-0 forward convolutions: E0a (4->16), E0b and D0b (16->16), 27 taps, 512 cells per brick
        conv_flops = {"k_unet_conv_l0": 2.0 * 27 * 512 * nb * (4 * 16 + 16 * 16 + 16 * 16) / 3.0,
                      "k_unet_conv_l1": 2.0 * 27 * 64 * nb * (16 * 32 + 32 * 32 + 32 * 32 + 32 * 16) / 4.0,
                      "k_unet_conv_l2": 2.0 * 27 * 8 * nb * (32 * 64 + 64 * 64 + 64 * 32) / 3.0}
        fl = {"kernel": fdom, "launches": flaunch, "avg_us": round(favg * 1e6, 2)}
        if fdom in conv_flops and favg == favg:
            ach = conv_flops[fdom] / favg / 1e12
            fl.update({"bound": "mfma", "achieved": round(ach, 1), "peak": 2500.0, "unit": "TFLOP/s",
                       "frac": round(ach / 2500.0, 4), "algorithmic_flops": int(conv_flops[fdom])})
        # every level's convolution from the instrumented step; level 0 also against HBM (each launch reads 16 BF16
        # channels and writes 16 per cell, 512 cells per brick)
        fl["convs"] = {}
        for k, flops in conv_flops.items():
            cl, cms = fprof.get(k, (0, 0.0))
            if cl and cms > 0:
                cavg = cms / 1e3 / cl
                fl["convs"][k] = {"launches": cl, "avg_us": round(cavg * 1e6, 2),
                                  "mfma_frac": round(flops / cavg / 1e12 / 2500.0, 4)}
                if k == "k_unet_conv_l0":
                    cbytes = nb * 512 * (16 + 16) * 2
                    fl["convs"][k]["hbm_frac"] = round(cbytes / cavg / 1e9 / HBM_PEAK_GBS, 4)
        # the learned preconditioner's evaluation scene: the dam break at 256^3
        c4 = learned_on_dam_break_256(unet.load(wpath), mg)
        learned = {"steps_per_s": round(fsps, 3), "speedup_vs_mgpcg": round(fsps / sps, 3),
                   "mean_iterations": round(sum(p[0] for p in fper) / len(fper), 2),
                   "fallback_steps": sum(1 for p in fper if p[3] == 1), "steps": fcg_steps, "warmup": fcg_warmup,
                   "converged_steps": sum(p[2] for p in fper), "max_rel_residual": max(p[1] for p in fper),
                   "roofline": fl, "weights": os.path.relpath(wpath, ROOT), "dam_break_256": c4}
    # the same workload with the reference's V-cycle, for comparison
    ref_cycle = None
    if args.solver == "mgpcg" and not args.mg_reference_cycle and not args.no_reference_cycle_run:
        rsps, rel_, rper, _, _, _, rsim = run(abi.SOLVER_MGPCG, None, args.steps, args.warmup, reference_cycle)
        rsim.close()
        ref_cycle = {"steps_per_s": round(rsps, 3),
                     "mean_solver_iterations": round(sum(p[0] for p in rper) / len(rper), 2),
                     "max_rel_residual": max(p[1] for p in rper), "converged_steps": sum(p[2] for p in rper),
                     "warm_start": "reference"}
    line = {
        "metric": "steps/s @256^3,16M particles; learned-FCG vs plain MGPCG at 1e-6 residual",
        "value": round(sps, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload], "solver": f"{args.solver} to ||r||/||b|| <= 1e-6",
                   "timed_window": f"steps {args.warmup + 1}-{args.warmup + args.steps} of the simulation",
                   "value_solver": args.solver, "particles": res.live_particles, "grid": list(grid),
                   "fluid_cells": nf, "allocated_bricks": nb, "fluid_bricks": ns // 512,
                   "mean_solver_iterations": round(sum(iters) / len(iters), 2), "max_rel_residual": max(rels),
                   "converged_steps": conv, "particles_per_s": round(sps * res.live_particles, 1),
                   "mg_preconditioner": {"fluid_first_coarse_types": bool(mg["flags"] & abi.FLAG_MG_FLUID_COARSENING),
                                         "coarse_sweeps": max(mg["mg_coarse_sweeps"], 1)},
                   "warm_start": "extrapolated" if mg["flags"] & abi.FLAG_WARM_START_EXTRAPOLATE else "reference",
                   "mg_reference_cycle": ref_cycle,
                   "learned_fcg": learned},
        "roofline": roofline,
    }
    if args.solver == "mgpcg":
        # the whole step against HBM: the per-step byte model at the timed steps' mean iteration count, over the
        # measured time per step
        sb = step_bytes(res.live_particles, nc, nf, sum(iters) / len(iters))
        ach = sb / (el / args.steps) / 1e9
        line["step_roofline"] = {"bound": "hbm", "bytes_per_step": int(sb), "achieved": round(ach, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                                 "model": "328 Np + 258 Nc + 108 Nf + k (136 Nf + 8 Nc): Np particles, Nc "
                                          "allocated cells, Nf fluid cells, k mean iterations per step"}
    if args.kernel_report:
        for name, table in (("mgpcg", prof), ("learned fcg", fprof if learned is not None else {})):
            if table:
                print(f"-- {name}: one instrumented step", file=sys.stderr)
            for k, (l, ms) in sorted(table.items(), key=lambda kv: -kv[1][1]):
                if l:
                    print(f"{k:22s} {l:6d} launches {ms:9.3f} ms", file=sys.stderr)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
