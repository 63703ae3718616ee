"""Training set of the learned preconditioner: pressure systems cut as 64^3 brick windows from seeded steps of the
dam break (64^3 cells, fluid block x < 0.4, y < 0.6, 8 jittered particles per cell, gravity -9.8 y, dt 1/120,
MGPCG to 1e-6) and of the sloshing tank (256^3 cells in 8^3 bricks, 16M particles, tank-frame gravity, dt 1/240).

A system is what the step's pressure solve sees inside the window: the cell types (0 air, 1 fluid, 2 solid), the
operator diagonal and the divergence, the right-hand side being b = -divergence. Its matrix, (diag + 1e-4) on the
diagonal and -1 between fluid neighbours inside the window, is a principal submatrix of the step's operator, so it
is symmetric and strictly diagonally dominant with a positive diagonal: SPD. slosh_load_system puts a system into a
context (cells outside the window count as solid, which keeps the original diagonal), where the library's solvers
and its trainer run on it.

Shards are compressed .npz files of up to `shard` systems: types uint8 [k, 64, 64, 64] (z, y, x), diag and
divergence float32 (zero off the fluid cells), and per system its scene, seed, step and window origin (x, y, z in
cells). The same seeds give the same systems (the steps are deterministic).

    python -m slosh_amd.training_set --out DIR [--systems 32768]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import abi, scenes

WINDOW = 64
DAM_DT = 1.0 / 120.0
TANK_DT = 1.0 / 240.0


def window_systems(sim, scene, seed, step, window=WINDOW, stride=WINDOW, min_fluid=512):
    """The systems of the current step of `sim` (a Sim, or the oracle's OracleSim): every window of window^3 cells
    at origins that are multiples of `stride` and hold at least `min_fluid` fluid cells."""
    types = (sim.read_grid(abi.FIELD_CELL_FLAGS) & 3).astype(np.uint8)
    diag = sim.read_grid(abi.FIELD_CELL_DIAG).astype(np.float32)
    div = sim.read_grid(abi.FIELD_DIVERGENCE).astype(np.float32)
    nz, ny, nx = types.shape
    out = []
    for z0 in range(0, max(nz - window, 0) + 1, stride):
        for y0 in range(0, max(ny - window, 0) + 1, stride):
            for x0 in range(0, max(nx - window, 0) + 1, stride):
                t = types[z0:z0 + window, y0:y0 + window, x0:x0 + window]
                fluid = t == 1
                if int(fluid.sum()) < min_fluid:
                    continue
                out.append({
                    "types": np.ascontiguousarray(t),
                    "diag": np.where(fluid, diag[z0:z0 + window, y0:y0 + window, x0:x0 + window], 0.0).astype(np.float32),
                    "divergence": np.where(fluid, div[z0:z0 + window, y0:y0 + window, x0:x0 + window],
                                           0.0).astype(np.float32),
                    "scene": scene, "seed": int(seed), "step": int(step), "origin": (x0, y0, z0),
                })
    return out


def dam_break_systems(make_sim, seed, steps=100, every=1, n=64):
    """Systems of one seeded dam break (the whole 64^3 grid is one window), steps 1..steps."""
    grid, h, reqs = scenes.dam_break(n, seed=seed)
    s = abi.make_settings(grid, h, scenes.particle_count(reqs) + 16, solver=abi.SOLVER_MGPCG, tolerance=1e-6)
    sim = make_sim(s)
    for r in reqs:
        sim.emit(r)
    for k in range(1, steps + 1):
        sim.step(DAM_DT, (0.0, -9.8, 0.0))
        if k % every == 0:
            yield from window_systems(sim, "dam_break_64", seed, k)
    sim.close()


def sloshing_systems(make_sim, seed, steps=200, every=4, n=256):
    """Systems of one seeded sloshing tank: 64^3 windows of every `every`-th step."""
    grid, h, reqs = scenes.sloshing_tank(n, seed=seed)
    s = abi.make_settings(grid, h, scenes.particle_count(reqs) + 16, solver=abi.SOLVER_MGPCG, tolerance=1e-6,
                          flags=abi.FLAG_MG_FLUID_COARSENING | abi.FLAG_WARM_START_EXTRAPOLATE, mg_coarse_sweeps=3)
    sim = make_sim(s)
    for r in reqs:
        sim.emit(r)
    t = 0.0
    for k in range(1, steps + 1):
        sim.step(TANK_DT, scenes.tank_gravity(t))
        t += TANK_DT
        if k % every == 0:
            yield from window_systems(sim, f"sloshing_tank_{n}", seed, k)
    sim.close()


def iter_systems(make_sim, systems, dam_fraction=0.5, dam_steps=100, tank_steps=200, tank_every=4, tank_n=256):
    """`systems` systems: dam breaks of seeds 0, 1, ... for the first dam_fraction of them, then sloshing tanks of
    seeds 1, 2, ..."""
    n_dam = int(round(systems * dam_fraction))
    made = 0
    seed = 0
    while made < n_dam:
        for sysm in dam_break_systems(make_sim, seed, dam_steps):
            if made >= n_dam:
                break
            yield sysm
            made += 1
        seed += 1
    seed = 1
    while made < systems:
        for sysm in sloshing_systems(make_sim, seed, tank_steps, tank_every, tank_n):
            if made >= systems:
                break
            yield sysm
            made += 1
        seed += 1


def write_shards(out_dir, system_iter, shard=64):
    """Writes the systems as compressed shards; returns (systems, bytes)."""
    os.makedirs(out_dir, exist_ok=True)
    buf, count, nbytes, index = [], 0, 0, 0

    def flush():
        nonlocal buf, nbytes, index
        if not buf:
            return
        path = os.path.join(out_dir, f"systems_{index:05d}.npz")
        np.savez_compressed(
            path,
            types=np.stack([b["types"] for b in buf]),
            diag=np.stack([b["diag"] for b in buf]),
            divergence=np.stack([b["divergence"] for b in buf]),
            scene=np.array([b["scene"] for b in buf]),
            seed=np.array([b["seed"] for b in buf], dtype=np.int32),
            step=np.array([b["step"] for b in buf], dtype=np.int32),
            origin=np.array([b["origin"] for b in buf], dtype=np.int32),
        )
        nbytes += os.path.getsize(path)
        index += 1
        buf = []

    for sysm in system_iter:
        buf.append(sysm)
        count += 1
        if len(buf) >= shard:
            flush()
    flush()
    return count, nbytes


def read_shards(out_dir):
    """Yields the systems of a directory of shards in shard order."""
    for name in sorted(os.listdir(out_dir)):
        if not (name.startswith("systems_") and name.endswith(".npz")):
            continue
        with np.load(os.path.join(out_dir, name), allow_pickle=False) as z:
            for k in range(len(z["seed"])):
                yield {"types": z["types"][k], "diag": z["diag"][k], "divergence": z["divergence"][k],
                       "scene": str(z["scene"][k]), "seed": int(z["seed"][k]), "step": int(z["step"][k]),
                       "origin": tuple(int(v) for v in z["origin"][k])}


def window_settings(solver=abi.SOLVER_MGPCG, flags=0, mg_coarse_sweeps=0, window=WINDOW, tolerance=1e-6):
    """Settings of a context that holds one window system (h = 1/64; the solve does not depend on h)."""
    return abi.make_settings((window, window, window), 1.0 / 64.0, 16, solver=solver, tolerance=tolerance,
                             flags=flags, mg_coarse_sweeps=mg_coarse_sweeps)


def matrix(sysm):
    """The system's matrix (scipy CSR over the fluid cells in dense order) and its right-hand side."""
    import scipy.sparse as sp

    t = sysm["types"]
    fluid = t == 1
    idx = -np.ones(t.shape, dtype=np.int64)
    idx[fluid] = np.arange(int(fluid.sum()))
    # the diagonal is the FP32 sum diag + 1e-4, as the solvers form it
    dg = (sysm["diag"][fluid].astype(np.float32) + np.float32(1e-4)).astype(np.float64)
    rows, cols, vals = [idx[fluid]], [idx[fluid]], [dg]
    for axis in range(3):
        a = np.moveaxis(fluid, axis, 0)
        ia = np.moveaxis(idx, axis, 0)
        pair = a[:-1] & a[1:]
        i0, i1 = ia[:-1][pair], ia[1:][pair]
        rows += [i0, i1]
        cols += [i1, i0]
        vals += [-np.ones(len(i0)), -np.ones(len(i0))]
    n = int(fluid.sum())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    return A, -sysm["divergence"][fluid].astype(np.float64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--systems", type=int, default=32768)
    ap.add_argument("--dam-fraction", type=float, default=0.5)
    ap.add_argument("--shard", type=int, default=64)
    ap.add_argument("--tank-n", type=int, default=256)
    a = ap.parse_args()
    from . import Sim

    t0 = time.time()
    last = [t0]

    def progress(it):
        for k, sysm in enumerate(it):
            if time.time() - last[0] > 30.0:
                print(f"{k} systems, {time.time() - t0:.0f} s", flush=True)
                last[0] = time.time()
            yield sysm

    count, nbytes = write_shards(a.out, progress(iter_systems(Sim, a.systems, a.dam_fraction, tank_n=a.tank_n)),
                                 a.shard)
    print(json.dumps({"systems": count, "bytes": nbytes, "seconds": round(time.time() - t0, 1), "out": a.out}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
