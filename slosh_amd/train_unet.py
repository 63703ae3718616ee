"""Trains the learned preconditioner on the GPU with the library's own forward/backward kernels.

The network is the second stage of the hybrid preconditioner M(r) = V(r) + N(r - A V(r)): it learns to correct what
the multigrid V-cycle leaves. Training data are pressure systems that the plain solver generates from seeded
scenes: each scene is stepped with MGPCG (with the V-cycle options the flexible CG uses), the residuals each V-cycle
leaves are captured (SLOSH_FLAG_CAPTURE_RESIDUALS | SLOSH_FLAG_CAPTURE_VCYCLE_RESIDUALS), and a few Adam steps
minimise ||A N(r') - r'|| / ||r'|| on them before the next step (= ||A M r - r|| / ||r - A V r||), so the network
sees the systems of whole trajectories (sloshing tanks with random fill depth, wave amplitude and tank
acceleration; dam breaks). Usage:

    python -m slosh_amd.train_unet --out slosh_amd/weights/unet_slosh.npz --seconds 300

With --dataset DIR the systems come from a training set written by slosh_amd.training_set (64^3 windows of seeded
dam breaks and sloshing tanks): each is loaded into one 64^3 context (slosh_load_system), solved with MGPCG and the
capture flags (slosh_solve), and trained on like a step's solve.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import Sim, abi, scenes, unet

# V-cycle options of the training solves: those the flexible CG of the benchmark uses (fluid-first coarse types,
# three red-black pairs on the coarse levels)
TRAIN_MG_FLAGS = abi.FLAG_MG_FLUID_COARSENING
TRAIN_MG_SWEEPS = 3


def scene(kind, n, rng):
    """(grid, h, requests, dt, gravity(t)) of a random training scene."""
    if kind == "slosh":
        mean = float(rng.uniform(0.08, 0.3))
        amp = float(rng.uniform(0.0, 0.3)) * mean
        grid, h, reqs = scenes.wave_columns(n, int(rng.integers(1, 1 << 30)), mean, amp)
        acc = float(rng.uniform(0.5, 3.0))
        freq = float(rng.uniform(0.5, 2.0))
        return grid, h, reqs, 1.0 / 240.0, lambda t: (-acc * math.sin(2.0 * math.pi * freq * t), -9.8, 0.0)
    fx, fy = float(rng.uniform(0.2, 0.6)), float(rng.uniform(0.3, 0.8))
    grid, h, reqs = scenes.dam_break(n, seed=int(rng.integers(1, 1 << 30)), fill=(fx, fy, 1.0))
    return grid, h, reqs, 1.0 / 120.0, lambda t: (0.0, -9.8, 0.0)


def train(out, seconds, sizes=(64, 128), steps_per_scene=24, iters_per_step=6, batch=4, lr=1e-3, seed=0,
          init=None, log=None):
    rng = np.random.default_rng(seed)
    w = unet.load(init) if init else unet.he_init(seed, head_scale=0.1)
    t0 = time.time()
    history = []
    nscene = 0
    while time.time() - t0 < seconds:
        kind = "slosh" if nscene % 3 != 2 else "dam"
        n = int(sizes[nscene % len(sizes)])
        grid, h, reqs, dt, grav = scene(kind, n, rng)
        cap = sum(int(np.prod(r.cell_dims)) * 8 for r in reqs) + 1024
        s = abi.make_settings(grid, h, cap, solver=abi.SOLVER_MGPCG, tolerance=1e-6,
                              flags=abi.FLAG_CAPTURE_RESIDUALS | abi.FLAG_CAPTURE_VCYCLE_RESIDUALS | TRAIN_MG_FLAGS,
                              mg_coarse_sweeps=TRAIN_MG_SWEEPS)
        sim = Sim(s)
        for r in reqs:
            sim.emit(r)
        sim.unet_set_weights(w)
        t = 0.0
        for k in range(steps_per_scene):
            sim.step(dt, grav(t))
            t += dt
            step_lr = lr * (0.3 if time.time() - t0 > 0.7 * seconds else 1.0)
            rep = sim.unet_train(iters_per_step, batch=batch, learning_rate=step_lr, offset=int(rng.integers(0, 64)))
            history.append({"scene": nscene, "kind": kind, "n": n, "step": k, "first_loss": rep["first_loss"],
                            "last_loss": rep["last_loss"], "systems": rep["systems"]})
            if time.time() - t0 >= seconds:
                break
        w = sim.unet_get_weights()
        sim.close()
        last = history[-1]
        msg = (f"scene {nscene} {kind} {n}^3: loss {history[-steps_per_scene if len(history) >= steps_per_scene else 0]['first_loss']:.4f}"
               f" -> {last['last_loss']:.4f}  ({time.time() - t0:.0f} s)")
        print(msg, flush=True)
        if log:
            with open(log, "a") as f:
                f.write(msg + "\n")
        nscene += 1
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    unet.save(out, w, seconds=seconds, scenes=nscene, final_loss=history[-1]["last_loss"] if history else -1.0)
    return w, history


def train_dataset(out, dataset, seconds, iters_per_system=6, batch=4, lr=1e-3, seed=0, init=None, log=None,
                  epochs=1000):
    """Trains on the systems of a training set, in a seeded random order of its shards, until `seconds` pass."""
    from . import training_set

    rng = np.random.default_rng(seed)
    w = unet.load(init) if init else unet.he_init(seed, head_scale=0.1)
    s = training_set.window_settings(flags=abi.FLAG_CAPTURE_RESIDUALS | abi.FLAG_CAPTURE_VCYCLE_RESIDUALS |
                                     TRAIN_MG_FLAGS, mg_coarse_sweeps=TRAIN_MG_SWEEPS)
    sim = Sim(s)
    sim.unet_set_weights(w)
    shards = sorted(n for n in os.listdir(dataset) if n.startswith("systems_") and n.endswith(".npz"))
    t0 = time.time()
    history = []
    done = False
    for epoch in range(epochs):
        for name in [shards[k] for k in rng.permutation(len(shards))]:
            with np.load(os.path.join(dataset, name), allow_pickle=False) as z:
                types, diag, div = z["types"], z["diag"], z["divergence"]
                for k in rng.permutation(len(types)):
                    sim.load_system(types[k], diag[k], div[k])
                    sim.solve()
                    step_lr = lr * (0.3 if time.time() - t0 > 0.7 * seconds else 1.0)
                    rep = sim.unet_train(iters_per_system, batch=batch, learning_rate=step_lr,
                                         offset=int(rng.integers(0, 64)))
                    history.append({"shard": name, "system": int(k), "first_loss": rep["first_loss"],
                                    "last_loss": rep["last_loss"], "systems": rep["systems"]})
                    if time.time() - t0 >= seconds:
                        done = True
                        break
            msg = f"epoch {epoch} {name}: loss {history[-1]['last_loss']:.4f} ({time.time() - t0:.0f} s, {len(history)} systems)"
            print(msg, flush=True)
            if log:
                with open(log, "a") as f:
                    f.write(msg + "\n")
            if done:
                break
        if done:
            break
    w = sim.unet_get_weights()
    sim.close()
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    unet.save(out, w, seconds=seconds, systems=len(history), final_loss=history[-1]["last_loss"] if history else -1.0)
    return w, history


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=unet.default_weights_path())
    ap.add_argument("--seconds", type=float, default=300.0)
    ap.add_argument("--sizes", default="64,128")
    ap.add_argument("--steps-per-scene", type=int, default=24)
    ap.add_argument("--iters-per-step", type=int, default=6)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None)
    ap.add_argument("--log", default=None)
    ap.add_argument("--history", default=None, help="write the per-step loss history (JSON) here")
    ap.add_argument("--dataset", default=None, help="train on a training set written by slosh_amd.training_set")
    a = ap.parse_args()
    if a.dataset:
        _, hist = train_dataset(a.out, a.dataset, a.seconds, a.iters_per_step, a.batch, a.lr, a.seed, a.init, a.log)
    else:
        _, hist = train(a.out, a.seconds, tuple(int(x) for x in a.sizes.split(",")), a.steps_per_scene,
                        a.iters_per_step, a.batch, a.lr, a.seed, a.init, a.log)
    if a.history:
        with open(a.history, "w") as f:
            json.dump(hist, f)
    print(json.dumps({"weights": a.out, "items": len(hist), "final_loss": hist[-1]["last_loss"] if hist else None}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
