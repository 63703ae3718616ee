"""Writes the committed fixtures of tests/test_golden.py.

pcg_hash.json: the PCG hash (Jarzynski & Olano 2020) of 64 inputs, computed here in numpy from the published
  formula (not by the oracle), the jitter generator of the reference's FillCells emission.
scene12.npz: a 12^3 dam break (emit seed 5) stepped twice by the oracle: after step 1 with the pressure solve off
  (every non-solver kernel, bit-exact for any implementation of the same semantics) the sorted particles, brick
  table, cell flags, face velocities and flags; after a further step with MGPCG to 1e-6, the FP64 pressure and the
  oracle's FP64 reference solve of that step's system.
Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from slosh_amd import abi, scenes  # noqa: E402


def np_pcg_hash(v):
    v = np.asarray(v, dtype=np.uint32)
    state = v * np.uint32(747796405) + np.uint32(2891336453)
    word = ((state >> ((state >> np.uint32(28)) + np.uint32(4))) ^ state) * np.uint32(277803737)
    return (word >> np.uint32(22)) ^ word


def scene12():
    grid, h, reqs = scenes.dam_break(12, seed=5, fill=(0.5, 0.6, 1.0))
    return grid, h, reqs


def run_scene(sim_factory):
    grid, h, reqs = scene12()
    out = {}
    s = abi.make_settings(grid, h, 4000, solver=abi.SOLVER_NONE)
    sim = sim_factory(s)
    for r in reqs:
        sim.emit(r)
    sim.step(1.0 / 120.0, (0.0, -9.8, 0.0))
    out["particles"] = sim.read_particles().view(np.uint32).reshape(-1, 16)
    out["brick_table"] = sim.read_grid(abi.FIELD_BRICK_TABLE)
    out["cell_flags"] = sim.read_grid(abi.FIELD_CELL_FLAGS)
    out["vel"] = sim.read_grid(abi.FIELD_VEL)
    out["face_flags"] = sim.read_grid(abi.FIELD_FACE_FLAGS)
    s = abi.make_settings(grid, h, 4000, solver=abi.SOLVER_MGPCG, tolerance=1e-6)
    sim = sim_factory(s)
    for r in reqs:
        sim.emit(r)
    sim.step(1.0 / 120.0, (0.0, -9.8, 0.0))
    sim.step(1.0 / 120.0, (0.0, -9.8, 0.0))
    out["pressure_mgpcg"] = sim.read_grid(abi.FIELD_PRESSURE_F64)
    return out, sim


def main():
    import oracle_py as O

    xs = [0, 1, 2, 3, 7, 42, 255, 256, 1000, 7919, 65535, 65536, 123456789, 2 ** 31 - 1, 2 ** 31, 2 ** 32 - 1]
    xs += [int(x) for x in np.random.default_rng(0).integers(0, 2 ** 32, 48, dtype=np.uint64)]
    with open(os.path.join(HERE, "pcg_hash.json"), "w") as f:
        json.dump({"input": xs, "output": [int(v) for v in np_pcg_hash(np.array(xs, dtype=np.uint64).astype(np.uint32))]},
                  f, indent=0)
    out, sim = run_scene(O.OracleSim)
    ref, _, _ = sim.solve_fp64(1e-13, 400)
    out["pressure_fp64_reference"] = ref
    np.savez_compressed(os.path.join(HERE, "scene12.npz"), **out)
    print("wrote", sorted(out), len(out["particles"]), "particles")


if __name__ == "__main__":
    main()
