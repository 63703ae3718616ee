"""What the trained network gains: on a held-out sloshing tank (a scene kind and seed not used in training), flexible
CG with the hybrid preconditioner M(r) = V(r) + N(r - A V(r)) and the shipped weights
(slosh_amd/weights/unet_slosh.npz) finishes every solve itself (no MGPCG fallback) and takes fewer iterations than
MGPCG with the same V-cycle (fluid-first coarse types, three coarse red-black pairs) to the same FP64 true residual.
One application of the network also removes most of the residual the V-cycle leaves: on the captured residuals
r' = r - A V(r) of the solves, min over s of ||r' - s A N(r')|| / ||r'|| stays well below 1."""
import os

import numpy as np
import pytest

import mg_numpy as M
from slosh_amd import Sim, abi, scenes, unet

pytestmark = pytest.mark.gpu

MG = dict(flags=abi.FLAG_MG_FLUID_COARSENING, mg_coarse_sweeps=3)


def tank(solver, flags=0):
    grid, h, reqs = scenes.sloshing_tank(64, seed=7)
    s = abi.make_settings(grid, h, scenes.particle_count(reqs), solver=solver, tolerance=1e-6,
                          flags=MG["flags"] | flags, mg_coarse_sweeps=MG["mg_coarse_sweeps"])
    sim = Sim(s)
    for r in reqs:
        sim.emit(r)
    return sim


def test_hybrid_preconditioner_takes_fewer_iterations_than_mgpcg():
    wpath = unet.default_weights_path()
    assert os.path.exists(wpath)
    print("case held-out sloshing tank 64 seed 7, FCG vs MGPCG", flush=True)
    fcg, mg = tank(abi.SOLVER_FCG_UNET), tank(abi.SOLVER_MGPCG)
    fcg.unet_set_weights(unet.load(wpath))
    it_f, it_m = [], []
    t = 0.0
    for _ in range(8):
        g = scenes.tank_gravity(t)
        fcg.step(1.0 / 240.0, g)
        mg.step(1.0 / 240.0, g)
        t += 1.0 / 240.0
        rf, rm = fcg.results(), mg.results()
        assert rf.fallback_used == 0 and rf.converged == 1 and rf.true_residual_rel <= 1e-6, rf.as_dict()
        assert rm.converged == 1
        it_f.append(rf.solver_iterations)
        it_m.append(rm.solver_iterations)
    print("flexible CG", it_f, "MGPCG", it_m, flush=True)
    assert np.mean(it_f) < np.mean(it_m), (it_f, it_m)
    fcg.close()
    mg.close()


def scale_free_residual(A, r, z):
    Az = A.apply(z)
    c = float((r * Az).sum()) ** 2 / (float((r * r).sum()) * float((Az * Az).sum()))
    return float(np.sqrt(max(1.0 - c, 0.0)))


def test_network_removes_most_of_what_the_vcycle_leaves():
    wpath = unet.default_weights_path()
    print("case held-out sloshing tank 64 seed 7, captured V-cycle residuals", flush=True)
    sim = tank(abi.SOLVER_MGPCG, abi.FLAG_CAPTURE_RESIDUALS | abi.FLAG_CAPTURE_VCYCLE_RESIDUALS)
    t = 0.0
    for _ in range(4):
        sim.step(1.0 / 240.0, scenes.tank_gravity(t))
        t += 1.0 / 240.0
    sim.unet_set_weights(unet.load(wpath))
    t0 = (sim.read_grid(abi.FIELD_CELL_FLAGS) & 3).astype(np.int32)
    A = M.Level(t0, sim.read_grid(abi.FIELD_CELL_DIAG).astype(np.float64))
    _, n = sim.captured_residual(0)
    assert n >= 4
    res = []
    for k in range(n):
        r32, _ = sim.captured_residual(k)
        r = np.where(A.fluid, r32.astype(np.float64), 0.0)
        res.append(scale_free_residual(A, r, sim.unet_apply(r32).astype(np.float64)))
    sim.close()
    print("scale-free residual after the network", np.round(res, 3), flush=True)
    assert np.mean(res) < 0.6 and max(res) < 0.9, res
