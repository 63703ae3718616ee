"""What the trained network gains: on residuals that MGPCG meets in the solves of a held-out scene (a sloshing tank
with a seed not used in training), one application of the shipped network (slosh_amd/weights/unet_slosh.npz)
removes more of the residual than the Jacobi preconditioner z = r / diag. The measure is scale-free, as flexible CG
is (its step length rescales z): min over s of ||r - s A z|| / ||r|| = sqrt(1 - (r.Az)^2 / (|r|^2 |Az|^2)), with A
assembled in numpy from the step's cell types and diagonal."""
import os

import numpy as np
import pytest

import mg_numpy as M
from slosh_amd import Sim, abi, scenes, unet

pytestmark = pytest.mark.gpu


def scale_free_residual(A, r, z):
    Az = A.apply(z)
    c = float((r * Az).sum()) ** 2 / (float((r * r).sum()) * float((Az * Az).sum()))
    return float(np.sqrt(max(1.0 - c, 0.0)))


def test_trained_network_beats_jacobi_on_solver_residuals():
    wpath = unet.default_weights_path()
    assert os.path.exists(wpath)
    print("case sloshing tank 64 seed 7, captured residuals", flush=True)
    grid, h, reqs = scenes.sloshing_tank(64, seed=7)
    s = abi.make_settings(grid, h, scenes.particle_count(reqs), solver=abi.SOLVER_MGPCG, tolerance=1e-6,
                          flags=abi.FLAG_CAPTURE_RESIDUALS)
    sim = Sim(s)
    for r in reqs:
        sim.emit(r)
    t = 0.0
    for _ in range(4):
        sim.step(1.0 / 240.0, scenes.tank_gravity(t))
        t += 1.0 / 240.0
    sim.unet_set_weights(unet.load(wpath))
    t0 = (sim.read_grid(abi.FIELD_CELL_FLAGS) & 3).astype(np.int32)
    diag = sim.read_grid(abi.FIELD_CELL_DIAG).astype(np.float64)
    A = M.Level(t0, diag)
    _, n = sim.captured_residual(0)
    assert n >= 4
    net, jac = [], []
    for k in range(n):
        r32, _ = sim.captured_residual(k)
        r = np.where(A.fluid, r32.astype(np.float64), 0.0)
        z_net = sim.unet_apply(r32).astype(np.float64)
        z_jac = np.where(A.fluid, r / A.dg, 0.0)
        net.append(scale_free_residual(A, r, z_net))
        jac.append(scale_free_residual(A, r, z_jac))
    sim.close()
    print("network", np.round(net, 3), "jacobi", np.round(jac, 3), flush=True)
    assert np.mean(net) < np.mean(jac), (net, jac)
    assert all(v < 1.0 for v in net)
