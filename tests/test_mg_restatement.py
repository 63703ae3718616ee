"""The oracle's MGPCG against an independent numpy restatement of the V-cycle (tests/mg_numpy.py) on the system of a
real step: the same number of PCG iterations to 1e-6 (within 1: the oracle's cycle is FP32, the restatement FP64)
for the reference cycle and for each preconditioner option (fluid-first coarse types, extra red-black pairs on
levels 4 and coarser). A 128 x 64 x 64 dam break has levels 3, 4, 5 of 1024, 128 and 16 cells, so level 4 takes the
extra pairs there; on a shallow sloshing tank the options remove several iterations in both restatements alike."""
import numpy as np
import pytest

import mg_numpy as M
import oracle_py as O
from slosh_amd import abi, scenes

OPTIONS = [(0, 0), (abi.FLAG_MG_FLUID_COARSENING, 0), (abi.FLAG_MG_FLUID_COARSENING, 3), (0, 2)]


def system(grid, h, reqs, dt, g, flags, sweeps):
    cap = sum(int(np.prod(r.cell_dims)) * 8 for r in reqs)
    s = abi.make_settings(grid, h, cap, solver=abi.SOLVER_MGPCG, tolerance=1e-6, flags=flags, mg_coarse_sweeps=sweeps)
    o = O.OracleSim(s)
    for r in reqs:
        o.emit(r)
    o.step(dt, g)
    t0 = (o.read_grid(abi.FIELD_CELL_FLAGS) & 3).astype(np.int32)
    diag = o.read_grid(abi.FIELD_CELL_DIAG).astype(np.float64)
    b = np.where(t0 == M.FLUID, -o.read_grid(abi.FIELD_DIVERGENCE).astype(np.float64), 0.0)
    res = o.results()
    p = o.read_grid(abi.FIELD_PRESSURE_F64)
    o.close()
    return t0, diag, b, res, p


@pytest.mark.parametrize("flags,sweeps", OPTIONS)
def test_oracle_mgpcg_matches_numpy_vcycle(flags, sweeps):
    grid, h = (128, 64, 64), 1.0 / 64
    cmin, dims = scenes.snap_box((0.0, 0.0, 0.0), (1.1, 0.55, 1.0), grid, h)
    t0, diag, b, res, p = system(grid, h, [abi.fill_request(cmin, dims, 5)], 1.0 / 120.0, (0.0, -9.8, 0.0),
                                 flags, sweeps)
    it, x = M.mgpcg_iterations(t0, diag, b, fluid_first=bool(flags & abi.FLAG_MG_FLUID_COARSENING),
                               coarse_sweeps=max(sweeps, 1))
    assert res.converged == 1 and abs(res.solver_iterations - it) <= 1, (res.solver_iterations, it)
    # both stop at a 1e-6 relative residual: their solutions differ by the sum of two such errors
    assert np.linalg.norm(p - x) / np.linalg.norm(x) <= 2e-4


def test_options_change_the_shallow_tank_iterations():
    """On a shallow sloshing tank (the free surface dominates the coarse levels) the options cut the iteration
    count in both restatements alike, and the reference rule with one pair is the slowest."""
    grid, h, reqs = scenes.sloshing_tank(128, seed=1)
    counts = {}
    for flags, sweeps in [(0, 0), (abi.FLAG_MG_FLUID_COARSENING, 3)]:
        t0, diag, b, res, _ = system(grid, h, reqs, 1.0 / 240.0, scenes.tank_gravity(0.0), flags, sweeps)
        it, _ = M.mgpcg_iterations(t0, diag, b, fluid_first=bool(flags), coarse_sweeps=max(sweeps, 1))
        assert abs(res.solver_iterations - it) <= 1, (flags, sweeps, res.solver_iterations, it)
        counts[(flags, sweeps)] = it
    assert counts[(abi.FLAG_MG_FLUID_COARSENING, 3)] < counts[(0, 0)], counts
