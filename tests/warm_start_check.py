"""Shared check of the extrapolated warm start (SLOSH_FLAG_WARM_START_EXTRAPOLATE) for the oracle (CPU test) and the
library (GPU test): a dam break stepped with a different dt each step, then one more step stopped after P2G, whose
FP64 pressure field is the solver's starting guess. On a cell fluid in that step and the last two full steps it must
equal dt3 (q1 + (q1 - q2) dt3/dt2) with q1 = P2/dt2, q2 = P1/dt1 (P1, P2 the last two returned solutions); on any
other fluid cell the reference's P2 dt3/dt2; elsewhere 0. The expected values are formed in float64 with the same
operations, so the comparison is bit for bit; both branches are exercised (the fluid moves)."""
import numpy as np

from slosh_amd import abi, scenes

DTS = (1.0 / 120.0, 1.0 / 150.0, 1.0 / 110.0, 1.0 / 130.0, 1.0 / 100.0, 1.0 / 140.0)
G = (0.0, -9.8, 0.0)


def make_settings(n=32):
    grid, h, reqs = scenes.dam_break(n, seed=0)
    s = abi.make_settings(grid, h, 8 * n ** 3 // 2, solver=abi.SOLVER_MGPCG, tolerance=1e-6,
                          flags=abi.FLAG_WARM_START_EXTRAPOLATE)
    return s, reqs


def run_and_check(sim, reqs):
    for r in reqs:
        sim.emit(r)
    P, T = [], []
    for dt in DTS[:-1]:
        sim.step(dt, G)
        P.append(sim.read_grid(abi.FIELD_PRESSURE_F64).astype(np.float64))
        T.append(sim.read_grid(abi.FIELD_CELL_FLAGS) & 3)
    sim.step(DTS[-1], G, stop_after=abi.STAGE_P2G)
    guess = sim.read_grid(abi.FIELD_PRESSURE_F64)
    t3 = sim.read_grid(abi.FIELD_CELL_FLAGS) & 3
    d1, d2, d3 = (float(np.float32(v)) for v in DTS[-3:])
    warm = float(np.float32(DTS[-1]) / np.float32(DTS[-2]))
    fluid3, fluid2, fluid1 = t3 == abi.CELL_FLUID, T[-1] == abi.CELL_FLUID, T[-2] == abi.CELL_FLUID
    hist = fluid3 & fluid2 & fluid1
    q1 = P[-1] / d2
    q2 = P[-2] / d1
    expect = np.where(fluid3, P[-1] * warm, 0.0)
    expect = np.where(hist, d3 * (q1 + (q1 - q2) * (d3 / d2)), expect)
    assert hist.sum() > 1000 and (fluid3 & ~hist).sum() > 0, (hist.sum(), (fluid3 & ~hist).sum())
    assert np.array_equal(guess.view(np.uint64), expect.view(np.uint64)), \
        f"{np.count_nonzero(guess != expect)} cells differ"
    return int(hist.sum()), int((fluid3 & ~hist).sum())
