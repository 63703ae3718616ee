"""Batches of independent tanks (slosh_batch_*): every tank of a batch equals its single-tank run bit for bit, for
any worker count, solver chunk and graph setting, in both modes (each tank on its own stream; the MGPCG iterations
of all tanks in lockstep, each pass one launch over every tank) and for any workgroup shape of the level-0 passes;
a tank whose solve cannot join the lockstep (RB-SOR) solves on its own stream; the autotuner (which picks among
those settings) only changes time."""
import numpy as np
import pytest

from slosh_amd import abi, scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 240.0


def tanks():
    """Three tanks of uneven size and shape with their gravities."""
    out = []
    grid, h, reqs = scenes.dam_break(32, seed=3)
    out.append((grid, h, reqs, (0.0, -9.8, 0.0)))
    grid, h, reqs = scenes.sloshing_tank(64, seed=5)
    out.append((grid, h, reqs, (-1.5, -9.8, 0.0)))
    grid = (40, 24, 32)
    h = 1.0 / 40
    cmin, dims = scenes.snap_box((0.0, 0.0, 0.0), (0.5, 0.4, 0.8), grid, h)
    out.append((grid, h, [abi.fill_request(cmin, dims, 9)], (0.5, -9.8, 0.3)))
    return out


def make(spec, solver=abi.SOLVER_MGPCG):
    from slosh_amd import Sim

    grid, h, reqs, _ = spec
    cap = sum(int(np.prod(r.cell_dims)) * 8 for r in reqs) + 1024
    sim = Sim(abi.make_settings(grid, h, cap, solver=solver, tolerance=1e-6))
    for r in reqs:
        sim.emit(r)
    return sim


def state(sim):
    return sim.read_particles().tobytes(), sim.read_grid(abi.FIELD_PRESSURE_F64).tobytes()


def single_runs(specs, steps):
    out = []
    for spec in specs:
        sim = make(spec)
        for _ in range(steps):
            sim.step(DT, spec[3])
        out.append(state(sim))
        sim.close()
    return out


@pytest.mark.parametrize("workers,chunk,graphs,mode,col_bricks",
                         [(3, 4, 1, 0, 4), (2, 8, 0, 0, 4), (1, 2, 1, 0, 4), (2, 4, 1, 1, 4), (1, 8, 0, 1, 1),
                          (3, 2, 1, 1, 2)])
def test_batch_equals_single_tank_runs(workers, chunk, graphs, mode, col_bricks):
    from slosh_amd import Batch

    specs = tanks()
    ref = single_runs(specs, 3)
    sims = [make(s) for s in specs]
    b = Batch(sims)
    print(f"case batch of 3 tanks, workers {workers} chunk {chunk} graphs {graphs} mode {mode} "
          f"bricks/workgroup {col_bricks}", flush=True)
    b.set_config(workers, chunk, graphs, mode, col_bricks)
    grav = np.array([s[3] for s in specs], np.float32)
    for _ in range(3):
        b.step(DT, grav)
    for k, sim in enumerate(sims):
        assert state(sim) == ref[k], k
        assert sim.results().converged == 1
    b.close()
    for s in sims:
        s.close()


def test_autotuner_changes_time_not_results():
    from slosh_amd import Batch

    specs = tanks()
    sims = [make(s) for s in specs]
    b = Batch(sims)
    grav = np.array([s[3] for s in specs], np.float32)
    print("case autotuner over a batch of 3 tanks", flush=True)
    chosen = b.autotune(DT, grav, steps_per_candidate=1)
    assert chosen == b.config() and chosen["workers"] in (1, 2)
    b.step(DT, grav)
    # 14 candidates of one step each (own streams: workers 1, 2 x chunk 4, 8 x graphs off, on; lockstep: workers 2 x
    # chunk 4, 8 x 1, 2, 4 bricks per workgroup), then one step
    ref = single_runs(specs, 15)
    for k, sim in enumerate(sims):
        assert state(sim) == ref[k], k
    b.close()
    for s in sims:
        s.close()


def test_lockstep_batch_with_a_tank_outside_the_lockstep():
    from slosh_amd import Batch

    specs = tanks()
    solvers = [abi.SOLVER_MGPCG, abi.SOLVER_RBSOR, abi.SOLVER_MGPCG]
    print("case lockstep batch, the second tank RB-SOR", flush=True)
    ref = []
    for spec, sv in zip(specs, solvers):
        sim = make(spec, sv)
        for _ in range(2):
            sim.step(DT, spec[3])
        ref.append(state(sim))
        sim.close()
    sims = [make(spec, sv) for spec, sv in zip(specs, solvers)]
    b = Batch(sims)
    b.set_config(2, 4, 1, 1, 4)
    grav = np.array([s[3] for s in specs], np.float32)
    for _ in range(2):
        b.step(DT, grav)
    for k, sim in enumerate(sims):
        assert state(sim) == ref[k], k
        assert sim.results().converged == 1
    b.close()
    for s in sims:
        s.close()


def test_batch_config_is_validated():
    from slosh_amd import Batch, SloshError

    print("case batch configuration checks", flush=True)
    specs = tanks()[:2]
    sims = [make(s) for s in specs]
    b = Batch(sims)
    for bad in [(2, 4, 1, 2, 4), (2, 4, 1, 1, 3), (2, 3, 1, 1, 4), (0, 4, 1, 1, 4)]:
        with pytest.raises(SloshError):
            b.set_config(*bad)
    b.set_config(2, 4, 1, 1, 2)
    assert b.config() == {"workers": 2, "solver_chunk": 4, "graphs": 1, "mode": 1, "col_bricks": 2}
    b.close()
    for s in sims:
        s.close()
