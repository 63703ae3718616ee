"""Training set of the learned preconditioner (slosh_amd/training_set.py): 64^3 brick windows of seeded dam-break
and sloshing-tank steps, and slosh_load_system / slosh_solve, which put such a system into a context and solve it.

CPU (oracle-driven generator):
- every system's matrix is symmetric with a positive diagonal and strictly diagonally dominant (so SPD), and a
  dense Cholesky factorisation of sub-windows succeeds;
- the window's right-hand side is the oracle's divergence restricted to the window's fluid cells, zero elsewhere;
- shards read back exactly what was written.
GPU (marked gpu):
- the windows the library's steps give equal the oracle's: bit for bit at the first step (same particles), within
  the solve's tolerance after it;
- a loaded whole-grid window is the step's own system: slosh_solve returns the step's pressure bit for bit;
- loaded windows (dam break, sloshing tank) solve to ||b - A p|| / ||b|| <= 1e-6 with MGPCG, RB-SOR and the learned
  flexible CG (no fallback), and at 1e-7 within 1e-5 relative L2 of scipy's sparse direct solve of the same matrix.
"""
import numpy as np
import pytest

import oracle_py as O
from slosh_amd import abi, scenes
from slosh_amd import training_set as T


@pytest.fixture(scope="module")
def dam_systems():
    return list(T.dam_break_systems(O.OracleSim, seed=0, steps=3))


@pytest.fixture(scope="module")
def tank_systems():
    return list(T.sloshing_systems(O.OracleSim, seed=1, steps=1, every=1, n=128))


def check_spd(sysm):
    A, b = T.matrix(sysm)
    assert A.shape[0] > 0 and np.all(np.isfinite(b))
    assert abs(A - A.T).max() == 0.0
    d = A.diagonal()
    off = np.asarray(abs(A).sum(axis=1)).ravel() - np.abs(d)
    # strictly dominant by the 1e-4 term (an FP32 sum: for diagonals below 120 it keeps more than 0.9e-4)
    assert d.min() > 0.0 and (d - off).min() >= 0.9e-4


def test_dam_break_windows_are_spd_and_cover_the_grid(dam_systems):
    assert len(dam_systems) == 3 and [s["step"] for s in dam_systems] == [1, 2, 3]
    for sysm in dam_systems:
        assert sysm["types"].shape == (64, 64, 64) and sysm["origin"] == (0, 0, 0)
        check_spd(sysm)


def test_tank_windows_are_spd_and_hold_fluid(tank_systems):
    # a 128^3 tank filled to about half: the 8 windows of 64^3 include the lower four, each with fluid
    assert len(tank_systems) >= 4
    for sysm in tank_systems:
        assert int((sysm["types"] == 1).sum()) >= 512
        check_spd(sysm)


def test_dense_cholesky_of_sub_windows(dam_systems):
    sysm = dam_systems[1]
    t = sysm["types"]
    for z0, y0, x0 in [(8, 8, 8), (24, 16, 0), (40, 24, 16)]:
        sub = {k: sysm[k][z0:z0 + 12, y0:y0 + 12, x0:x0 + 12] for k in ("types", "diag", "divergence")}
        if int((sub["types"] == 1).sum()) < 10:
            continue
        A, _ = T.matrix(sub)
        np.linalg.cholesky(A.toarray())  # raises unless SPD
    assert int((t == 1).sum()) > 10000


def test_rhs_is_the_oracle_divergence():
    o = O.OracleSim(abi.make_settings((64, 64, 64), 1.0 / 64, 600000, solver=abi.SOLVER_MGPCG, tolerance=1e-6))
    grid, h, reqs = scenes.dam_break(64, seed=5)
    for r in reqs:
        o.emit(r)
    o.step(T.DAM_DT, (0.0, -9.8, 0.0))
    (sysm,) = T.window_systems(o, "dam_break_64", 5, 1)
    div = o.read_grid(abi.FIELD_DIVERGENCE)
    diag = o.read_grid(abi.FIELD_CELL_DIAG)
    flags = o.read_grid(abi.FIELD_CELL_FLAGS)
    fluid = (flags & 3) == 1
    assert np.array_equal(sysm["types"], (flags & 3).astype(np.uint8))
    assert np.array_equal(sysm["divergence"][fluid], div[fluid]) and not sysm["divergence"][~fluid].any()
    assert np.array_equal(sysm["diag"][fluid], diag[fluid])
    o.close()


def test_shards_round_trip(tmp_path, dam_systems):
    count, nbytes = T.write_shards(str(tmp_path), iter(dam_systems), shard=2)
    assert count == 3 and nbytes > 0 and len(list(tmp_path.iterdir())) == 2
    back = list(T.read_shards(str(tmp_path)))
    for a, b in zip(dam_systems, back):
        for k in ("types", "diag", "divergence"):
            assert np.array_equal(a[k], b[k])
        assert (a["scene"], a["seed"], a["step"], tuple(a["origin"])) == (b["scene"], b["seed"], b["step"],
                                                                           b["origin"])


def test_iter_systems_mixes_the_scenes():
    # dam breaks first (seeds 0, 1, ...), then sloshing tanks; a 64^3 tank keeps the test small
    got = list(T.iter_systems(O.OracleSim, 4, dam_fraction=0.5, dam_steps=1, tank_steps=1, tank_every=1, tank_n=64))
    assert [s["scene"] for s in got] == ["dam_break_64", "dam_break_64", "sloshing_tank_64", "sloshing_tank_64"]
    assert [s["seed"] for s in got[:2]] == [0, 1]


# ---- GPU --------------------------------------------------------------------------------------------------------
@pytest.mark.gpu
def test_gpu_windows_equal_the_oracle():
    from slosh_amd import Sim

    print("case training-set windows, dam break 64 seed 3, steps 1-3", flush=True)
    g = list(T.dam_break_systems(Sim, seed=3, steps=3))
    o = list(T.dam_break_systems(O.OracleSim, seed=3, steps=3))
    assert len(g) == len(o) == 3
    for k, (a, b) in enumerate(zip(g, o)):
        assert np.array_equal(a["types"], b["types"]), k
        if k == 0:
            assert np.array_equal(a["divergence"], b["divergence"]) and np.array_equal(a["diag"], b["diag"])
        else:
            scale = float(np.abs(b["divergence"]).max())
            assert float(np.abs(a["divergence"] - b["divergence"]).max()) <= 1e-4 * scale, k
            assert float(np.abs(a["diag"] - b["diag"]).max()) <= 1e-4, k


@pytest.mark.gpu
def test_gpu_loaded_window_is_the_steps_own_system():
    from slosh_amd import Sim

    print("case load_system of a whole-grid window = the step's system", flush=True)
    grid, h, reqs = scenes.dam_break(64, seed=2)
    s = abi.make_settings(grid, h, scenes.particle_count(reqs) + 16, solver=abi.SOLVER_MGPCG, tolerance=1e-6)
    sim = Sim(s)
    for r in reqs:
        sim.emit(r)
    sim.step(T.DAM_DT, (0.0, -9.8, 0.0))  # the first step starts its solve from p = 0, as slosh_solve does
    p_step = sim.read_grid(abi.FIELD_PRESSURE_F64)
    it_step = sim.results().solver_iterations
    (sysm,) = T.window_systems(sim, "dam_break_64", 2, 1)
    sim.close()
    w = Sim(T.window_settings())
    w.load_system(sysm["types"], sysm["diag"], sysm["divergence"])
    res = w.solve()
    p = w.read_grid(abi.FIELD_PRESSURE_F64)
    assert res.solver_iterations == it_step and res.converged == 1
    assert np.array_equal(p, p_step)
    w.close()


def tank_window_gpu():
    from slosh_amd import Sim

    grid, h, reqs = scenes.sloshing_tank(256, seed=1)
    s = abi.make_settings(grid, h, scenes.particle_count(reqs) + 16, solver=abi.SOLVER_MGPCG, tolerance=1e-6,
                          flags=abi.FLAG_MG_FLUID_COARSENING, mg_coarse_sweeps=3)
    sim = Sim(s)
    for r in reqs:
        sim.emit(r)
    for k in range(3):
        sim.step(T.TANK_DT, scenes.tank_gravity(k * T.TANK_DT))
    wins = T.window_systems(sim, "sloshing_tank_256", 1, 3)
    sim.close()
    # a window at the free surface (fluid, air and a wall)
    return max(wins, key=lambda w: int((w["types"] == 0).sum()) * int((w["types"] == 1).sum()))


@pytest.fixture(scope="module")
def loaded_windows():
    """The two windows of the solve tests (a dam break's whole grid, a sloshing tank's free-surface window), each
    with scipy's sparse direct solve of its matrix (computed once for the three solvers)."""
    import scipy.sparse.linalg as spla

    from slosh_amd import Sim

    dam = next(T.dam_break_systems(Sim, seed=4, steps=2, every=2))
    out = []
    for sysm in (dam, tank_window_gpu()):
        A, b = T.matrix(sysm)
        out.append((sysm, spla.spsolve(A.tocsc(), b)))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("solver", ["mgpcg", "rbsor", "fcg"])
def test_gpu_loaded_windows_solve_to_tolerance(solver, loaded_windows):
    from slosh_amd import Sim, unet

    print(f"case loaded windows, {solver}", flush=True)
    for sysm, ref in loaded_windows:
        # at 1e-6 every accepted solve meets the true residual; a solve from p = 0 to 1e-6 can still be 2e-5 from
        # the exact solution on these systems (condition number), so the solution is checked at 1e-7
        for tol in (1e-6, 1e-7):
            kind = {"mgpcg": abi.SOLVER_MGPCG, "rbsor": abi.SOLVER_RBSOR, "fcg": abi.SOLVER_FCG_UNET}[solver]
            s = T.window_settings(solver=kind, flags=abi.FLAG_MG_FLUID_COARSENING, mg_coarse_sweeps=3, tolerance=tol)
            s.max_iterations = 8000 if solver == "rbsor" else 300
            sim = Sim(s)
            if solver == "fcg":
                sim.unet_set_weights(unet.load(unet.default_weights_path()))
            sim.load_system(sysm["types"], sysm["diag"], sysm["divergence"])
            res = sim.solve()
            assert res.converged == 1 and res.true_residual_rel <= tol, (sysm["scene"], tol, res.as_dict())
            if solver == "fcg":
                assert res.fallback_used == 0, res.as_dict()
            p = sim.read_grid(abi.FIELD_PRESSURE_F64)[sysm["types"] == 1]
            err = float(np.linalg.norm(p - ref) / np.linalg.norm(ref))
            print(sysm["scene"], solver, tol, res.solver_iterations, "iterations, error", f"{err:.2e}", flush=True)
            if tol < 1e-6:
                assert err <= 1e-5, err
            sim.close()
