"""Two-sided density projection (Slosh.usf:743-854, driver SloshProxy.cpp:656-702) and the velocity density
correction (Slosh.usf:708-727).

CPU pins of the oracle, each against what the mathematics fixes:
- a regular lattice of 8 particles per cell has rest density exactly: zero right-hand side, potential and
  displacement;
- the right-hand side equals an independent numpy restatement of CellDensityInterior and the air rule;
- the face displacement is -(q_b - q_a), the negative face gradient of the potential;
- with many RB-SOR sweeps the potential solves A q = -rhs with the pressure operator (scipy sparse solve);
- a potential linear in x shifts interior particles by -slope cells along x, shortened to half a cell.
GPU parity (marked gpu): the right-hand side, potential and displacement bit for bit with a written potential,
the particles after the step bit for bit; the potential of the MGPCG and RB-SOR solves against the oracle's; the
velocity density correction (density_correction > 0) bit for bit in the divergence.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle_py as O
from slosh_amd import abi, scenes

G = (0.0, -9.8, 0.0)
DT = 1.0 / 120.0
FLAG = abi.FLAG_DENSITY_PROJECTION


def lattice(n, lo, hi):
    """8 particles per cell at the centres of the 2x2x2 sub-cells of cells lo..hi-1 (inclusive-exclusive), at rest."""
    h = 1.0 / n
    ax = [np.arange(2 * lo[a], 2 * hi[a]) * 0.5 + 0.25 for a in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    p = np.zeros(X.size, dtype=abi.PARTICLE_DTYPE)
    p["pos"] = (np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) * h).astype(np.float32)
    return p


def stepped(n, parts, solver, stop=abi.STAGE_ALL, flags=FLAG, **kw):
    s = abi.make_settings((n, n, n), 1.0 / n, len(parts) + 16, solver=solver, flags=flags, **kw)
    o = O.OracleSim(s)
    o.write_particles(parts)
    o.step(DT, (0.0, 0.0, 0.0), stop)
    return o


def test_rest_density_lattice_projects_to_nothing():
    n = 24
    o = stepped(n, lattice(n, (2, 2, 2), (20, 14, 20)), abi.SOLVER_MGPCG, stop=abi.STAGE_PROJECT)
    fluid = (o.read_grid(abi.FIELD_CELL_FLAGS) & 3) == 1
    assert fluid.sum() > 2000
    assert np.all(o.read_grid(abi.FIELD_DENSITY_RHS) == 0.0)
    assert np.all(o.read_grid(abi.FIELD_DENSITY_Q) == 0.0)
    assert np.all(o.read_grid(abi.FIELD_DENSITY_DISP) == 0.0)


def np_rhs(flags, fw, relax):
    """CellDensityInterior (Slosh.usf:753-773) and DensityProjectionRhsCS (:775-805) restated on dense arrays."""
    t = (flags & 3).astype(np.int64)
    nz, ny, nx = t.shape
    tp = np.full((nz + 2, ny + 2, nx + 2), 2, np.int64)
    tp[1:-1, 1:-1, 1:-1] = t
    out = np.zeros(t.shape, np.float32)
    for z, y, x in zip(*np.nonzero(t == 1)):
        ws, faces = np.float32(0.0), np.float32(0.0)
        air = False
        for a in range(3):
            d = [0, 0, 0]
            d[a] = 1
            lo = tp[z + 1 - d[2], y + 1 - d[1], x + 1 - d[0]]
            hi = tp[z + 1 + d[2], y + 1 + d[1], x + 1 + d[0]]
            if lo != 2:
                ws = np.float32(ws + fw[a, z, y, x])
                faces = np.float32(faces + 1)
            if hi != 2:
                ws = np.float32(ws + fw[a, z + d[2], y + d[1], x + d[0]])
                faces = np.float32(faces + 1)
            air = air or lo == 0 or hi == 0
        rho = np.float32(ws / np.float32(faces * np.float32(8.0))) if faces > 0 else np.float32(1.0)
        dev = np.float32(rho - np.float32(1.0))
        if air:
            dev = max(dev, np.float32(0.0))
        out[z, y, x] = np.float32(-np.float32(relax) * np.float32(min(max(dev, -0.5), 0.5)))
    return out


def compressed_scene(n=20, seed=3):
    """A block of fluid with an over-dense core (12 particles per cell there) and a low-density corner."""
    h = 1.0 / n
    rng = np.random.default_rng(seed)
    base = lattice(n, (2, 2, 2), (17, 12, 17))
    core = lattice(n, (7, 4, 7), (12, 9, 12))
    core["pos"] += rng.uniform(-0.2, 0.2, size=core["pos"].shape).astype(np.float32) * h
    keep = rng.random(len(base)) > 0.15 * ((base["pos"][:, 0] > 14 * h) & (base["pos"][:, 2] > 14 * h))
    p = np.concatenate([base[keep], core[: len(core) // 2]])
    p["age"] = 0.0
    return p


def test_rhs_matches_numpy_restatement():
    n = 20
    o = stepped(n, compressed_scene(n), abi.SOLVER_NONE, stop=abi.STAGE_PROJECT, density_relaxation=0.7)
    flags = o.read_grid(abi.FIELD_CELL_FLAGS)
    fw = o.read_grid(abi.FIELD_FACE_WEIGHT)
    rhs = o.read_grid(abi.FIELD_DENSITY_RHS)
    ref = np_rhs(flags, fw, 0.7)
    assert np.count_nonzero(ref) > 100 and ref.min() < 0.0
    assert np.array_equal(rhs.view(np.uint32), ref.view(np.uint32))


def test_displacement_is_the_negative_face_gradient_of_q():
    n = 20
    parts = compressed_scene(n)
    s = abi.make_settings((n, n, n), 1.0 / n, len(parts) + 16, solver=abi.SOLVER_NONE, flags=FLAG)
    o = O.OracleSim(s)
    o.write_particles(parts)
    o.step(DT, (0.0, 0.0, 0.0))  # allocates bricks
    q = np.random.default_rng(1).uniform(-0.3, 0.3, size=(n, n, n)).astype(np.float32)
    o.write_density_q(q)
    o.step(DT, (0.0, 0.0, 0.0), abi.STAGE_PROJECT)
    t = o.read_grid(abi.FIELD_CELL_FLAGS) & 3
    qk = np.where(t == 1, o.read_grid(abi.FIELD_DENSITY_Q).astype(np.float64), 0.0)
    disp = o.read_grid(abi.FIELD_DENSITY_DISP)
    checked = 0
    for a, axis in ((0, 2), (1, 1), (2, 0)):
        sl_b = [slice(1, None)] * 3
        sl_a = [slice(1, None)] * 3
        sl_a[axis] = slice(0, -1)
        tb, ta = t[tuple(sl_b)], t[tuple(sl_a)]
        ref = (-(qk[tuple(sl_b)] - qk[tuple(sl_a)])).astype(np.float32)
        ref[(ta == 2) | (tb == 2)] = 0.0
        got = disp[a][tuple(sl_b)]
        assert np.array_equal(got, ref), a
        checked += int(np.count_nonzero(ref))
    assert checked > 1000


def test_converged_potential_solves_the_pressure_operator():
    n = 16
    parts = compressed_scene(n, seed=5)
    o = stepped(n, parts, abi.SOLVER_RBSOR, stop=abi.STAGE_PROJECT, density_iterations=4000, max_iterations=4000)
    t = o.read_grid(abi.FIELD_CELL_FLAGS) & 3
    fluid = t == 1
    diag = o.read_grid(abi.FIELD_CELL_DIAG)
    rhs = o.read_grid(abi.FIELD_DENSITY_RHS).astype(np.float64)
    q = o.read_grid(abi.FIELD_DENSITY_Q).astype(np.float64)
    idx = -np.ones(t.shape, np.int64)
    idx[fluid] = np.arange(fluid.sum())
    rows, cols, vals = [], [], []
    for z, y, x in zip(*np.nonzero(fluid)):
        i = idx[z, y, x]
        rows.append(i)
        cols.append(i)
        vals.append(float(np.float32(diag[z, y, x]) + np.float32(1e-4)))
        for dz, dy, dx in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            zz, yy, xx = z + dz, y + dy, x + dx
            if fluid[zz, yy, xx]:
                rows.append(i)
                cols.append(idx[zz, yy, xx])
                vals.append(-1.0)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(fluid.sum(), fluid.sum()))
    ref = spla.spsolve(A, -rhs[fluid])
    assert np.abs(ref).max() > 1e-3
    assert np.linalg.norm(q[fluid] - ref) <= 1e-5 * np.linalg.norm(ref)


@pytest.mark.parametrize("slope", [0.3, -0.2, 0.9])
def test_linear_potential_shifts_particles_by_its_gradient(slope):
    """q = slope * x (cells) on fluid cells: interior faces carry -slope along x, so interior particles move by
    -slope cells in x on top of their advection, at most half a cell (SampleDensityDisplacement's limit)."""
    n = 24
    h = 1.0 / n
    parts = lattice(n, (2, 2, 2), (22, 12, 22))
    parts["vel"][:, 1] = -0.05
    runs = []
    for flags in (FLAG, 0):
        s = abi.make_settings((n, n, n), h, len(parts) + 16, solver=abi.SOLVER_NONE, flags=flags)
        o = O.OracleSim(s)
        o.write_particles(parts)
        o.step(DT, (0.0, 0.0, 0.0))  # allocates the bricks the potential is written to
        if flags:
            xs = np.arange(n, dtype=np.float32)
            o.write_density_q(np.broadcast_to(slope * xs[None, None, :], (n, n, n)).astype(np.float32))
        o.step(DT, (0.0, 0.0, 0.0))
        runs.append(o.read_particles_brick_order())
    with_p, without = runs
    # the records keep their emission order within a brick; match them by their positions without the shift
    assert len(with_p) == len(without)
    key = np.lexsort(without["pos"].T)
    pos0 = without["pos"][key]
    keyw = np.lexsort((with_p["pos"] - np.array([-min(abs(slope), 0.5) * np.sign(slope) * h, 0, 0],
                                                np.float32)).T)
    pos1 = with_p["pos"][keyw]
    inner = np.all((pos0 > 5 * h) & (pos0 < [17 * h, 8 * h, 17 * h]), axis=1)
    assert inner.sum() > 1000
    d = pos1[inner].astype(np.float64) - pos0[inner]
    expect = -min(abs(slope), 0.5) * np.sign(slope) * h
    assert np.allclose(d[:, 0], expect, atol=2e-6), (d[:, 0].min(), d[:, 0].max(), expect)
    assert np.allclose(d[:, 1:], 0.0, atol=2e-6)


# ---- GPU parity --------------------------------------------------------------------------------------------------
def gpu_pair(n, parts, solver, flags=FLAG, **kw):
    from slosh_amd import Sim

    s = abi.make_settings((n, n, n), 1.0 / n, len(parts) + 16, solver=solver, flags=flags, **kw)
    gpu, orc = Sim(s), O.OracleSim(s)
    for sim in (gpu, orc):
        sim.write_particles(parts)
    return gpu, orc


def same_bits(a, b, what):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), what


@pytest.mark.gpu
def test_gpu_density_projection_bitexact_with_a_written_potential():
    n = 20
    print("case density projection bit-exact", flush=True)
    gpu, orc = gpu_pair(n, compressed_scene(n), abi.SOLVER_NONE)
    for sim in (gpu, orc):
        sim.step(DT, G)
    rng = np.random.default_rng(2)
    q = rng.uniform(-0.4, 0.4, size=(n, n, n)).astype(np.float32)
    p = rng.uniform(-0.01, 0.01, size=(n, n, n)).astype(np.float32)
    for sim in (gpu, orc):
        sim.write_density_q(q)
        sim.write_pressure(p)
        sim.step(DT, G, stop_after=abi.STAGE_PROJECT)
    for f in (abi.FIELD_DENSITY_RHS, abi.FIELD_DENSITY_Q, abi.FIELD_DENSITY_DISP):
        same_bits(gpu.read_grid(f), orc.read_grid(f), f"field {f}")
    # a full step: G2P shifts every particle by the interpolated displacement
    gpu2, orc2 = gpu_pair(n, compressed_scene(n), abi.SOLVER_NONE)
    for sim in (gpu2, orc2):
        sim.step(DT, G)
        sim.write_density_q(q)
        sim.write_pressure(p)
        sim.step(DT, G)
    same_bits(gpu2.read_particles(), orc2.read_particles_brick_order(), "particles")


@pytest.mark.gpu
@pytest.mark.parametrize("solver,iters", [(abi.SOLVER_MGPCG, 30), (abi.SOLVER_RBSOR, 30), (abi.SOLVER_RBSOR, 7)])
def test_gpu_density_potential_matches_the_oracle(solver, iters):
    """The density projection's potential: two MGPCG iterations within 1e-5 of the oracle's (FP64 dots summed in
    another order), or `iters` RB-SOR iterations at omega 1.7 from q = 0 bit for bit (the same FP64 updates in the
    same order; an odd count ends in the library's second pressure array and is copied back)."""
    n = 32
    print(f"case density projection solve {solver}, {iters} iterations", flush=True)
    grid, h, reqs = scenes.dam_break(n, seed=4)
    from slosh_amd import Sim

    s = abi.make_settings(grid, h, 8 * n ** 3, solver=solver, flags=FLAG, max_iterations=3000,
                          density_iterations=iters)
    gpu, orc = Sim(s), O.OracleSim(s)
    for sim in (gpu, orc):
        for r in reqs:
            sim.emit(r)
        sim.step(DT, G)
    rg = gpu.results()
    assert rg.converged == 1 and rg.true_residual_rel <= 1e-6, rg.as_dict()
    same_bits(gpu.read_grid(abi.FIELD_DENSITY_RHS), orc.read_grid(abi.FIELD_DENSITY_RHS), "rhs")
    qg, qo = gpu.read_grid(abi.FIELD_DENSITY_Q), orc.read_grid(abi.FIELD_DENSITY_Q)
    assert np.abs(qo).max() > 1e-3
    err = np.linalg.norm(qg - qo) / np.linalg.norm(qo)
    print(f"potential difference {err:.2e}", flush=True)
    if solver == abi.SOLVER_RBSOR:
        same_bits(qg, qo, "potential")
    else:
        assert err <= 1e-5


@pytest.mark.gpu
def test_gpu_velocity_density_correction_bitexact():
    """density_correction > 0 without the projection: the divergence carries the velocity correction term
    min(c * max(rho - 1 - deadband, 0), 0.5) h / dt (Slosh.usf:722-727), bit for bit."""
    n = 20
    print("case velocity density correction", flush=True)
    gpu, orc = gpu_pair(n, compressed_scene(n), abi.SOLVER_NONE, flags=0, density_correction=0.15)
    for sim in (gpu, orc):
        sim.step(DT, G)
        sim.step(DT, G, stop_after=abi.STAGE_DIVERGENCE)
    same_bits(gpu.read_grid(abi.FIELD_DIVERGENCE), orc.read_grid(abi.FIELD_DIVERGENCE), "divergence")
    # the term is active: the same scene without it differs
    ref, _ = gpu_pair(n, compressed_scene(n), abi.SOLVER_NONE, flags=0)
    ref.step(DT, G)
    ref.step(DT, G, stop_after=abi.STAGE_DIVERGENCE)
    assert not np.array_equal(ref.read_grid(abi.FIELD_DIVERGENCE), gpu.read_grid(abi.FIELD_DIVERGENCE))
