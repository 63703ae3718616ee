"""Pins the CPU oracle against what the mathematics fixes, independently of its own code.

The reference has no test suite, golden vectors or runnable CPU path, so each oracle function is checked against
a closed form, a brute-force or vectorised numpy restatement written separately, an invariant, or a textbook
routine (dense LU), chosen so that a dropped term, a wrong sign or index, or a transposed operand fails.
"""
import numpy as np
import pytest

import oracle_py as O
from slosh_amd import abi, scenes

G0 = (0.0, 0.0, 0.0)
G = (0.0, -9.8, 0.0)
DT = 1.0 / 120.0


def np_pcg_hash(v):
    """PCG-RXS-M-XS 32-bit output of one LCG step (Jarzynski & Olano 2020), vectorised in numpy uint32."""
    v = np.asarray(v, dtype=np.uint32)
    state = v * np.uint32(747796405) + np.uint32(2891336453)
    word = ((state >> ((state >> np.uint32(28)) + np.uint32(4))) ^ state) * np.uint32(277803737)
    return (word >> np.uint32(22)) ^ word


def test_pcg_hash_matches_numpy_and_is_bijective():
    xs = np.arange(0, 1 << 16, dtype=np.uint32) * np.uint32(65537) + np.uint32(12345)
    ref = np_pcg_hash(xs)
    got = np.array([O.pcg_hash(int(x)) for x in xs[:4096]], dtype=np.uint32)
    assert np.array_equal(got, ref[:4096])
    # the output permutation of a PCG step is a bijection of uint32: no collisions on a dense block of inputs
    blk = np_pcg_hash(np.arange(1 << 20, dtype=np.uint32))
    assert np.unique(blk).size == blk.size
    # low 24 bits are close to uniform
    u = (blk & 0xFFFFFF).astype(np.float64) / 2 ** 24
    assert abs(u.mean() - 0.5) < 2e-3 and abs(u.var() - 1 / 12) < 2e-3


def test_hash_to_unit():
    lib = O.load()
    for h in [0, 1, 0xFFFFFF, 0x1000000, 0xFFFFFFFF, 0x12345678]:
        assert lib.oracle_hash_to_unit(h) == np.float32((h & 0xFFFFFF) / 2 ** 24)


def fill_scene(n, cell_min, dims, seed):
    s = abi.make_settings((n, n, n), 1.0 / n, 8 * n ** 3, solver=abi.SOLVER_NONE)
    o = O.OracleSim(s)
    o.emit(abi.fill_request(cell_min, dims, seed))
    return o


def test_seeding_fill_cells():
    """Slosh.usf:295-311: particle I goes to cell I/8, sub-cell I%8, offset (sub + 0.5 + 0.8 (rnd - 0.5)) / 2."""
    n, cmin, dims, seed = 16, (2, 3, 1), (5, 4, 7), 42
    o = fill_scene(n, cmin, dims, seed)
    p = o.read_particles()
    count = dims[0] * dims[1] * dims[2] * 8
    assert len(p) == count
    h = np.float32(1.0 / n)
    i = np.arange(count, dtype=np.uint32)
    cl, sub = i // 8, i % 8
    local = np.stack([cl % dims[0], (cl // dims[0]) % dims[1], cl // (dims[0] * dims[1])], axis=1)
    cell = local + np.array(cmin)
    so = np.stack([sub & 1, (sub >> 1) & 1, (sub >> 2) & 1], axis=1)
    # every particle lies strictly inside its own octant of its cell, 0.05 cell away from the octant's faces
    frac = p["pos"].astype(np.float64) * n - cell - 0.5 * so
    assert np.all(frac >= 0.05 - 1e-5) and np.all(frac <= 0.45 + 1e-5)
    # exact positions from the published hash, evaluated in float32 in the HLSL order
    h0 = np_pcg_hash(i * np.uint32(3) + np.uint32(seed) * np.uint32(7919))
    h1 = np_pcg_hash(h0)
    h2 = np_pcg_hash(h1)
    rnd = np.stack([(hh & 0xFFFFFF).astype(np.float32) / np.float32(2 ** 24) for hh in (h0, h1, h2)], axis=1)
    in_cell = ((so.astype(np.float32) + np.float32(0.5)) + np.float32(0.8) * (rnd - np.float32(0.5))) * np.float32(0.5)
    pos = (cell.astype(np.float32) + in_cell) * h
    assert np.array_equal(p["pos"], pos)
    assert np.all(p["vel"] == 0) and np.all(p["c"] == 0) and np.all(p["age"] == 0)


def random_particles(n, count, seed, dead=0.05):
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    p = np.zeros(count, dtype=abi.PARTICLE_DTYPE)
    p["pos"] = rng.uniform(1.01 * h, (n - 1.01) * h, size=(count, 3)).astype(np.float32)
    p["vel"] = rng.uniform(-1, 1, size=(count, 3)).astype(np.float32)
    p["c"] = rng.uniform(-3, 3, size=(count, 9)).astype(np.float32)
    p["age"] = rng.uniform(0, 1, size=count).astype(np.float32)
    p["age"][rng.random(count) < dead] = -1.0
    return p


def brick_keys(pos, n):
    c = np.clip(np.floor(pos * np.float32(n)).astype(np.int64), 0, n - 1)
    b = c >> 3
    bd = (n + 7) // 8
    return b[:, 0] + bd * (b[:, 1] + bd * b[:, 2])


def test_sort_is_stable_by_brick_and_drops_dead():
    n = 24
    parts = random_particles(n, 20000, 3)
    s = abi.make_settings((n, n, n), 1.0 / n, 20000, solver=abi.SOLVER_NONE)
    o = O.OracleSim(s)
    o.write_particles(parts)
    o.step(DT, G, stop_after=abi.STAGE_SORT)
    out = o.read_particles()
    live = parts[parts["age"] >= 0]
    order = np.argsort(brick_keys(live["pos"], n), kind="stable")
    assert np.array_equal(out.view(np.uint8), live[order].view(np.uint8))


def test_sort_by_brick_helper_matches_the_sort():
    """oracle_py.sort_by_brick (how the GPU tests order the oracle's records after a full step) gives exactly the
    order of the oracle's sort: a 20 x 28 x 12 grid (partial bricks), dead records, positions clamped at the walls."""
    grid = (20, 28, 12)
    rng = np.random.default_rng(8)
    h = 1.0 / 28
    parts = np.zeros(6000, dtype=abi.PARTICLE_DTYPE)
    for a in range(3):
        parts["pos"][:, a] = rng.uniform(-0.5 * h, (grid[a] + 0.5) * h, size=6000).astype(np.float32)
    parts["vel"] = rng.uniform(-1, 1, size=(6000, 3)).astype(np.float32)
    parts["age"] = rng.uniform(0.0, 1.0, size=6000).astype(np.float32)
    parts["age"][rng.random(6000) < 0.1] = -1.0
    s = abi.make_settings(grid, h, 6000, solver=abi.SOLVER_NONE)
    o = O.OracleSim(s)
    o.write_particles(parts)
    o.step(DT, G, stop_after=abi.STAGE_SORT)
    out = o.read_particles()
    ref = O.sort_by_brick(parts, grid, h)
    assert len(ref) == int((parts["age"] >= 0).sum())
    assert np.array_equal(out.view(np.uint8), ref.view(np.uint8))


def test_brick_pool_is_dilated_touch_set():
    from scipy.ndimage import binary_dilation

    n = 40
    parts = random_particles(n, 300, 5, dead=0.0)
    s = abi.make_settings((n, n, n), 1.0 / n, 1000, solver=abi.SOLVER_NONE)
    o = O.OracleSim(s)
    o.write_particles(parts)
    o.step(DT, G, stop_after=abi.STAGE_POOL)
    table = o.read_grid(abi.FIELD_BRICK_TABLE)
    bd = 5
    touched = np.zeros(bd ** 3, bool)
    touched[brick_keys(parts["pos"], n)] = True
    want = binary_dilation(touched.reshape(bd, bd, bd), structure=np.ones((3, 3, 3), bool)).reshape(-1)
    assert np.array_equal(table != 0, want)
    # first allocation: slots 1..k in ascending brick order
    assert np.array_equal(table[want], np.arange(1, want.sum() + 1))


def test_brick_pool_overflow_keeps_lowest_bricks():
    n = 40
    parts = random_particles(n, 300, 6, dead=0.0)
    s = abi.make_settings((n, n, n), 1.0 / n, 1000, solver=abi.SOLVER_NONE, budget=0.2)
    o = O.OracleSim(s)
    o.write_particles(parts)
    o.step(DT, G, stop_after=abi.STAGE_POOL)
    table = o.read_grid(abi.FIELD_BRICK_TABLE)
    slots = int(np.ceil(float(np.float32(0.2)) * 125))  # the FP32 budget widened to double (SloshProxy.cpp:195)
    assert (table != 0).sum() == slots
    assert o.results().pool_overflow_bricks > 0


def test_classification_walls_sphere_and_counts():
    n = 20
    grid, h, reqs = scenes.dam_break(n, seed=1)
    s = abi.make_settings(grid, h, 40000, solver=abi.SOLVER_NONE)
    o = O.OracleSim(s)
    for r in reqs:
        o.emit(r)
    col = abi.sphere_collider((0.5, 0.3, 0.45), 0.17)
    o.set_colliders([col])
    o.step(DT, G, stop_after=abi.STAGE_P2G)
    flags = o.read_grid(abi.FIELD_CELL_FLAGS)
    table = o.read_grid(abi.FIELD_BRICK_TABLE).reshape(3, 3, 3)
    alloc = np.kron(table != 0, np.ones((8, 8, 8), bool))[:n, :n, :n]  # z, y, x
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    wall = (x == 0) | (y == 0) | (z == 0) | (x == n - 1) | (y == n - 1) | (z == n - 1)
    ctr = np.stack([(x + 0.5) * h, (y + 0.5) * h, (z + 0.5) * h], -1)
    inside = np.linalg.norm(ctr - np.array([0.5, 0.3, 0.45]), axis=-1) < 0.17 - 1e-6
    t = flags & 3
    assert np.all(t[wall & alloc] == abi.CELL_SOLID)
    assert np.all((flags[wall & alloc] >> 8) == 0)
    core = inside & ~wall & alloc
    assert core.any() and np.all(t[core] == abi.CELL_SOLID) and np.all((flags[core] >> 8) == 1)
    # fluid = non-solid allocated cell holding a particle; dense flag at >= 4 particles
    p = o.read_particles()
    cell = np.clip(np.floor(p["pos"] * np.float32(n)).astype(int), 0, n - 1)
    cnt = np.zeros((n, n, n), int)
    np.add.at(cnt, (cell[:, 2], cell[:, 1], cell[:, 0]), 1)
    solid = t == abi.CELL_SOLID
    assert np.array_equal((t == abi.CELL_FLUID) & alloc, (cnt > 0) & ~solid & alloc)
    assert np.array_equal(((flags & 4) != 0) & ~solid & alloc, (cnt >= 4) & ~solid & alloc)


def linear_field_particles(n, A, b, count, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    p = np.zeros(count, dtype=abi.PARTICLE_DTYPE)
    p["pos"] = rng.uniform(0.3, 0.7, size=(count, 3)).astype(np.float32)
    p["vel"] = (p["pos"].astype(np.float64) @ A.T + b).astype(np.float32)
    p["c"] = np.tile(A.reshape(-1), (count, 1)).astype(np.float32)
    return p


def test_apic_p2g_reproduces_affine_field():
    """APIC transfer is exact for affine velocity fields (Jiang et al. 2015): with u(x) = A x + b and C = A, every
    face that received weight carries u(face) up to the 2^-20 fixed-point quantisation."""
    n = 24
    A = np.array([[0.3, -0.8, 0.5], [0.2, 0.1, -0.6], [-0.4, 0.7, 0.25]])
    b = np.array([0.1, -0.2, 0.05])
    parts = linear_field_particles(n, A, b, 200000, 1)
    s = abi.make_settings((n, n, n), 1.0 / n, 200000, solver=abi.SOLVER_NONE)
    o = O.OracleSim(s)
    o.write_particles(parts)
    o.step(DT, G0, stop_after=abi.STAGE_P2G)
    acc = o.read_grid(abi.FIELD_ACCUM).astype(np.float64)
    h = 1.0 / n
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    for a in range(3):
        wt = acc[3 + a]
        ok = wt > 0.5 * 2 ** 20  # faces with at least half a particle's weight
        pos = [(x + (0.0 if a == 0 else 0.5)) * h, (y + (0.0 if a == 1 else 0.5)) * h, (z + (0.0 if a == 2 else 0.5)) * h]
        u_exact = A[a, 0] * pos[0] + A[a, 1] * pos[1] + A[a, 2] * pos[2] + b[a]
        u = acc[a][ok] / wt[ok]
        assert ok.sum() > 500
        assert np.max(np.abs(u - u_exact[ok])) < 2e-5, a


def test_p2g_single_particle_brute_force():
    """One particle: every node's momentum and weight against a float64 evaluation of the trilinear stencil."""
    n = 16
    h = 1.0 / n
    p = np.zeros(1, dtype=abi.PARTICLE_DTYPE)
    p["pos"] = [[0.4137, 0.5621, 0.3309]]
    p["vel"] = [[0.7, -1.3, 0.4]]
    p["c"] = [[1.5, -2.0, 0.5, 0.3, 0.8, -1.1, -0.6, 0.2, 2.4]]
    s = abi.make_settings((n, n, n), h, 10, solver=abi.SOLVER_NONE)
    o = O.OracleSim(s)
    o.write_particles(p)
    o.step(DT, G0, stop_after=abi.STAGE_P2G)
    acc = o.read_grid(abi.FIELD_ACCUM).astype(np.int64)
    xp = p["pos"][0].astype(np.float64)
    for a in range(3):
        off = np.full(3, 0.5)
        off[a] = 0.0
        g = xp / h - off
        base = np.floor(g).astype(int)
        fr = g - base
        mom = np.zeros((n, n, n))
        wt = np.zeros((n, n, n))
        for o3 in range(8):
            od = np.array([o3 & 1, (o3 >> 1) & 1, (o3 >> 2) & 1])
            node = base + od
            w = np.prod(np.where(od == 1, fr, 1 - fr))
            xn = (node + off) * h
            mom[node[2], node[1], node[0]] = w * (p["vel"][0][a] + p["c"][0][3 * a:3 * a + 3].astype(np.float64) @ (xn - xp))
            wt[node[2], node[1], node[0]] = w
        assert np.max(np.abs(acc[a] - mom * 2 ** 20)) <= 3
        assert np.max(np.abs(acc[3 + a] - wt * 2 ** 20)) <= 2
        assert acc[a].sum() != 0


def test_normalise_ratio_gravity_and_walls():
    n = 20
    grid, h, reqs = scenes.dam_break(n, seed=2)
    s = abi.make_settings(grid, h, 50000, solver=abi.SOLVER_NONE)
    o = O.OracleSim(s)
    for r in reqs:
        o.emit(r)
    o.step(DT, G)  # particles now carry velocity
    o.step(DT, G, stop_after=abi.STAGE_P2G)
    acc = o.read_grid(abi.FIELD_ACCUM)
    o2 = O.OracleSim(s)
    for r in reqs:
        o2.emit(r)
    o2.step(DT, G)
    o2.step(DT, G, stop_after=abi.STAGE_NORMALIZE)
    vel, ff, fw = o2.read_grid(abi.FIELD_VEL), o2.read_grid(abi.FIELD_FACE_FLAGS), o2.read_grid(abi.FIELD_FACE_WEIGHT)
    for a in range(3):
        wt, mom = acc[3 + a], acc[a]
        valid_fluid = (ff[a] & 1) == 1
        solid = (ff[a] & 2) == 2
        has = wt > 0
        assert np.array_equal((ff[a] & 4) != 0, has)
        assert np.array_equal(fw[a][has], (wt[has].astype(np.float32) * np.float32(2 ** -20)))
        sel = valid_fluid & ~solid & has
        exp = mom[sel].astype(np.float32) / wt[sel].astype(np.float32) + np.float32(np.float32(G[a]) * np.float32(DT))
        assert np.array_equal(vel[a][sel], exp)
        assert np.all(vel[a][solid] == 0.0)  # static walls


def test_ghost_theta_closed_form():
    for rf, ra in [(1.0, 0.2), (0.9, 0.6), (0.55, 0.5), (1.2, 0.0), (0.51, 0.5099), (0.3, 0.1), (2.0, 1.0)]:
        d = np.float32(rf) - np.float32(ra)
        exp = 1.0 if d <= np.float32(1e-4) else min(max((np.float32(rf) - np.float32(0.5)) / d, np.float32(0.05)), 1.0)
        assert O.ghost_theta(rf, ra) == np.float32(exp)


def test_diagonal_without_ghost_fluid_counts_open_neighbours():
    n = 20
    grid, h, reqs = scenes.dam_break(n, seed=4)
    s = abi.make_settings(grid, h, 50000, solver=abi.SOLVER_NONE, ghost_fluid=0)
    o = O.OracleSim(s)
    for r in reqs:
        o.emit(r)
    o.step(DT, G, stop_after=abi.STAGE_DIVERGENCE)
    t = o.read_grid(abi.FIELD_CELL_FLAGS) & 3
    diag = o.read_grid(abi.FIELD_CELL_DIAG)
    tp = np.pad(t, 1, constant_values=abi.CELL_SOLID)
    cnt = np.zeros_like(t)
    for ax in range(3):
        for sgn in (-1, 1):
            cnt += np.roll(tp, sgn, axis=ax)[1:-1, 1:-1, 1:-1] != abi.CELL_SOLID
    fluid = t == abi.CELL_FLUID
    assert np.array_equal(diag[fluid], cnt[fluid].astype(np.float32))
    assert np.all(diag[~fluid] == 0)


def test_divergence_matches_face_differences():
    n = 20
    grid, h, reqs = scenes.dam_break(n, seed=4)
    s = abi.make_settings(grid, h, 50000, solver=abi.SOLVER_NONE)
    o = O.OracleSim(s)
    for r in reqs:
        o.emit(r)
    o.step(DT, G)
    o.step(DT, G, stop_after=abi.STAGE_DIVERGENCE)
    vel = o.read_grid(abi.FIELD_VEL).astype(np.float64)
    div = o.read_grid(abi.FIELD_DIVERGENCE).astype(np.float64)
    t = o.read_grid(abi.FIELD_CELL_FLAGS) & 3
    d = np.zeros_like(div)
    for a, ax in zip(range(3), (2, 1, 0)):
        up = np.roll(vel[a], -1, axis=ax)
        d += up - vel[a]
    fluid = t == abi.CELL_FLUID
    assert np.max(np.abs(d[fluid] - div[fluid])) < 1e-6


def dense_system(o, n):
    """The pressure system read back from the oracle, as a dense matrix: (diag + 1e-4) on fluid cells, -1 between
    fluid neighbours, rhs = -divergence."""
    t = o.read_grid(abi.FIELD_CELL_FLAGS) & 3
    diag = o.read_grid(abi.FIELD_CELL_DIAG)
    div = o.read_grid(abi.FIELD_DIVERGENCE)
    fl = np.argwhere(t == abi.CELL_FLUID)
    idx = -np.ones(t.shape, int)
    idx[tuple(fl.T)] = np.arange(len(fl))
    A = np.zeros((len(fl), len(fl)))
    for k, (z, y, x) in enumerate(fl):
        A[k, k] = np.float64(np.float32(diag[z, y, x]) + np.float32(1e-4))
        for dz, dy, dx in [(0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)]:
            zz, yy, xx = z + dz, y + dy, x + dx
            if 0 <= zz < n and 0 <= yy < n and 0 <= xx < n and idx[zz, yy, xx] >= 0:
                A[k, idx[zz, yy, xx]] = -1.0
    rhs = -div[tuple(fl.T)].astype(np.float64)
    return A, rhs, fl


def test_fp64_solve_matches_dense_lu_and_mgpcg_reaches_tolerance():
    n = 14
    grid, h, reqs = scenes.dam_break(n, seed=7, fill=(0.6, 0.7, 1.0))
    s = abi.make_settings(grid, h, 40000, solver=abi.SOLVER_MGPCG, tolerance=1e-6)
    o = O.OracleSim(s)
    for r in reqs:
        o.emit(r)
    o.step(DT, G)
    A, rhs, fl = dense_system(o, n)
    assert np.allclose(A, A.T)  # symmetric
    assert np.all(np.linalg.eigvalsh(A) > 0)  # positive definite
    x = np.linalg.solve(A, rhs)
    ref, iters, rel = o.solve_fp64(1e-13, 500)
    assert rel <= 1e-12
    assert np.max(np.abs(ref[tuple(fl.T)] - x)) <= 1e-10 * np.max(np.abs(x))
    r = o.results()
    assert r.converged == 1 and r.residual_rel <= 1e-6 and r.solver_iterations > 0
    p = o.read_grid(abi.FIELD_PRESSURE_F64)[tuple(fl.T)]
    assert np.linalg.norm(A @ p - rhs) / np.linalg.norm(rhs) <= 1e-6
    assert np.linalg.norm(p - x) / np.linalg.norm(x) <= 1e-5


def test_rbsor_reaches_same_solution():
    n = 12
    grid, h, reqs = scenes.dam_break(n, seed=8, fill=(0.6, 0.7, 1.0))
    s = abi.make_settings(grid, h, 30000, solver=abi.SOLVER_RBSOR, tolerance=1e-6, max_iterations=5000)
    o = O.OracleSim(s)
    for r in reqs:
        o.emit(r)
    o.step(DT, G)
    A, rhs, fl = dense_system(o, n)
    x = np.linalg.solve(A, rhs)
    r = o.results()
    assert r.converged == 1 and r.residual_rel <= 1e-6
    p = o.read_grid(abi.FIELD_PRESSURE_F64)[tuple(fl.T)]
    assert np.linalg.norm(p - x) / np.linalg.norm(x) <= 1e-5


def test_hydrostatic_pressure_gradient():
    """A still pool: below the surface, scaled pressure grows by g dt per cell of depth (p_hat = p dt / (rho h)),
    and the projection brings the vertical velocity back to rest. The operator's 1e-4 diagonal term (a weak
    leak, SloshProxy.cpp:845) bends the profile by ~1e-4 p per cell, hence the 0.5% tolerance."""
    n = 24
    grid, h, reqs = scenes.dam_break(n, seed=9, fill=(1.0, 0.5, 1.0))
    s = abi.make_settings(grid, h, 200000, solver=abi.SOLVER_MGPCG, tolerance=1e-8, max_iterations=400)
    o = O.OracleSim(s)
    for r in reqs:
        o.emit(r)
    o.step(DT, G, stop_after=abi.STAGE_PROJECT)
    p = o.read_grid(abi.FIELD_PRESSURE_F64)
    col = p[n // 2, 2:9, n // 2]  # z, y, x: a vertical column in the lower half
    dp = -np.diff(col)
    assert np.allclose(dp, 9.8 * DT, rtol=1.5e-2), dp
    assert np.all(col[:-1] > col[1:])  # pressure grows with depth everywhere in the column
    v = o.read_grid(abi.FIELD_VEL)[1]
    assert np.max(np.abs(v[n // 2, 2:9, n // 2])) <= 1.5e-2 * 9.8 * DT


def test_projection_divergence_equals_minus_residual():
    """After projection, div_c = -(b - A p)_c - 1e-4 p_c in every fluid cell (the ghost-fluid faces included)."""
    n = 16
    grid, h, reqs = scenes.dam_break(n, seed=10, fill=(0.6, 0.7, 1.0))
    s = abi.make_settings(grid, h, 60000, solver=abi.SOLVER_MGPCG, tolerance=1e-6)
    o = O.OracleSim(s)
    for r in reqs:
        o.emit(r)
    o.step(DT, G)
    o.step(DT, G, stop_after=abi.STAGE_SOLVE)
    A, rhs, fl = dense_system(o, n)
    p = o.read_grid(abi.FIELD_PRESSURE_F64)[tuple(fl.T)]
    res = rhs - A @ p
    o.step2 = None
    # rerun the same step to the projection (same inputs, deterministic)
    o2 = O.OracleSim(s)
    for r in reqs:
        o2.emit(r)
    o2.step(DT, G)
    o2.step(DT, G, stop_after=abi.STAGE_PROJECT)
    vel = o2.read_grid(abi.FIELD_VEL).astype(np.float64)
    d = np.zeros((n, n, n))
    for a, ax in zip(range(3), (2, 1, 0)):
        d += np.roll(vel[a], -1, axis=ax) - vel[a]
    post = d[tuple(fl.T)]
    expect = -res - np.float32(1e-4) * p
    assert np.max(np.abs(post - expect)) <= 1e-5 * np.max(np.abs(rhs)) + 1e-7


def test_extrapolation_matches_numpy_restatement():
    n = 20
    grid, h, reqs = scenes.dam_break(n, seed=11)
    s = abi.make_settings(grid, h, 50000, solver=abi.SOLVER_MGPCG)
    sims = []
    for stop in (abi.STAGE_PROJECT, abi.STAGE_EXTRAPOLATE):
        o = O.OracleSim(s)
        for r in reqs:
            o.emit(r)
        o.step(DT, G)
        o.step(DT, G, stop_after=stop)
        sims.append(o)
    vel = sims[0].read_grid(abi.FIELD_VEL).copy()
    ff = sims[0].read_grid(abi.FIELD_FACE_FLAGS).astype(np.uint32).copy()
    table = sims[0].read_grid(abi.FIELD_BRICK_TABLE).reshape(3, 3, 3)
    alloc = np.kron(table != 0, np.ones((8, 8, 8), bool))[:n, :n, :n]
    for layer in range(2):
        known = 1 | sum(1 << (4 + k) for k in range(layer))
        new = vel.copy()
        newf = ff.copy()
        for a in range(3):
            kn = ((ff[a] & known) != 0) & ((ff[a] & 2) == 0)
            vals = np.where(kn, vel[a], 0.0).astype(np.float32)
            ssum = np.zeros((n, n, n), np.float32)
            cnt = np.zeros((n, n, n), np.float32)
            for ax in range(3):
                for sh in (1, -1):
                    # neighbour in direction -sh along axis ax (rolled into place), zero outside the grid;
                    # accumulate in the order -x, +x, -y, +y, -z, +z
                    pass
            order = [(2, 1), (2, -1), (1, 1), (1, -1), (0, 1), (0, -1)]
            for ax, sh in order:
                v = np.roll(vals, sh, axis=ax)
                k = np.roll(kn, sh, axis=ax)
                edge = [slice(None)] * 3
                edge[ax] = 0 if sh == 1 else -1
                v[tuple(edge)] = 0
                k[tuple(edge)] = False
                ssum = np.where(k, ssum + v, ssum)
                cnt = cnt + k
            unknown = ((ff[a] & known) == 0) & alloc
            fill = unknown & (cnt > 0)
            new[a] = np.where(fill, ssum / np.maximum(cnt, 1), np.where(unknown, 0.0, vel[a]))
            newf[a] = np.where(fill, ff[a] | (1 << (4 + layer)), ff[a])
        vel, ff = new.astype(np.float32), newf
    got = sims[1].read_grid(abi.FIELD_VEL)
    gotf = sims[1].read_grid(abi.FIELD_FACE_FLAGS)
    assert np.array_equal(gotf, ff)
    assert np.array_equal(got[:, alloc], vel[:, alloc])


def test_g2p_affine_field_and_translation():
    """G2P of an affine grid field returns the field at the particle and C = its gradient; with no pressure and no
    gravity a uniform flow translates particles by v dt."""
    n = 24
    A = np.array([[0.3, -0.8, 0.5], [0.2, 0.1, -0.6], [-0.4, 0.7, 0.25]]) * 0.1
    b = np.array([0.4, -0.3, 0.2])
    parts = linear_field_particles(n, A, b, 200000, 2)
    s = abi.make_settings((n, n, n), 1.0 / n, 200000, solver=abi.SOLVER_NONE, extrapolation_layers=2)
    o = O.OracleSim(s)
    o.write_particles(parts)
    o.step(1e-4, G0)
    out = o.read_particles()
    inner = np.all((out["pos"] > 0.35) & (out["pos"] < 0.65), axis=1)
    assert inner.sum() > 1000
    c = out["c"][inner].reshape(-1, 3, 3)
    assert np.max(np.abs(c - A[None])) < 5e-4
    v_exact = out["pos"][inner].astype(np.float64) @ A.T + b
    assert np.max(np.abs(out["vel"][inner] - v_exact)) < 1e-4
    # uniform translation
    A0 = np.zeros((3, 3))
    parts = linear_field_particles(n, A0, b, 100000, 3)
    o = O.OracleSim(s)
    o.write_particles(parts)
    o.step(1e-3, G0)
    out = o.read_particles()
    order = np.argsort(brick_keys(parts["pos"], n), kind="stable")
    moved = out["pos"].astype(np.float64) - parts["pos"][order].astype(np.float64)
    assert np.max(np.abs(moved - b * 1e-3)) < 2e-6


def test_wall_clamp_removes_inward_velocity():
    n = 16
    h = 1.0 / n
    # a slab of particles two cells from the -x wall moving at it: the gathered velocity carries them through
    rng = np.random.default_rng(0)
    p = np.zeros(4096, dtype=abi.PARTICLE_DTYPE)
    p["pos"][:, 0] = rng.uniform(1.9 * h, 3.0 * h, 4096)
    p["pos"][:, 1:] = rng.uniform(0.3, 0.7, (4096, 2))
    p["vel"] = [[-30.0, 0.0, 0.0]] * 4096
    s = abi.make_settings((n, n, n), h, 4096, solver=abi.SOLVER_NONE)
    o = O.OracleSim(s)
    o.write_particles(p)
    o.step(1.0 / 60.0, G0)
    out = o.read_particles()
    lo = np.float32(np.float32(h) * np.float32(1.01))
    assert np.all(out["pos"][:, 0] >= lo)
    hit = out["pos"][:, 0] == lo
    assert hit.sum() > 100
    assert np.all(out["vel"][hit, 0] == 0.0)  # inward normal velocity removed (max(v, 0))

def test_warm_start_extrapolation_guess():
    """SLOSH_FLAG_WARM_START_EXTRAPOLATE in the oracle: the solver's starting guess, bit for bit, from the two
    previous solutions (tests/warm_start_check.py)."""
    import warm_start_check as W

    s, reqs = W.make_settings()
    o = O.OracleSim(s)
    W.run_and_check(o, reqs)


def test_warm_start_extrapolation_saves_iterations():
    """On a smoothly sloshing tank the extrapolated start takes fewer MGPCG iterations than the reference's scaled
    last solution (the solves meet the same tolerance)."""
    from slosh_amd import scenes

    grid, h, reqs = scenes.sloshing_tank(48, seed=1)
    cap = sum(int(np.prod(r.cell_dims)) * 8 for r in reqs) + 1024
    total = {}
    for flags in (0, abi.FLAG_WARM_START_EXTRAPOLATE):
        o = O.OracleSim(abi.make_settings(grid, h, cap, solver=abi.SOLVER_MGPCG, tolerance=1e-6, flags=flags))
        for r in reqs:
            o.emit(r)
        t, its = 0.0, []
        for _ in range(8):
            o.step(1.0 / 240.0, scenes.tank_gravity(t))
            t += 1.0 / 240.0
            r = o.results()
            assert r.converged == 1 and r.residual_rel <= 1e-6
            its.append(r.solver_iterations)
        total[flags] = sum(its[2:])
        o.close()
    assert total[abi.FLAG_WARM_START_EXTRAPOLATE] < total[0], total


def test_ref48_packing_known_answers():
    """SloshCommon.ush:71-88: binary16 bit patterns from the IEEE format (1.0 = 0x3C00, 2.0 = 0x4000, ...), halves
    packed low first in the order Cx.xyz, Cy.xyz, Cz.xyz, the tenth half zero."""
    p = np.zeros(1, dtype=abi.PARTICLE_DTYPE)
    p["pos"] = (0.25, 0.5, 0.75)
    p["age"] = 1.5
    p["vel"] = (-1.0, 2.0, 3.0)
    p["c"] = np.arange(1, 10, dtype=np.float32)
    q = O.pack_ref48(p)
    assert [int(v) for v in q["c"][0]] == [0x40003C00, 0x44004200, 0x46004500, 0x48004700, 0x00004880]
    assert np.array_equal(q["pos"], p["pos"]) and np.array_equal(q["vel"], p["vel"]) and q["age"][0] == 1.5


def test_ref48_clamp_and_rounding():
    p = np.zeros(1, dtype=abi.PARTICLE_DTYPE)
    # +-1e9 clamp to +-65000, which rounds to the nearest half 64992 (spacing 32 above 32768); 1 + 2^-11 is a tie
    # between 1 and 1 + 2^-10 and rounds to the even one (1.0); 1 + 3 * 2^-11 rounds up to 1 + 2^-9; the smallest
    # subnormal half is 2^-24
    p["c"] = np.array([1e9, -1e9, 1 + 2 ** -11, 1 + 3 * 2 ** -11, 2 ** -24, -0.0, 65000.0, 0.1, -3.0], np.float32)
    halves = []
    for w in O.pack_ref48(p)["c"][0]:
        halves += [int(w) & 0xFFFF, int(w) >> 16]
    assert halves[:9] == [0x7BEF, 0xFBEF, 0x3C00, 0x3C02, 0x0001, 0x8000, 0x7BEF, 0x2E66, 0xC200]
    assert halves[9] == 0


def test_ref48_round_trip():
    rng = np.random.default_rng(3)
    p = np.zeros(1000, dtype=abi.PARTICLE_DTYPE)
    p["pos"] = rng.uniform(0, 1, size=(1000, 3))
    p["vel"] = rng.normal(size=(1000, 3))
    # values exactly representable in binary16 survive the round trip unchanged
    p["c"] = (rng.integers(-2048, 2048, size=(1000, 9)) / 64.0).astype(np.float32)
    assert O.unpack_ref48(O.pack_ref48(p)).tobytes() == p.tobytes()
    # others within half an ulp of binary16 (relative 2^-11)
    p["c"] = rng.normal(scale=10.0, size=(1000, 9)).astype(np.float32)
    back = O.unpack_ref48(O.pack_ref48(p))["c"]
    assert np.all(np.abs(back - p["c"]) <= np.abs(p["c"]) * 2.0 ** -11 + 2.0 ** -25)
