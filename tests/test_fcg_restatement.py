"""Flexible CG (Notay 2000, FCG(1)) as the library runs it, restated in numpy on the system of a real step, pinned by
the two properties that fix its formulas: with M = A^-1 it converges in one iteration, and with a fixed SPD
preconditioner (the V-cycle of tests/mg_numpy.py) its iterates are PCG's, so it stops at PCG's iteration count with
PCG's solution. The update is the one of level0.hip (k_fcg_zq, k_fcg_direction, k_fcg_update) and solver.hip
(k_finalize kFinBetaF, kFinAlphaF): beta = -(z . q_prev) / (d_prev . q_prev), d = z + beta d_prev, q = A d,
alpha = (d . r) / (d . q), x += alpha d, r -= alpha q. A flipped sign of beta breaks the PCG equivalence (29 against
15 iterations, checked on the CPU when the test was written); forms equal to these in exact arithmetic (z . r for
d . r, z . q for d . q) pass, as they should."""
import numpy as np

import mg_numpy as M
import oracle_py as O
from slosh_amd import abi, scenes


def step_system(n=12, seed=2):
    grid, h, reqs = scenes.dam_break(n, seed=seed)
    s = abi.make_settings(grid, h, 8 * n ** 3, solver=abi.SOLVER_NONE)
    o = O.OracleSim(s)
    for r in reqs:
        o.emit(r)
    o.step(1.0 / 120.0, (0.0, -9.8, 0.0))
    t0 = (o.read_grid(abi.FIELD_CELL_FLAGS) & 3).astype(np.int32)
    diag = o.read_grid(abi.FIELD_CELL_DIAG).astype(np.float64)
    b = np.where(t0 == M.FLUID, -o.read_grid(abi.FIELD_DIVERGENCE).astype(np.float64), 0.0)
    o.close()
    return t0, diag, b


def fcg(A0, precond, b, tol=1e-6, maxit=200):
    """Iterations until ||r|| <= tol ||b||, and the solution."""
    x = np.zeros_like(b)
    r = b.copy()
    bn = np.sqrt((b * b).sum())
    d_prev = q_prev = None
    for it in range(1, maxit + 1):
        z = precond(r)
        if d_prev is None:
            d = z
        else:
            beta = -(z * q_prev).sum() / (d_prev * q_prev).sum()
            d = z + beta * d_prev
        q = A0.apply(d)
        alpha = (d * r).sum() / (d * q).sum()
        x = x + alpha * d
        r = r - alpha * q
        if np.sqrt((r * r).sum()) <= tol * bn:
            return it, x
        d_prev, q_prev = d, q
    return maxit, x


def test_fcg_with_exact_inverse_converges_in_one_iteration():
    t0, diag, b = step_system()
    A0 = M.Level(t0, diag)
    idx = np.flatnonzero(A0.fluid)
    A = np.zeros((idx.size, idx.size))
    for k, i in enumerate(idx):
        e = np.zeros(t0.size)
        e[i] = 1.0
        A[:, k] = A0.apply(e.reshape(t0.shape)).ravel()[idx]
    Ainv = np.linalg.inv(A)

    def exact(r):
        z = np.zeros(t0.size)
        z[idx] = Ainv @ r.ravel()[idx]
        return z.reshape(t0.shape)

    assert idx.size > 100
    it, x = fcg(A0, exact, b, tol=1e-10)
    assert it == 1
    assert np.linalg.norm((b - A0.apply(x))[A0.fluid]) <= 1e-10 * np.linalg.norm(b)


def test_fcg_with_the_v_cycle_is_pcg():
    t0, diag, b = step_system(16, seed=4)
    A0 = M.Level(t0, diag)
    V = M.VCycle(t0, diag)
    it_f, x_f = fcg(A0, V, b)
    it_p, x_p = M.mgpcg_iterations(t0, diag, b)
    assert it_f == it_p and it_f > 3
    assert np.linalg.norm(x_f - x_p) <= 1e-9 * np.linalg.norm(x_p)
