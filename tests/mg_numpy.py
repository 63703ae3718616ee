"""Independent numpy restatement of the MGPCG preconditioner on dense grids (test infrastructure).

Follows SloshPressureMG.usf as the oracle does (MGBuildTypeCS :256-278, MGSmoothCS :311-338, MGRestrictCS :343-381,
MGProlongCS :418-433; V-cycle SloshProxy.cpp:904-959) but with whole-array numpy operations in FP64, so it shares no
code with the oracle: coarse types from the 8 children, RB Gauss-Seidel pairs (red then black down, black then red
up), b_c = 1/4 of the children's residuals, piecewise-constant prolongation onto fluid cells with a fluid parent,
an exact solve of the coarsest level (<= 64 cells or 8 levels). Options: fluid-first coarse types and a number of
red-black pairs on levels 4 and coarser.
"""
import numpy as np

AIR, FLUID, SOLID = 0, 1, 2
EPS = 1e-4
NB = [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]  # (dz, dy, dx): -x +x -y +y -z +z


def shift(a, dz, dy, dx, fill):
    """out[z, y, x] = a[z + dz, y + dy, x + dx], `fill` outside."""
    out = np.full_like(a, fill)
    Z, Y, X = a.shape

    def sl(d, n):
        return slice(max(-d, 0), n - max(d, 0)), slice(max(d, 0), n - max(-d, 0))

    (oz, iz), (oy, iy), (ox, ix) = sl(dz, Z), sl(dy, Y), sl(dx, X)
    out[oz, oy, ox] = a[iz, iy, ix]
    return out


class Level:
    def __init__(self, t, diag0=None):
        self.t = t
        self.fluid = t == FLUID
        cnt = np.zeros(t.shape)
        self.nbf = []
        for dz, dy, dx in NB:
            tn = shift(t, dz, dy, dx, SOLID)
            self.nbf.append((dz, dy, dx, (tn == FLUID) & self.fluid))
            cnt += tn != SOLID
        self.dg = (diag0 + EPS) if diag0 is not None else cnt + EPS
        z, y, x = np.indices(t.shape)
        self.red = ((x + y + z) % 2 == 0) & self.fluid
        self.black = ((x + y + z) % 2 == 1) & self.fluid

    def nbsum(self, v):
        s = np.zeros_like(v)
        for dz, dy, dx, m in self.nbf:
            s += np.where(m, shift(v, dz, dy, dx, 0.0), 0.0)
        return s

    def apply(self, v):
        return np.where(self.fluid, self.dg * v - self.nbsum(v), 0.0)

    def sweep(self, x, b, mask):
        x[mask] = ((b + self.nbsum(x)) / self.dg)[mask]


def coarse_types(t, fluid_first):
    cshape = tuple((n + 1) // 2 for n in t.shape)
    air = np.zeros(cshape, bool)
    fl = np.zeros(cshape, bool)
    for oz in (0, 1):
        for oy in (0, 1):
            for ox in (0, 1):
                sub = t[oz::2, oy::2, ox::2]
                s = tuple(slice(0, n) for n in sub.shape)
                air[s] |= sub == AIR
                fl[s] |= sub == FLUID
    c = np.full(cshape, SOLID, np.int32)
    if fluid_first:
        c[air] = AIR
        c[fl] = FLUID
    else:
        c[fl] = FLUID
        c[air] = AIR
    return c


def restrict_sum(r, cshape):
    out = np.zeros(cshape)
    for oz in (0, 1):
        for oy in (0, 1):
            for ox in (0, 1):
                sub = r[oz::2, oy::2, ox::2]
                out[tuple(slice(0, n) for n in sub.shape)] += sub
    return out


def prolong(xc, fshape):
    return np.repeat(np.repeat(np.repeat(xc, 2, 0), 2, 1), 2, 2)[:fshape[0], :fshape[1], :fshape[2]]


class VCycle:
    def __init__(self, t0, diag0, fluid_first=False, coarse_sweeps=1):
        self.L = [Level(t0, diag0)]
        while self.L[-1].t.size > 64 and len(self.L) < 8:
            self.L.append(Level(coarse_types(self.L[-1].t, fluid_first)))
        m = len(self.L)
        self.pairs = [coarse_sweeps if 4 <= l < m - 1 else 1 for l in range(m)]
        last = self.L[-1]
        self.cidx = np.flatnonzero(last.fluid)
        n = len(self.cidx)
        A = np.zeros((n, n))
        for k, i in enumerate(self.cidx):
            e = np.zeros(last.t.size)
            e[i] = 1.0
            A[:, k] = last.apply(e.reshape(last.t.shape)).ravel()[self.cidx]
        self.Ainv = np.linalg.inv(A) if n else None

    def __call__(self, r):
        L, m = self.L, len(self.L)
        xs, bs = [None] * m, [None] * m
        bs[0] = np.where(L[0].fluid, r, 0.0)
        for l in range(m - 1):
            x = np.zeros_like(bs[l])
            for _ in range(self.pairs[l]):
                L[l].sweep(x, bs[l], L[l].red)
                L[l].sweep(x, bs[l], L[l].black)
            xs[l] = x
            res = np.where(L[l].fluid, bs[l] - L[l].apply(x), 0.0)
            bs[l + 1] = np.where(L[l + 1].fluid, 0.25 * restrict_sum(res, L[l + 1].t.shape), 0.0)
        xc = np.zeros(L[-1].t.size)
        if len(self.cidx):
            xc[self.cidx] = self.Ainv @ bs[-1].ravel()[self.cidx]
        xs[-1] = xc.reshape(L[-1].t.shape)
        for l in range(m - 2, -1, -1):
            parent_fluid = prolong(L[l + 1].fluid, L[l].t.shape)
            x = xs[l] + np.where(L[l].fluid & parent_fluid, prolong(xs[l + 1], L[l].t.shape), 0.0)
            for _ in range(self.pairs[l]):
                L[l].sweep(x, bs[l], L[l].black)
                L[l].sweep(x, bs[l], L[l].red)
            xs[l] = x
        return xs[0]


def mgpcg_iterations(t0, diag0, b, tol=1e-6, maxit=200, **opts):
    """Iterations of PCG with the V-cycle preconditioner until ||r|| <= tol ||b||, and the solution."""
    A0 = Level(t0, diag0)
    M = VCycle(t0, diag0, **opts)
    x = np.zeros_like(b)
    r = b.copy()
    bn = np.sqrt((b * b).sum())
    z = M(r)
    d = z.copy()
    rz = (r * z).sum()
    for it in range(1, maxit + 1):
        q = A0.apply(d)
        a = rz / (d * q).sum()
        x += a * d
        r -= a * q
        if np.sqrt((r * r).sum()) <= tol * bn:
            return it, x
        z = M(r)
        rz2 = (r * z).sum()
        d = z + (rz2 / rz) * d
        rz = rz2
    return maxit, x
