"""Scene builders for the benchmark workloads, as emit requests of the C ABI.

Fill boxes are snapped to cells the way the reference's volume actor does (SloshVolume.cpp:120-146,
QueueFillLocalBox): the min corner rounds down, the max corner rounds up, both clamped to [1, N-1) so the wall
layer stays empty. Curved fills are emitted one x-column at a time with the cells whose centre lies below the
surface. Each request takes the next emit seed, as the reference's per-request EmitSeed++ (SloshProxy.cpp:529).
"""
import math

from . import abi


def snap_box(lo, hi, grid, h):
    """Cell box [cmin, cmin + dims) of the metric box [lo, hi] (SloshVolume.cpp:125-135)."""
    cmin, dims = [], []
    for a in range(3):
        n = grid[a]
        mn = min(max(math.floor(lo[a] / h), 1), n - 1)
        mx = min(max(math.ceil(hi[a] / h), 1), n - 1)
        cmin.append(mn)
        dims.append(mx - mn)
    return cmin, dims


def dam_break(n, seed=0, fill=(0.4, 0.6, 1.0), extent=1.0):
    """Dam break on a unit cube of n^3 cells: fluid block x < 0.4, y < 0.6 (all z), v = 0, y up."""
    h = extent / n
    grid = (n, n, n)
    cmin, dims = snap_box((0.0, 0.0, 0.0), (fill[0] * extent, fill[1] * extent, fill[2] * extent), grid, h)
    return grid, h, [abi.fill_request(cmin, dims, seed)]


def wave_columns(n, seed, mean, amp, extent=1.0):
    """Tank of n^3 cells filled below y = mean + amp sin(2 pi x), one request per x column of cells."""
    h = extent / n
    reqs = []
    k = 0
    for i in range(1, n - 1):
        xc = (i + 0.5) * h
        ys = mean + amp * math.sin(2.0 * math.pi * xc / extent)
        top = min(max(math.ceil(ys / h - 0.5), 1), n - 1)
        if top <= 1:
            continue
        reqs.append(abi.fill_request((i, 1, 1), (1, top - 1, n - 2), seed + k))
        k += 1
    return (n, n, n), h, reqs


def sloshing_tank(n=256, seed=1, literal=False):
    """Sloshing tank: y < 0.5 + 0.1 sin(2 pi x) (literal, ~65.5M particles at 256^3) or the 16M-particle point
    y < 0.125 + 0.025 sin(2 pi x) that the steps/s metric names."""
    if literal:
        return wave_columns(n, seed, 0.5, 0.1)
    return wave_columns(n, seed, 0.125, 0.025)


def tank_gravity(t):
    """Tank-frame gravity of the sloshing workload: horizontal 2 m/s^2 at 1 Hz plus -9.8 y."""
    return (-2.0 * math.sin(2.0 * math.pi * t), -9.8, 0.0)


def particle_count(reqs):
    return sum(int(r.cell_dims[0]) * int(r.cell_dims[1]) * int(r.cell_dims[2]) * 8 for r in reqs)


def mask_requests(mask, seed):
    """FillCells requests covering the cells of a boolean mask (z, y, x): contiguous y runs per (x, z) column,
    merged along z while the runs repeat. Each request takes the next seed."""
    import numpy as np

    m = np.asarray(mask, bool)
    nz, ny, nx = m.shape
    reqs, k = [], 0
    for x in range(nx):
        col = m[:, :, x]  # (z, y)
        if not col.any():
            continue
        d = np.diff(np.concatenate([np.zeros((nz, 1), np.int8), col.astype(np.int8), np.zeros((nz, 1), np.int8)], axis=1), axis=1)
        runs = [tuple(zip(np.nonzero(d[z] == 1)[0], np.nonzero(d[z] == -1)[0])) for z in range(nz)]
        z = 0
        while z < nz:
            if not runs[z]:
                z += 1
                continue
            z1 = z + 1
            while z1 < nz and runs[z1] == runs[z]:
                z1 += 1
            for y0, y1 in runs[z]:
                reqs.append(abi.fill_request((x, int(y0), z), (1, int(y1 - y0), z1 - z), seed + k))
                k += 1
            z = z1
    return reqs


def _interior(shape):
    import numpy as np

    m = np.zeros(shape, bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


def spheres_pool(n=512, seed=2, nspheres=64, depth=0.15):
    """Mixed-precision workload: 64 fluid spheres of radius U(0.03, 0.08) above a pool of the given depth in a unit
    cube of n^3 cells; cells whose centre lies inside a shape are filled (8 particles per cell)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    mask = _interior((n, n, n))
    mask &= (c[None, :, None] < depth)
    fill = mask.copy()
    for _ in range(nspheres):
        r = rng.uniform(0.03, 0.08)
        ctr = rng.uniform(r + 0.02, 1.0 - r - 0.02, 3)
        ctr[1] = rng.uniform(depth + r + 0.02, 1.0 - r - 0.02)
        lo = np.maximum(((ctr - r) / h).astype(int), 0)
        hi = np.minimum(((ctr + r) / h).astype(int) + 2, n)
        zs, ys, xs = np.ogrid[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        inside = ((c[xs] - ctr[0]) ** 2 + (c[ys] - ctr[1]) ** 2 + (c[zs] - ctr[2]) ** 2) <= r * r
        fill[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] |= inside
    fill &= _interior((n, n, n))
    return (n, n, n), h, mask_requests(fill, 1)


def tank_batch(count=64, seed=4):
    """Batched workload: `count` tanks, each axis uniform in {32, 40, ..., 128} cells (h = 1/64 m), 2-6 static
    sphere colliders each, half-filled below a random tilted plane. Returns [(grid, h, requests, colliders)]."""
    import numpy as np

    rng = np.random.default_rng(seed)
    h = 1.0 / 64
    out = []
    for t in range(count):
        grid = tuple(int(v) for v in rng.choice(np.arange(32, 129, 8), 3))
        nx, ny, nz = grid
        ext = np.array([nx, ny, nz]) * h
        cols = []
        for _ in range(int(rng.integers(2, 7))):
            r = float(rng.uniform(0.08, 0.2)) * float(ext.min())
            ctr = [float(rng.uniform(r + 2 * h, e - r - 2 * h)) for e in ext]
            cols.append(abi.sphere_collider(ctr, r))
        a, b = rng.uniform(-0.3, 0.3, 2)
        xc = (np.arange(nx) + 0.5) * h
        yc = (np.arange(ny) + 0.5) * h
        zc = (np.arange(nz) + 0.5) * h
        surf = 0.5 * ext[1] + a * (xc[None, :] - 0.5 * ext[0]) + b * (zc[:, None] - 0.5 * ext[2])  # (z, x)
        mask = yc[None, :, None] < surf[:, None, :]
        for col in cols:
            ctr, r = np.array(col.center[:]), col.extent[0]
            mask &= ((xc[None, None, :] - ctr[0]) ** 2 + (yc[None, :, None] - ctr[1]) ** 2 +
                     (zc[:, None, None] - ctr[2]) ** 2) > r * r
        mask &= _interior((nz, ny, nx))
        out.append(((nx, ny, nz), h, mask_requests(mask, 1 + 100000 * t), cols))
    return out
