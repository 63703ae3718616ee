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
