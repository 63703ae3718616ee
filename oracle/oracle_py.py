"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so). TEST INFRASTRUCTURE ONLY.

The oracle is the checker of the HIP solver: the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` load it; the product library and its Python wrapper never do.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from slosh_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None

CELL_COUNT_FIELD = 8  # oracle-only field: particles per cell after P2G


def build():
    """Compiles the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    lib.osim_create.restype = vp
    lib.osim_create.argtypes = [C.POINTER(abi.Settings)]
    lib.osim_destroy.argtypes = [vp]
    lib.osim_destroy.restype = None
    lib.osim_emit.argtypes = [vp, C.POINTER(abi.EmitRequest)]
    lib.osim_set_colliders.argtypes = [vp, C.POINTER(abi.Collider), C.c_uint32]
    lib.osim_set_manipulators.argtypes = [vp, C.POINTER(abi.Manipulator), C.c_uint32]
    lib.osim_write_particles.argtypes = [vp, vp, C.c_uint32]
    lib.osim_write_pressure.argtypes = [vp, vp]
    lib.osim_write_density_q.argtypes = [vp, vp]
    lib.osim_step_until.argtypes = [vp, C.c_float, C.POINTER(C.c_float), C.c_int]
    lib.osim_num_particles.argtypes = [vp]
    lib.osim_num_particles.restype = C.c_uint32
    lib.osim_read_particles.argtypes = [vp, vp, C.c_uint32]
    lib.osim_read_grid.argtypes = [vp, C.c_int, vp, C.c_uint64]
    lib.osim_get_results.argtypes = [vp, C.POINTER(abi.Results)]
    lib.osim_solve_fp64.argtypes = [vp, C.c_double, C.c_int, vp, C.POINTER(C.c_double)]
    lib.osim_true_residual.argtypes = [vp, vp]
    lib.osim_true_residual.restype = C.c_double
    lib.osim_true_residual_d.argtypes = [vp, vp]
    lib.osim_true_residual_d.restype = C.c_double
    lib.oracle_pcg_hash.argtypes = [C.c_uint32]
    lib.oracle_pcg_hash.restype = C.c_uint32
    lib.oracle_hash_to_unit.argtypes = [C.c_uint32]
    lib.oracle_hash_to_unit.restype = C.c_float
    lib.oracle_collider_sdf.argtypes = [C.POINTER(abi.Collider), C.POINTER(C.c_float)]
    lib.oracle_collider_sdf.restype = C.c_float
    lib.oracle_ghost_theta.argtypes = [C.c_float, C.c_float]
    lib.oracle_ghost_theta.restype = C.c_float
    f3 = C.POINTER(C.c_float)
    lib.oracle_collider_normal.argtypes = [C.POINTER(abi.Collider), f3, C.c_float, f3]
    lib.oracle_collider_normal.restype = None
    lib.oracle_collider_velocity.argtypes = [C.POINTER(abi.Collider), f3, f3]
    lib.oracle_collider_velocity.restype = None
    lib.oracle_collider_response.argtypes = [C.POINTER(abi.Collider), C.c_uint32, C.c_float, C.c_float, C.c_float,
                                             f3, f3]
    lib.oracle_collider_response.restype = None
    lib.osim_coupling_raw.restype = C.c_int
    lib.osim_coupling_raw.argtypes = [vp, vp, C.c_uint32]
    lib.oracle_set_threads.argtypes = [C.c_int]
    lib.oracle_set_threads.restype = None
    lib.oracle_threads.argtypes = []
    lib.oracle_threads.restype = C.c_int
    _lib = lib
    return lib


def _check(status, what):
    if status != 0:
        raise RuntimeError(f"oracle {what} failed: {abi.STATUS_NAMES.get(status, status)}")


class OracleSim:
    """CPU restatement of one solver domain, with the same call surface as slosh_amd.Sim."""

    def __init__(self, settings):
        self.lib = load()
        self.settings = settings
        self.dims = tuple(int(settings.grid_dim[a]) for a in range(3))
        self.h = C.c_void_p(self.lib.osim_create(C.byref(settings)))
        if not self.h.value:
            raise RuntimeError("osim_create failed")

    def close(self):
        if self.h is not None and self.h.value:
            self.lib.osim_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def emit(self, req):
        _check(self.lib.osim_emit(self.h, C.byref(req)), "emit")

    def set_colliders(self, colliders):
        arr = (abi.Collider * max(len(colliders), 1))(*colliders)
        _check(self.lib.osim_set_colliders(self.h, arr, len(colliders)), "set_colliders")

    def set_manipulators(self, manips):
        arr = (abi.Manipulator * max(len(manips), 1))(*manips)
        _check(self.lib.osim_set_manipulators(self.h, arr, len(manips)), "set_manipulators")

    def write_particles(self, parts):
        parts = np.ascontiguousarray(parts, dtype=abi.PARTICLE_DTYPE)
        _check(self.lib.osim_write_particles(self.h, parts.ctypes.data, len(parts)), "write_particles")

    def write_pressure(self, dense):
        dense = np.ascontiguousarray(dense, dtype=np.float32)
        _check(self.lib.osim_write_pressure(self.h, dense.ctypes.data), "write_pressure")

    def write_density_q(self, dense):
        dense = np.ascontiguousarray(dense, dtype=np.float32)
        _check(self.lib.osim_write_density_q(self.h, dense.ctypes.data), "write_density_q")

    def step(self, dt, gravity, stop_after=abi.STAGE_ALL):
        g = (C.c_float * 3)(*[float(x) for x in gravity])
        _check(self.lib.osim_step_until(self.h, C.c_float(dt), g, int(stop_after)), "step")

    def num_particles(self):
        return int(self.lib.osim_num_particles(self.h))

    def read_particles(self):
        n = self.num_particles()
        out = np.zeros(n, dtype=abi.PARTICLE_DTYPE)
        self.lib.osim_read_particles(self.h, out.ctypes.data, n)
        return out

    def read_particles_brick_order(self):
        """The records in the order the next step's sort gives them (the library's order after a full step)."""
        return sort_by_brick(self.read_particles(), self.dims, self.settings.cell_size)

    def read_grid(self, field):
        nx, ny, nz = self.dims
        if field == abi.FIELD_BRICK_TABLE:
            nb = ((nx + 7) // 8) * ((ny + 7) // 8) * ((nz + 7) // 8)
            out = np.zeros(nb, dtype=np.uint32)
            _check(self.lib.osim_read_grid(self.h, field, out.ctypes.data, out.nbytes), "read_grid")
            return out
        if field == CELL_COUNT_FIELD:
            out = np.zeros((nz, ny, nx), dtype=np.int32)
            _check(self.lib.osim_read_grid(self.h, field, out.ctypes.data, out.nbytes), "read_grid")
            return out
        dt, planes = abi.FIELD_INFO[field]
        out = np.zeros((planes, nz, ny, nx), dtype=dt)
        _check(self.lib.osim_read_grid(self.h, field, out.ctypes.data, out.nbytes), "read_grid")
        return out if planes > 1 else out[0]

    def results(self):
        r = abi.Results()
        self.lib.osim_get_results(self.h, C.byref(r))
        return r

    def coupling_raw(self):
        """The last step's collider coupling sums, int32 [colliders][16] in 1/1024 units."""
        out = np.zeros((64, 16), dtype=np.int32)
        n = self.lib.osim_coupling_raw(self.h, out.ctypes.data, 64)
        return out[:n].copy()

    def collider_feedback(self):
        return abi.decode_coupling(self.coupling_raw())

    def solve_fp64(self, rel_tol=1e-13, max_iter=500):
        nx, ny, nz = self.dims
        out = np.zeros((nz, ny, nx), dtype=np.float64)
        rel = C.c_double(0.0)
        iters = self.lib.osim_solve_fp64(self.h, rel_tol, max_iter, out.ctypes.data, C.byref(rel))
        return out, iters, rel.value

    def true_residual(self, dense_p=None):
        """FP64 relative residual of dense_p (float32 dense field), or of the solver's own FP64 pressure."""
        if dense_p is None:
            return float(self.lib.osim_true_residual(self.h, None))
        if np.asarray(dense_p).dtype == np.float64:
            dense_p = np.ascontiguousarray(dense_p)
            return float(self.lib.osim_true_residual_d(self.h, dense_p.ctypes.data))
        dense_p = np.ascontiguousarray(dense_p, dtype=np.float32)
        return float(self.lib.osim_true_residual(self.h, dense_p.ctypes.data))


def sort_by_brick(parts, grid, cell_size):
    """`parts` (PARTICLE_DTYPE records) in the stable order by 8^3 brick key of the clamped cell, dead records
    dropped: the order the reference's sort at the start of the next step gives (SloshSort.usf:65-135, stable as
    restated in oracle.c). The library sorts by the new keys at the end of a step, so after a full step its records
    are the oracle's records in this order."""
    grid = [int(v) for v in grid]
    inv_h = np.float32(1.0) / np.float32(cell_size)
    pos = parts["pos"].astype(np.float32)
    cell = np.floor(pos * inv_h).astype(np.int64)
    for a in range(3):
        cell[:, a] = np.clip(cell[:, a], 0, grid[a] - 1)
    bd = [(g + 7) // 8 for g in grid]
    key = (cell[:, 0] >> 3) + bd[0] * ((cell[:, 1] >> 3) + bd[1] * (cell[:, 2] >> 3))
    live = parts["age"] >= 0.0
    order = np.argsort(np.where(live, key, bd[0] * bd[1] * bd[2]), kind="stable")
    return parts[order[: int(live.sum())]]


def set_threads(n):
    """Threads of the oracle's parallel loops (results do not depend on the count)."""
    load().oracle_set_threads(int(n))


def threads():
    return int(load().oracle_threads())


def pcg_hash(v):
    return int(load().oracle_pcg_hash(int(v) & 0xFFFFFFFF))


def ghost_theta(rf, ra):
    return float(load().oracle_ghost_theta(rf, ra))


def collider_sdf(col, p):
    pp = (C.c_float * 3)(*[float(x) for x in p])
    return float(load().oracle_collider_sdf(C.byref(col), pp))


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def collider_normal(col, p, eps):
    out = _f3((0, 0, 0))
    load().oracle_collider_normal(C.byref(col), _f3(p), C.c_float(eps), out)
    return np.array(out[:], dtype=np.float32)


def collider_velocity(col, p):
    out = _f3((0, 0, 0))
    load().oracle_collider_velocity(C.byref(col), _f3(p), out)
    return np.array(out[:], dtype=np.float32)


def collider_response(cols, x, v, radius, friction, cell_size):
    """The G2P collider response of one particle: returns the new (position, velocity)."""
    arr = (abi.Collider * max(len(cols), 1))(*cols)
    xx, vv = _f3(x), _f3(v)
    load().oracle_collider_response(arr, len(cols), C.c_float(radius), C.c_float(friction), C.c_float(cell_size),
                                    xx, vv)
    return np.array(xx[:], dtype=np.float32), np.array(vv[:], dtype=np.float32)


def _half_bits(x):
    """IEEE binary16 bits of FP32 values, round to nearest even, after the reference's clamp to +-65000."""
    c = np.clip(np.asarray(x, dtype=np.float32), np.float32(-65000.0), np.float32(65000.0))
    return c.astype(np.float16).view(np.uint16).astype(np.uint32)


def pack_ref48(parts):
    """The reference's 48-byte record of 64-byte FP32 records: SloshCommon.ush:71-88 (PackHalf2, PackParticle). The
    APIC rows Cx, Cy, Cz are stored as nine halves in five uints, two per uint with the first in the low 16 bits, the
    last high half 0; position, age and velocity stay FP32."""
    parts = np.asarray(parts, dtype=abi.PARTICLE_DTYPE)
    out = np.zeros(len(parts), dtype=abi.PARTICLE_REF48_DTYPE)
    out["pos"] = parts["pos"]
    out["age"] = parts["age"]
    out["vel"] = parts["vel"]
    h = _half_bits(parts["c"])
    h = np.concatenate([h, np.zeros((len(parts), 1), dtype=np.uint32)], axis=1)
    out["c"] = h[:, 0::2] | (h[:, 1::2] << np.uint32(16))
    return out


def unpack_ref48(packed):
    """SloshCommon.ush:90-105 (UnpackParticle): the 48-byte record back to 64-byte FP32 records."""
    packed = np.asarray(packed, dtype=abi.PARTICLE_REF48_DTYPE)
    out = np.zeros(len(packed), dtype=abi.PARTICLE_DTYPE)
    out["pos"] = packed["pos"]
    out["age"] = packed["age"]
    out["vel"] = packed["vel"]
    lo = (packed["c"] & np.uint32(0xFFFF)).astype(np.uint16).view(np.float16)
    hi = (packed["c"] >> np.uint32(16)).astype(np.uint16).view(np.float16)
    halves = np.empty((len(packed), 10), dtype=np.float16)
    halves[:, 0::2] = lo
    halves[:, 1::2] = hi
    out["c"] = halves[:, :9].astype(np.float32)
    return out
