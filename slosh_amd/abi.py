"""ctypes mirror of include/slosh.h (the C ABI of the solver library).

The structures here must match the header byte for byte; ``tests/test_abi.py`` checks their sizes against the
library's own ``static_assert``s by calling ``slosh_default_settings`` and comparing field values.
"""
import ctypes as C

import numpy as np

SLOSH_OK = 0
SLOSH_EINVAL = 1
SLOSH_ENOMEM = 2
SLOSH_EDEVICE = 3
SLOSH_ECAPACITY = 4
SLOSH_ESTATE = 5

STATUS_NAMES = {0: "OK", 1: "EINVAL", 2: "ENOMEM", 3: "EDEVICE", 4: "ECAPACITY", 5: "ESTATE"}

SOLVER_MGPCG = 0
SOLVER_RBSOR = 1
SOLVER_NONE = 2
SOLVER_FCG_UNET = 3

FLAG_CAPTURE_RESIDUALS = 1
FLAG_MG_FP16 = 2
FLAG_MG_FLUID_COARSENING = 4
FLAG_WARM_START_EXTRAPOLATE = 8
FLAG_REPORT_FP64_ERROR = 16
FLAG_CAPTURE_VCYCLE_RESIDUALS = 32
FLAG_DENSITY_PROJECTION = 64
FLAG_BINARY_COUPLING = 128
COLLIDER_TWO_WAY = 1

EMIT_FILL_CELLS = 0
EMIT_SPHERE = 1
EMIT_BOX = 2

LAYOUT_F32_64 = 0
LAYOUT_REF_48 = 1

FIELD_CELL_FLAGS = 0
FIELD_PRESSURE = 1
FIELD_DIVERGENCE = 2
FIELD_CELL_DIAG = 3
FIELD_VEL = 4
FIELD_FACE_FLAGS = 5
FIELD_FACE_WEIGHT = 6
FIELD_ACCUM = 7
FIELD_BRICK_TABLE = 9
FIELD_PRESSURE_F64 = 10
FIELD_DENSITY_Q = 11
FIELD_DENSITY_RHS = 12
FIELD_DENSITY_DISP = 13

# (numpy dtype, planes) of each dense grid field
FIELD_INFO = {
    FIELD_CELL_FLAGS: (np.uint32, 1),
    FIELD_PRESSURE: (np.float32, 1),
    FIELD_DIVERGENCE: (np.float32, 1),
    FIELD_CELL_DIAG: (np.float32, 1),
    FIELD_VEL: (np.float32, 3),
    FIELD_FACE_FLAGS: (np.uint32, 3),
    FIELD_FACE_WEIGHT: (np.float32, 3),
    FIELD_ACCUM: (np.int32, 6),
    FIELD_PRESSURE_F64: (np.float64, 1),
    FIELD_DENSITY_Q: (np.float32, 1),
    FIELD_DENSITY_RHS: (np.float32, 1),
    FIELD_DENSITY_DISP: (np.float32, 3),
}

ORDER_DENSE = 0

STAGE_SORT = 1
STAGE_POOL = 2
STAGE_P2G = 3
STAGE_NORMALIZE = 4
STAGE_DIVERGENCE = 5
STAGE_SOLVE = 6
STAGE_PROJECT = 7
STAGE_EXTRAPOLATE = 8
STAGE_ALL = 9

MAX_COLLIDERS = 64

CELL_AIR, CELL_FLUID, CELL_SOLID = 0, 1, 2


class Settings(C.Structure):
    _fields_ = [
        ("grid_dim", C.c_int32 * 3),
        ("cell_size", C.c_float),
        ("particle_capacity", C.c_uint32),
        ("sparse_storage_budget", C.c_float),
        ("pressure_solver", C.c_int32),
        ("tolerance", C.c_float),
        ("max_iterations", C.c_int32),
        ("sor_weight", C.c_float),
        ("ghost_fluid", C.c_int32),
        ("extrapolation_layers", C.c_int32),
        ("boundary_friction", C.c_float),
        ("max_speed", C.c_float),
        ("particle_radius_cells", C.c_float),
        ("density_correction", C.c_float),
        ("density_deadband", C.c_float),
        ("flags", C.c_uint32),
        ("mg_coarse_sweeps", C.c_int32),
        ("density_relaxation", C.c_float),
        ("density_projection_iterations", C.c_int32),
        ("density_projection_interval", C.c_int32),
        ("kinematic_viscosity", C.c_float),
        ("viscosity_iterations", C.c_int32),
        ("fluid_density", C.c_float),
        ("reserved", C.c_int32 * 7),
    ]


class EmitRequest(C.Structure):
    _fields_ = [
        ("shape", C.c_uint32),
        ("seed", C.c_uint32),
        ("center", C.c_float * 3),
        ("extent", C.c_float * 3),
        ("velocity", C.c_float * 3),
        ("rotation", C.c_float * 4),
        ("velocity_spread", C.c_float),
        ("speed_jitter", C.c_float),
        ("cell_min", C.c_int32 * 3),
        ("cell_dims", C.c_int32 * 3),
        ("count", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class Collider(C.Structure):
    _fields_ = [
        ("center", C.c_float * 3),
        ("type", C.c_uint32),
        ("rotation", C.c_float * 4),
        ("extent", C.c_float * 3),
        ("flags", C.c_uint32),
        ("linear_velocity", C.c_float * 3),
        ("pad0", C.c_float),
        ("angular_velocity", C.c_float * 3),
        ("pad1", C.c_float),
    ]


class Results(C.Structure):
    _fields_ = [
        ("sim_frame", C.c_uint64),
        ("live_particles", C.c_uint32),
        ("fluid_cells", C.c_uint32),
        ("active_bricks", C.c_uint32),
        ("pool_overflow_bricks", C.c_uint32),
        ("max_particle_speed", C.c_float),
        ("max_divergence", C.c_float),
        ("mean_divergence", C.c_float),
        ("clamped_particles", C.c_uint32),
        ("solver_iterations", C.c_int32),
        ("converged", C.c_int32),
        ("residual_rel", C.c_double),
        ("rhs_norm", C.c_double),
        ("step_ms", C.c_float),
        ("solve_ms", C.c_float),
        ("fallback_used", C.c_int32),
        ("true_residual_rel", C.c_float),
        ("fp64_error_rel", C.c_double),
        ("drained_particles", C.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class ColliderFeedback(C.Structure):
    """slosh_collider_feedback (FSloshColliderFeedback, SloshTypes.h:197-206), 64 bytes."""
    _fields_ = [
        ("pressure_force", C.c_float * 3),
        ("pressure_torque", C.c_float * 3),
        ("mean_fluid_velocity", C.c_float * 3),
        ("submerged_centroid_offset", C.c_float * 3),
        ("wetted_faces", C.c_int32),
        ("submerged_cells", C.c_int32),
        ("solid_cells", C.c_int32),
        ("reserved", C.c_int32),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if name != "reserved":
                out[name] = list(v) if isinstance(v, C.Array) else v
        return out


def decode_coupling(raw):
    """The decode of slosh_get_collider_feedback (SloshProxy.cpp:755-768, one substep) applied to raw int32 sums
    [n][16] in 1/1024 units, with the library's FP32 operations."""
    out = []
    inv = np.float32(1.0 / 1024.0)
    for row in np.asarray(raw, dtype=np.int32).reshape(-1, 16):
        f = ColliderFeedback()
        c = row.astype(np.float32)
        for a in range(3):
            f.pressure_force[a] = float(c[a] * inv)
            f.pressure_torque[a] = float(c[3 + a] * inv)
            f.mean_fluid_velocity[a] = float(c[8 + a] * (inv / c[6])) if row[6] > 0 else 0.0
            f.submerged_centroid_offset[a] = float(c[12 + a] * (inv / c[7])) if row[7] > 0 else 0.0
        f.wetted_faces = int(row[6])
        f.submerged_cells = int(row[7])
        f.solid_cells = int(row[11])
        out.append(f)
    return out


class Manipulator(C.Structure):
    _fields_ = [
        ("center", C.c_float * 3),
        ("kind", C.c_uint32),
        ("rotation", C.c_float * 4),
        ("extent", C.c_float * 3),
        ("shape", C.c_uint32),
        ("axis", C.c_float * 3),
        ("strength", C.c_float),
        ("falloff", C.c_float),
        ("param0", C.c_float),
        ("param1", C.c_float),
        ("pad", C.c_uint32),
    ]


MANIP_DRAIN, MANIP_DIRECTIONAL, MANIP_RADIAL, MANIP_SWIRL, MANIP_DRAG = 0, 1, 2, 3, 4
SHAPE_SPHERE, SHAPE_BOX, SHAPE_CAPSULE, SHAPE_EVERYWHERE = 0, 1, 2, 3
MAX_MANIPULATORS = 32


def manipulator(kind, shape, center=(0.0, 0.0, 0.0), extent=(0.0, 0.0, 0.0), strength=0.0, axis=(0.0, 0.0, 1.0),
                falloff=0.0, param0=0.0, rotation=(0.0, 0.0, 0.0, 1.0)):
    m = Manipulator()
    m.kind = int(kind)
    m.shape = int(shape)
    for a in range(3):
        m.center[a] = float(center[a])
        m.extent[a] = float(extent[a])
        m.axis[a] = float(axis[a])
    for k in range(4):
        m.rotation[k] = float(rotation[k])
    m.strength = float(strength)
    m.falloff = float(falloff)
    m.param0 = float(param0)
    return m


def shape_request(shape, center, extent, count, seed, velocity=(0.0, 0.0, 0.0), rotation=(0.0, 0.0, 0.0, 1.0),
                  spread=0.0, jitter=0.0):
    """Sphere (shape 1, radius extent[0]) or box (shape 2, half extents) emission of `count` particles."""
    r = EmitRequest()
    r.shape = int(shape)
    r.seed = int(seed) & 0xFFFFFFFF
    for a in range(3):
        r.center[a] = float(center[a])
        r.extent[a] = float(extent[a])
        r.velocity[a] = float(velocity[a])
    for k in range(4):
        r.rotation[k] = float(rotation[k])
    r.velocity_spread = float(spread)
    r.speed_jitter = float(jitter)
    r.count = int(count)
    return r


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint32), ("pad", C.c_uint32), ("total_ms", C.c_double)]


class BatchConfig(C.Structure):
    _fields_ = [("workers", C.c_uint32), ("solver_chunk", C.c_uint32), ("graphs", C.c_uint32),
                ("mode", C.c_uint32), ("col_bricks", C.c_uint32)]


class UNetTrainConfig(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("batch", C.c_uint32), ("learning_rate", C.c_float), ("offset", C.c_uint32)]


class UNetTrainReport(C.Structure):
    _fields_ = [("first_loss", C.c_double), ("last_loss", C.c_double), ("systems", C.c_uint32),
                ("adam_steps", C.c_uint32)]


PARTICLE_DTYPE = np.dtype([("pos", np.float32, 3), ("age", np.float32), ("vel", np.float32, 3), ("c", np.float32, 9)])
PARTICLE_REF48_DTYPE = np.dtype([("pos", np.float32, 3), ("age", np.float32), ("vel", np.float32, 3), ("c", np.uint32, 5)])

assert C.sizeof(Collider) == 80
assert C.sizeof(Settings) == 128
assert C.sizeof(EmitRequest) == 100
assert C.sizeof(Results) == 96
assert PARTICLE_DTYPE.itemsize == 64 and PARTICLE_REF48_DTYPE.itemsize == 48


def make_settings(grid, cell_size, capacity, solver=SOLVER_MGPCG, tolerance=1e-6, max_iterations=200, sor_weight=1.8,
                  budget=1.0, extrapolation_layers=2, ghost_fluid=1, friction=0.0, max_speed=20.0, flags=0,
                  mg_coarse_sweeps=0, density_correction=0.0, density_relaxation=0.5, density_iterations=30,
                  density_interval=1, viscosity=0.0, viscosity_iterations=12, fluid_density=1000.0):
    """Parity-profile settings (APIC, ghost fluid, no density correction) with the given grid and solver."""
    s = Settings()
    for a in range(3):
        s.grid_dim[a] = int(grid[a])
    s.cell_size = float(cell_size)
    s.particle_capacity = int(capacity)
    s.sparse_storage_budget = float(budget)
    s.pressure_solver = int(solver)
    s.tolerance = float(tolerance)
    s.max_iterations = int(max_iterations)
    s.sor_weight = float(sor_weight)
    s.ghost_fluid = int(ghost_fluid)
    s.extrapolation_layers = int(extrapolation_layers)
    s.boundary_friction = float(friction)
    s.max_speed = float(max_speed)
    s.particle_radius_cells = 0.25
    s.density_correction = float(density_correction)
    s.density_deadband = 0.05
    s.flags = int(flags)
    s.mg_coarse_sweeps = int(mg_coarse_sweeps)
    s.density_relaxation = float(density_relaxation)
    s.density_projection_iterations = int(density_iterations)
    s.density_projection_interval = int(density_interval)
    s.kinematic_viscosity = float(viscosity)
    s.viscosity_iterations = int(viscosity_iterations)
    s.fluid_density = float(fluid_density)
    return s


def fill_request(cell_min, cell_dims, seed, velocity=(0.0, 0.0, 0.0)):
    r = EmitRequest()
    r.shape = EMIT_FILL_CELLS
    r.seed = int(seed) & 0xFFFFFFFF
    for a in range(3):
        r.cell_min[a] = int(cell_min[a])
        r.cell_dims[a] = int(cell_dims[a])
        r.velocity[a] = float(velocity[a])
    r.rotation[3] = 1.0
    return r


def sphere_collider(center, radius, linear_velocity=(0.0, 0.0, 0.0)):
    c = Collider()
    for a in range(3):
        c.center[a] = float(center[a])
        c.linear_velocity[a] = float(linear_velocity[a])
    c.type = 0
    c.rotation[3] = 1.0
    c.extent[0] = float(radius)
    return c


def collider(kind, center, extent, rotation=(0.0, 0.0, 0.0, 1.0), linear_velocity=(0.0, 0.0, 0.0),
             angular_velocity=(0.0, 0.0, 0.0), flags=0):
    """Analytic collider (FSloshColliderGPU): kind 0 sphere (extent[0] = radius), 1 box (half extents), 2 capsule
    (extent[0] = radius, extent[1] = half length of its segment along the local z axis); rotation is the unit
    quaternion (x, y, z, w) from local to sim space."""
    c = Collider()
    c.type = int(kind)
    for a in range(3):
        c.center[a] = float(center[a])
        c.extent[a] = float(extent[a]) if a < len(extent) else 0.0
        c.linear_velocity[a] = float(linear_velocity[a])
        c.angular_velocity[a] = float(angular_velocity[a])
    for k in range(4):
        c.rotation[k] = float(rotation[k])
    c.flags = int(flags)
    return c
