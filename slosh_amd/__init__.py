"""slosh_amd -- MI355X-native per-step liquid solver (FLIP/APIC on sparse 8^3 bricks) behind a C ABI.

The solver is ``libslosh.so`` (hand-written gfx950 HIP kernels + a C++ host library, ``include/slosh.h``). This
module is a thin ctypes binding of that ABI for tests and benchmarks; it never falls back to any CPU code: if the
library is missing or the device is not a gfx950, ``Sim`` raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libslosh.so")

_lib = None


class SloshError(RuntimeError):
    def __init__(self, what, status):
        super().__init__(f"{what}: {abi.STATUS_NAMES.get(status, status)}")
        self.status = status


def load_library(path=LIB_PATH):
    """Loads libslosh.so and declares its C signatures. Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    vp = C.c_void_p
    st = C.c_int
    lib.slosh_abi_version.restype = C.c_uint32
    lib.slosh_build_info.restype = C.c_char_p
    lib.slosh_status_string.restype = C.c_char_p
    lib.slosh_status_string.argtypes = [st]
    lib.slosh_default_settings.argtypes = [C.POINTER(abi.Settings)]
    lib.slosh_default_settings.restype = None
    lib.slosh_create.argtypes = [C.POINTER(abi.Settings), C.POINTER(vp)]
    lib.slosh_create.restype = st
    lib.slosh_destroy.argtypes = [vp]
    lib.slosh_destroy.restype = None
    lib.slosh_reset.argtypes = [vp]
    lib.slosh_reset.restype = st
    lib.slosh_emit.argtypes = [vp, C.POINTER(abi.EmitRequest)]
    lib.slosh_emit.restype = st
    lib.slosh_set_colliders.argtypes = [vp, C.POINTER(abi.Collider), C.c_uint32]
    lib.slosh_set_colliders.restype = st
    lib.slosh_set_manipulators.argtypes = [vp, C.POINTER(abi.Manipulator), C.c_uint32]
    lib.slosh_set_manipulators.restype = st
    lib.slosh_step.argtypes = [vp, C.c_float, C.POINTER(C.c_float)]
    lib.slosh_step.restype = st
    lib.slosh_step_until.argtypes = [vp, C.c_float, C.POINTER(C.c_float), C.c_int32]
    lib.slosh_step_until.restype = st
    lib.slosh_read_particles.argtypes = [vp, vp, C.c_uint64, C.POINTER(C.c_uint32), C.c_int32]
    lib.slosh_read_particles.restype = st
    lib.slosh_read_particles_noalloc.argtypes = [vp, vp, C.c_uint64, C.POINTER(C.c_uint32), C.c_int32]
    lib.slosh_read_particles_noalloc.restype = st
    lib.slosh_read_grid.argtypes = [vp, C.c_int32, vp, C.c_uint64, C.c_int32]
    lib.slosh_read_grid.restype = st
    lib.slosh_get_results.argtypes = [vp, C.POINTER(abi.Results)]
    lib.slosh_load_system.restype = st
    lib.slosh_load_system.argtypes = [vp, vp, vp, vp]
    lib.slosh_solve.restype = st
    lib.slosh_solve.argtypes = [vp]
    lib.slosh_get_collider_feedback.restype = st
    lib.slosh_get_collider_feedback.argtypes = [vp, C.POINTER(abi.ColliderFeedback), C.c_uint32,
                                                C.POINTER(C.c_uint32)]
    lib.slosh_get_results.restype = st
    lib.slosh_device_particles.argtypes = [vp, C.c_int32, C.POINTER(vp), C.POINTER(C.c_uint32)]
    lib.slosh_device_particles.restype = st
    lib.slosh_write_particles.argtypes = [vp, vp, C.c_uint32]
    lib.slosh_write_particles.restype = st
    lib.slosh_write_grid.argtypes = [vp, C.c_int32, vp, C.c_uint64, C.c_int32]
    lib.slosh_write_grid.restype = st
    lib.slosh_set_kernel_timing.argtypes = [vp, C.c_uint64]
    lib.slosh_set_kernel_timing.restype = st
    lib.slosh_kernel_times.argtypes = [vp, C.POINTER(abi.KernelTime), C.c_uint32, C.POINTER(C.c_uint32)]
    lib.slosh_kernel_times.restype = st
    lib.slosh_unet_param_count.argtypes = []
    lib.slosh_unet_param_count.restype = C.c_uint64
    lib.slosh_unet_set_weights.argtypes = [vp, vp, C.c_uint64]
    lib.slosh_unet_set_weights.restype = st
    lib.slosh_unet_get_weights.argtypes = [vp, vp, C.c_uint64]
    lib.slosh_unet_get_weights.restype = st
    lib.slosh_unet_apply.argtypes = [vp, vp, vp]
    lib.slosh_unet_apply.restype = st
    lib.slosh_unet_loss_grad.argtypes = [vp, vp, C.POINTER(C.c_double), vp]
    lib.slosh_unet_loss_grad.restype = st
    lib.slosh_batch_create.argtypes = [C.POINTER(vp), C.c_uint32, C.POINTER(vp)]
    lib.slosh_batch_create.restype = st
    lib.slosh_batch_destroy.argtypes = [vp]
    lib.slosh_batch_destroy.restype = None
    lib.slosh_batch_step.argtypes = [vp, C.c_float, C.POINTER(C.c_float)]
    lib.slosh_batch_step.restype = st
    lib.slosh_batch_set_config.argtypes = [vp, C.POINTER(abi.BatchConfig)]
    lib.slosh_batch_set_config.restype = st
    lib.slosh_batch_get_config.argtypes = [vp, C.POINTER(abi.BatchConfig)]
    lib.slosh_batch_get_config.restype = st
    lib.slosh_batch_autotune.argtypes = [vp, C.c_float, C.POINTER(C.c_float), C.c_uint32, C.POINTER(abi.BatchConfig)]
    lib.slosh_batch_autotune.restype = st
    lib.slosh_unet_read_capture.argtypes = [vp, C.c_uint32, vp, C.c_uint64, C.POINTER(C.c_uint32)]
    lib.slosh_unet_read_capture.restype = st
    lib.slosh_unet_train.argtypes = [vp, C.POINTER(abi.UNetTrainConfig), C.POINTER(abi.UNetTrainReport)]
    lib.slosh_unet_train.restype = st
    _lib = lib
    return lib


def _check(status, what):
    if status != abi.SLOSH_OK:
        raise SloshError(what, status)


def default_settings():
    s = abi.Settings()
    load_library().slosh_default_settings(C.byref(s))
    return s


class Sim:
    """One solver domain on the current HIP device (the C ABI's slosh_ctx)."""

    def __init__(self, settings):
        self.lib = load_library()
        self.settings = settings
        self.dims = tuple(int(settings.grid_dim[a]) for a in range(3))
        h = C.c_void_p()
        _check(self.lib.slosh_create(C.byref(settings), C.byref(h)), "slosh_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.slosh_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        _check(self.lib.slosh_reset(self.h), "slosh_reset")

    def emit(self, req):
        _check(self.lib.slosh_emit(self.h, C.byref(req)), "slosh_emit")

    def set_colliders(self, colliders):
        arr = (abi.Collider * max(len(colliders), 1))(*colliders)
        _check(self.lib.slosh_set_colliders(self.h, arr, len(colliders)), "slosh_set_colliders")

    def set_manipulators(self, manips):
        arr = (abi.Manipulator * max(len(manips), 1))(*manips)
        _check(self.lib.slosh_set_manipulators(self.h, arr, len(manips)), "slosh_set_manipulators")

    def write_particles(self, parts):
        parts = np.ascontiguousarray(parts, dtype=abi.PARTICLE_DTYPE)
        _check(self.lib.slosh_write_particles(self.h, parts.ctypes.data, len(parts)), "slosh_write_particles")

    def write_pressure(self, dense):
        dense = np.ascontiguousarray(dense, dtype=np.float32)
        _check(self.lib.slosh_write_grid(self.h, abi.FIELD_PRESSURE, dense.ctypes.data, dense.nbytes, abi.ORDER_DENSE),
               "slosh_write_grid")

    def write_density_q(self, dense):
        """The density projection's potential q (dense (z, y, x)); a step with SOLVER_NONE keeps it."""
        dense = np.ascontiguousarray(dense, dtype=np.float32)
        _check(self.lib.slosh_write_grid(self.h, abi.FIELD_DENSITY_Q, dense.ctypes.data, dense.nbytes, abi.ORDER_DENSE),
               "slosh_write_grid")

    def step(self, dt, gravity, stop_after=abi.STAGE_ALL):
        g = (C.c_float * 3)(*[float(x) for x in gravity])
        if stop_after == abi.STAGE_ALL:
            _check(self.lib.slosh_step(self.h, C.c_float(dt), g), "slosh_step")
        else:
            _check(self.lib.slosh_step_until(self.h, C.c_float(dt), g, int(stop_after)), "slosh_step_until")

    def results(self):
        r = abi.Results()
        _check(self.lib.slosh_get_results(self.h, C.byref(r)), "slosh_get_results")
        return r

    def load_system(self, types, diag, divergence):
        """Loads a pressure system in dense (z, y, x) arrays over the grid: types 0 air / 1 fluid / 2 solid, the
        operator diagonal and the divergence (slosh_load_system)."""
        shape = tuple(int(self.settings.grid_dim[a]) for a in (2, 1, 0))
        t = np.ascontiguousarray(types, dtype=np.uint8).reshape(shape)
        d = np.ascontiguousarray(diag, dtype=np.float32).reshape(shape)
        v = np.ascontiguousarray(divergence, dtype=np.float32).reshape(shape)
        _check(self.lib.slosh_load_system(self.h, t.ctypes.data, d.ctypes.data, v.ctypes.data), "slosh_load_system")

    def solve(self):
        """Solves the current system with the configured solver (slosh_solve); returns the results."""
        _check(self.lib.slosh_solve(self.h), "slosh_solve")
        return self.results()

    def collider_feedback(self):
        """Per-collider feedback of the last step (slosh_get_collider_feedback), a list of abi.ColliderFeedback."""
        buf = (abi.ColliderFeedback * 64)()
        n = C.c_uint32()
        _check(self.lib.slosh_get_collider_feedback(self.h, buf, 64, C.byref(n)), "slosh_get_collider_feedback")
        return [buf[k] for k in range(n.value)]

    def num_particles(self):
        return int(self.results().live_particles)

    def read_particles(self, layout=abi.LAYOUT_F32_64):
        n = self.num_particles()
        dt = abi.PARTICLE_DTYPE if layout == abi.LAYOUT_F32_64 else abi.PARTICLE_REF48_DTYPE
        out = np.zeros(max(n, 1), dtype=dt)
        cnt = C.c_uint32(0)
        _check(self.lib.slosh_read_particles(self.h, out.ctypes.data, out.nbytes, C.byref(cnt), layout),
               "slosh_read_particles")
        return out[: cnt.value]

    def read_particles_noalloc(self, layout=abi.LAYOUT_F32_64):
        """slosh_read_particles_noalloc: the same records, staged through the library's fixed device buffer."""
        n = self.num_particles()
        dt = abi.PARTICLE_DTYPE if layout == abi.LAYOUT_F32_64 else abi.PARTICLE_REF48_DTYPE
        out = np.zeros(max(n, 1), dtype=dt)
        cnt = C.c_uint32(0)
        _check(self.lib.slosh_read_particles_noalloc(self.h, out.ctypes.data, out.nbytes, C.byref(cnt), layout),
               "slosh_read_particles_noalloc")
        return out[: cnt.value]

    def read_grid(self, field):
        nx, ny, nz = self.dims
        if field == abi.FIELD_BRICK_TABLE:
            nb = ((nx + 7) // 8) * ((ny + 7) // 8) * ((nz + 7) // 8)
            out = np.zeros(nb, dtype=np.uint32)
        else:
            dt, planes = abi.FIELD_INFO[field]
            out = np.zeros((planes, nz, ny, nx), dtype=dt)
        _check(self.lib.slosh_read_grid(self.h, field, out.ctypes.data, out.nbytes, abi.ORDER_DENSE), "slosh_read_grid")
        if field != abi.FIELD_BRICK_TABLE and out.shape[0] == 1:
            return out[0]
        return out

    def set_kernel_timing(self, mask):
        """Times the selected kernels (bit i = kernel i of kernel_times(); ~0 all, 0 none) with HIP events."""
        _check(self.lib.slosh_set_kernel_timing(self.h, C.c_uint64(mask & 0xFFFFFFFFFFFFFFFF)), "slosh_set_kernel_timing")

    def kernel_times(self, reset=False):
        """{kernel name: (launches, total ms)} since the last reset; reset=True clears the totals after reading."""
        n = C.c_uint32(0)
        arr = (abi.KernelTime * 64)()
        _check(self.lib.slosh_kernel_times(self.h, arr, 64, C.byref(n)), "slosh_kernel_times")
        out = {arr[k].name.decode(): (int(arr[k].launches), float(arr[k].total_ms)) for k in range(n.value)}
        if reset:
            _check(self.lib.slosh_kernel_times(self.h, None, 0, C.byref(n)), "slosh_kernel_times")
        return out

    def kernel_index(self, name):
        return list(self.kernel_times().keys()).index(name)

    # ---- learned preconditioner ----
    def unet_param_count(self):
        return int(self.lib.slosh_unet_param_count())

    def unet_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        _check(self.lib.slosh_unet_set_weights(self.h, w.ctypes.data, w.size), "slosh_unet_set_weights")

    def unet_get_weights(self):
        w = np.zeros(self.unet_param_count(), np.float32)
        _check(self.lib.slosh_unet_get_weights(self.h, w.ctypes.data, w.size), "slosh_unet_get_weights")
        return w

    def unet_apply(self, r_dense):
        """Preconditioned residual z = M(r) for a dense residual (z, y, x) on the system of the last step."""
        r = np.ascontiguousarray(r_dense, dtype=np.float32)
        z = np.zeros_like(r)
        _check(self.lib.slosh_unet_apply(self.h, r.ctypes.data, z.ctypes.data), "slosh_unet_apply")
        return z

    def unet_loss_grad(self, r_dense):
        r = np.ascontiguousarray(r_dense, dtype=np.float32)
        g = np.zeros(self.unet_param_count(), np.float32)
        loss = C.c_double(0.0)
        _check(self.lib.slosh_unet_loss_grad(self.h, r.ctypes.data, C.byref(loss), g.ctypes.data), "slosh_unet_loss_grad")
        return loss.value, g

    def captured_residual(self, k):
        """Captured residual k of the last solve (dense (z, y, x)), and the number captured."""
        nx, ny, nz = self.dims
        out = np.zeros((nz, ny, nx), np.float32)
        n = C.c_uint32(0)
        _check(self.lib.slosh_unet_read_capture(self.h, int(k), out.ctypes.data, out.nbytes, C.byref(n)),
               "slosh_unet_read_capture")
        return out, n.value

    def unet_train(self, iterations, batch=4, learning_rate=1e-3, offset=0):
        cfg = abi.UNetTrainConfig(int(iterations), int(batch), float(learning_rate), int(offset))
        rep = abi.UNetTrainReport()
        _check(self.lib.slosh_unet_train(self.h, C.byref(cfg), C.byref(rep)), "slosh_unet_train")
        return {"first_loss": rep.first_loss, "last_loss": rep.last_loss, "systems": rep.systems,
                "adam_steps": rep.adam_steps}

    def device_particles(self, layout=abi.LAYOUT_F32_64):
        ptr = C.c_void_p()
        cnt = C.c_uint32(0)
        _check(self.lib.slosh_device_particles(self.h, layout, C.byref(ptr), C.byref(cnt)), "slosh_device_particles")
        return ptr.value, cnt.value


class Batch:
    """Independent tanks (Sim objects) stepped together by the library (slosh_batch_*)."""

    def __init__(self, sims):
        self.lib = load_library()
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.h for s in self.sims])
        self.h = C.c_void_p()
        _check(self.lib.slosh_batch_create(arr, len(self.sims), C.byref(self.h)), "slosh_batch_create")

    def close(self):
        if self.h:
            self.lib.slosh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _grav(self, gravity):
        g = np.asarray(gravity, np.float32)
        if g.shape == (3,):
            g = np.tile(g, (len(self.sims), 1))
        g = np.ascontiguousarray(g.reshape(len(self.sims), 3))
        return g, g.ctypes.data_as(C.POINTER(C.c_float))

    def step(self, dt, gravity):
        g, ptr = self._grav(gravity)
        _check(self.lib.slosh_batch_step(self.h, float(dt), ptr), "slosh_batch_step")

    def set_config(self, workers, solver_chunk=4, graphs=1, mode=0, col_bricks=4):
        cfg = abi.BatchConfig(int(workers), int(solver_chunk), int(graphs), int(mode), int(col_bricks))
        _check(self.lib.slosh_batch_set_config(self.h, C.byref(cfg)), "slosh_batch_set_config")

    def config(self):
        cfg = abi.BatchConfig()
        _check(self.lib.slosh_batch_get_config(self.h, C.byref(cfg)), "slosh_batch_get_config")
        return {"workers": cfg.workers, "solver_chunk": cfg.solver_chunk, "graphs": cfg.graphs, "mode": cfg.mode,
                "col_bricks": cfg.col_bricks}

    def autotune(self, dt, gravity, steps_per_candidate=1):
        g, ptr = self._grav(gravity)
        cfg = abi.BatchConfig()
        _check(self.lib.slosh_batch_autotune(self.h, float(dt), ptr, int(steps_per_candidate), C.byref(cfg)),
               "slosh_batch_autotune")
        return {"workers": cfg.workers, "solver_chunk": cfg.solver_chunk, "graphs": cfg.graphs, "mode": cfg.mode,
                "col_bricks": cfg.col_bricks}
