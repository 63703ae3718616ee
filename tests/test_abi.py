"""The C ABI boundary without a GPU: the library loads, exports every entry point include/slosh.h declares, its
structures have the sizes the ctypes mirror assumes, and settings are validated before any device call."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

import oracle_py as O
from slosh_amd import abi, load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "slosh.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(slosh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = load_library()
    names = declared_functions()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), f"{n} declared in slosh.h but not exported"


def test_struct_sizes_match_header():
    src = """
#include <stdio.h>
#include "slosh.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(slosh_settings), sizeof(slosh_emit_request), sizeof(slosh_collider),
         sizeof(slosh_results), sizeof(slosh_particle), sizeof(slosh_particle_ref48), sizeof(slosh_kernel_time));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "sz.c")
        exe = os.path.join(d, "sz")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    sizes = [int(x) for x in out]
    assert sizes == [C.sizeof(abi.Settings), C.sizeof(abi.EmitRequest), C.sizeof(abi.Collider),
                     C.sizeof(abi.Results), abi.PARTICLE_DTYPE.itemsize, abi.PARTICLE_REF48_DTYPE.itemsize,
                     C.sizeof(abi.KernelTime)]
    assert sizes[2] == 80 and sizes[4] == 64 and sizes[5] == 48


def test_default_settings_are_the_parity_profile():
    lib = load_library()
    s = abi.Settings()
    lib.slosh_default_settings(C.byref(s))
    assert s.pressure_solver == abi.SOLVER_MGPCG and abs(s.tolerance - 1e-6) < 1e-12
    assert s.ghost_fluid == 1 and s.extrapolation_layers == 2 and s.density_correction == 0.0
    assert list(s.grid_dim) == [64, 64, 64] and s.sparse_storage_budget == 1.0
    assert lib.slosh_abi_version() == 2
    assert b"gfx950" in lib.slosh_build_info()


@pytest.mark.parametrize("field,value", [("cell_size", 0.0), ("sparse_storage_budget", 1.5), ("sor_weight", 2.5),
                                         ("extrapolation_layers", 9), ("pressure_solver", 7), ("tolerance", -1.0),
                                         ("particle_capacity", 0)])
def test_invalid_settings_rejected_before_device(field, value):
    lib = load_library()
    s = abi.make_settings((16, 16, 16), 1 / 16, 1000)
    setattr(s, field, value)
    h = C.c_void_p()
    assert lib.slosh_create(C.byref(s), C.byref(h)) == abi.SLOSH_EINVAL
    assert not h.value
    s = abi.make_settings((16, 16, 16), 1 / 16, 1000)
    s.grid_dim[1] = 3
    assert lib.slosh_create(C.byref(s), C.byref(h)) == abi.SLOSH_EINVAL


def test_create_without_gpu_reports_device_error():
    """Here (no GPU) creation fails cleanly with SLOSH_EDEVICE; on a GPU box it succeeds."""
    lib = load_library()
    s = abi.make_settings((16, 16, 16), 1 / 16, 1000)
    h = C.c_void_p()
    st = lib.slosh_create(C.byref(s), C.byref(h))
    assert st in (abi.SLOSH_OK, abi.SLOSH_EDEVICE)
    if st == abi.SLOSH_OK:
        lib.slosh_destroy(h)
    else:
        assert not h.value
    lib.slosh_destroy(None)  # tolerated
    assert lib.slosh_step(None, C.c_float(0.01), (C.c_float * 3)()) == abi.SLOSH_EINVAL


def test_status_strings():
    lib = load_library()
    for code in range(6):
        assert lib.slosh_status_string(code)


def test_oracle_library_loads_and_rejects_bad_settings():
    lib = O.load()
    s = abi.make_settings((16, 16, 16), 0.0, 1000)
    assert not lib.osim_create(C.byref(s))
