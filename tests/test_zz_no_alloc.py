"""slosh_read_particles_noalloc against slosh_read_particles, through the C ABI.

The allocation-free readback stages the records through one fixed device buffer per GPU instead of the context's
growable staging buffer; its output must equal the allocating entry point's byte for byte, in both record layouts,
on consecutive calls (the second with fewer particles, so stale bytes of the first would show), and a request
larger than the buffer must be refused without copying.
"""
import numpy as np
import pytest

from slosh_amd import Sim, SloshError, abi, scenes

pytestmark = pytest.mark.gpu

G = (0.0, -9.8, 0.0)
DT = 1.0 / 120.0
STAGING_BYTES = 64 << 20


def stepped(n, seed, steps=2):
    grid, h, reqs = scenes.dam_break(n, seed=seed)
    s = abi.make_settings(grid, h, scenes.particle_count(reqs) + 1024, solver=abi.SOLVER_MGPCG)
    sim = Sim(s)
    for r in reqs:
        sim.emit(r)
    for _ in range(steps):
        sim.step(DT, G)
    return sim


def same_bytes(a, b, what):
    assert a.shape == b.shape, (what, a.shape, b.shape)
    assert a.tobytes() == b.tobytes(), what


def test_noalloc_readback_equals_the_allocating_readback():
    print("case noalloc 40 then 24", flush=True)
    big = stepped(40, 3)
    small = stepped(24, 5)
    assert small.num_particles() < big.num_particles()
    for layout in (abi.LAYOUT_F32_64, abi.LAYOUT_REF_48):
        # two calls in a row, the second with fewer particles
        same_bytes(big.read_particles_noalloc(layout), big.read_particles(layout), f"big layout {layout}")
        same_bytes(small.read_particles_noalloc(layout), small.read_particles(layout), f"small layout {layout}")
    # after another step of the small tank the buffer holds its new records, not the big tank's
    small.step(DT, G)
    same_bytes(small.read_particles_noalloc(), small.read_particles(), "small after a step")
    big.close()
    small.close()


def test_noalloc_readback_refuses_requests_beyond_its_buffer():
    print("case noalloc over the staging size", flush=True)
    n = 64
    s = abi.make_settings((n, n, n), 1.0 / n, 1 << 21, solver=abi.SOLVER_NONE)
    sim = Sim(s)
    sim.emit(abi.fill_request((1, 1, 1), (62, 62, 35), 7))  # 1,076,320 particles
    sim.step(DT, G)
    np_ = sim.num_particles()
    assert np_ * 64 > STAGING_BYTES >= np_ * 48
    with pytest.raises(SloshError) as e:
        sim.read_particles_noalloc(abi.LAYOUT_F32_64)
    assert e.value.status == abi.SLOSH_ECAPACITY
    same_bytes(sim.read_particles_noalloc(abi.LAYOUT_REF_48), sim.read_particles(abi.LAYOUT_REF_48), "48-byte")
    sim.close()
