"""Two-way collider coupling: pressure force and torque, wetted faces and their fluid velocity, seepage pressure on
contact faces, and the submerged-cell estimate (Slosh.usf:909-1216, 1316-1370; decode SloshProxy.cpp:755-768).

CPU pins of the oracle against closed forms (z up, gravity -9.8 along z, walls one cell thick, tanks of 32^3 at
h = 1/32 filled to half height):
- a box held under water feels the buoyancy rho g V of its cells within 5% with the smoothed estimator, no lateral
  force or torque, and every one of its cells and faces counts as submerged / wetted;
- the binary estimator reads each face's pressure at the fluid cell's centre, half a cell outside the body, which
  scales the buoyancy of a box H cells tall by (H + 1) / H (the reference's own note, Slosh.usf:1092-1093);
- a box resting on the floor gets its lift only from the seepage pressure of its contact faces, the pressure at
  the centre of its lowest row: rho g A (H - 1/2) h for H rows above the wall; with only a puddle on top, its
  contact is dry, none of its cells counts as submerged and the puddle presses it down by less than its weight;
- a half-submerged box tilted about x feels rho g V_sub through the centre of buoyancy: torque_x = y_b F_z, with
  V_sub and y_b integrated independently by Monte Carlo;
- a body moving with a uniform stream reports the stream's velocity as its mean fluid velocity;
- the force is linear in fluid_density; one-way colliders get no force but keep their cell counts.
GPU parity (marked gpu): with a written pressure (no solve), the fixed-point sums of every collider, decoded, are
bit for bit the oracle's for the smoothed and the binary estimator over sphere, rotated box, capsule, body-body
and body-wall contacts; with MGPCG at 1e-8 the forces agree to 1e-4 of their scale and the counts exactly.
"""
import math

import numpy as np
import pytest

import oracle_py as O
from slosh_amd import abi

N = 32
H = 1.0 / N
RHO_G = 1000.0 * 9.8
G = (0.0, 0.0, -9.8)
DT = 1.0 / 240.0


def settings(solver=abi.SOLVER_MGPCG, flags=0, **kw):
    return abi.make_settings((N, N, N), H, 400000, solver=solver, tolerance=1e-8, max_iterations=400, flags=flags,
                             **kw)


def run(cols, fill_min=(0, 0, 0), fill_dims=(N, N, 16), steps=6, vel=(0.0, 0.0, 0.0), g=G, **kw):
    o = O.OracleSim(settings(**kw))
    o.emit(abi.fill_request(fill_min, fill_dims, 1, velocity=vel))
    o.set_colliders(cols)
    for _ in range(steps):
        o.step(DT, g)
    fb = o.collider_feedback()
    o.close()
    return fb


def box(center, half, rotation=(0.0, 0.0, 0.0, 1.0), flags=abi.COLLIDER_TWO_WAY, **kw):
    return abi.collider(1, center, half, rotation=rotation, flags=flags, **kw)


SUBMERGED = box((0.5, 0.5, 0.25), (0.125, 0.125, 0.0625))  # 8 x 8 x 4 cells, 6 below the surface


def test_submerged_box_feels_its_buoyancy():
    f = run([SUBMERGED])[0]
    V = 0.25 * 0.25 * 0.125
    assert abs(f.pressure_force[2] / (RHO_G * V) - 1.0) < 0.05, f.as_dict()
    assert max(abs(f.pressure_force[0]), abs(f.pressure_force[1])) < 1e-3 * f.pressure_force[2]
    assert max(abs(t) for t in f.pressure_torque) < 1e-2
    assert f.solid_cells == f.submerged_cells == 256 and f.wetted_faces == 2 * 64 + 4 * 32
    assert list(f.submerged_centroid_offset) == [0.0, 0.0, 0.0]


def test_binary_faces_read_the_pressure_half_a_cell_outside():
    f = run([SUBMERGED], flags=abi.FLAG_BINARY_COUPLING)[0]
    V, rows = 0.25 * 0.25 * 0.125, 4
    assert abs(f.pressure_force[2] / (RHO_G * V * (rows + 1) / rows) - 1.0) < 0.02, f.as_dict()
    smooth = run([SUBMERGED])[0]
    assert f.wetted_faces == smooth.wetted_faces and f.submerged_cells == smooth.submerged_cells


def test_box_on_the_floor_gets_seepage_lift():
    # cells z = 0..3; z = 0 is the wall, so 3 body rows stand on it and the contact pressure is that of row z = 1
    f = run([box((0.5, 0.5, 0.0625), (0.125, 0.125, 0.0625))])[0]
    A, rows = 0.25 * 0.25, 3
    assert abs(f.pressure_force[2] / (RHO_G * A * (rows - 0.5) * H) - 1.0) < 0.05, f.as_dict()
    assert f.solid_cells == 64 * rows and f.wetted_faces == 64 + 4 * 8 * rows
    assert abs(f.submerged_centroid_offset[2] - 0.5 * H) < 1e-6  # rows 1..3 about a centre at 2 cells


def test_puddle_on_a_dry_body_presses_down_and_does_not_submerge_it():
    # two layers of liquid on top of a box standing on the floor: every horizontal walk from the body ends in air,
    # so the floor contact is dry (no lift) and no cell counts as submerged (liquid on top must not buoy the body,
    # Slosh.usf:1333-1335); the puddle's weight pushes down, at most rho g V_puddle
    f = run([box((0.5, 0.5, 0.0625), (0.125, 0.125, 0.0625))], fill_min=(12, 12, 4), fill_dims=(8, 8, 2), steps=2)[0]
    assert f.solid_cells == 192 and f.submerged_cells == 0 and f.wetted_faces == 64, f.as_dict()
    assert -RHO_G * 128 * H ** 3 < f.pressure_force[2] < -0.25 * RHO_G * 128 * H ** 3, f.as_dict()
    assert max(abs(f.pressure_force[0]), abs(f.pressure_force[1])) < 1e-3 * abs(f.pressure_force[2])


def test_tilted_box_torque_is_the_buoyancy_through_the_centre_of_buoyancy():
    th = 0.4
    half = np.array([0.15, 0.2, 0.1])
    f = run([box((0.5, 0.5, 0.5), half, rotation=(math.sin(th / 2), 0.0, 0.0, math.cos(th / 2)))])[0]
    rng = np.random.default_rng(0)
    p = rng.uniform(-1.0, 1.0, (1000000, 3)) * half
    y = p[:, 1] * math.cos(th) - p[:, 2] * math.sin(th)
    z = p[:, 1] * math.sin(th) + p[:, 2] * math.cos(th)
    under = z < 0.0  # the centre sits on the surface at z = 0.5
    v_sub = 8.0 * half.prod() * under.mean()
    y_b = y[under].mean()
    F = RHO_G * v_sub
    assert abs(f.pressure_force[2] / F - 1.0) < 0.06, (f.as_dict(), F)
    assert y_b < 0.0 and abs(f.pressure_torque[0] / (y_b * F) - 1.0) < 0.10, (f.as_dict(), y_b * F)
    assert abs(f.pressure_torque[1]) < 0.02 * abs(f.pressure_torque[0])


def test_mean_fluid_velocity_of_a_body_in_a_uniform_stream():
    sphere = abi.collider(0, (0.5, 0.5, 0.5), (0.1, 0.0, 0.0), linear_velocity=(0.5, 0.0, 0.0),
                          flags=abi.COLLIDER_TWO_WAY)
    f = run([sphere], fill_min=(8, 8, 8), fill_dims=(16, 16, 16), steps=2, vel=(0.5, 0.0, 0.0),
            g=(0.0, 0.0, 0.0))[0]
    assert f.wetted_faces > 100
    assert abs(f.mean_fluid_velocity[0] - 0.5) < 0.025 and max(abs(v) for v in f.mean_fluid_velocity[1:]) < 0.01


def test_force_is_linear_in_the_fluid_density_and_one_way_bodies_get_none():
    a = run([SUBMERGED], steps=2)[0]
    b = run([SUBMERGED], steps=2, fluid_density=500.0)[0]
    # each cell's contribution is rounded to 1/1024 N: a few hundred of them leave at most 0.1 N
    assert abs(b.pressure_force[2] - 0.5 * a.pressure_force[2]) < 0.1
    one_way = box((0.5, 0.5, 0.25), (0.125, 0.125, 0.0625), flags=0)
    c = run([one_way], steps=2)[0]
    assert list(c.pressure_force) == [0.0, 0.0, 0.0] and c.wetted_faces == 0
    assert c.solid_cells == c.submerged_cells == 256


def test_feedback_decode_matches_the_reference_rules():
    raw = np.zeros((1, 16), dtype=np.int32)
    raw[0, :6] = [1024, -2048, 512, 3, -5, 7]
    raw[0, 6], raw[0, 8:11] = 4, [4096, -1024, 0]
    raw[0, 7], raw[0, 11], raw[0, 12:15] = 2, 9, [1024, 0, -512]
    f = abi.decode_coupling(raw)[0]
    assert list(f.pressure_force) == [1.0, -2.0, 0.5]
    assert list(f.pressure_torque) == [3 / 1024, -5 / 1024, 7 / 1024]
    assert list(f.mean_fluid_velocity) == [1.0, -0.25, 0.0]
    assert list(f.submerged_centroid_offset) == [0.5, 0.0, -0.25]
    assert (f.wetted_faces, f.submerged_cells, f.solid_cells) == (4, 2, 9)


# ---- GPU parity ----------------------------------------------------------------------------------------------
def parity_colliders():
    q = (0.0, 0.0, math.sin(0.3), math.cos(0.3))
    return [
        abi.collider(0, (0.3, 0.3, 0.45), (0.09, 0.0, 0.0), linear_velocity=(0.2, 0.0, 0.0),
                     flags=abi.COLLIDER_TWO_WAY),                                        # sphere at the surface
        box((0.65, 0.35, 0.09), (0.1, 0.07, 0.06), rotation=q),                          # rotated box on the floor
        box((0.65, 0.58, 0.09), (0.1, 0.1, 0.06)),                                       # touching it (body contact)
        abi.collider(2, (0.3, 0.7, 0.25), (0.05, 0.1, 0.0), rotation=(math.sin(0.5), 0.0, 0.0, math.cos(0.5)),
                     angular_velocity=(0.0, 0.0, 1.0), flags=abi.COLLIDER_TWO_WAY),    # capsule under water
        box((0.8, 0.8, 0.3), (0.06, 0.06, 0.2), flags=0),                                # one-way post
    ]


def feedback_bytes(fb):
    return b"".join(bytes(f) for f in fb)


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [0, abi.FLAG_BINARY_COUPLING])
def test_gpu_coupling_bitexact_with_a_written_pressure(flags):
    from slosh_amd import Sim

    print(f"case coupling bit-exact flags {flags}", flush=True)
    s = settings(solver=abi.SOLVER_NONE, flags=flags)
    gpu, orc = Sim(s), O.OracleSim(s)
    z = (np.arange(N) + 0.5) * H
    rng = np.random.default_rng(11)
    pres = (np.maximum(0.5 - z, 0.0)[:, None, None] * 0.04 + rng.uniform(-1e-3, 1e-3, (N, N, N))).astype(np.float32)
    for sim in (gpu, orc):
        sim.emit(abi.fill_request((0, 0, 0), (N, N, 16), 1, velocity=(0.1, -0.05, 0.0)))
        sim.set_colliders(parity_colliders())
        sim.step(DT, G)
        sim.write_pressure(pres)
        sim.step(DT, G)
    a, b = gpu.collider_feedback(), orc.collider_feedback()
    assert len(a) == len(b) == 5
    for k, (x, y) in enumerate(zip(a, b)):
        assert bytes(x) == bytes(y), (k, x.as_dict(), y.as_dict())
    assert b[1].pressure_force[2] != 0.0 and b[2].wetted_faces > 0 and b[4].solid_cells > 0
    gpu.close()
    orc.close()


@pytest.mark.gpu
def test_gpu_coupling_with_mgpcg_matches_the_oracle():
    from slosh_amd import Sim

    print("case coupling MGPCG 1e-8", flush=True)
    s = settings()
    gpu, orc = Sim(s), O.OracleSim(s)
    for sim in (gpu, orc):
        sim.emit(abi.fill_request((0, 0, 0), (N, N, 16), 1))
        sim.set_colliders(parity_colliders())
        sim.step(DT, G)
    a, b = gpu.collider_feedback(), orc.collider_feedback()
    scale = max(abs(v) for f in b for v in f.pressure_force)
    for k, (x, y) in enumerate(zip(a, b)):
        assert (x.wetted_faces, x.submerged_cells, x.solid_cells) == (y.wetted_faces, y.submerged_cells,
                                                                     y.solid_cells), k
        assert np.allclose(x.pressure_force, y.pressure_force, rtol=0.0, atol=1e-4 * scale), (k, x.as_dict(),
                                                                                               y.as_dict())
    gpu.close()
    orc.close()
