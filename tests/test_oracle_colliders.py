"""Pins the oracle's collider handling against closed forms written independently of its code.

Covered (reference in parentheses): the sphere, box and capsule signed distances (SloshCommon.ush:178-195), the
finite-difference surface normal (SloshCommon.ush:197-205), a moving collider's material velocity
(SloshCommon.ush:232-235) and the solid-face velocity the normalise pass writes (Slosh.usf:232-242), and the G2P
collider response: push-out to the particle radius, removal of the inward relative normal velocity and tangential
friction (Slosh.usf:1578-1591). Rotations are built by scipy from the same quaternions, distances by projection
onto the shape or by brute-force sampling, so a sign, axis or operand-order slip in the oracle fails.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle_py as O
from slosh_amd import abi


def rot_matrix(q):
    return Rotation.from_quat(np.asarray(q, dtype=np.float64)).as_matrix()  # scalar-last (x, y, z, w)


def random_quat(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    return tuple(np.float32(q))


def local(col, p):
    """Sim-space point p in the collider's local frame (float64)."""
    R = rot_matrix(col.rotation[:])
    return R.T @ (np.asarray(p, dtype=np.float64) - np.asarray(col.center[:], dtype=np.float64))


def box_sdf_ref(e, l):
    """Distance to the box [-e, e]: to its nearest point outside, minus the distance to the nearest face inside."""
    e = np.asarray(e, dtype=np.float64)
    nearest = np.clip(l, -e, e)
    if np.any(np.abs(l) > e):
        return float(np.linalg.norm(l - nearest))
    return -float(np.min(e - np.abs(l)))


def capsule_sdf_brute(r, hl, l):
    """Distance to the segment from (0,0,-hl) to (0,0,hl), sampled densely, minus the radius."""
    t = np.linspace(-hl, hl, 20001)
    seg = np.stack([np.zeros_like(t), np.zeros_like(t), t], axis=1)
    return float(np.min(np.linalg.norm(seg - l[None, :], axis=1)) - r)


def test_sphere_sdf_is_distance_to_centre_minus_radius():
    rng = np.random.default_rng(1)
    col = abi.collider(0, (0.4, 0.5, 0.6), (0.15, 0.0, 0.0), rotation=random_quat(rng))
    for p in rng.uniform(0.0, 1.0, size=(200, 3)).astype(np.float32):
        ref = np.linalg.norm(p.astype(np.float64) - np.array([0.4, 0.5, 0.6])) - 0.15
        assert abs(O.collider_sdf(col, p) - ref) < 2e-6


def test_rotated_box_sdf_matches_projection():
    rng = np.random.default_rng(2)
    for _ in range(5):
        q = random_quat(rng)
        ext = (0.22, 0.07, 0.12)
        col = abi.collider(1, (0.5, 0.45, 0.55), ext, rotation=q)
        for p in rng.uniform(0.15, 0.85, size=(200, 3)).astype(np.float32):
            ref = box_sdf_ref(ext, local(col, p))
            assert abs(O.collider_sdf(col, p) - ref) < 3e-6, (q, p)


def test_box_rotation_direction():
    """30 degrees about z maps the box's long local x axis to (cos 30, sin 30, 0) in sim space, not its mirror."""
    c = np.cos(np.radians(15.0))
    s = np.sin(np.radians(15.0))
    col = abi.collider(1, (0.5, 0.5, 0.5), (0.3, 0.02, 0.02), rotation=(0.0, 0.0, s, c))
    a = np.radians(30.0)
    tip = np.array([0.5 + 0.29 * np.cos(a), 0.5 + 0.29 * np.sin(a), 0.5], dtype=np.float32)
    mirror = np.array([0.5 + 0.29 * np.cos(a), 0.5 - 0.29 * np.sin(a), 0.5], dtype=np.float32)
    assert O.collider_sdf(col, tip) < 0.0
    assert O.collider_sdf(col, mirror) > 0.05


def test_rotated_capsule_sdf_matches_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(4):
        q = random_quat(rng)
        r, hl = 0.06, 0.18
        col = abi.collider(2, (0.5, 0.5, 0.5), (r, hl, 0.0), rotation=q)
        for p in rng.uniform(0.2, 0.8, size=(60, 3)).astype(np.float32):
            ref = capsule_sdf_brute(r, hl, local(col, p))
            assert abs(O.collider_sdf(col, p) - ref) < 1e-5, (q, p)


def analytic_normal(col, p):
    """Outward unit normal at the surface point nearest to p (p away from edges and axes)."""
    R = rot_matrix(col.rotation[:])
    l = local(col, p)
    if col.type == 0:
        n = l
    elif col.type == 1:
        e = np.asarray(col.extent[:], dtype=np.float64)
        k = int(np.argmax(np.abs(l) - e))
        n = np.zeros(3)
        n[k] = np.sign(l[k])
    else:
        hl = col.extent[1]
        n = l - np.array([0.0, 0.0, np.clip(l[2], -hl, hl)])
    return R @ (n / np.linalg.norm(n))


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_fd_normal_matches_analytic_normal(kind):
    rng = np.random.default_rng(10 + kind)
    h = 1.0 / 64.0
    q = random_quat(rng)
    ext = {0: (0.12, 0, 0), 1: (0.15, 0.1, 0.12), 2: (0.08, 0.15, 0)}[kind]
    col = abi.collider(kind, (0.5, 0.5, 0.5), ext, rotation=q)
    checked = 0
    for _ in range(400):
        p = (0.5 + rng.uniform(-0.3, 0.3, size=3)).astype(np.float32)
        l = local(col, p)
        if kind == 1:
            # stay two eps away from the box's edges, where the nearest face changes within the stencil
            d = np.abs(l) - np.asarray(ext)
            srt = np.sort(d)
            if srt[2] - srt[1] < 2.5 * h or np.sum(d > 0) > 1:
                continue
        if kind == 2 and np.linalg.norm(l[:2]) < 0.02:
            continue
        if abs(O.collider_sdf(col, p)) > 0.06:
            continue
        n = O.collider_normal(col, p, 0.25 * h)
        ref = analytic_normal(col, p)
        assert abs(np.linalg.norm(n) - 1.0) < 1e-5
        assert np.linalg.norm(n - ref) < 5e-3, (p, n, ref)
        checked += 1
    assert checked > 20


def test_fd_normal_degenerate_point_falls_back_to_z():
    col = abi.collider(0, (0.5, 0.5, 0.5), (0.1, 0, 0))
    # at the centre every central difference cancels: the reference returns (0, 0, 1)
    n = O.collider_normal(col, (0.5, 0.5, 0.5), 0.001)
    assert np.array_equal(n, np.array([0.0, 0.0, 1.0], dtype=np.float32))


def test_collider_velocity_is_linear_plus_omega_cross_r():
    rng = np.random.default_rng(4)
    v = np.array([0.3, -0.2, 0.5])
    w = np.array([1.5, -2.0, 0.7])
    col = abi.collider(1, (0.4, 0.5, 0.6), (0.1, 0.1, 0.1), linear_velocity=v, angular_velocity=w)
    for p in rng.uniform(0, 1, size=(50, 3)).astype(np.float32):
        r = p.astype(np.float64) - np.array([0.4, 0.5, 0.6])
        ref = v + np.array([w[1] * r[2] - w[2] * r[1], w[2] * r[0] - w[0] * r[2], w[0] * r[1] - w[1] * r[0]])
        assert np.allclose(O.collider_velocity(col, p), ref, atol=1e-6)
    # the cross product order: a spin about +z moves the point on +x towards +y
    spin = abi.collider(0, (0.5, 0.5, 0.5), (0.1, 0, 0), angular_velocity=(0.0, 0.0, 2.0))
    assert np.allclose(O.collider_velocity(spin, (0.6, 0.5, 0.5)), [0.0, 0.2, 0.0], atol=1e-7)


def test_response_pushes_out_to_the_radius_along_the_normal():
    h = 1.0 / 64.0
    radius = 0.25 * h
    col = abi.collider(0, (0.5, 0.4, 0.5), (0.1, 0, 0))
    rng = np.random.default_rng(5)
    for _ in range(50):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        depth = rng.uniform(-0.02, radius * 0.99)
        x0 = (np.array([0.5, 0.4, 0.5]) + d * (0.1 + depth)).astype(np.float32)
        x1, _ = O.collider_response([col], x0, (0.0, 0.0, 0.0), radius, 0.0, h)
        # the new position sits at the particle radius from the surface, moved along the analytic normal
        assert abs(O.collider_sdf(col, x1) - radius) < 2e-6
        move = x1.astype(np.float64) - x0
        d0 = x0.astype(np.float64) - np.array([0.5, 0.4, 0.5])
        d0 /= np.linalg.norm(d0)
        assert np.linalg.norm(np.cross(move, d0)) < 1e-3 * max(np.linalg.norm(move), 1e-3)
        assert np.dot(move, d0) > 0.0
    # a particle outside the contact band is left alone
    x0 = np.array([0.5, 0.4 + 0.1 + 2 * radius, 0.5], dtype=np.float32)
    x1, v1 = O.collider_response([col], x0, (0.1, -1.0, 0.2), radius, 0.5, h)
    assert np.array_equal(x1, x0) and np.array_equal(v1, np.array([0.1, -1.0, 0.2], dtype=np.float32))


@pytest.mark.parametrize("friction", [0.0, 0.3, 1.0])
def test_response_removes_inward_velocity_and_applies_friction(friction):
    h = 1.0 / 64.0
    radius = 0.25 * h
    vs_lin = np.array([0.2, 0.1, -0.3])
    col = abi.collider(0, (0.5, 0.4, 0.5), (0.1, 0, 0), linear_velocity=vs_lin)
    # particle just inside the top of the sphere: the normal is +y
    x0 = np.array([0.5, 0.4 + 0.1 - 0.001, 0.5], dtype=np.float32)
    n = np.array([0.0, 1.0, 0.0])
    v_in = np.array([0.7, -2.0, 0.4], dtype=np.float32)
    x1, v1 = O.collider_response([col], x0, v_in, radius, friction, h)
    rel = v_in.astype(np.float64) - vs_lin
    vn = rel @ n
    assert vn < 0.0
    tang = rel - vn * n
    ref = vs_lin + (1.0 - friction) * tang
    assert np.allclose(v1, ref, atol=2e-5), (v1, ref)
    # separating relative motion keeps its velocity whatever the friction
    v_out = np.array([0.7, 2.0, 0.4], dtype=np.float32)
    _, v2 = O.collider_response([col], x0, v_out, radius, friction, h)
    assert np.allclose(v2, v_out, atol=1e-6)


def test_response_uses_the_moving_surface_velocity_at_the_new_position():
    """A spinning sphere: the removed normal velocity is measured against v + w x (x_new - c)."""
    h = 1.0 / 64.0
    radius = 0.25 * h
    w = np.array([0.0, 0.0, 3.0])
    col = abi.collider(0, (0.5, 0.5, 0.5), (0.1, 0, 0), angular_velocity=w)
    x0 = np.array([0.5 + 0.1 - 0.002, 0.5, 0.5], dtype=np.float32)  # inside, normal +x
    v_in = np.array([-1.0, 0.0, 0.0], dtype=np.float32)
    x1, v1 = O.collider_response([col], x0, v_in, radius, 0.0, h)
    assert abs(x1[0] - (0.6 + radius)) < 2e-6
    vs = np.cross(w, x1.astype(np.float64) - 0.5)
    assert vs[1] > 0.25
    rel = v_in - vs
    ref = vs + (rel - (rel @ [1, 0, 0]) * np.array([1.0, 0, 0]))
    assert np.allclose(v1, ref, atol=2e-5), (v1, ref)


def test_solid_faces_take_the_moving_collider_velocity():
    """Normalise (Slosh.usf:604-651 with SolidFaceVelocity :232-241): every face touching a collider cell gets the
    collider's v + w x (face centre - centre) component; wall faces get 0."""
    n = 24
    h = 1.0 / n
    s = abi.make_settings((n, n, n), h, 200000, solver=abi.SOLVER_NONE)
    o = O.OracleSim(s)
    o.emit(abi.fill_request((1, 1, 1), (22, 10, 22), 3))
    v = np.array([0.3, -0.4, 0.25])
    w = np.array([1.2, -0.8, 2.0])
    rng = np.random.default_rng(6)
    ctr = np.array([0.45, 0.42, 0.55])
    col = abi.collider(1, ctr, (0.17, 0.09, 0.13), rotation=random_quat(rng), linear_velocity=v, angular_velocity=w)
    o.set_colliders([col])
    o.step(1.0 / 120.0, (0.0, -9.8, 0.0), abi.STAGE_NORMALIZE)
    flags = o.read_grid(abi.FIELD_CELL_FLAGS)
    vel = o.read_grid(abi.FIELD_VEL)
    solid = (flags & 3) == 2
    cid = (flags >> 8) & 0xFF
    checked = 0
    for a in range(3):
        for z in range(1, n - 1):
            for y in range(1, n - 1):
                for x in range(1, n - 1):
                    c = [x, y, z]
                    ca = list(c)
                    ca[a] -= 1
                    sa, sb = solid[ca[2], ca[1], ca[0]], solid[z, y, x]
                    ida, idb = cid[ca[2], ca[1], ca[0]], cid[z, y, x]
                    if not (sa or sb):
                        continue
                    # the reference takes the velocity of the solid cell (the cell itself when both are solid)
                    which = idb if sb else ida
                    fc = (np.array([x + 0.5, y + 0.5, z + 0.5]))
                    fc[a] = c[a]
                    fc = fc * h
                    if which == 0:
                        ref = 0.0
                    else:
                        ref = (v + np.cross(w, fc - ctr))[a]
                    assert abs(vel[a, z, y, x] - ref) < 2e-6, (a, x, y, z, vel[a, z, y, x], ref)
                    checked += int(which != 0)
    assert checked > 500
