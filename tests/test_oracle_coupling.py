"""Pins of the oracle's coupling restatement (world.rs:400-773) by closed
forms, brute force and invariants, independently of the GPU.

Each check is chosen so that a plausible slip in the restatement fails it:
a wrong quaternion formula fails the rotation-matrix check, a dropped
velocity projection or a wrong support point fails the push-out closed
forms, a wrong impulse or buoyancy term fails the single-contact body
velocity, a wrong neighbourhood or weight fails the brute-force couple_points
check, and a wrong interpolation or cell index fails the linear-field and
per-cell source checks.
"""
import numpy as np
import pytest

import oracle as O

G = np.array([0.0, -9.81, 0.0])


def _world(pos, vel=None, mass=0.02, **kw):
    p = O.Params.defaults()
    for k, v in kw.items():
        if k in ("gravity", "bounds_min", "bounds_max"):
            for a in range(3):
                getattr(p, k)[a] = v[a]
        else:
            setattr(p, k, v)
    w = O.World(p)
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    w.set_particles(pos, None if vel is None else np.asarray(vel, dtype=np.float64).reshape(-1, 3),
                    np.full(len(pos), mass))
    return w


def _rotmat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_quaternion_rotation_matches_rotation_matrix():
    rng = np.random.default_rng(1)
    for _ in range(200):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        v = rng.normal(size=3) * 3.0
        assert np.allclose(O.quat_rotate(q, v), _rotmat(q) @ v, rtol=0, atol=1e-14)


def test_push_out_of_a_static_sphere_lands_on_the_surface():
    c = np.array([0.3, -0.2, 0.1])
    r = 0.5
    p0 = c + np.array([0.1, 0.2, -0.05])
    v0 = np.array([-1.0, 0.5, 2.0])
    w = _world([p0], [v0])
    vb = w.couple_bodies([[*c, 1, 0, 0, 0]], [[0, 0, 0]], [0.0], [O.SHAPES["sphere"]], [[r, 0, 0]], 0.01)
    assert np.array_equal(vb, np.zeros((1, 3)))
    p1, v1 = w.get("pos")[0], w.get("vel")[0]
    n = (p0 - c) / np.linalg.norm(p0 - c)
    assert np.allclose(p1, c + r * n, atol=1e-15)
    vn = v0 @ n
    expect = v0 - n * vn if vn < 0 else v0
    assert np.allclose(v1, expect, atol=1e-15)


def test_push_out_of_a_static_box_goes_to_the_corner_of_the_local_octant():
    # support_local of a box is the corner with the direction's signs
    half = np.array([1.0, 0.5, 2.0])
    c = np.array([0.0, 1.0, 0.0])
    p0 = c + np.array([0.3, -0.1, 0.0])  # dir.z == 0 takes +half.z
    v0 = np.array([0.0, 3.0, 0.0])
    w = _world([p0], [v0])
    w.couple_bodies([[*c, 1, 0, 0, 0]], [[0, 0, 0]], [0.0], [O.SHAPES["box"]], [half], 0.01)
    corner = c + np.array([half[0], -half[1], half[2]])
    assert np.allclose(w.get("pos")[0], corner, atol=1e-15)
    n = np.array([half[0], -half[1], half[2]]) / np.linalg.norm(half)
    vn = v0 @ n
    assert vn < 0
    assert np.allclose(w.get("vel")[0], v0 - n * vn, atol=1e-15)


def test_dynamic_sphere_single_contact_impulse_and_buoyancy():
    r, m_p, m_b, dt = 0.5, 0.02, 3.0, 0.01
    p0 = np.array([0.1, 0.0, 0.0])
    v0 = np.array([-1.0, 0.0, 0.0])
    vb0 = np.array([0.2, 0.0, 0.0])
    p = O.Params.defaults()
    w = O.World(p)
    w.set_particles(p0[None], v0[None], [m_p])
    vb = w.couple_bodies([[0, 0, 0, 1, 0, 0, 0]], [vb0], [1 / m_b], [O.SHAPES["sphere"]], [[r, 0, 0]], dt)[0]
    n = np.array([1.0, 0.0, 0.0])
    v_rel_n = (v0 - vb0) @ n
    j = -v_rel_n / (1 / m_p + 1 / m_b)
    vb_contact = vb0 - n * j / m_b
    vol = 4.0 / 3.0 * np.pi * r ** 3
    buoy = -np.array(p.gravity) * (p.rest_density * vol)
    assert np.allclose(vb, vb_contact + buoy * dt / m_b, rtol=1e-14, atol=1e-15)
    # the particle ends with push_out's result (relative to the body's new
    # velocity); the contact impulse itself does not reach it (world.rs:488)
    v_rel = v0 - vb_contact
    expect = v0 - n * (v_rel @ n)
    assert np.allclose(w.get("vel")[0], expect, atol=1e-15)
    assert np.allclose(w.get("pos")[0], [r, 0, 0], atol=1e-15)


def test_no_contact_means_no_buoyancy_and_no_change():
    w = _world([[5.0, 5.0, 5.0]], [[1.0, 0, 0]])
    vb = w.couple_bodies([[0, 0, 0, 1, 0, 0, 0]], [[0.1, 0.2, 0.3]], [0.5], [O.SHAPES["box"]],
                         [[1, 1, 1]], 0.01)
    assert np.array_equal(vb, [[0.1, 0.2, 0.3]])
    assert np.array_equal(w.get("pos"), [[5.0, 5.0, 5.0]])


def test_rotated_box_equals_box_with_swapped_extents():
    rng = np.random.default_rng(2)
    pos = rng.uniform(-1.2, 1.2, size=(400, 3))
    vel = rng.normal(size=(400, 3))
    a = _world(pos, vel)
    b = _world(pos, vel)
    s = np.sqrt(0.5)
    # 90 degrees about y: local x <- world -z... the box (1, 0.5, 0.3) rotated
    # has world extents (0.3, 0.5, 1)
    a.couple_bodies([[0, 0, 0, s, 0, s, 0]], [[0, 0, 0]], [0.0], [O.SHAPES["box"]], [[1.0, 0.5, 0.3]], 0.01)
    b.couple_bodies([[0, 0, 0, 1, 0, 0, 0]], [[0, 0, 0]], [0.0], [O.SHAPES["box"]], [[0.3, 0.5, 1.0]], 0.01)
    moved_a = np.any(a.get("pos") != pos, axis=1)
    moved_b = np.any(b.get("pos") != pos, axis=1)
    assert moved_a.sum() > 10
    assert np.array_equal(moved_a, moved_b)
    # both push every contained particle onto a corner of the same world box
    pa = np.abs(a.get("pos")[moved_a])
    assert np.allclose(pa, [0.3, 0.5, 1.0], atol=1e-15)


def test_capsule_containment_and_support():
    hh, r = 0.6, 0.25
    pts = np.array([[0.0, 0.8, 0.0], [0.2, 0.0, 0.1], [0.0, -0.86, 0.0], [0.3, 0.0, 0.0]])
    w = _world(pts, np.zeros_like(pts))
    w.couple_bodies([[0, 0, 0, 1, 0, 0, 0]], [[0, 0, 0]], [0.0], [O.SHAPES["capsule"]], [[hh, r, 0]], 0.01)
    out = w.get("pos")
    # inside: the top cap point goes to the top end + r along its direction
    assert np.allclose(out[0], [0.0, hh + r, 0.0], atol=1e-15)
    # dir.y == 0: the support point sits on the segment's midpoint circle
    d = pts[1] / np.linalg.norm(pts[1])
    assert np.allclose(out[1], r * d, atol=1e-15)
    # outside: untouched
    assert np.array_equal(out[2], pts[2]) and np.array_equal(out[3], pts[3])


def _keys(x, h):
    return np.floor(x * (1.0 / h)).astype(np.int64)


def test_couple_points_matches_brute_force_neighbourhood():
    rng = np.random.default_rng(3)
    n = 600
    pos = rng.uniform(-0.6, 0.6, size=(n, 3))
    vel = rng.normal(size=(n, 3))
    mass = rng.uniform(0.01, 0.03, size=n)
    p = O.Params.defaults()
    w = O.World(p)
    w.set_particles(pos, vel, mass)
    q = np.array([[0.05, -0.1, 0.2]])
    qv = np.array([[0.3, -0.2, 0.1]])
    m, dt, drag, soft = 0.05, 0.01, 1.5, 800.0
    f = w.couple_points(q, qv, [m], np.zeros((1, 3)), dt, drag, soft)[0]
    kq = _keys(q[0], p.h)
    nb = np.all(np.abs(_keys(pos, p.h) - kq) <= 1, axis=1)
    v_f = (vel[nb] * mass[nb, None]).sum(0) / mass[nb].sum()
    fb = -np.array(p.gravity) * (p.rest_density * (m / soft))
    expect = fb + (v_f - qv[0]) * (drag * m)
    assert np.allclose(f, expect, rtol=1e-13, atol=1e-13)
    dv = w.get("vel") - vel
    assert np.allclose(dv[~nb], 0.0, atol=0)
    assert np.allclose(dv[nb], -f * dt * (mass[nb] / mass[nb].sum())[:, None], rtol=1e-12, atol=1e-15)


def test_couple_points_outside_bounds_is_untouched_and_equal_masses_share_evenly():
    pos = np.array([[0.0, 0.0, 0.0], [0.05, 0.0, 0.0], [0.0, 0.1, 0.0]])
    w = _world(pos, np.zeros_like(pos), bounds_min=[-1, -1, -1], bounds_max=[1, 1, 1])
    f = w.couple_points([[2.0, 0.0, 0.0]], [[0, 0, 0]], [1.0], [[7.0, 8.0, 9.0]], 0.01, 1.0, 1000.0)
    assert np.array_equal(f, [[7.0, 8.0, 9.0]])
    assert np.array_equal(w.get("vel"), np.zeros((3, 3)))
    f = w.couple_points([[0.01, 0.01, 0.0]], [[0, 0, 0]], [1.0], [[0, 0, 0]], 0.01, 0.0, 1000.0)[0]
    # buoyancy only (drag 0, fluid at rest): every particle gets -f dt / 3
    assert np.allclose(f, [0.0, 9.81, 0.0], rtol=1e-15)
    assert np.allclose(w.get("vel"), np.tile(-f * 0.01 / 3, (3, 1)), rtol=1e-15)


def test_couple_heat_linear_field_and_sources():
    rng = np.random.default_rng(4)
    n = 500
    pos = rng.uniform(-3.0, 3.0, size=(n, 3))
    pos[:5] = [[10, 0, 0], [0, -9, 0], [4.9, 4.9, 4.9], [-4.0, 0, 0], [0, 0, 4.2]]
    vel = rng.normal(size=(n, 3))
    w = _world(pos, vel, gravity=G)
    nx, ny, nz, dx = 9, 7, 8, 1.0
    origin = np.array([-4.0, -3.5, -3.7])
    xs = origin[0] + dx * np.arange(nx)
    ys = origin[1] + dx * np.arange(ny)
    zs = origin[2] + dx * np.arange(nz)
    a, b, c, d = 0.3, 0.7, -0.2, 0.45
    u = a + b * xs[None, None, :] + c * ys[None, :, None] + d * zs[:, None, None]
    src0 = rng.uniform(0, 1, size=(nz, ny, nx))
    t_ref, beta, gain, dt = 0.25, 0.5, 0.1, 0.01
    src = w.couple_heat(u, origin, dx, src0, dt, t_ref, beta, gain)
    ba = w.get("body_acc")
    hi = origin + dx * (np.array([nx, ny, nz]) - 1)
    inb = np.all((pos >= -5) & (pos <= 5), axis=1)
    ing = np.all((pos >= origin) & (pos <= hi), axis=1)
    temp = np.where(ing, a + b * pos[:, 0] + c * pos[:, 1] + d * pos[:, 2], t_ref)
    expect = -G[None, :] * beta * (temp - t_ref)[:, None]
    expect[~inb] = 0.0
    assert np.allclose(ba, expect, rtol=1e-12, atol=1e-12)
    exp_src = src0.copy()
    sel = inb & ing
    cell = np.floor((pos[sel] - origin) / dx).astype(int)
    q = gain * np.linalg.norm(vel[sel], axis=1) * dt
    np.add.at(exp_src, (cell[:, 2], cell[:, 1], cell[:, 0]), q)
    assert np.allclose(src, exp_src, rtol=1e-13, atol=1e-15)


def test_nearest_velocity_brute_force_ties_and_nan():
    rng = np.random.default_rng(5)
    pos = np.round(rng.uniform(-1, 1, size=(300, 3)), 1)  # many exact ties
    vel = rng.normal(size=(300, 3))
    w = _world(pos, vel)
    q = np.round(rng.uniform(-2, 2, size=(200, 3)), 1)
    q[0] = [np.nan, 0, 0]
    q[1] = [np.inf, 0, 0]
    got = w.nearest_velocity(q)
    d = ((q[:, None, :] - pos[None, :, :]) ** 2).sum(-1)
    # first strict minimum below +inf; particle 0 when there is none
    idx = np.argmin(np.where(np.isnan(d), np.inf, d), 1)
    assert np.array_equal(got, vel[idx])
