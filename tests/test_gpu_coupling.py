"""GPU coupling passes (world.rs:400-773) against the oracle on the same FP32
state, and the reference's own coupling tests restated
(crates/phy-fluid/src/sph/tests.rs:98-176, 179-368, 470-548).

The coupling arithmetic is f64 on the FP32 particle state, rounded once when
stored, so: body velocities, point forces and heat sources are compared in
f64 (relative 1e-12: the only differences are summation orders); particle
positions, velocities and body_acc are FP32 roundings of the oracle's f64
values (norm-wise 1e-6); sampled velocities are bit-identical (the same
nearest particle, ties included). The reference's Boussinesq closed-loop test
needs the heat field's own diffusion-advection solver, which this library
does not host.
"""
import numpy as np
import pytest

import oracle as O
from physics_amd import ffi, scenes

pytestmark = pytest.mark.gpu

G = (0.0, -9.81, 0.0)


def _params(cls, bounds=3.0, gravity=G):
    p = cls.defaults()
    for a in range(3):
        p.bounds_min[a] = -bounds
        p.bounds_max[a] = bounds
        p.gravity[a] = gravity[a]
    return p


def _settled_block(n=10, steps=12, seed=7):
    g = ffi.FluidWorld(_params(ffi.PhySphParams))
    pos = scenes.lattice_block(n, n, n, scenes.S32, (-0.5, -0.6, -0.5), 0.01, seed)
    m = np.float32(scenes.lattice_mass(scenes.H32, scenes.S32))
    g.set_particles_f32(pos, None, np.full(len(pos), m, np.float32))
    g.step(float(np.float32(0.0025)), steps)
    return g


def _oracle_of(g, bounds=3.0, gravity=G):
    w = O.World(_params(O.Params, bounds, gravity))
    w.set_particles(g.get("pos"), g.get("vel"), g.get("mass"), None, g.body_acc())
    return w


def _normwise(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _bodies():
    s = np.sqrt(0.5)
    c, sn = np.cos(0.4), np.sin(0.4)
    pose = np.array([
        [0.0, -0.55, 0.0, 1, 0, 0, 0],          # static floor slab
        [0.05, 0.0, 0.02, 1, 0, 0, 0],          # dynamic sphere in the block
        [-0.2, 0.1, 0.1, c, 0, sn, 0],          # dynamic capsule, rotated about y
        [0.15, -0.1, -0.1, s, s, 0, 0],         # static box rotated about x, overlapping the sphere
        [2.5, 2.5, 2.5, 1, 0, 0, 0],            # dynamic sphere out of the fluid
    ])
    vel = np.array([[0, 0, 0], [0.1, -0.4, 0.0], [0.0, 0.3, -0.2], [0, 0, 0], [0.0, -1.0, 0.0]], float)
    inv_mass = np.array([0.0, 1 / 300.0, 1 / 80.0, 0.0, 1 / 5.0])
    shape = np.array([1, 0, 2, 1, 0], np.int32)
    dims = np.array([[1.0, 0.06, 1.0], [0.22, 0, 0], [0.15, 0.08, 0], [0.12, 0.05, 0.2], [0.3, 0, 0]])
    return pose, vel, inv_mass, shape, dims


def test_couple_bodies_matches_oracle():
    g = _settled_block()
    w = _oracle_of(g)
    pose, vel, im, sh, dm = _bodies()
    vg = g.couple_bodies(pose, vel, im, sh, dm, 0.0025, 0.1)
    vo = w.couple_bodies(pose, vel, im, sh, dm, 0.0025, 0.1)
    print("body velocities", vg, "max |gpu - oracle|", np.abs(vg - vo).max(), flush=True)
    assert np.abs(vg - vo).max() <= 1e-12 * np.abs(vo).max()
    # the dynamic bodies in the fluid take contacts and buoyancy; the one
    # outside it keeps its velocity
    assert np.any(vg[1] != vel[1]) and np.any(vg[2] != vel[2])
    assert np.array_equal(vg[4], vel[4])
    for f in ("pos", "vel"):
        e = _normwise(g.get(f), w.get(f))
        print(f"{f}: norm-wise {e:.2e}", flush=True)
        assert e < 1e-6
    assert np.abs(g.get("pos") - w.get("pos")).max() > 0  # FP32 storage, but the same moves
    moved = np.any(g.get("pos") != _settled_block().get("pos"), axis=1)
    assert moved.sum() > 20


def test_couple_bodies_repeated_with_steps_tracks_oracle():
    """Couple after each of 10 steps, the oracle re-fed the GPU state each step."""
    g = _settled_block(8, 6, seed=3)
    pose, vel, im, sh, dm = _bodies()
    for it in range(10):
        g.step(0.0025)
        w = _oracle_of(g)
        vo = w.couple_bodies(pose, vel, im, sh, dm, 0.0025)
        vel_g = g.couple_bodies(pose, vel, im, sh, dm, 0.0025)
        assert np.abs(vel_g - vo).max() <= 1e-12 * max(np.abs(vo).max(), 1.0)
        assert _normwise(g.get("pos"), w.get("pos")) < 1e-6
        assert _normwise(g.get("vel"), w.get("vel")) < 1e-6
        vel = vel_g
        pose[:, :3] += vel * 0.0025


def test_couple_points_matches_oracle():
    g = _settled_block()
    w = _oracle_of(g)
    rng = np.random.default_rng(11)
    n = 64
    q = rng.uniform(-0.7, 0.5, size=(n, 3))
    q[:8] = q[0] + rng.normal(scale=0.05, size=(8, 3))   # overlapping neighbourhoods
    q[8] = [3.5, 0.0, 0.0]                                 # outside the bounds
    q[9] = [2.0, 2.0, 2.0]                                 # inside, no fluid near
    qv = rng.normal(size=(n, 3))
    qm = rng.uniform(0.01, 0.05, size=n)
    f0 = np.full((n, 3), 7.0)
    fg = g.couple_points(q, qv, qm, f0, 0.0025, 1.3, 900.0)
    fo = w.couple_points(q, qv, qm, f0, 0.0025, 1.3, 900.0)
    assert np.array_equal(fg[8], [7.0, 7.0, 7.0])
    assert np.abs(fg - fo).max() <= 1e-12 * np.abs(fo).max()
    e = _normwise(g.get("vel"), w.get("vel"))
    print(f"vel norm-wise {e:.2e}", flush=True)
    assert e < 1e-6
    assert np.array_equal(g.get("pos"), w.get("pos").astype(np.float32))


def _heat_grid(rng, nx=9, ny=7, nz=8):
    origin = np.array([-0.8, -0.75, -0.45])
    dx = 0.15
    u = rng.uniform(-1, 2, size=(nz, ny, nx))
    src = rng.uniform(0, 1, size=(nz, ny, nx))
    return u, origin, dx, src


def test_couple_heat_matches_oracle():
    g = _settled_block()
    rng = np.random.default_rng(12)
    b0 = rng.normal(size=(len(g), 3))
    g.set_body_acc(b0)
    w = _oracle_of(g)
    u, origin, dx, src = _heat_grid(rng)
    sg = g.couple_heat(u, origin, dx, src, 0.0025, 0.3, 0.5, 0.8)
    so = w.couple_heat(u, origin, dx, src, 0.0025, 0.3, 0.5, 0.8)
    assert np.abs(sg - so).max() <= 1e-12 * np.abs(so).max()
    assert np.any(sg != src) and np.any(sg == src)  # some cells hold particles, some not
    e = _normwise(g.body_acc(), w.get("body_acc"))
    print(f"body_acc norm-wise {e:.2e}", flush=True)
    assert e < 1e-6
    # heat_gain 0: no sources, the grid is not touched
    s2 = g.couple_heat(u, origin, dx, src, 0.0025, 0.3, 0.5, 0.0)
    assert np.array_equal(s2, src)
    # the next step consumes body_acc
    g.step(0.0025)
    assert np.array_equal(g.body_acc(), np.zeros((len(g), 3)))


def test_couple_heat_source_sums_are_deterministic():
    g = _settled_block()
    rng = np.random.default_rng(13)
    u, origin, dx, src = _heat_grid(rng, 3, 3, 3)
    a = g.couple_heat(u, origin, dx, src, 0.0025, 0.0, 0.0, 1.0)
    b = g.couple_heat(u, origin, dx, src, 0.0025, 0.0, 0.0, 1.0)
    assert np.array_equal(a, b)


def test_sample_velocity_is_the_oracles_nearest_particle():
    g = ffi.FluidWorld(_params(ffi.PhySphParams))
    pos = scenes.lattice_block(7, 7, 7, scenes.S32, (-0.3, -0.3, -0.3), 0.0, 1)  # exact lattice: ties
    rng = np.random.default_rng(14)
    vel = rng.normal(size=pos.shape).astype(np.float32)
    g.set_particles_f32(pos, vel, np.full(len(pos), 0.02, np.float32))
    w = _oracle_of(g)
    p = g.get("pos")
    q = np.concatenate([
        (p[:40] + p[1:41]) / 2,                       # midpoints: two particles tie
        p[100:120],                                   # on a particle
        rng.uniform(-0.5, 0.5, size=(100, 3)),
        rng.uniform(-2.9, 2.9, size=(40, 3)),         # far from the block
        [[np.nan, 0, 0], [np.inf, 0, 0], [0, -np.inf, 1]],
    ])
    vg = g.sample_velocity(q)
    vo = w.nearest_velocity(q)
    assert np.array_equal(vg, vo)


# ---- the reference's coupling tests (sph/tests.rs), restated ----

def _fill(lo, hi, spacing, margin, gravity=(0.0, 0.0, 0.0)):
    g = ffi.FluidWorld(_params(ffi.PhySphParams, 3.0, gravity))
    g.fill_box(lo, hi, spacing, margin)
    return g


def test_static_body_blocks_fluid():
    """sph/tests.rs:98-131 (default gravity)."""
    g = _fill((-1.5, -1.5, -0.5), (1.5, -1.0, 0.5), 0.1, 0.05, G)
    pose = [[0, 0, 0, 1, 0, 0, 0]]
    for _ in range(100):
        g.step(0.0025)
        g.couple_bodies(pose, [[0, 0, 0]], [0.0], [1], [[1.0, 0.5, 1.0]], 0.0025, 0.0)
    loc = np.abs(g.get("pos"))
    assert not np.any((loc[:, 0] <= 1.0) & (loc[:, 1] <= 0.5) & (loc[:, 2] <= 1.0))


def test_dynamic_body_gets_buoyancy_upward():
    """sph/tests.rs:133-176."""
    g = _fill((-0.7, -0.7, -0.7), (0.7, 0.7, 0.7), 0.1, 0.05, G)
    g.step(0.0025, 15)
    vol = 4.0 / 3.0 * np.pi * 0.6 ** 3
    m_b = 1000.0 * vol * 0.25
    pos = np.zeros(3)
    vel = np.zeros(3)
    for _ in range(120):
        vel = vel + np.array(G) * 0.0025
        pos = pos + vel * 0.0025
        g.step(0.0025)
        vel = g.couple_bodies([[*pos, 1, 0, 0, 0]], [vel], [1 / m_b], [0], [[0.6, 0, 0]], 0.0025, 0.0)[0]
    assert pos[1] > 0.0


def test_couple_points_buoyancy_upward():
    """sph/tests.rs:179-215."""
    g = _fill((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 0.1, 0.05, G)
    g.step(0.0025, 15)
    m = g.get("mass")
    p_before = (g.get("vel") * m[:, None]).sum(0)
    f = g.couple_points([[0, 0, 0]], [[0, 0, 0]], [0.02], [[0, 0, 0]], 0.0025, 1.0, 1000.0)[0]
    assert f[1] > 0.0
    p_after = (g.get("vel") * m[:, None]).sum(0)
    assert p_after[1] - p_before[1] < 0.0


def test_couple_heat_warmer_fluid_rises():
    """sph/tests.rs:217-257."""
    g = ffi.FluidWorld(_params(ffi.PhySphParams))
    g.set_particles([[0.0, 1.0, 0.0]], None, [0.05])
    nx = 16
    ys = np.arange(nx) * 1.0 - nx / 2.0
    u = np.broadcast_to(np.where(ys > 0, 1.0, 0.0)[None, :, None], (nx, nx, nx))
    g.couple_heat(u, (-8.0, -8.0, -8.0), 1.0, np.zeros((nx, nx, nx)), 0.01, 0.0, 0.5, 0.0)
    assert G[1] + g.body_acc()[0, 1] > -9.81


def test_couple_heat_injects_source_into_moving_region():
    """sph/tests.rs:259-284 (up to the heat field's own diffusion step)."""
    g = ffi.FluidWorld(_params(ffi.PhySphParams, gravity=(0.0, 0.0, 0.0)))
    g.set_particles([[0.0, 0.0, 0.0]], [[2.0, 0.0, 0.0]], [0.05])
    nx = 16
    src = g.couple_heat(np.zeros((nx, nx, nx)), (0.0, 0.0, 0.0), 1.0, np.zeros((nx, nx, nx)), 0.01, 0.0, 0.0, 0.1)
    assert src[0, 0, 0] == pytest.approx(0.1 * 2.0 * 0.01, rel=1e-15)
    assert np.count_nonzero(src) == 1


def test_water_entry_energy_bounded():
    """sph/tests.rs:470-548 (b0_water_entry_energy_bounded)."""
    g = _fill((-1.0, -1.0, -1.0), (1.0, 0.5, 1.0), 0.12, 0.05, G)
    dt = 0.0025
    g.step(dt, 80)
    vol = 4.0 / 3.0 * np.pi * 0.4 ** 3
    m_b = 1000.0 * vol * 2.0
    pos = np.array([0.0, 2.0, 0.0])
    vel = np.array([0.0, -3.0, 0.0])
    m = g.get("mass")

    def ke():
        v = g.get("vel")
        return float((0.5 * m * (v * v).sum(1)).sum() + 0.5 * m_b * vel @ vel)

    e0 = ke()
    peak, last = e0, e0
    for _ in range(400):
        vel = vel + np.array(G) * dt
        pos = pos + vel * dt
        g.step(dt)
        vel = g.couple_bodies([[*pos, 1, 0, 0, 0]], [vel], [1 / m_b], [0], [[0.4, 0, 0]], dt, 0.1)[0]
        e = ke()
        assert np.isfinite(e)
        peak = max(peak, e)
        last = e
    print(f"KE e0 {e0:.4g} peak {peak:.4g} last {last:.4g}", flush=True)
    assert peak <= e0 * 50.0
    assert last < peak * 0.5
