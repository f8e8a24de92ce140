"""Pins the f64 CPU oracle (oracle/sph_oracle.cpp) to the reference solver.

The reference (Rust) cannot be built or run here and its tests hold no numeric
golden vectors, so each oracle function is pinned by checks that a plausible
mistake fails:
  * kernels          Mueller-2003 normalisation and derivative identities by
                     quadrature / finite differences (not by retyping formulas)
  * lattice_mass     three values at h = 0.2 + the interior-density identity rho = rho0
  * fill_box         closed-form counts of the reference's x += d loop
  * neighbour grid   O(N^2) brute force with the reference predicate, exact sets
  * density          brute force, single-particle closed form, lattice interior
  * forces           hand-derived two-particle closed forms, pairwise
                     antisymmetry (momentum), lattice symmetry
  * power law        two-particle closed form with unequal masses, velocity
                     scaling, Galilean invariance, the floor at rest, the
                     pair identity, the out-of-table material rule
  * force sets       O(N^2) brute force of the force-pass predicate
  * integrate/bounds README dam break (343 free particles) in closed form
  * whole step       the reference's own test assertions (crates/phy-fluid/src/
                     sph/tests.rs, crates/phy-demo/tests/stability.rs)
  * determinism      1 vs 8 threads bitwise
and test_mutations_are_caught plants each known mistake in the oracle and
requires at least one of these checks to fail.
"""
import math

import numpy as np
import pytest

import oracle as O
from physics_amd import scenes

H = 0.2


def test_kernel_coefficients_match_closed_forms():
    c, _ = O.kernels(0.2)
    # c6 = 315 / (64 pi h^9), cs = -45 / (pi h^6), cv = 45 / (pi h^6) at f64 h = 0.2
    # and at h = (double)0.2f
    assert c[0] == pytest.approx(3059924.7481657104, rel=1e-15)
    assert c[1] == pytest.approx(-223811.63872297775, rel=1e-15)
    assert c[2] == pytest.approx(223811.63872297775, rel=1e-15)
    c32, _ = O.kernels(float(np.float32(0.2)))
    assert c32[0] == pytest.approx(3059924.337797856, rel=1e-15)
    assert c32[2] == pytest.approx(223811.61871265905, rel=1e-15)


def _quad(f, a, b, n=20001):
    x = np.linspace(a, b, n)
    y = f(x)
    return float(np.trapezoid(y, x))


@pytest.mark.parametrize("h", [0.2, 0.37, 1.0])
def test_poly6_is_normalised(h):
    # integral over the ball of W_poly6 = 1 (Mueller 2003 eq. 20)
    vals = np.array([O.kernels(h, r)[1][0] for r in np.linspace(0, h, 2001)])
    r = np.linspace(0, h, 2001)
    integral = float(np.trapezoid(4 * math.pi * r * r * vals, r))
    assert integral == pytest.approx(1.0, rel=1e-5)


@pytest.mark.parametrize("h", [0.2, 0.5])
def test_spiky_gradient_is_derivative_of_normalised_spiky(h):
    # W_spiky = 15/(pi h^6) (h - r)^3 integrates to 1; |grad W| = dW/dr
    w = lambda r: 15.0 / (math.pi * h ** 6) * (h - r) ** 3
    r = np.linspace(0, h, 4001)
    assert float(np.trapezoid(4 * math.pi * r * r * w(r), r)) == pytest.approx(1.0, rel=1e-5)
    for rr in [0.1 * h, 0.5 * h, 0.9 * h]:
        e = 1e-6 * h
        fd = (w(rr + e) - w(rr - e)) / (2 * e)
        assert O.kernels(h, rr)[1][1] == pytest.approx(fd, rel=1e-6)
    assert O.kernels(h, 0.0)[1][1] == 0.0      # r <= 0 branch (kernels.rs:51)
    assert O.kernels(h, h)[1][1] == 0.0        # r >= h branch


@pytest.mark.parametrize("h", [0.2, 0.5])
def test_viscosity_laplacian_is_laplacian_of_viscosity_kernel(h):
    # Mueller's W_visc = 15/(2 pi h^3) (-r^3/(2h^3) + r^2/h^2 + h/(2r) - 1);
    # laplacian = W'' + 2 W'/r = 45/(pi h^6) (h - r)
    w = lambda r: 15.0 / (2 * math.pi * h ** 3) * (-r ** 3 / (2 * h ** 3) + r ** 2 / h ** 2 + h / (2 * r) - 1)
    for rr in [0.2 * h, 0.5 * h, 0.8 * h]:
        e = 1e-4 * h
        d1 = (w(rr + e) - w(rr - e)) / (2 * e)
        d2 = (w(rr + e) - 2 * w(rr) + w(rr - e)) / (e * e)
        assert O.kernels(h, rr)[1][2] == pytest.approx(d2 + 2 * d1 / rr, rel=1e-5)


def test_lattice_mass_values_and_interior_density():
    w = O.World()
    assert w.lattice_mass(0.1) == pytest.approx(0.99031946198009, rel=1e-13)
    assert w.lattice_mass(0.3) == pytest.approx(5.10633472583484, rel=1e-13)
    assert w.lattice_mass(0.18) == pytest.approx(4.904495133126164, rel=1e-13)
    # the mass is defined so a perfect lattice interior has rho == rho0
    s = 0.1
    n = 9
    g = np.arange(n) * s
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    w.set_particles(pos, mass=w.lattice_mass(s))
    w.density()
    rho = w.get("rho").reshape(n, n, n)
    assert rho[4, 4, 4] == pytest.approx(1000.0, rel=1e-12)


def _fill_count(lo, hi, d):
    c = 0
    x = lo
    while x <= hi:
        c += 1
        x += d
    return c


@pytest.mark.parametrize("mn,mx,sp,mg,expect", [
    ((-1, 1, -1), (1, 3, 1), 0.3, 0.1, 343),            # README / config 1
    ((-4, -4, -1), (0, 4, 1), 0.18, 0.05, 10648),       # phy_world_create_fluid
    ((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 0.1, 0.05, 729),
])
def test_fill_box_counts(mn, mx, sp, mg, expect):
    w = O.World()
    w.fill_box(mn, mx, sp, mg)
    per = [_fill_count(mn[a] + mg, mx[a] - mg, sp) for a in range(3)]
    assert len(w) == per[0] * per[1] * per[2] == expect
    pos = w.get("pos")
    # index order: x outer, y, z inner (world.rs:130-143)
    assert np.all(np.diff(pos[:, 0]) >= 0)
    assert pos[1, 2] > pos[0, 2] and pos[1, 0] == pos[0, 0]


def _brute_neighbors(pos, h):
    # reference predicate: r = sqrt((dx*dx + dy*dy) + dz*dz) < h, f64
    d = pos[:, None, :] - pos[None, :, :]
    s = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    return np.sqrt(s) < h


def _jittered_block(n, s, jit, seed, origin=(0.0, 0.0, 0.0)):
    return scenes.lattice_block(n, n, n, s, origin, jit, seed).astype(np.float64)


@pytest.mark.parametrize("seed,jit,origin", [(1, 0.01, (0, 0, 0)), (2, 0.3, (-0.35, 0.1, -0.05)),
                                             (3, 0.5, (-1.0, -1.0, -1.0))])
def test_neighbor_sets_equal_brute_force(seed, jit, origin):
    pos = _jittered_block(7, 0.1, jit, seed, origin)
    w = O.World()
    w.set_particles(pos, mass=1.0)
    adj = _brute_neighbors(pos, H)
    for i in range(len(pos)):
        got = np.sort(w.neighbors(i))
        want = np.nonzero(adj[i])[0]
        assert np.array_equal(got, want), i
    c, s, x = w.neighbor_stats()
    assert np.array_equal(c, adj.sum(1))
    assert np.array_equal(s, (adj * np.arange(len(pos))[None, :]).sum(1).astype(np.uint64))


def _mix64(j):
    z = (np.asarray(j, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _stats_of(adj):
    idx = np.arange(adj.shape[1], dtype=np.uint64)
    h = _mix64(idx)
    c = adj.sum(1)
    s = np.array([idx[row].sum(dtype=np.uint64) for row in adj], dtype=np.uint64)
    x = np.array([np.bitwise_xor.reduce(h[row]) if row.any() else np.uint64(0) for row in adj],
                 dtype=np.uint64)
    return c, s, x


def test_force_neighbor_sets_equal_brute_force():
    # world.rs:292-299: j != i, skip r >= h and r <= 0, so coincident points
    # drop out and a NaN distance stays in (for candidates of the 27 cells: a
    # NaN coordinate has reference key 0 on that axis, grid.rs:29-31)
    with np.errstate(invalid="ignore", over="ignore"):
        pos = _jittered_block(6, 0.1, 0.3, 8, (-0.3, -0.3, -0.3))
        pos = np.concatenate([pos, pos[[5, 5, 40]], [[np.nan, 0.02, 0.03], [0.05, 0.01, np.nan]]])
        w = O.World()
        w.set_particles(pos, mass=1.0)
        d = pos[None, :, :] - pos[:, None, :]
        s2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
        r = np.sqrt(s2)
        key = np.where(np.isnan(pos), 0, np.floor(np.nan_to_num(pos) * 5.0)).astype(np.int64)
        adjacent = (np.abs(key[:, None, :] - key[None, :, :]) <= 1).all(-1)
        adj = adjacent & ~(r >= H) & ~(r <= 0.0) & ~np.eye(len(pos), dtype=bool)
    c, s, x = w.force_neighbor_stats()
    bc, bs, bx = _stats_of(adj)
    assert np.array_equal(c, bc) and np.array_equal(s, bs) and np.array_equal(x, bx)
    n0 = len(pos) - 5
    assert c[n0 + 3] > 0                       # the NaN particle is a force neighbour
    dc, _, _ = w.neighbor_stats()
    assert dc[n0 + 3] == 0                     # ... but no density neighbour
    assert c[5] == c[n0] and dc[5] == dc[n0]   # coincident copies


def test_density_matches_brute_force_and_single_particle():
    pos = _jittered_block(6, 0.1, 0.2, 7)
    m = np.linspace(0.5, 1.5, len(pos))
    w = O.World()
    w.set_particles(pos, mass=m)
    w.density()
    adj = _brute_neighbors(pos, H)
    c6 = 315.0 / (64 * math.pi * H ** 9)
    d = pos[:, None, :] - pos[None, :, :]
    r = np.sqrt((d[..., 0] ** 2 + d[..., 1] ** 2) + d[..., 2] ** 2)
    wk = np.where(adj, c6 * (H * H - r * r) ** 3, 0.0)
    rho = (wk * m[None, :]).sum(1)
    np.testing.assert_allclose(w.get("rho"), rho, rtol=1e-12)
    p = np.where(rho - 1000.0 > 0, 250.0 * (rho - 1000.0), 0.0)
    np.testing.assert_allclose(w.get("p"), p, rtol=1e-9, atol=1e-9)
    # a lone particle: rho = m W(0) = m c6 h^6
    w1 = O.World()
    w1.set_particles([[0.3, 0.2, 0.1]], mass=2.0)
    w1.density()
    assert w1.get("rho")[0] == pytest.approx(2.0 * c6 * H ** 6, rel=1e-14)


def _two_particle_world(r, m, u, dt, gravity=False):
    p = O.Params.defaults()
    if not gravity:
        p.gravity[1] = 0.0
    w = O.World(p)
    pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    vel = np.array([[0.0, u, 0.0], [0.0, -u, 0.0]])
    w.set_particles(pos, vel=vel, mass=m)
    w.step(dt)
    return w


@pytest.mark.parametrize("r", [0.05, 0.12, 0.19])
def test_two_particle_forces_closed_form(r):
    m, u, dt = 10.0, 0.7, 1e-3
    w = _two_particle_world(r, m, u, dt)
    c6 = 315.0 / (64 * math.pi * H ** 9)
    cs = -45.0 / (math.pi * H ** 6)
    cv = 45.0 / (math.pi * H ** 6)
    W = lambda q: c6 * (H * H - q * q) ** 3
    rho = m * (W(0.0) + W(r))
    p = 250.0 * max(0.0, rho - 1000.0)
    # pressure: A is pushed away from B along -x (Mueller 2003 symmetric form)
    ax = 2.0 * m * m * p / rho ** 2 * cs * (H - r) ** 2 / rho
    # viscosity: mu * m m / rho_j * (v_j - v_i) * cv (h - r) / rho_i
    ay = 3.5 * (m * m / rho) * (-2.0 * u) * cv * (H - r) / rho
    xs_y = (-2.0 * u) * (m / rho) * W(r)
    acc = w.get("acc")
    assert acc[0, 0] == pytest.approx(ax, rel=1e-12)
    assert acc[0, 1] == pytest.approx(ay, rel=1e-12)
    assert acc[1, 0] == pytest.approx(-ax, rel=1e-12)
    assert acc[1, 1] == pytest.approx(-ay, rel=1e-12)
    assert w.get("xsph")[0, 1] == pytest.approx(xs_y, rel=1e-12)
    assert w.get("rho")[0] == pytest.approx(rho, rel=1e-13)
    assert w.get("mu_eff")[0] == 3.5
    # semi-implicit Euler + XSPH (world.rs:356-360)
    v_y = u + ay * dt + 0.5 * xs_y
    assert w.get("vel")[0, 1] == pytest.approx(v_y, rel=1e-12)
    assert w.get("pos")[0, 1] == pytest.approx(v_y * dt, rel=1e-12)
    assert w.get("vel")[0, 0] == pytest.approx(ax * dt, rel=1e-12)


def test_pressure_forces_are_pairwise_antisymmetric():
    # at rest (no viscosity / XSPH contribution) sum_i rho_i * acc_i = 0
    pos = _jittered_block(6, 0.1, 0.3, 11)
    m = np.linspace(0.8, 1.3, len(pos))
    p = O.Params.defaults()
    p.gravity[1] = 0.0
    p.stiffness = 2000.0
    w = O.World(p)
    w.set_particles(pos, mass=m * 1.7)
    w.step(1e-3)
    f = w.get("acc") * w.get("rho")[:, None]
    scale = np.abs(f).sum()
    assert scale > 0
    assert np.abs(f.sum(0)).max() < 1e-12 * scale


def test_lattice_interior_has_zero_net_force():
    n, s = 9, 0.1
    g = np.arange(n) * s
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    p = O.Params.defaults()
    p.gravity[1] = 0.0
    w = O.World(p)
    w.set_particles(pos, mass=1.3)   # heavier than lattice mass: p > 0 everywhere
    w.step(1e-3)
    acc = w.get("acc").reshape(n, n, n, 3)
    assert np.abs(acc[4, 4, 4]).max() < 1e-9 * np.abs(acc).max()
    assert np.abs(acc).max() > 1e-2


def _closed_form_c1(steps, dt=0.01):
    """README dam break: 7^3 free particles, free fall + damped floor reflection."""
    lo_y = -5.0
    w0 = O.World()
    w0.fill_box((-1, 1, -1), (1, 3, 1), 0.3, 0.1)
    pos = w0.get("pos").copy()
    y = pos[:, 1].copy()
    v = np.zeros_like(y)
    traj = []
    for _ in range(steps):
        a = (0.0 + -9.81) + 0.0
        v = v + a * dt
        v = v + 0.0 * 0.5
        y = y + v * dt
        hit = y < lo_y
        y = np.where(hit, lo_y, y)
        v = np.where(hit, -v * 0.4, v)
        p = pos.copy()
        p[:, 1] = y
        traj.append((p, v.copy()))
    return traj


def test_readme_dam_break_is_closed_form():
    traj = _closed_form_c1(200)
    w = O.World()
    w.fill_box((-1, 1, -1), (1, 3, 1), 0.3, 0.1)
    assert len(w) == 343
    rho0 = w.lattice_mass(0.3) * (315.0 / (64 * math.pi * H ** 3))
    checked = 0
    for k in range(200):
        p_cf, v_cf = traj[k]
        # the closed form holds while no two particles come within h
        if k > 0:
            d = p_prev[:, None, :] - p_prev[None, :, :]
            r = np.sqrt((d ** 2).sum(-1)) + np.eye(len(p_prev)) * 9
            if r.min() < H:
                break
        w.step(0.01)
        assert np.array_equal(w.get("pos"), p_cf)
        assert np.array_equal(w.get("vel")[:, 1], v_cf)
        np.testing.assert_allclose(w.get("rho"), rho0, rtol=1e-14)
        p_prev = p_cf
        checked += 1
    assert checked >= 110


# ---- the reference's own assertions (crates/phy-fluid/src/sph/tests.rs) -----

def _test_params():
    p = O.Params.defaults()
    for a in range(3):
        p.bounds_min[a] = -3.0
        p.bounds_max[a] = 3.0
    return p


def test_ref_fill_box_count_matches_grid():                 # tests.rs:19-29
    w = O.World(_test_params())
    w.fill_box((-1, -1, -1), (1, 1, 1), 0.2, 0.05)
    assert 500 < len(w) < 2000


def test_ref_density_converges_near_rest():                  # tests.rs:32-56
    p = _test_params()
    p.gravity[1] = 0.0
    w = O.World(p)
    w.fill_box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 0.1, 0.05)
    for _ in range(15):
        w.step(0.002)
    avg = w.get("rho").mean()
    assert abs(avg - 1000.0) / 1000.0 < 0.25


def test_ref_dam_break_stays_in_bounds():                    # tests.rs:75-95
    w = O.World(_test_params())
    w.fill_box((-1.5, -1.5, -0.5), (-0.5, 1.5, 0.5), 0.1, 0.05)
    n0 = len(w)
    for _ in range(120):
        w.step(0.0025)
        x = w.get("pos")
        assert np.isfinite(x).all()
        assert (x >= -3.0 - 1e-6).all() and (x <= 3.0 + 1e-6).all()
    assert len(w) == n0


def _shear_world(n, dv):                                     # tests.rs:371-390
    p = _test_params()
    p.gravity[1] = 0.0
    p.set_visc([3.5], [n])
    w = O.World(p)
    pos, vel = [], []
    for iy in range(2):
        vy = -dv if iy == 0 else dv
        for ix in range(5):
            pos.append([(ix - 2.0) * 0.08, iy * 0.06, 0.0])
            vel.append([vy, 0.0, 0.0])
    w.set_particles(pos, vel=vel, mass=0.05)
    w.step(0.001)
    return w


def test_ref_non_newtonian_shear_thinning_and_thickening():  # tests.rs:393-404
    mt = _shear_world(0.5, 300.0).get("mu_eff")[0]
    mn = _shear_world(1.0, 300.0).get("mu_eff")[0]
    mk = _shear_world(1.5, 300.0).get("mu_eff")[0]
    assert np.isfinite([mt, mn, mk]).all()
    assert mt < mn < mk


# ---- the power-law branch (world.rs:313-327) ---------------------------------
#
# shear_i = sum_j |v_j - v_i| / (r + 1e-4) * (m_j / rho_j), mu_i = k_m *
# max'(shear_i, shear_min)^(n_m - 1) with m = material_i if it indexes visc_k,
# else 0. Pinned by a two-particle closed form with unequal masses (which of
# m_i / rho_i and m_j / rho_j enters), relations the formula implies without
# restating it (velocity scaling, Galilean invariance, the floor at rest, the
# pair identity through the oracle's own densities), and the reference's
# out-of-table material rule. max'(a, b) = a if a > b else b equals the ">="
# form for every input (equal arguments give the same value either way; NaN
# fails both), so that variant is not a distinct mistake.

def _pl_world(pos, vel, mass, k=(3.5,), n=(1.0,), shear_min=0.01, material=None):
    p = _test_params()
    p.gravity[1] = 0.0
    p.shear_min = shear_min
    p.set_visc(list(k), list(n))
    w = O.World(p)
    w.set_particles(pos, vel=vel, mass=mass, material=material)
    w.step(1e-4)
    return w


@pytest.mark.parametrize("nexp", [0.7, 1.3])
@pytest.mark.parametrize("r", [0.03, 0.15])
def test_power_law_two_particle_closed_form(nexp, r):
    ma, mb, u, k = 0.05, 0.08, 0.9, 3.5
    w = _pl_world([[0, 0, 0], [r, 0, 0]], [[0, -u, 0], [0, u, 0]], [ma, mb], k=(k,), n=(nexp,))
    c6 = 315.0 / (64 * math.pi * H ** 9)
    W = lambda q: c6 * (H * H - q * q) ** 3
    rho_a = ma * W(0.0) + mb * W(r)
    rho_b = mb * W(0.0) + ma * W(r)
    sh_a = 2 * u / (r + 1e-4) * (mb / rho_b)
    sh_b = 2 * u / (r + 1e-4) * (ma / rho_a)
    mu = w.get("mu_eff")
    assert mu[0] == pytest.approx(k * max(sh_a, 0.01) ** (nexp - 1.0), rel=1e-12)
    assert mu[1] == pytest.approx(k * max(sh_b, 0.01) ** (nexp - 1.0), rel=1e-12)


def _pl_block(seed):
    pos = _jittered_block(4, 0.07, 0.3, seed)
    n = len(pos)
    vel = scenes.rng_f64(seed + 50, 3 * n).reshape(n, 3) - 0.5
    mass = 0.03 + 0.04 * scenes.rng_f64(seed + 90, n)
    return pos, vel, mass


def test_power_law_scales_with_velocity_and_is_galilean():
    nexp, lam = 0.6, 3.0
    pos, vel, mass = _pl_block(3)
    mu1 = _pl_world(pos, vel, mass, n=(nexp,), shear_min=1e-6).get("mu_eff")
    mu2 = _pl_world(pos, lam * vel, mass, n=(nexp,), shear_min=1e-6).get("mu_eff")
    mu3 = _pl_world(pos, vel + np.array([4.0, -2.0, 1.0]), mass, n=(nexp,), shear_min=1e-6).get("mu_eff")
    # shear is linear in the velocity differences and blind to a common velocity
    np.testing.assert_allclose(mu2 / mu1, lam ** (nexp - 1.0), rtol=1e-12)
    np.testing.assert_allclose(mu3, mu1, rtol=1e-12)


def test_power_law_floor_at_rest_and_pair_identity():
    nexp, k, smin = 1.4, 2.0, 0.03
    pos, _, mass = _pl_block(4)
    mu = _pl_world(pos, np.zeros_like(pos), mass, k=(k,), n=(nexp,), shear_min=smin).get("mu_eff")
    assert (mu == k * smin ** (nexp - 1.0)).all()
    # two particles: shear_a rho_b / m_b = shear_b rho_a / m_a = |dv| / (r + 1e-4),
    # with the densities the oracle itself computed
    ma, mb, r, u = 0.05, 0.11, 0.08, 1.3
    w = _pl_world([[0, 0, 0], [0, r, 0]], [[u, 0, 0], [-u, 0, 0]], [ma, mb], k=(k,), n=(nexp,),
                  shear_min=1e-9)
    rho = w.get("rho")
    sh = (w.get("mu_eff") / k) ** (1.0 / (nexp - 1.0))
    want = 2 * u / (r + 1e-4)
    assert sh[0] * rho[1] / mb == pytest.approx(want, rel=1e-11)
    assert sh[1] * rho[0] / ma == pytest.approx(want, rel=1e-11)


def test_power_law_material_outside_the_table_uses_material_0():
    pos, vel, mass = _pl_block(5)
    n = len(pos)
    mat = np.where(np.arange(n) % 2 == 0, 7, 1).astype(np.uint64)
    w = _pl_world(pos, vel, mass, k=(3.5, 10.0), n=(0.8, 1.2), material=mat)
    w0 = _pl_world(pos, vel, mass, k=(3.5, 10.0), n=(0.8, 1.2), material=np.where(mat == 7, 0, 1))
    assert np.array_equal(w.get("mu_eff"), w0.get("mu_eff"))


def test_ref_material_tag_distinguishes_viscosity():         # tests.rs:407-427
    p = _test_params()
    p.gravity[1] = 0.0
    p.set_visc([3.5, 10.0], [1.0, 1.0])
    w = O.World(p)
    w.set_particles([[0, 0, 0], [0, 0.06, 0]], vel=[[-300, 0, 0], [300, 0, 0]], mass=0.05,
                    material=[0, 1])
    w.step(0.001)
    mu = w.get("mu_eff")
    assert abs(mu[0] - 3.5) < 1e-6 and abs(mu[1] - 10.0) < 1e-6


def test_ref_closed_system_conserves_energy_and_momentum():  # stability.rs:55-106
    p = O.Params.defaults()
    p.gravity[1] = 0.0
    p.bounds_min[0], p.bounds_min[1], p.bounds_min[2] = -7.0, -1.0, -7.0
    p.bounds_max[0], p.bounds_max[1], p.bounds_max[2] = 7.0, 13.0, 7.0
    half = 5.0
    spacing = (8.0 * half ** 3 / 800) ** (1.0 / 3.0)
    p.h = spacing * 2.0
    w = O.World(p)
    w.fill_box((-half, 1.0, -half), (half, 1.0 + 2 * half, half), spacing, 0.1)
    e0 = w.kinetic_energy()
    for _ in range(1000):
        w.step(0.01)
        assert w.validate() == 0
    assert abs(w.kinetic_energy() - e0) <= 1e-2
    assert np.linalg.norm(w.total_momentum()) <= 1e-3


def test_ref_watchdog_catches_nan():                         # extreme_parameters.rs:160-186
    w = O.World()
    w.fill_box((-4, -4, -1), (0, 4, 1), 0.18, 0.05)
    for _ in range(5):
        w.step(0.005)
        assert w.validate() == 0
    pos = w.get("pos")
    pos[0, 0] = np.nan
    w.set_particles(pos, vel=w.get("vel"), mass=w.get("mass"))
    w.step(0.005)
    assert w.validate() == 1


def test_oracle_is_thread_count_independent():
    sc = scenes.dam_break((12, 12, 12), seed=5)
    res = []
    for t in (1, 8):
        O.set_threads(t)
        p = O.Params.defaults()
        p.h = sc.h
        for a in range(3):
            p.bounds_min[a], p.bounds_max[a] = sc.bounds_min[a], sc.bounds_max[a]
        w = O.World(p)
        w.set_particles(sc.pos.astype(np.float64), mass=sc.mass)
        w.step(0.01, 20)
        res.append((w.get("pos"), w.get("vel"), w.get("rho")))
    O.set_threads(8)
    for a, b in zip(*res):
        assert np.array_equal(a, b)


# ---- planted mistakes must be caught by the checks above --------------------

def _checks():
    """The cheap pinning checks, each returning True when it passes."""
    out = {}

    def run(name, fn):
        try:
            fn()
            out[name] = True
        except AssertionError:
            out[name] = False

    run("two_particle", lambda: [test_two_particle_forces_closed_form(r) for r in (0.05, 0.12)])
    run("antisym", test_pressure_forces_are_pairwise_antisymmetric)
    run("density", test_density_matches_brute_force_and_single_particle)
    run("interior", test_lattice_mass_values_and_interior_density)
    run("c1", test_readme_dam_break_is_closed_form)
    run("material", test_ref_material_tag_distinguishes_viscosity)
    run("power_law_closed_form", lambda: [test_power_law_two_particle_closed_form(n, r)
                                          for n in (0.7, 1.3) for r in (0.03, 0.15)])
    run("power_law_scaling", test_power_law_scales_with_velocity_and_is_galilean)
    run("power_law_floor", test_power_law_floor_at_rest_and_pair_identity)
    run("power_law_material", test_power_law_material_outside_the_table_uses_material_0)

    def bounce():
        # a particle driven through the floor reflects with damping 0.4
        w = O.World()
        w.set_particles([[0.0, -4.99, 0.0]], vel=[[0.0, -3.0, 0.0]], mass=0.02)
        w.step(0.01)
        v = (-3.0 + -9.81 * 0.01)
        assert w.get("pos")[0, 1] == -5.0
        assert w.get("vel")[0, 1] == pytest.approx(-v * 0.4, rel=1e-14)

    run("bounce", bounce)

    def low_density_pressure():
        # isolated pair below rest density: p must be clamped to 0
        w = O.World()
        w.set_particles([[0, 0, 0], [0.15, 0, 0]], mass=0.5)
        w.step(1e-3)
        assert (w.get("p") == 0.0).all()

    run("clamp", low_density_pressure)
    return out


POWER_LAW_MUTATIONS = ("shear_no_reps", "shear_mass_i", "power_exp_n", "shear_no_floor",
                       "shear_dv_sq", "mat_oor_last")


@pytest.mark.parametrize("mutation", POWER_LAW_MUTATIONS)
def test_power_law_mutations_pass_the_ordering_test(mutation):
    # the reference's own power-law test (thin < Newtonian < thick) cannot see
    # these mistakes; the checks above must (test_mutations_are_caught)
    O.set_mutation(mutation)
    try:
        test_ref_non_newtonian_shear_thinning_and_thickening()
    finally:
        O.set_mutation("none")


@pytest.mark.parametrize("mutation", [m for m in O.MUTATIONS if m not in ("none", "self_in_force")])
def test_mutations_are_caught(mutation):
    base = _checks()
    assert all(base.values()), base
    O.set_mutation(mutation)
    try:
        res = _checks()
    finally:
        O.set_mutation("none")
    assert not all(res.values()), f"planted mistake {mutation} passed every check"
