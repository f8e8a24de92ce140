"""GPU parity of the HIP step against the f64 oracle, through the C ABI.

Bars: neighbour sets bit-exact; every per-step field within 1e-5 norm-wise
relative of the oracle on identical FP32 inputs (parity.TOL_STEP); repeat runs
bit-identical; the watchdog and the reference's FFI contract as in
crates/phy-ffi/src/lib.rs:906-1001.
"""
import numpy as np
import pytest

import oracle as O
import parity as P
from physics_amd import ffi, scenes

pytestmark = pytest.mark.gpu


def _block(n, seed, jit=0.01, s=scenes.S32, origin=(0.0, 0.0, 0.0)):
    return scenes.lattice_block(n, n, n, s, origin, jit, seed)


def _print(tag, err):
    print(tag, " ".join(f"{k}={v:.2e}" for k, v in err.items()), flush=True)


# ---- neighbour sets ---------------------------------------------------------

@pytest.mark.parametrize("seed,jit", [(1, 0.01), (2, 0.3), (3, 0.5)])
def test_neighbor_sets_bit_exact(seed, jit):
    pos = _block(12, seed, jit, origin=(-0.6, -0.6, -0.6))
    g, o = P.make_pair(pos, bounds=((-5, -5, -5), (5, 5, 5)))
    gc, gs, gx = g.neighbor_stats()
    oc, os_, ox = o.neighbor_stats()
    assert np.array_equal(gc, oc)
    assert np.array_equal(gs, os_)
    assert np.array_equal(gx, ox)


def test_neighbor_sets_at_the_support_boundary():
    # pairs placed within a few ulps of r = h exercise the exact f64 fallback
    h = scenes.H32
    rng = np.random.default_rng(7)
    base = rng.uniform(-1, 1, (400, 3)).astype(np.float32)
    d = rng.normal(size=(400, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    eps = rng.integers(-4, 5, 400) * 1e-8
    other = (base.astype(np.float64) + d * (h + eps)[:, None]).astype(np.float32)
    pos = np.concatenate([base, other])
    g, o = P.make_pair(pos)
    gc, gs, gx = g.neighbor_stats()
    oc, os_, ox = o.neighbor_stats()
    assert np.array_equal(gc, oc) and np.array_equal(gs, os_) and np.array_equal(gx, ox)
    # the band really is populated: some pairs in, some out
    assert (oc[:400] == 2).any() and (oc[:400] == 1).any()


# ---- one-step parity --------------------------------------------------------

@pytest.mark.parametrize("n,seed", [(8, 1), (16, 2), (24, 3)])
def test_one_step_parity_column(n, seed):
    sc = scenes.dam_break((n, n, n), seed=seed)
    g, o = P.make_pair(sc.pos, mass=sc.mass, bounds=(sc.bounds_min, sc.bounds_max))
    x0 = P.step_both(g, o)
    err = P.compare_step(g, o, x0)
    _print(f"column n={n}", err)
    for k, v in err.items():
        assert v <= P.TOL_STEP, (k, v)


def test_one_step_parity_moving_compressed_multimaterial():
    # velocities, heavier particles (p > 0 everywhere), two non-Newtonian materials
    pos = _block(14, 9, jit=0.2, origin=(-0.7, -0.7, -0.7))
    n = len(pos)
    vel = (scenes.rng_f64(4, 3 * n).reshape(n, 3) - 0.5).astype(np.float32)
    mat = (np.arange(n) % 3).astype(np.uint32)
    g, o = P.make_pair(pos, vel, mass=1.6, material=mat, visc_k=(3.5, 10.0, 1.0),
                       visc_n=(1.0, 0.6, 1.4))
    x0 = P.step_both(g, o)
    err = P.compare_step(g, o, x0)
    _print("moving", err)
    for k, v in err.items():
        assert v <= P.TOL_STEP, (k, v)


@pytest.mark.parametrize("rest,varied", [(1000.0, False), (6000.0, False), (1e6, False), (1000.0, True),
                                         (6000.0, True)])
def test_compressed_column_waves_parity_and_sets(rest, varied):
    # 27 particles per half cell (spacing h/6), so whole waves sit within three
    # half cells of one column and the density pass tries FP32 terms first:
    # rho ~ 27 rho0 keeps them at rho0 = 1000; at 6000 lanes near the block's
    # faces fall below rho - rho0 >= 0.75 rho and their waves sum again in f64;
    # at 1e6 every wave does (p = 0), so rho carries only the f64 sum's final
    # rounding to FP32. Equal masses take the terms with the common mass
    # factored out, varied masses (0.8 .. 1.2 x 0.99) the per-candidate masses.
    pos = _block(12, 5, jit=0.05, s=float(np.float32(scenes.H32 / 6)), origin=(0.31, 0.47, 0.23))
    n = len(pos)
    vel = (0.2 * (scenes.rng_f64(6, 3 * n).reshape(n, 3) - 0.5)).astype(np.float32)
    mass = (0.99 * (0.8 + 0.4 * scenes.rng_f64(8, n))).astype(np.float32) if varied else 0.99
    g, o = P.make_pair(pos, vel, mass=mass, rest_density=rest)
    for which, ofn in (("density", o.neighbor_stats), ("force", o.force_neighbor_stats)):
        c, s, x = g.gather_stats(which)
        oc, os_, ox = ofn()
        assert np.array_equal(c, oc) and np.array_equal(s, os_) and np.array_equal(x, ox), which
    x0 = P.step_both(g, o)
    err = P.compare_step(g, o, x0)
    _print(f"compressed rest={rest:g} varied={varied}", err)
    for k, v in err.items():
        assert v <= P.TOL_STEP, (k, v)
    if rest == 1e6:
        rg, ro = g.get("rho"), o.get("rho")
        rel = np.abs(rg - ro) / np.abs(ro)
        assert rel.max() <= 2.0 ** -24 * (1 + 1e-6), rel.max()



def test_compressed_column_waves_at_the_support_boundary():
    # the compressed block of the test above (column waves: candidate-major
    # density, uniform force scan) with partners placed within a few ulps of
    # r = h of 200 block particles, so the density's band vote and exact
    # predicate and the force pass's pair test decide those pairs
    h = scenes.H32
    pos = _block(12, 5, jit=0.05, s=float(np.float32(h / 6)), origin=(0.31, 0.47, 0.23))
    rng = np.random.default_rng(11)
    idx = rng.choice(len(pos), 200, replace=False)
    d = rng.normal(size=(200, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    eps = rng.integers(-4, 5, 200) * 1e-8
    other = (pos[idx].astype(np.float64) + d * (h + eps)[:, None]).astype(np.float32)
    allp = np.concatenate([pos, other])
    g, o = P.make_pair(allp, mass=0.99)
    for which, ofn in (("density", o.neighbor_stats), ("force", o.force_neighbor_stats)):
        c, s, x = g.gather_stats(which)
        oc, os_, ox = ofn()
        assert np.array_equal(c, oc) and np.array_equal(s, os_) and np.array_equal(x, ox), which
    # the band is populated: some placed pairs are neighbours, some are not
    r = np.linalg.norm(other.astype(np.float64) - pos[idx].astype(np.float64), axis=1)
    assert (r < h).any() and (r >= h).any()
    x0 = P.step_both(g, o)
    err = P.compare_step(g, o, x0)
    _print("compressed boundary", err)
    for k, v in err.items():
        assert v <= P.TOL_STEP, (k, v)

def test_per_step_parity_over_a_collapse():
    # the oracle is re-fed the GPU state every step: the per-step bar holds along
    # a trajectory that compresses the column against the floor
    sc = scenes.dam_break((12, 12, 12), seed=4)
    g, o = P.make_pair(sc.pos, mass=sc.mass, bounds=(sc.bounds_min, sc.bounds_max))
    worst = {}
    for k in range(60):
        P.sync_oracle_to_gpu(g, o)
        x0 = P.step_both(g, o)
        err = P.compare_step(g, o, x0)
        for f, v in err.items():
            worst[f] = max(worst.get(f, 0.0), v)
    _print("per-step worst", worst)
    for f, v in worst.items():
        assert v <= P.TOL_STEP, (f, v)


def test_trajectory_parity_first_20_steps():
    # independent runs (no re-feeding); reported per step, bounded loosely
    sc = scenes.dam_break((20, 20, 20), seed=1)
    g, o = P.make_pair(sc.pos, mass=sc.mass, bounds=(sc.bounds_min, sc.bounds_max))
    for k in range(20):
        g.step(P.DT32)
        o.step(P.DT32)
    e_pos = P.normwise(g.get("pos"), o.get("pos"))
    e_vel = P.normwise(g.get("vel"), o.get("vel"))
    e_rho = P.normwise(g.get("rho"), o.get("rho"))
    print(f"20-step trajectory: pos={e_pos:.2e} vel={e_vel:.2e} rho={e_rho:.2e}", flush=True)
    assert e_pos < 1e-5 and e_vel < 1e-3 and e_rho < 1e-4


# ---- edge cases of the reference semantics -----------------------------------

def test_overflow_and_nan_particles_match_oracle():
    # one particle far outside the dense table (overflow segment), one NaN
    # position (reference key 0 on that axis, poisons force sums of neighbours)
    pos = _block(8, 5, jit=0.1, origin=(-0.4, -0.4, -0.4)).astype(np.float32)
    far = np.array([[3.0e6, 0.01, 0.02]], dtype=np.float32)
    nanp = np.array([[np.nan, 0.03, 0.05]], dtype=np.float32)
    pos = np.concatenate([pos, far, nanp])
    g, o = P.make_pair(pos, mass=1.2, bounds=((-5, -5, -5), (5, 5, 5)))
    g.step(P.DT32)
    o.step(P.DT32)
    for f in ("pos", "vel", "rho"):
        a, b = g.get(f), o.get(f)
        assert np.array_equal(np.isnan(a), np.isnan(b)), f
        fin = np.isfinite(b)
        assert P.normwise(np.where(fin, a, 0), np.where(fin, b, 0)) <= P.TOL_STEP, f


def test_crowded_cell_matches_oracle():
    # 2,600 particles in a cube of side h/8 inside one half cell: the cell holds
    # more members than the reorder stages in LDS (2048 slots), so their
    # within-cell ranks are counted from global memory
    h = float(scenes.H32)
    rng = np.random.default_rng(11)
    pos = (0.1 * h + rng.random((2600, 3)) * (h / 8)).astype(np.float32)
    g, o = P.make_pair(pos, mass=0.01)
    x0 = P.step_both(g, o)
    err = P.compare_step(g, o, x0)
    _print("crowded cell", err)
    for k, v in err.items():
        assert v <= P.TOL_STEP, (k, v)


def test_large_overflow_segment_matches_oracle():
    # 2,100 particles far outside the bounds box (the table is clipped to the
    # box, so they form the overflow segment) beside a dense block: the
    # segment is longer than the reorder's LDS window
    blk = _block(8, 6, jit=0.1, origin=(-0.4, -0.4, -0.4)).astype(np.float32)
    far = np.zeros((2100, 3), np.float32)
    far[:, 0] = 3.0e6 + np.arange(2100, dtype=np.float32)
    far[:, 1] = 0.01
    far[:, 2] = 0.02
    pos = np.concatenate([blk, far[::-1]])
    g, o = P.make_pair(pos, mass=1.2, bounds=((-5, -5, -5), (5, 5, 5)))
    x0 = P.step_both(g, o)
    err = P.compare_step(g, o, x0)
    _print("overflow segment", err)
    for k, v in err.items():
        assert v <= P.TOL_STEP, (k, v)


def test_coincident_single_and_empty():
    g, o = P.make_pair(np.array([[0.1, 0.2, 0.3], [0.1, 0.2, 0.3], [0.15, 0.2, 0.3]], np.float32),
                       mass=2.0)
    x0 = P.step_both(g, o)
    err = P.compare_step(g, o, x0)
    for k, v in err.items():
        assert v <= P.TOL_STEP, (k, v)
    gp, _ = P.params_pair()
    e = ffi.FluidWorld(gp)
    e.step(0.01)                       # empty world: no-op (world.rs:184-186)
    assert len(e) == 0 and e.time() == pytest.approx(0.01)


# ---- determinism --------------------------------------------------------------

def test_repeat_runs_are_bit_identical():
    sc = scenes.dam_break((24, 24, 24), seed=2)
    outs = []
    for _ in range(2):
        gp, _ = P.params_pair(bounds=(sc.bounds_min, sc.bounds_max))
        g = ffi.FluidWorld(gp)
        g.set_particles_f32(sc.pos, None, np.full(len(sc.pos), sc.mass, np.float32))
        g.step(P.DT32, 150)
        outs.append((g.get("pos", f32=True), g.get("vel", f32=True), g.get("rho", f32=True)))
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---- the reference FFI contract (crates/phy-ffi/src/lib.rs:906-1001) -----------

def test_ffi_fluid_lifecycle_and_readback():
    w = ffi.World.create_fluid()
    assert w.fluid_count() == 10648
    t0 = w.time()
    for _ in range(200):
        w.step(0.005)
    assert abs(w.time() - t0 - 1.0) < 1e-6
    assert np.isfinite(w.positions()).all()


def test_ffi_checked_step_and_f32_channel():
    w = ffi.World.create_fluid()
    for _ in range(200):
        assert w.step_checked(0.005) == 0
    L = ffi.lib()
    n = w.fluid_count()
    b64 = np.zeros(3 * n)
    b32 = np.zeros(3 * n, dtype=np.float32)
    assert L.phy_world_get_fluid_positions(w.handle, ffi._dp(b64), b64.size) == n
    assert L.phy_world_get_fluid_positions_f32(w.handle, ffi._fp(b32), b32.size) == n
    assert np.array_equal(b32, b64.astype(np.float32))
    # truncation at len / 3 particles
    assert L.phy_world_get_fluid_positions(w.handle, ffi._dp(b64), 3 * 5 + 2) == 5


def test_watchdog_catches_injected_nan():
    gp = ffi.PhySphParams.defaults()
    g = ffi.FluidWorld(gp)
    g.fill_box((-4, -4, -1), (0, 4, 1), 0.18, 0.05)
    for _ in range(5):
        assert g.step_checked(0.005) == 0
    pos, vel, m = g.get("pos"), g.get("vel"), g.get("mass")
    pos[0, 0] = np.nan
    g.set_particles(pos, vel, m)
    assert g.step_checked(0.005) == 2


def _sph_world_with(gp, n):
    # crates/phy-demo/tests/stability.rs:22-37: spacing from n in a 10 m box
    half = 5.0
    spacing = (8.0 * half ** 3 / n) ** (1.0 / 3.0)
    gp.h = spacing * 2.0
    g = ffi.FluidWorld(gp)
    g.fill_box((-half, 1.0, -half), (half, 1.0 + 2 * half, half), spacing, 0.1)
    return g


def test_extreme_stiffness_detected_or_bounded():       # stability.rs:140-173
    gp = ffi.PhySphParams.defaults()
    gp.stiffness = 1.0e6
    g = _sph_world_with(gp, 400)
    saw_err, e_max = False, 0.0
    for _ in range(50):
        if g.step_checked(0.01) != 0:
            saw_err = True
            break
        e = g.diagnostics()[4]
        if np.isfinite(e):
            e_max = max(e_max, e)
    assert saw_err or np.isfinite(e_max)


def test_large_dt_does_not_crash():                      # stability.rs:196-220
    g = _sph_world_with(ffi.PhySphParams.defaults(), 400)
    for _ in range(50):
        assert g.step_checked(0.5) in (0, 2)


def test_zero_mass_fires_the_watchdog_like_the_oracle():
    # mass 0: rho = 0, then m_j / rho_j = 0/0 in the force sums -> NaN velocity
    pos = _block(6, 2, jit=0.1)
    g, o = P.make_pair(pos, mass=0.0)
    rc = g.step_checked(P.DT32)
    o.step(P.DT32)
    # x' = x + v' dt turns NaN too, and validate checks pos first: NonFinite{pos}
    assert o.validate() == 1
    assert rc == 2
    for f in ("pos", "vel"):
        assert np.array_equal(np.isnan(g.get(f)), np.isnan(o.get(f))), f


# ---- FP64 diagnostics -------------------------------------------------------------

def test_fp64_diagnostics_match_oracle():
    pos = _block(16, 3, jit=0.1)
    n = len(pos)
    vel = (scenes.rng_f64(8, 3 * n).reshape(n, 3) - 0.5).astype(np.float32)
    g, o = P.make_pair(pos, vel, mass=0.9)
    d = g.diagnostics()
    assert d[0] == pytest.approx(n * float(np.float32(0.9)), rel=1e-12)
    np.testing.assert_allclose(d[1:4], o.total_momentum(), rtol=1e-10, atol=1e-10)
    assert d[4] == pytest.approx(o.kinetic_energy(), rel=1e-10)


# ---- full size (1M particles): the configuration the throughput is quoted on -------

def test_full_size_column_one_step_parity_and_invariants():
    sc = scenes.dam_break((100, 100, 100), seed=1)
    g, o = P.make_pair(sc.pos, mass=sc.mass, bounds=(sc.bounds_min, sc.bounds_max))
    # neighbour sets on the identical input state
    gc, gs, gx = g.neighbor_stats()
    oc, os_, ox = o.neighbor_stats()
    assert np.array_equal(gc, oc) and np.array_equal(gs, os_) and np.array_equal(gx, ox)
    x0 = P.step_both(g, o)
    err = P.compare_step(g, o, x0)
    _print("1M one step", err)
    for k, v in err.items():
        assert v <= P.TOL_STEP, (k, v)
    m0 = g.diagnostics()[0]
    g.step(P.DT32, 30)
    x = g.get("pos")
    assert len(x) == 1_000_000 and np.isfinite(x).all()
    assert (x >= 0.0).all() and (x <= sc.bounds_max[0] * (1 + 1e-7)).all()
    assert g.diagnostics()[0] == m0           # mass is carried, never recomputed
