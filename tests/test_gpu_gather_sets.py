"""The neighbour sets the density and force kernels sum over, checked against
the reference's sets, through the C ABI.

phy_fluid_gather_stats runs the step's own gathers (the same column-wave and
per-lane scans, row trimming, column-local filter, lists and exact accept
tests as k_density / k_force, with the world's current tuning) and records per
particle the count, index sum and hashed-index xor of the slots each pass's
exact test accepts. The oracle computes the same statistics from the
reference's 27-cell visit:
  density  j with r < h, self included (world.rs:220-230);
  force    j != i with 0 < r < h, a NaN distance kept (world.rs:292-299).
Bar: bit-exact (count, sum and xor of every particle).
"""
import numpy as np
import pytest

import oracle as O
import parity as P
from physics_amd import ffi, scenes

pytestmark = pytest.mark.gpu


def _assert_sets(g, o, tag=""):
    for which, ofn in (("density", o.neighbor_stats), ("force", o.force_neighbor_stats)):
        gc, gs, gx = g.gather_stats(which)
        oc, os_, ox = ofn()
        bad = np.flatnonzero((gc != oc) | (gs != os_) | (gx != ox))
        assert bad.size == 0, (tag, which, bad[:10], gc[bad[:5]], oc[bad[:5]])
        print(f"{tag} {which}: {len(gc)} particles, mean {gc.mean():.1f} neighbours, "
              f"max {gc.max()}", flush=True)


def _state_pair(g, op):
    """An oracle world holding the GPU world's current FP32 state."""
    o = O.World(op)
    o.set_particles(g.get("pos"), g.get("vel"), g.get("mass"))
    return o


def test_gather_sets_at_the_support_boundary():
    # pairs placed within a few ulps of r = h exercise the exact f64 decisions
    # of both passes
    h = scenes.H32
    rng = np.random.default_rng(7)
    base = rng.uniform(-1, 1, (400, 3)).astype(np.float32)
    d = rng.normal(size=(400, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    eps = rng.integers(-4, 5, 400) * 1e-8
    other = (base.astype(np.float64) + d * (h + eps)[:, None]).astype(np.float32)
    g, o = P.make_pair(np.concatenate([base, other]))
    _assert_sets(g, o, "support boundary")
    oc, _, _ = o.force_neighbor_stats()
    assert (oc[:400] == 1).any() and (oc[:400] == 0).any()   # the band is populated


def test_gather_sets_with_nan_coincident_and_overflow_particles():
    # a NaN position (reference key 0 on that axis: a force neighbour of the
    # particles in adjacent cells, never a density neighbour), coincident
    # points (density neighbours, never force neighbours), a particle far
    # outside the dense table (overflow segment)
    pos = scenes.lattice_block(8, 8, 8, scenes.S32, (-0.4, -0.4, -0.4), 0.1, 5).astype(np.float32)
    extra = np.array([[np.nan, 0.03, 0.05], [0.01, np.nan, np.nan], [3.0e6, 0.01, 0.02],
                      pos[10], pos[10], pos[200]], dtype=np.float32)
    g, o = P.make_pair(np.concatenate([pos, extra]), mass=1.2)
    _assert_sets(g, o, "nan/coincident/overflow")
    oc, _, _ = o.neighbor_stats()
    fc, _, _ = o.force_neighbor_stats()
    n0 = len(pos)
    assert oc[n0] == 0 and fc[n0] > 0          # the NaN particle poisons force sums only
    assert oc[n0 + 3] == oc[10] and fc[n0 + 3] == fc[10]   # coincident copies


@pytest.mark.parametrize("seed,jit", [(1, 0.01), (2, 0.3), (3, 0.5)])
def test_gather_sets_jittered_blocks(seed, jit):
    pos = scenes.lattice_block(12, 12, 12, scenes.S32, (-0.6, -0.6, -0.6), jit, seed)
    g, o = P.make_pair(pos)
    _assert_sets(g, o, f"block jit={jit}")


def test_gather_sets_column_after_60_steps():
    sc = scenes.dam_break((12, 12, 12), seed=4)
    g, o = P.make_pair(sc.pos, mass=sc.mass, bounds=(sc.bounds_min, sc.bounds_max))
    g.step(P.DT32, 60)
    _, op = P.params_pair(bounds=(sc.bounds_min, sc.bounds_max))
    _assert_sets(g, _state_pair(g, op), "12^3 column, step 60")


@pytest.mark.parametrize("block", [256, 128])
@pytest.mark.parametrize("cell_div", [2, 1])
@pytest.mark.parametrize("dl", [8, 16, 32])
@pytest.mark.parametrize("fl", [8, 16, 32])
@pytest.mark.parametrize("split", [0, 1])
def test_gather_sets_under_every_tuning(dl, fl, split, cell_div, block):
    # a compressed 20^3 column (column waves with trimmed rows, per-lane
    # waves at its edges) under every list length, launch form, cell size and
    # workgroup size
    sc = scenes.dam_break((20, 20, 20), seed=2)
    g, o = P.make_pair(sc.pos, mass=sc.mass, bounds=(sc.bounds_min, sc.bounds_max))
    g.step(P.DT32, 120)
    g.set_tuning(dl, fl, split, cell_div, block)
    _, op = P.params_pair(bounds=(sc.bounds_min, sc.bounds_max))
    _assert_sets(g, _state_pair(g, op), f"20^3 step 120 lists {dl}/{fl} split {split} cells h/{cell_div} "
                 f"block {block}")


@pytest.mark.parametrize("cell_div", [2, 1])
def test_gather_sets_with_cells_of_h_and_h2_at_rest_and_collapsed(cell_div):
    sc = scenes.dam_break((40, 40, 40), seed=3)
    gp, op = P.params_pair(bounds=(sc.bounds_min, sc.bounds_max))
    g = ffi.FluidWorld(gp)
    g.set_particles_f32(sc.pos, None, np.full(len(sc.pos), sc.mass, np.float32))
    g.set_tuning(16, 32, 0, cell_div, 256)
    _assert_sets(g, _state_pair(g, op), f"40^3 at rest, cells h/{cell_div}")
    g.step(P.DT32, 150)
    _assert_sets(g, _state_pair(g, op), f"40^3 step 150, cells h/{cell_div}")


def test_gather_sets_of_the_shear_and_recorded_variants():
    # power-law materials and a recording tape launch the force kernel that
    # also sums the shear rate: same gather, checked the same way
    pos = scenes.lattice_block(14, 14, 14, scenes.S32, (-0.7, -0.7, -0.7), 0.2, 9)
    n = len(pos)
    vel = (scenes.rng_f64(4, 3 * n).reshape(n, 3) - 0.5).astype(np.float32)
    mat = (np.arange(n) % 3).astype(np.uint32)
    g, o = P.make_pair(pos, vel, mass=1.6, material=mat, visc_k=(3.5, 10.0, 1.0),
                       visc_n=(1.0, 0.6, 1.4))
    _assert_sets(g, o, "power law")
    gp, op = P.params_pair()
    g2 = ffi.FluidWorld(gp)
    g2.set_particles_f32(pos, vel, np.full(n, 1.6, np.float32))
    g2.tape_begin(4)
    g2.step(P.DT32, 3)
    _assert_sets(g2, _state_pair(g2, op), "recording")
    g2.tape_end()


def test_gather_sets_batch_scenes_are_scene_local():
    scs = scenes.batch_scenes(6, seed=3, lo=500, hi=3000)
    counts = [len(sc.pos) for sc in scs]
    b = ffi.FluidBatch([scenes.apply_scene(ffi.PhySphParams.defaults(), sc) for sc in scs])
    b.set_particles_f32(counts, np.concatenate([sc.pos for sc in scs]),
                        np.concatenate([sc.vel for sc in scs]),
                        np.concatenate([np.full(len(sc.pos), sc.mass, np.float32) for sc in scs]))
    b.step(P.DT32, 40)
    off = b.offsets()
    x, v, m = b.get("pos"), b.get("vel"), b.get("mass")
    for which in ("density", "force"):
        gc, gs, gx = b.gather_stats(which)
        for s, sc in enumerate(scs):
            lo, hi = off[s], off[s + 1]
            o = O.World(scenes.apply_scene(O.Params.defaults(), sc))
            o.set_particles(x[lo:hi], v[lo:hi], m[lo:hi])
            oc, os_, ox = o.neighbor_stats() if which == "density" else o.force_neighbor_stats()
            assert np.array_equal(gc[lo:hi], oc), (which, s)
            assert np.array_equal(gs[lo:hi], os_), (which, s)
            assert np.array_equal(gx[lo:hi], ox), (which, s)
    b.close()


# ---- the 1M column: at rest (per-lane scans) and collapsed (trimmed column scans)

def test_gather_sets_full_size_column_at_step_0_and_collapsed():
    sc = scenes.dam_break((100, 100, 100), seed=1)
    g, o = P.make_pair(sc.pos, mass=sc.mass, bounds=(sc.bounds_min, sc.bounds_max))
    _assert_sets(g, o, "1M step 0")
    del o
    g.step(P.DT32, 200)
    _, op = P.params_pair(bounds=(sc.bounds_min, sc.bounds_max))
    o = _state_pair(g, op)
    c, _, _ = o.force_neighbor_stats()
    # the collapsed sheet: hundreds of neighbours per particle
    assert c.mean() > 300, c.mean()
    _assert_sets(g, o, "1M step 200")
