"""Flat export (phy_fluid_flat_export): SphFlatData / Grid::to_flat
(crates/phy-fluid/src/sph/world.rs:782-859, grid.rs:62-175) against the
oracle's restatement of to_flat on the same FP32 state (exact: integer keys,
prefix sums and index order)."""
import numpy as np
import pytest

import oracle as O
from physics_amd import ffi, scenes
from tests.parity import DT32

pytestmark = pytest.mark.gpu


def _check(g, params_o):
    ex = g.flat_export()
    x = g.get("pos")
    o = O.World(params_o)
    o.set_particles(x, g.get("vel"), g.get("mass"))
    mn, nc, cs, so = o.flat_grid()
    assert np.array_equal(ex["grid_min"], mn) and np.array_equal(ex["nc"], nc)
    assert np.array_equal(ex["cell_start"], cs)
    assert np.array_equal(ex["sorted"], so)
    n = ex["n"]
    if n:
        assert np.array_equal(ex["pos"][:, :3], g.get("pos", f32=True), equal_nan=True)
        assert np.array_equal(ex["vel"][:, :3], g.get("vel", f32=True))
        assert (ex["pos"][:, 3] == 0).all()
        assert np.array_equal(ex["scalar"][:, 2], g.get("mass", f32=True))
    return ex


def test_flat_export_matches_to_flat_through_a_collapse():
    sc = scenes.dam_break((16, 16, 16), seed=2)
    gp = scenes.apply_scene(ffi.PhySphParams.defaults(), sc)
    op = scenes.apply_scene(O.Params.defaults(), sc)
    g = ffi.FluidWorld(gp)
    g.set_particles_f32(sc.pos, None, np.full(len(sc.pos), sc.mass, np.float32))
    for steps in (0, 1, 120):
        g.step(DT32, steps) if steps else None
        ex = _check(g, op)
        if steps:
            assert np.array_equal(ex["scalar"][:, 0], g.get("rho", f32=True))
            assert np.array_equal(ex["scalar"][:, 1], g.get("p", f32=True))
    assert ex["h"] == np.float32(sc.h) and ex["stiffness"] == 250.0
    assert list(ex["visc_k"]) == [3.5] and list(ex["gravity"]) == [0.0, np.float32(-9.81), 0.0]


def test_flat_export_edge_cases():
    p_g, p_o = ffi.PhySphParams.defaults(), O.Params.defaults()
    g = ffi.FluidWorld(p_g)                       # empty world: one cell
    ex = g.flat_export()
    assert ex["n"] == 0 and list(ex["nc"]) == [1, 1, 1] and list(ex["cell_start"]) == [0, 0]
    # NaN particle -> key 0 (to_i64 failure), negative keys, materials in scalar.w
    pos = np.array([[0.31, -0.45, 1.2], [np.nan, 0.1, 0.1], [-0.7, 0.05, 0.33], [0.32, -0.44, 1.21]],
                   np.float32)
    g.set_particles_f32(pos, None, np.full(4, 0.5, np.float32), np.array([0, 3, 1, 2], np.uint32))
    ex = _check(g, p_o)
    assert list(ex["scalar"][:, 3]) == [0.0, 3.0, 1.0, 2.0]
