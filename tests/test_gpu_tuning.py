"""Autotuner and launch choices of the step.

Every list length, launch form, cell size (h / 2, h) and workgroup size
(256, 128) must keep one-step parity with the f64 oracle; the density pass's
summation order does not depend on its list length (bit-identical rho) and no
result depends on the workgroup size; the autotuner must leave the state
untouched and its cached choice must be reapplied without timing.
"""
import os

import numpy as np
import pytest

from physics_amd import ffi, scenes
from tests.parity import DT32, TOL_STEP, compare_step, make_pair, step_both

pytestmark = pytest.mark.gpu

LISTS = (8, 16, 32)


def _compressed_block(seed=21):
    pos = scenes.lattice_block(14, 9, 14, 0.07, (-0.5, -2.0, -0.5), 0.2, seed)
    vel = (0.2 * (scenes.rng_f64(seed + 1, pos.size) - 0.5)).reshape(-1, 3).astype(np.float32)
    return pos, vel


@pytest.mark.parametrize("block", (256, 128))
@pytest.mark.parametrize("cell_div", (2, 1))
@pytest.mark.parametrize("split", (0, 1))
@pytest.mark.parametrize("dl", LISTS)
@pytest.mark.parametrize("fl", LISTS)
def test_every_launch_choice_keeps_one_step_parity(dl, fl, split, cell_div, block):
    pos, vel = _compressed_block()
    g, o = make_pair(pos, vel, mass=0.9)
    g.set_tuning(dl, fl, split, cell_div, block)
    assert g.tuning() == (dl, fl, split, cell_div, block)
    x0 = step_both(g, o)
    err = compare_step(g, o, x0)
    print(f"density_list {dl} force_list {fl} split {split} cell_div {cell_div} block {block}: {err}",
          flush=True)
    assert max(err.values()) < TOL_STEP


@pytest.mark.parametrize("cell_div", (2, 1))
def test_workgroup_size_does_not_change_results(cell_div):
    """128- and 256-lane workgroups walk the same lanes through the same scans."""
    sc = scenes.dam_break((20, 20, 20), seed=9)
    p = scenes.apply_scene(ffi.PhySphParams.defaults(), sc)
    outs = []
    for block in (256, 128):
        g = ffi.FluidWorld(p)
        g.set_particles_f32(sc.pos, None, np.full(len(sc.pos), sc.mass, np.float32))
        g.set_tuning(16, 32, 0, cell_div, block)
        g.step(DT32, 120)
        outs.append((g.get("pos", f32=True), g.get("vel", f32=True)))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_cell_size_switches_keep_per_step_parity_through_a_collapse():
    """Cells of h and of h / 2 alternate along a collapse (tables resized each
    time); every step keeps one-step parity with the oracle fed the GPU state."""
    from tests.parity import sync_oracle_to_gpu
    sc = scenes.dam_break((14, 14, 14), seed=5)
    g, o = make_pair(sc.pos, mass=sc.mass, bounds=(sc.bounds_min, sc.bounds_max))
    worst = 0.0
    for k in range(80):
        g.set_tuning(16, 32, 0, 1 if (k // 10) % 2 else 2, 256)
        sync_oracle_to_gpu(g, o)
        x0 = step_both(g, o)
        worst = max(worst, max(compare_step(g, o, x0).values()))
    print(f"worst one-step error over 80 steps switching cell sizes: {worst:.2e}", flush=True)
    assert worst < TOL_STEP


@pytest.mark.parametrize("fl", LISTS)
def test_launch_form_does_not_change_results(fl):
    """The split and combined launches build the same lists in the same order."""
    sc = scenes.dam_break((20, 20, 20), seed=9)
    p = scenes.apply_scene(ffi.PhySphParams.defaults(), sc)
    outs = []
    for split in (0, 1):
        g = ffi.FluidWorld(p)
        g.set_particles_f32(sc.pos, None, np.full(len(sc.pos), sc.mass, np.float32))
        g.set_tuning(16, fl, split)
        g.step(DT32, 120)   # through the collapse: uniform and mixed waves
        outs.append((g.get("pos", f32=True), g.get("vel", f32=True)))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_density_sum_order_does_not_depend_on_list_length():
    pos, vel = _compressed_block(22)
    rhos = []
    for dl in LISTS:
        g, _ = make_pair(pos, vel, mass=0.9)
        g.set_tuning(dl, 32)
        g.step(DT32)
        rhos.append(g.get("rho", f32=True))
    assert all(np.array_equal(r, rhos[0]) for r in rhos[1:])


def test_autotune_restores_state_and_caches(tmp_path):
    sc = scenes.dam_break((24, 24, 24), seed=3)
    p = scenes.apply_scene(ffi.PhySphParams.defaults(), sc)
    m = np.full(len(sc.pos), sc.mass, np.float32)
    g = ffi.FluidWorld(p)
    g.set_particles_f32(sc.pos, None, m)
    g.step(DT32, 30)
    x, v = g.get("pos", f32=True), g.get("vel", f32=True)
    cache = str(tmp_path / "tune.txt")
    cached, choice, ms = g.autotune(DT32, 4, cache)
    print("choice", choice, "ms", ms, flush=True)
    assert not cached and choice[0] in LISTS and choice[1] in LISTS and choice[2] in (0, 1)
    assert choice[3] in (1, 2) and choice[4] in (128, 256)
    assert np.isfinite(ms).all()
    assert np.array_equal(g.get("pos", f32=True), x) and np.array_equal(g.get("vel", f32=True), v)
    assert os.path.exists(cache)
    # the tuned world steps exactly like a world given the same choice by hand
    h = ffi.FluidWorld(p)
    h.set_particles_f32(sc.pos, None, m)
    h.step(DT32, 30)
    h.set_tuning(*choice)
    g.step(DT32, 5)
    h.step(DT32, 5)
    assert np.array_equal(g.get("pos", f32=True), h.get("pos", f32=True))
    # a second world of the same class takes the stored choice without timing
    k = ffi.FluidWorld(p)
    k.set_particles_f32(sc.pos, None, m)
    cached2, choice2, ms2 = k.autotune(DT32, 4, cache)
    assert cached2 and choice2 == choice and np.isnan(ms2).all()


def test_batch_autotune_keeps_scenes_exact(tmp_path):
    counts = [700, 2100, 300]
    scs = [scenes.batch_scene(n, 60 + s) for s, n in enumerate(counts)]
    params = [scenes.apply_scene(ffi.PhySphParams.defaults(), sc) for sc in scs]
    pos = np.concatenate([sc.pos for sc in scs])
    vel = np.concatenate([sc.vel for sc in scs])
    m = np.concatenate([np.full(n, sc.mass, np.float32) for n, sc in zip(counts, scs)])
    b = ffi.FluidBatch(params)
    b.set_particles_f32(counts, pos, vel, m)
    b.step(DT32, 3)
    x = b.get("pos")
    cached, choice, ms = b.autotune(DT32, 3, str(tmp_path / "t.txt"))
    assert not cached and np.array_equal(b.get("pos"), x)
    b.step(DT32, 4)
    off = b.offsets()
    for s, sc in enumerate(scs):
        w = ffi.FluidWorld(params[s])
        w.set_particles_f32(sc.pos, sc.vel, np.full(counts[s], sc.mass, np.float32))
        w.step(DT32, 3)
        w.set_tuning(*choice)
        w.step(DT32, 4)
        assert np.array_equal(b.get("pos")[off[s]:off[s + 1]], w.get("pos", f32=True))
