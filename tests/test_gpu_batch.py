"""Batched scenes: many independent FluidWorlds in one set of device arrays,
stepped with one launch per stage and replayed from a captured hipGraph.

Every scene of a batch must reproduce, bit for bit, the same scene stepped
alone as a FluidWorld (the per-particle gathers see the same neighbours in the
same order; only the scene lookup differs), so the single-world parity tests
against the f64 oracle carry over; each scene is also checked against the
oracle directly for one step.
"""
import numpy as np
import pytest

import oracle as O
from physics_amd import ffi, scenes
from tests.parity import DT32, TOL_STEP, normwise

pytestmark = pytest.mark.gpu

COUNTS = [300, 1000, 37, 2500, 0, 777, 1, 4096]


def _scene_params(cls, sc, k=250.0, mu=3.5):
    p = scenes.apply_scene(cls.defaults(), sc)
    p.stiffness = k
    p.set_visc([mu], [1.0])
    return p


def _build(counts=COUNTS, seed=3):
    scs = [scenes.batch_scene(n, seed * 1000 + s) if n else None for s, n in enumerate(counts)]
    params = []
    for s, sc in enumerate(scs):
        if sc is None:
            params.append(ffi.PhySphParams.defaults())
        else:
            # uneven parameters across scenes
            params.append(_scene_params(ffi.PhySphParams, sc, k=200.0 + 20 * s, mu=1.75 + 0.5 * s))
    pos = np.concatenate([sc.pos for sc in scs if sc is not None])
    vel = np.concatenate([sc.vel for sc in scs if sc is not None])
    mass = np.concatenate([np.full(sc.pos.shape[0], sc.mass, np.float32) for sc in scs if sc is not None])
    return scs, params, pos, vel, mass


def _single(sc, p):
    w = ffi.FluidWorld(p)
    w.set_particles_f32(sc.pos, sc.vel, np.full(len(sc.pos), sc.mass, np.float32))
    return w


def test_batch_scenes_match_single_worlds_bitwise():
    scs, params, pos, vel, mass = _build()
    b = ffi.FluidBatch(params)
    b.set_particles_f32(COUNTS, pos, vel, mass)
    off = b.offsets()
    assert list(np.diff(off)) == COUNTS
    b.step(DT32, 12)
    xb, vb, rb = b.get("pos"), b.get("vel"), b.get("rho")
    for s, sc in enumerate(scs):
        if sc is None:
            continue
        w = _single(sc, params[s])
        w.step(DT32, 12)
        lo, hi = off[s], off[s + 1]
        print(f"scene {s}: n={hi - lo}", flush=True)
        assert np.array_equal(xb[lo:hi], w.get("pos", f32=True))
        assert np.array_equal(vb[lo:hi], w.get("vel", f32=True))
        assert np.array_equal(rb[lo:hi], w.get("rho", f32=True))


def test_graph_replays_equal_eager_steps():
    scs, params, pos, vel, mass = _build(seed=4)
    a = ffi.FluidBatch(params)
    e = ffi.FluidBatch(params)
    e.set_graph(False)
    for x in (a, e):
        x.set_particles_f32(COUNTS, pos, vel, mass)
    for k in range(6):
        a.step(DT32, 1 + k)      # graph replays
        e.step(DT32, 1 + k)      # eager launches
        # eager host work between replays
        assert np.array_equal(a.get("pos"), e.get("pos")), f"after round {k}"
        assert np.array_equal(a.get("vel"), e.get("vel"))
        assert np.array_equal(a.get("rho"), e.get("rho"))
    # a new dt recaptures
    a.step(DT32 / 2, 3)
    e.step(DT32 / 2, 3)
    assert np.array_equal(a.get("pos"), e.get("pos"))


def test_single_world_graph_mode_is_bit_identical():
    sc = scenes.batch_scene(3000, 11)
    p = _scene_params(ffi.PhySphParams, sc)
    g, e = _single(sc, p), _single(sc, p)
    g.set_graph(True)
    for k in range(4):
        g.step(DT32, 5)
        e.step(DT32, 5)
        assert np.array_equal(g.get("pos", f32=True), e.get("pos", f32=True))
        assert np.array_equal(g.get("vel", f32=True), e.get("vel", f32=True))


def test_batch_one_step_parity_per_scene():
    scs, params, pos, vel, mass = _build(seed=5)
    b = ffi.FluidBatch(params)
    b.set_particles_f32(COUNTS, pos, vel, mass)
    b.step(DT32, 5)
    off = b.offsets()
    x0, v0 = b.get("pos"), b.get("vel")
    b.step(DT32)
    x1, v1, r1 = b.get("pos"), b.get("vel"), b.get("rho")
    for s, sc in enumerate(scs):
        if sc is None:
            continue
        lo, hi = off[s], off[s + 1]
        q = _scene_params(O.Params, sc, k=200.0 + 20 * s, mu=1.75 + 0.5 * s)
        o = O.World(q)
        o.set_particles(x0[lo:hi].astype(np.float64), v0[lo:hi].astype(np.float64),
                        np.full(hi - lo, float(np.float32(sc.mass))))
        o.step(DT32)
        # positions: one FP32 ulp of x' is storage rounding; only the excess
        # over it counts, relative to the step's displacement
        xo = o.get("pos")
        ulp = np.abs(np.spacing(xo.astype(np.float32)).astype(np.float64))
        excess = np.maximum(np.abs(x1[lo:hi] - xo) - ulp, 0.0)
        disp = np.sqrt(((xo - x0[lo:hi]) ** 2).sum(1)).max()
        e = {"rho": normwise(r1[lo:hi], o.get("rho")), "vel": normwise(v1[lo:hi], o.get("vel")),
             "dx": float(np.sqrt((excess ** 2).sum(1)).max() / disp)}
        print(f"scene {s} n={hi - lo}: {e}", flush=True)
        assert max(e.values()) < TOL_STEP


def test_batch_watchdog_is_per_scene():
    scs, params, pos, vel, mass = _build(seed=6)
    pos = pos.copy()
    b = ffi.FluidBatch(params)
    # poison one particle of scene 3 (offset of scene 3 = sum of earlier counts)
    k = sum(COUNTS[:3]) + 17
    pos[k, 1] = np.nan
    b.set_particles_f32(COUNTS, pos, vel, mass)
    st = b.step_checked(DT32)
    print("status", st, flush=True)
    assert st[3] == 1
    assert all(st[s] == 0 for s in range(len(COUNTS)) if s != 3)


def test_batch_diagnostics_and_neighbor_stats_per_scene():
    scs, params, pos, vel, mass = _build(seed=8)
    b = ffi.FluidBatch(params)
    b.set_particles_f32(COUNTS, pos, vel, mass)
    b.step(DT32, 3)
    off = b.offsets()
    d = b.diagnostics()
    x, v, m = b.get("pos").astype(np.float64), b.get("vel").astype(np.float64), b.get("mass").astype(np.float64)
    c, sm, xr = b.neighbor_stats()
    for s, sc in enumerate(scs):
        lo, hi = off[s], off[s + 1]
        ref = np.array([m[lo:hi].sum(), *(v[lo:hi] * m[lo:hi, None]).sum(0),
                        0.5 * ((v[lo:hi] ** 2).sum(1) * m[lo:hi]).sum()])
        assert np.allclose(d[s], ref, rtol=1e-12, atol=1e-12), (s, d[s], ref)
        if sc is None:
            continue
        o = O.World(_scene_params(O.Params, sc))
        o.set_particles(x[lo:hi], v[lo:hi], m[lo:hi])
        oc, osm, oxr = o.neighbor_stats()
        assert np.array_equal(c[lo:hi], oc) and np.array_equal(sm[lo:hi], osm) and np.array_equal(xr[lo:hi], oxr)


def test_batch_adjoint_matches_single_worlds():
    counts = [500, 64, 1200]
    scs = [scenes.batch_scene(n, 40 + s) for s, n in enumerate(counts)]
    params = []
    for s, sc in enumerate(scs):
        p = _scene_params(ffi.PhySphParams, sc, mu=2.0 + s)
        for a in range(3):   # walls out of reach
            p.bounds_min[a] = -100.0
            p.bounds_max[a] = 100.0
        params.append(p)
    pos = np.concatenate([sc.pos for sc in scs])
    vel = np.concatenate([sc.vel for sc in scs])
    mass = np.concatenate([np.full(n, sc.mass, np.float32) for n, sc in zip(counts, scs)])
    b = ffi.FluidBatch(params)
    b.set_particles_f32(counts, pos, vel, mass)
    b.tape_begin(4)
    b.step(DT32, 10)
    off = b.offsets()
    x, rho = b.get("pos"), b.get("rho")
    xbar = np.zeros_like(x)
    rhobar = np.zeros_like(rho)
    for s in range(3):
        lo, hi = off[s], off[s + 1]
        _, xbar[lo:hi], rhobar[lo:hi] = scenes.rollout_loss_cotangents(x[lo:hi], rho[lo:hi])
    x0b, v0b, dk, dmu = b.tape_backward(xbar, None, rhobar)
    for s, sc in enumerate(scs):
        lo, hi = off[s], off[s + 1]
        w = _single(sc, params[s])
        w.tape_begin(4)
        w.step(DT32, 10)
        gx0, gv0, gk, gmu = w.tape_backward(xbar[lo:hi].astype(np.float64), None,
                                            rhobar[lo:hi].astype(np.float64))
        print(f"scene {s}: dk {dk[s]} vs {gk}, dmu {dmu[s]} vs {gmu}", flush=True)
        assert np.array_equal(x0b[lo:hi], gx0.astype(np.float32))
        assert np.array_equal(v0b[lo:hi], gv0.astype(np.float32))
        assert dk[s] == gk and dmu[s] == gmu
