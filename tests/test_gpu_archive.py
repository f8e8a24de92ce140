"""JSON world archive (phy_world_save / phy_world_load): the reference's
phy-io format (crates/phy-io/src/archive.rs) for fluid worlds.

Pinned by the reference's own tests: ffi_save_load_roundtrip (crates/phy-ffi/
src/lib.rs:935-952: 50 steps of 0.005, save, load, equal subsystem counts) and
save_reload_replay_is_deterministic (crates/phy-demo/tests/determinism.rs:
65-98: the reloaded world replays the next 50 steps within 1e-6; here the state
is FP32 on both paths, so the replay is bit-identical). The committed fixture
tests/golden/fluid_archive_small.json is written by tests/golden/
make_archive_fixture.py in the reference layout.
"""
import json
import os
import struct

import numpy as np
import pytest

import oracle as O
from physics_amd import ffi
from tests.parity import TOL_STEP, normwise

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "fluid_archive_small.json")


def _f64(bits):
    return struct.unpack("<d", struct.pack("<Q", bits))[0]


def test_save_load_roundtrip_and_replay(tmp_path):
    w = ffi.World.create_fluid()
    for _ in range(50):
        assert w.step(0.005) is None
    path = str(tmp_path / "world.json")
    w.save(path)
    w2 = ffi.World.load(path)
    assert w2.sub_count() == w.sub_count() == 1            # lib.rs:948
    assert w2.fluid_count() == w.fluid_count() == 10648
    assert w2.time() == w.time()
    assert np.array_equal(w2.positions(), w.positions())
    assert np.array_equal(w2.velocities(), w.velocities())
    # determinism.rs:65-98: the reloaded world replays the next 50 steps
    for _ in range(50):
        w.step(0.005)
        w2.step(0.005)
    assert np.array_equal(w2.positions(), w.positions())
    assert np.array_equal(w2.velocities(), w.velocities())
    assert w2.time() == w.time()


def test_saved_file_has_the_reference_layout(tmp_path):
    w = ffi.World.create_fluid()
    w.step(0.005, 3)
    path = str(tmp_path / "w.json")
    w.save(path)
    d = json.load(open(path))
    assert d["version"] == 1 and _f64(d["time"]["__f64"]) == w.time()
    fl = d["subsystems"][0]["Fluid"]
    assert list(fl) == ["world", "thermal_expansion", "heat_gain", "t_ref", "drag", "friction"]
    assert _f64(fl["drag"]["__f64"]) == 1.0 and _f64(fl["friction"]["__f64"]) == 0.1
    parts = fl["world"]["particles"]
    x = np.array([[_f64(b) for b in q["pos"]] for q in parts])
    assert np.array_equal(x, w.positions())
    rho = np.array([_f64(q["rho"]["__f64"]) for q in parts])
    assert np.isfinite(rho).all() and rho.max() > 0.0


def test_load_reference_layout_fixture_then_step_matches_oracle():
    d = json.load(open(FIXTURE))
    parts = d["subsystems"][0]["Fluid"]["world"]["particles"]
    x64 = np.array([[_f64(b) for b in q["pos"]] for q in parts])
    v64 = np.array([[_f64(b) for b in q["vel"]] for q in parts])
    m64 = np.array([_f64(q["mass"]["__f64"]) for q in parts])
    mat = np.array([q["material"] for q in parts], dtype=np.uint64)
    body = np.array([[_f64(b) for b in q["body_acc"]] for q in parts])
    w = ffi.World.load(FIXTURE)
    assert w.fluid_count() == 27 and w.time() == 0.25
    # positions are rounded to FP32 on load
    assert np.array_equal(w.positions(), x64.astype(np.float32).astype(np.float64))
    # one step against the oracle fed the same FP32 state, materials and body_acc
    p = O.Params.defaults()
    p.set_visc([3.5, 10.0], [1.0, 1.0])
    o = O.World(p)
    o.set_particles(x64.astype(np.float32).astype(np.float64), v64.astype(np.float32).astype(np.float64),
                    m64.astype(np.float32).astype(np.float64), mat, body)
    w.step(0.01)
    o.step(float(np.float32(0.01)))
    assert normwise(w.velocities(), o.get("vel")) < TOL_STEP
    assert w.time() == 0.26


def test_resave_keeps_particle_state_bits(tmp_path):
    w = ffi.World.load(FIXTURE)
    path = str(tmp_path / "again.json")
    w.save(path)
    a = json.load(open(FIXTURE))["subsystems"][0]["Fluid"]["world"]
    b = json.load(open(path))["subsystems"][0]["Fluid"]["world"]
    assert a["params"] == b["params"]
    for qa, qb in zip(a["particles"], b["particles"]):
        # FP32 rounding of the loaded state: compare through float32
        for f in ("pos", "vel"):
            assert [np.float32(_f64(v)) for v in qa[f]] == [np.float32(_f64(v)) for v in qb[f]]
        assert qa["material"] == qb["material"]
        assert qa["body_acc"] == qb["body_acc"]
