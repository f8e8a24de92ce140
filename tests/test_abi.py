"""C-ABI checks that need no GPU: libphy_ffi.so loads, exports every symbol of
include/phy_ffi.h, and honours the reference's null-handle contract
(crates/phy-ffi/src/lib.rs:954-961, 1003-1015)."""
import ctypes as C
import math
import os
import re
import subprocess

import pytest

from physics_amd import ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "phy_ffi.h")


def _header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(phy_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_and_binding_agree():
    assert _header_functions() == sorted(ffi.EXPORTS)


def test_library_exports_every_header_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", ffi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in _header_functions() if n not in exported]
    assert not missing, missing


def test_abi_version_and_null_handles():
    L = ffi.lib()
    assert L.phy_ffi_abi_version() == 2
    assert L.phy_world_step(None, 0.01) == -1
    assert L.phy_world_step_checked(None, 0.01) == -1
    assert L.phy_world_step_f32(None, 0.01) == -1
    assert math.isnan(L.phy_world_time(None))
    assert L.phy_world_fluid_count(None) == 0
    assert L.phy_world_sub_count(None) == 0
    assert L.phy_world_get_fluid_positions(None, None, 0) == 0
    assert L.phy_world_get_fluid_positions_f32(None, None, 0) == 0
    assert L.phy_world_get_rigid_transforms_f32(None, None, 0) == 0
    L.phy_world_destroy(None)  # null-safe no-op


def test_out_of_scope_factories_return_null():
    L = ffi.lib()
    for name in ["phy_world_create_rigid", "phy_world_create_rigid_empty",
                 "phy_world_create_granular", "phy_world_create_coupled",
                 "phy_world_create_soft_empty", "phy_world_create_field_empty",
                 "phy_world_create_optics_empty", "phy_world_create_solid_empty"]:
        assert getattr(L, name)() is None
        assert "not provided" in ffi.last_error()


def test_archive_error_paths(tmp_path):
    """phy_world_save / phy_world_load fail with their documented values, and
    the archive parser rejects what the reference's loader rejects, before any
    world (GPU) is created."""
    L = ffi.lib()
    assert L.phy_world_save(None, str(tmp_path / "x.json").encode()) == -1
    assert "null world" in ffi.last_error()
    assert L.phy_world_load(str(tmp_path / "missing.json").encode()) is None
    assert "cannot open" in ffi.last_error()
    cases = {
        "truncated": '{"version": 1, "time": {"__f64": 0}, "subsystems": [',
        "not json": "hello",
        "rigid": '{"version": 1, "time": {"__f64": 0}, "subsystems": [{"Rigid": {}}]}',
        "nan time": '{"version": 1, "time": {"__f64": %d}, "subsystems": []}' % 0x7FF8000000000000,
        "missing field": '{"version": 1, "subsystems": []}',
    }
    for name, text in cases.items():
        f = tmp_path / f"{name.replace(' ', '_')}.json"
        f.write_text(text)
        assert L.phy_world_load(str(f).encode()) is None, name
        assert "archive" in ffi.last_error(), (name, ffi.last_error())


def test_archive_nesting_limit(tmp_path):
    """serde_json stops at 128 nested levels with an error (its recursion
    limit), so phy_world_load returns NULL for a deeply nested file rather than
    exhausting the stack; 128 levels still parse (and then fail on content)."""
    L = ffi.lib()
    deep = tmp_path / "deep.json"
    deep.write_text("[" * 200000 + "]" * 200000)
    assert L.phy_world_load(str(deep).encode()) is None
    assert "nesting too deep" in ffi.last_error()
    ok = tmp_path / "ok.json"
    ok.write_text('{"version": 1, "time": ' + "[" * 127 + "]" * 127 + ', "subsystems": []}')
    assert L.phy_world_load(str(ok).encode()) is None
    assert "nesting" not in ffi.last_error(), ffi.last_error()


def test_archive_fixture_layout():
    """The committed fixture follows the reference layout (tests/golden/
    make_archive_fixture.py): tagged subsystems, {"__f64": bits} scalars, u64
    bit-pattern vectors, integer material tags."""
    import json, struct
    d = json.load(open(os.path.join(ROOT, "tests", "golden", "fluid_archive_small.json")))
    assert d["version"] == 1 and list(d["subsystems"][0]) == ["Fluid"]
    fl = d["subsystems"][0]["Fluid"]
    assert list(fl) == ["world", "thermal_expansion", "heat_gain", "t_ref", "drag", "friction"]
    p = fl["world"]["params"]
    assert list(p) == ["rest_density", "stiffness", "viscosity", "visc_k", "visc_n", "shear_min",
                       "mass", "h", "gravity", "bounds_min", "bounds_max", "boundary_damp", "xsph_eps"]
    q = fl["world"]["particles"][0]
    assert list(q) == ["pos", "vel", "acc", "body_acc", "rho", "p", "mass", "material"]
    h = struct.unpack("<d", struct.pack("<Q", p["h"]["__f64"]))[0]
    assert h == 0.2 and isinstance(q["material"], int) and len(q["pos"]) == 3


def test_params_defaults_match_reference():
    p = ffi.PhySphParams.defaults()                       # params.rs:54-85
    assert (p.rest_density, p.stiffness, p.viscosity) == (1000.0, 250.0, 3.5)
    assert (p.n_visc_k, p.visc_k[0], p.n_visc_n, p.visc_n[0]) == (1, 3.5, 1, 1.0)
    assert (p.shear_min, p.mass, p.h) == (0.01, 0.02, 0.2)
    assert list(p.gravity) == [0.0, -9.81, 0.0]
    assert list(p.bounds_min) == [-5.0] * 3 and list(p.bounds_max) == [5.0] * 3
    assert (p.boundary_damp, p.xsph_eps) == (0.4, 0.5)


def _has_gpu():
    # no torch here: it bundles a second HIP runtime
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_has_gpu(), reason="checks the no-GPU failure path")
def test_no_gpu_fails_loudly():
    # no CPU fallback: creation fails with a message instead of computing on the host
    with pytest.raises(ffi.PhyError, match="no HIP device"):
        ffi.World.create_fluid()
