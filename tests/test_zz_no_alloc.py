"""phy_fluid_diagnostics_f64_noalloc: the FP64 diagnostics of caller-owned
device arrays with no allocation (per-tile partials in a file-scope device
array), checked bit for bit against phy_fluid_diagnostics_f64 of a world
holding the same particles in the same order, on two calls in a row (the
second smaller, so stale partials of the first would show), plus the size
limit. The inputs and outputs are device buffers the test allocates itself
through the HIP runtime the library is linked against (plumbing)."""
import ctypes as C

import numpy as np
import pytest

from physics_amd import ffi, scenes

pytestmark = pytest.mark.gpu


def _state(n, seed):
    pos = scenes.lattice_block(n, n, n, scenes.S32, (0.1, 0.2, 0.3), 0.3, seed)
    m = len(pos)
    vel = (scenes.rng_f64(seed + 1, 3 * m).reshape(m, 3) - 0.5).astype(np.float32)
    mass = (0.5 + scenes.rng_f64(seed + 2, m)).astype(np.float32)
    return pos, vel, mass


def _world_diagnostics(pos, vel, mass):
    g = ffi.FluidWorld(ffi.PhySphParams.defaults())
    g.set_particles_f32(pos, vel, mass)     # index order: no step, so no reordering
    d = g.diagnostics()
    g.close()
    return d


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libphy_ffi.so loaded."""

    def __init__(self):
        ffi.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 16)) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
        return p

    def download(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0   # device to host
        return out

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc(hip, pos, vel, mass):
    n = len(pos)
    p4 = np.concatenate([pos, mass[:, None]], 1).astype(np.float32)
    v4 = np.concatenate([vel, np.zeros((n, 1), np.float32)], 1).astype(np.float32)
    dp, dv = hip.upload(p4), hip.upload(v4)
    out = hip.upload(np.full(5, np.nan))
    rc = ffi.lib().phy_fluid_diagnostics_f64_noalloc(dp, dv, n, out, None)
    assert rc == 0, ffi.last_error()
    hip.sync()
    return hip.download(out, 5, np.float64)


def test_noalloc_diagnostics_match_the_world_path_on_two_calls():
    hip = _Hip()
    try:
        for n_side, seed in ((40, 3), (13, 4)):        # 64,000 then 2,197 particles
            pos, vel, mass = _state(n_side, seed)
            want = _world_diagnostics(pos, vel, mass)
            got = _noalloc(hip, pos, vel, mass)
            print(f"n={len(pos)}: world {want.tolist()} noalloc {got.tolist()}", flush=True)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
            # the f64 sums themselves
            assert got[0] == pytest.approx(float(mass.astype(np.float64).sum()), rel=1e-12)
    finally:
        hip.free()


def test_noalloc_diagnostics_reject_oversized_input_and_empty_is_zero():
    hip = _Hip()
    try:
        out = hip.upload(np.full(5, np.nan))
        L = ffi.lib()
        # larger than the scratch array: refused before any launch (pointers unused)
        rc = L.phy_fluid_diagnostics_f64_noalloc(out, out, (1 << 24) + 1, out, None)
        assert rc == -1 and "2^24" in ffi.last_error()
        rc = L.phy_fluid_diagnostics_f64_noalloc(None, None, 0, out, None)
        assert rc == 0, ffi.last_error()
        hip.sync()
        assert (hip.download(out, 5, np.float64) == 0.0).all()
    finally:
        hip.free()
