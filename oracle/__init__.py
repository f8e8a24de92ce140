"""ctypes binding of the f64 CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this package; it is the checker, never the product path.  The oracle
restates ``FluidWorld<f64>::step`` of crates/phy-fluid/src/sph/world.rs:183-397
(see sph_oracle.cpp for the per-function citations).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_MAX_MAT = 16

MUTATIONS = {
    "none": 0,
    "drop_xsph": 1,
    "visc_sign": 2,
    "no_pressure_clamp": 3,
    "self_in_force": 4,
    "density_no_self": 5,
    "bounds_no_damp": 6,
    "poly6_h2": 7,
    "dt_in_xsph": 8,
    "shear_no_reps": 9,
    "shear_mass_i": 10,
    "power_exp_n": 11,
    "shear_no_floor": 12,
    "shear_dv_sq": 13,
    "mat_oor_last": 14,
}


class Params(C.Structure):
    """Mirror of ``SphParams<f64>`` (crates/phy-fluid/src/sph/params.rs:12-50)."""

    _fields_ = [
        ("rest_density", C.c_double),
        ("stiffness", C.c_double),
        ("viscosity", C.c_double),
        ("visc_k", C.c_double * _MAX_MAT),
        ("visc_n", C.c_double * _MAX_MAT),
        ("n_visc_k", C.c_int32),
        ("n_visc_n", C.c_int32),
        ("shear_min", C.c_double),
        ("mass", C.c_double),
        ("h", C.c_double),
        ("gravity", C.c_double * 3),
        ("bounds_min", C.c_double * 3),
        ("bounds_max", C.c_double * 3),
        ("boundary_damp", C.c_double),
        ("xsph_eps", C.c_double),
    ]

    @classmethod
    def defaults(cls) -> "Params":
        p = cls()
        lib().or_params_defaults(C.byref(p))
        return p

    def set_visc(self, k, n) -> None:
        k = list(k)
        n = list(n)
        assert 1 <= len(k) <= _MAX_MAT and 1 <= len(n) <= _MAX_MAT
        for i, v in enumerate(k):
            self.visc_k[i] = v
        for i, v in enumerate(n):
            self.visc_n[i] = v
        self.n_visc_k = len(k)
        self.n_visc_n = len(n)

    def copy(self) -> "Params":
        q = Params()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(Params))
        return q


_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (g++, OpenMP, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    vp, dp, sz = C.c_void_p, C.POINTER(C.c_double), C.c_size_t
    L.or_params_size.restype = C.c_int
    L.or_params_defaults.argtypes = [C.POINTER(Params)]
    L.or_set_mutation.argtypes = [C.c_int]
    L.or_set_threads.argtypes = [C.c_int]
    L.or_get_threads.restype = C.c_int
    L.or_kernels.argtypes = [C.c_double, dp, C.c_double, dp]
    L.or_world_new.argtypes = [C.POINTER(Params)]
    L.or_world_new.restype = vp
    L.or_world_free.argtypes = [vp]
    L.or_world_clone.argtypes = [vp]
    L.or_world_clone.restype = vp
    L.or_world_set_params.argtypes = [vp, C.POINTER(Params)]
    L.or_world_len.argtypes = [vp]
    L.or_world_len.restype = sz
    L.or_world_lattice_mass.argtypes = [vp, C.c_double]
    L.or_world_lattice_mass.restype = C.c_double
    L.or_world_fill_box.argtypes = [vp, dp, dp, C.c_double, C.c_double]
    L.or_world_set_particles.argtypes = [vp, sz, dp, dp, dp, C.POINTER(C.c_uint64), dp]
    L.or_world_get.argtypes = [vp, C.c_int, dp]
    L.or_world_step.argtypes = [vp, C.c_double]
    L.or_world_step_n.argtypes = [vp, C.c_double, C.c_int]
    L.or_world_density.argtypes = [vp]
    L.or_world_forces.argtypes = [vp]
    L.or_world_validate.argtypes = [vp]
    L.or_world_validate.restype = C.c_int
    L.or_world_momentum.argtypes = [vp, dp]
    L.or_world_kinetic_energy.argtypes = [vp]
    L.or_world_kinetic_energy.restype = C.c_double
    L.or_world_neighbors.argtypes = [vp, sz, C.POINTER(C.c_uint64), sz]
    L.or_world_neighbors.restype = sz
    L.or_world_neighbor_stats.argtypes = [
        vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.or_world_force_neighbor_stats.argtypes = [
        vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.or_world_time_steps.argtypes = [vp, C.c_double, C.c_int]
    L.or_world_time_steps.restype = C.c_double
    L.or_world_flat_grid.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_int32), C.c_uint64, C.POINTER(C.c_int32), C.c_uint64]
    L.or_world_flat_grid.restype = C.c_int
    i32p = C.POINTER(C.c_int32)
    L.or_world_couple_bodies.argtypes = [vp, sz, dp, dp, dp, i32p, dp, C.c_double]
    L.or_world_couple_points.argtypes = [vp, sz, dp, dp, dp, dp, C.c_double, C.c_double, C.c_double]
    L.or_world_couple_heat.argtypes = [vp, sz, sz, sz, dp, C.c_double, dp, dp, C.c_double, C.c_double,
                                       C.c_double, C.c_double]
    L.or_world_nearest_velocity.argtypes = [vp, sz, dp, dp]
    L.or_quat_rotate.argtypes = [dp, dp, dp]
    assert L.or_params_size() == C.sizeof(Params), "Params layout mismatch"
    _lib = L
    return L


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def kernels(h: float, r: float = 0.0):
    """(poly6_coeff, spiky_grad_coeff, visc_lap_coeff), (poly6(r), spiky(r), lap(r))."""
    c = np.zeros(3)
    v = np.zeros(3)
    lib().or_kernels(h, _dptr(c), r, _dptr(v))
    return c, v


def quat_rotate(q_wijk, v) -> np.ndarray:
    """nalgebra's UnitQuaternion * Vector3 as the oracle restates it."""
    q = np.ascontiguousarray(q_wijk, dtype=np.float64)
    a = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros(3)
    lib().or_quat_rotate(_dptr(q), _dptr(a), _dptr(out))
    return out


SHAPES = {"sphere": 0, "box": 1, "capsule": 2}


def set_mutation(name: str) -> None:
    lib().or_set_mutation(MUTATIONS[name])


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def get_threads() -> int:
    return lib().or_get_threads()


_VEC_FIELDS = {"pos": 0, "vel": 1, "acc": 2, "body_acc": 3, "xsph": 4}
_SCALAR_FIELDS = {"rho": 5, "p": 6, "mass": 7, "mu_eff": 8}


class World:
    """The oracle's ``FluidWorld<f64>``."""

    def __init__(self, params: Params | None = None, _handle=None):
        self.params = (params or Params.defaults()).copy()
        L = lib()
        self._h = _handle if _handle is not None else L.or_world_new(C.byref(self.params))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().or_world_free(h)
            self._h = None

    def clone(self) -> "World":
        return World(self.params, _handle=lib().or_world_clone(self._h))

    def set_params(self, params: Params) -> None:
        self.params = params.copy()
        lib().or_world_set_params(self._h, C.byref(self.params))

    def __len__(self) -> int:
        return int(lib().or_world_len(self._h))

    def lattice_mass(self, s: float) -> float:
        return lib().or_world_lattice_mass(self._h, s)

    def fill_box(self, mn, mx, spacing: float, margin: float) -> None:
        a = np.ascontiguousarray(mn, dtype=np.float64)
        b = np.ascontiguousarray(mx, dtype=np.float64)
        lib().or_world_fill_box(self._h, _dptr(a), _dptr(b), spacing, margin)

    def set_particles(self, pos, vel=None, mass=None, material=None, body_acc=None) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        args = [pos]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 3)
        mass = None if mass is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,)))
        material = None if material is None else np.ascontiguousarray(material, dtype=np.uint64)
        body_acc = None if body_acc is None else np.ascontiguousarray(
            body_acc, dtype=np.float64).reshape(n, 3)
        self._keep = (args, vel, mass, material, body_acc)
        lib().or_world_set_particles(
            self._h, n, _dptr(pos),
            _dptr(vel) if vel is not None else None,
            _dptr(mass) if mass is not None else None,
            material.ctypes.data_as(C.POINTER(C.c_uint64)) if material is not None else None,
            _dptr(body_acc) if body_acc is not None else None)

    def get(self, field: str) -> np.ndarray:
        n = len(self)
        if field in _VEC_FIELDS:
            out = np.empty((n, 3))
            lib().or_world_get(self._h, _VEC_FIELDS[field], _dptr(out))
        else:
            out = np.empty(n)
            lib().or_world_get(self._h, _SCALAR_FIELDS[field], _dptr(out))
        return out

    def step(self, dt: float, n: int = 1) -> None:
        if n == 1:
            lib().or_world_step(self._h, dt)
        else:
            lib().or_world_step_n(self._h, dt, int(n))

    def density(self) -> None:
        lib().or_world_density(self._h)

    def forces(self) -> None:
        lib().or_world_forces(self._h)

    def validate(self) -> int:
        return lib().or_world_validate(self._h)

    def total_momentum(self) -> np.ndarray:
        out = np.zeros(3)
        lib().or_world_momentum(self._h, _dptr(out))
        return out

    def kinetic_energy(self) -> float:
        return lib().or_world_kinetic_energy(self._h)

    def neighbors(self, i: int) -> np.ndarray:
        cap = 4096
        out = np.zeros(cap, dtype=np.uint64)
        n = lib().or_world_neighbors(self._h, i, out.ctypes.data_as(C.POINTER(C.c_uint64)), cap)
        assert n <= cap
        return out[:n].astype(np.int64)

    def neighbor_stats(self):
        n = len(self)
        c = np.zeros(n, dtype=np.uint32)
        s = np.zeros(n, dtype=np.uint64)
        x = np.zeros(n, dtype=np.uint64)
        lib().or_world_neighbor_stats(
            self._h, c.ctypes.data_as(C.POINTER(C.c_uint32)),
            s.ctypes.data_as(C.POINTER(C.c_uint64)), x.ctypes.data_as(C.POINTER(C.c_uint64)))
        return c, s, x

    def force_neighbor_stats(self):
        """Count / index sum / hashed-index xor of each particle's force-pass
        neighbours (j != i, 0 < r < h or r NaN; world.rs:292-299)."""
        n = len(self)
        c = np.zeros(n, dtype=np.uint32)
        s = np.zeros(n, dtype=np.uint64)
        x = np.zeros(n, dtype=np.uint64)
        lib().or_world_force_neighbor_stats(
            self._h, c.ctypes.data_as(C.POINTER(C.c_uint32)),
            s.ctypes.data_as(C.POINTER(C.c_uint64)), x.ctypes.data_as(C.POINTER(C.c_uint64)))
        return c, s, x

    def time_steps(self, dt: float, n: int) -> float:
        return lib().or_world_time_steps(self._h, dt, int(n))

    def flat_grid(self):
        """Grid::to_flat of the current positions (grid.rs:62-175):
        (grid_min[3], nc[3], cell_start[ncells + 1], sorted[n])."""
        mn = (C.c_int64 * 3)()
        nc = (C.c_uint64 * 3)()
        lib().or_world_flat_grid(self._h, mn, nc, None, 0, None, 0)
        ncell = nc[0] * nc[1] * nc[2]
        cs = np.zeros(ncell + 1, dtype=np.int32)
        so = np.zeros(len(self), dtype=np.int32)
        rc = lib().or_world_flat_grid(self._h, mn, nc, cs.ctypes.data_as(C.POINTER(C.c_int32)), cs.size,
                                      so.ctypes.data_as(C.POINTER(C.c_int32)), so.size)
        assert rc == 0
        return np.array(list(mn), dtype=np.int64), np.array(list(nc), dtype=np.uint64), cs, so

    # ---- coupling (world.rs:400-773) ----

    def couple_bodies(self, pose7, vel3, inv_mass, shape, dims3, dt: float, friction: float = 0.0):
        """FluidWorld::couple_bodies on bodies given as arrays (pose = pos xyz +
        quaternion w i j k). Returns the bodies' new velocities; the particles
        change in place. ``friction`` is unused by the reference."""
        pose = np.ascontiguousarray(pose7, dtype=np.float64).reshape(-1, 7)
        nb = pose.shape[0]
        vel = np.array(vel3, dtype=np.float64).reshape(nb, 3)
        im = np.ascontiguousarray(inv_mass, dtype=np.float64).reshape(nb)
        sh = np.ascontiguousarray(shape, dtype=np.int32).reshape(nb)
        dm = np.ascontiguousarray(dims3, dtype=np.float64).reshape(nb, 3)
        lib().or_world_couple_bodies(self._h, nb, _dptr(pose), _dptr(vel), _dptr(im),
                                     sh.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(dm), dt)
        return vel

    def couple_points(self, pos3, vel3, mass, force3, dt: float, drag: float, soft_density: float):
        """FluidWorld::couple_points; returns the points' forces (unchanged for
        points outside the bounds)."""
        pos = np.ascontiguousarray(pos3, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        vel = np.ascontiguousarray(vel3, dtype=np.float64).reshape(n, 3)
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,)))
        f = np.array(force3, dtype=np.float64).reshape(n, 3)
        lib().or_world_couple_points(self._h, n, _dptr(pos), _dptr(vel), _dptr(m), _dptr(f), dt, drag,
                                     soft_density)
        return f

    def couple_heat(self, u, origin, dx: float, src, dt: float, t_ref: float, beta: float,
                    heat_gain: float):
        """FluidWorld::couple_heat against a temperature grid u[z, y, x] (x
        fastest); returns the updated source grid, body_acc changes in place."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        nz, ny, nx = u.shape
        o = np.ascontiguousarray(origin, dtype=np.float64)
        out = np.array(src, dtype=np.float64).reshape(nz, ny, nx)
        lib().or_world_couple_heat(self._h, nx, ny, nz, _dptr(o), dx, _dptr(u), _dptr(out), dt, t_ref, beta,
                                   heat_gain)
        return out

    def nearest_velocity(self, q3) -> np.ndarray:
        """The velocity sampler couple_heat installs (world.rs:614-643)."""
        q = np.ascontiguousarray(q3, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(q)
        lib().or_world_nearest_velocity(self._h, q.shape[0], _dptr(q), _dptr(out))
        return out
