"""ctypes binding of libphy_ffi.so (include/phy_ffi.h).

Python mirror of the reference's fluid path, called only through the C ABI:
``FluidWorld`` follows crates/phy-fluid/src/sph/world.rs (new, fill_box,
lattice_mass, step, particle access) and ``World`` the World/step_checked
contract of crates/phy-core/src/world.rs:224-293. There is no CPU fallback:
without the built library or a GPU every constructor raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphy_ffi.so")
MAX_MAT = 16
TAPE_PARAMS = 1 + 2 * MAX_MAT  # PHY_TAPE_PARAMS: stiffness, visc_k[16], visc_n[16]

FIELDS = {"pos": 0, "vel": 1, "acc": 2, "xsph": 3, "rho": 4, "p": 5, "mass": 6, "mu_eff": 7}
VEC_FIELDS = {"pos", "vel", "acc", "xsph"}

# every symbol include/phy_ffi.h declares
EXPORTS = [
    "phy_ffi_abi_version", "phy_world_create_fluid", "phy_world_create_rigid",
    "phy_world_create_rigid_empty", "phy_world_create_granular", "phy_world_create_coupled",
    "phy_world_create_soft_empty", "phy_world_create_field_empty", "phy_world_create_optics_empty",
    "phy_world_create_solid_empty", "phy_world_step", "phy_world_step_checked", "phy_world_time",
    "phy_world_sub_count", "phy_world_fluid_count", "phy_world_get_fluid_positions",
    "phy_world_get_fluid_velocities", "phy_world_rigid_count", "phy_world_get_rigid_transforms",
    "phy_world_save", "phy_world_load", "phy_world_destroy", "phy_world_step_f32",
    "phy_world_get_fluid_positions_f32", "phy_world_get_fluid_velocities_f32",
    "phy_world_get_rigid_transforms_f32", "phy_world_rigid_add_body",
    "phy_world_rigid_apply_force", "phy_world_rigid_apply_torque", "phy_world_rigid_get_velocity",
    "phy_world_rigid_get_angular_velocity", "phy_last_error", "phy_sph_params_defaults",
    "phy_world_create_fluid_with", "phy_fluid_set_params", "phy_fluid_get_params",
    "phy_fluid_fill_box", "phy_fluid_lattice_mass", "phy_fluid_set_particles",
    "phy_fluid_set_particles_f32", "phy_fluid_get_field", "phy_fluid_get_field_f32",
    "phy_fluid_set_keep_aux", "phy_fluid_set_body_acc", "phy_fluid_diagnostics_f64",
    "phy_fluid_gather_stats", "phy_fluid_batch_gather_stats", "phy_device_mem_stats",
    "phy_device_mem_reset_peak", "phy_fluid_diagnostics_f64_noalloc",
    "phy_fluid_neighbor_stats", "phy_world_step_n", "phy_fluid_sync", "phy_fluid_profile_begin",
    "phy_fluid_profile_end", "phy_fluid_tape_begin", "phy_fluid_tape_steps", "phy_fluid_tape_backward",
    "phy_fluid_tape_end", "phy_fluid_set_graph", "phy_fluid_batch_create", "phy_fluid_batch_destroy",
    "phy_fluid_batch_scenes", "phy_fluid_batch_set_params", "phy_fluid_batch_get_params",
    "phy_fluid_batch_set_particles_f32", "phy_fluid_batch_offsets", "phy_fluid_batch_step",
    "phy_fluid_batch_step_checked", "phy_fluid_batch_set_graph", "phy_fluid_batch_set_keep_aux",
    "phy_fluid_batch_sync", "phy_fluid_batch_get_field_f32", "phy_fluid_batch_diagnostics_f64",
    "phy_fluid_batch_neighbor_stats", "phy_fluid_batch_tape_begin", "phy_fluid_batch_tape_steps",
    "phy_fluid_batch_tape_backward", "phy_fluid_batch_tape_end", "phy_fluid_batch_profile_begin",
    "phy_fluid_batch_profile_end", "phy_device_mem_info", "phy_fluid_get_tuning",
    "phy_fluid_set_tuning", "phy_fluid_autotune", "phy_fluid_batch_get_tuning",
    "phy_fluid_batch_set_tuning", "phy_fluid_batch_autotune", "phy_fluid_flat_export",
    "phy_fluid_tape_backward_full", "phy_fluid_batch_tape_backward_full", "phy_fluid_get_body_acc",
    "phy_fluid_couple_bodies", "phy_fluid_couple_points", "phy_fluid_couple_heat",
    "phy_fluid_sample_velocity",
]


class PhySphParams(C.Structure):
    """``PhySphParams`` = SphParams<f64> (crates/phy-fluid/src/sph/params.rs:12-50)."""

    _fields_ = [
        ("rest_density", C.c_double),
        ("stiffness", C.c_double),
        ("viscosity", C.c_double),
        ("visc_k", C.c_double * MAX_MAT),
        ("visc_n", C.c_double * MAX_MAT),
        ("n_visc_k", C.c_uint32),
        ("n_visc_n", C.c_uint32),
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
    def defaults(cls) -> "PhySphParams":
        p = cls()
        lib().phy_sph_params_defaults(C.byref(p))
        return p

    def set_visc(self, k, n) -> None:
        k, n = list(k), list(n)
        for i, v in enumerate(k):
            self.visc_k[i] = v
        for i, v in enumerate(n):
            self.visc_n[i] = v
        self.n_visc_k, self.n_visc_n = len(k), len(n)

    def copy(self) -> "PhySphParams":
        q = PhySphParams()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(PhySphParams))
        return q


class PhyFlatInfo(C.Structure):
    """Shape and parameters of a flat export (SphFlatData, sph/gpu_flat.rs:14-30)."""

    _fields_ = [
        ("n", C.c_uint64), ("ncells", C.c_uint64), ("grid_min", C.c_int64 * 3), ("nc", C.c_uint64 * 3),
        ("h", C.c_float), ("rest_density", C.c_float), ("stiffness", C.c_float), ("shear_min", C.c_float),
        ("gravity", C.c_float * 3), ("n_visc", C.c_uint32), ("visc_k", C.c_float * MAX_MAT),
        ("visc_n", C.c_float * MAX_MAT),
    ]


SHAPES = {"sphere": 0, "box": 1, "capsule": 2}


class PhyCoupleBody(C.Structure):
    """Body<f64> as couple_bodies reads it: pose = position + quaternion w i j k."""

    _fields_ = [("pose", C.c_double * 7), ("vel", C.c_double * 3), ("inv_mass", C.c_double),
                ("shape", C.c_int32), ("_pad", C.c_int32), ("dims", C.c_double * 3)]


class PhyCouplePoint(C.Structure):
    """CouplePoint<f64> (sph/couple.rs:10-19)."""

    _fields_ = [("pos", C.c_double * 3), ("vel", C.c_double * 3), ("mass", C.c_double),
                ("force", C.c_double * 3)]


class PhyHeatGrid(C.Structure):
    """Temperature grid (x fastest) and its source accumulator."""

    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("_pad", C.c_uint32),
                ("origin", C.c_double * 3), ("dx", C.c_double), ("u", C.POINTER(C.c_double)),
                ("src", C.POINTER(C.c_double))]


class PhyTuning(C.Structure):
    """Per-lane LDS neighbour-list lengths of the density and force passes, the
    launch form (split: uniform-cell waves and the others in two launches), the
    grid's cells per h (cell_div 2: h / 2, 1: h) and the gathers' workgroup size."""

    _fields_ = [("density_list", C.c_int32), ("force_list", C.c_int32), ("split", C.c_int32),
                ("cell_div", C.c_int32), ("block", C.c_int32)]

    def as_tuple(self):
        return (int(self.density_list), int(self.force_list), int(self.split), int(self.cell_div),
                int(self.block))


_lib = None


class PhyError(RuntimeError):
    pass


def lib():
    """Load libphy_ffi.so; raises when it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PhyError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    H = C.c_void_p
    dp, fp, up = C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_size_t
    u32p, u64p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
    sig = {
        "phy_ffi_abi_version": ([], C.c_uint32),
        "phy_world_create_fluid": ([], H),
        "phy_world_create_rigid": ([], H),
        "phy_world_create_rigid_empty": ([], H),
        "phy_world_create_granular": ([], H),
        "phy_world_create_coupled": ([], H),
        "phy_world_create_soft_empty": ([], H),
        "phy_world_create_field_empty": ([], H),
        "phy_world_create_optics_empty": ([], H),
        "phy_world_create_solid_empty": ([], H),
        "phy_world_step": ([H, C.c_double], C.c_int32),
        "phy_world_step_n": ([H, C.c_double, C.c_uint32], C.c_int32),
        "phy_world_step_checked": ([H, C.c_double], C.c_int32),
        "phy_world_step_f32": ([H, C.c_float], C.c_int32),
        "phy_world_time": ([H], C.c_double),
        "phy_world_sub_count": ([H], up),
        "phy_world_fluid_count": ([H], up),
        "phy_world_get_fluid_positions": ([H, dp, up], up),
        "phy_world_get_fluid_velocities": ([H, dp, up], up),
        "phy_world_get_fluid_positions_f32": ([H, fp, up], up),
        "phy_world_get_fluid_velocities_f32": ([H, fp, up], up),
        "phy_world_rigid_count": ([H], up),
        "phy_world_get_rigid_transforms": ([H, dp, up], up),
        "phy_world_get_rigid_transforms_f32": ([H, fp, up], up),
        "phy_world_rigid_add_body": ([H, C.c_int32, C.c_double, dp, dp], C.c_int64),
        "phy_world_rigid_apply_force": ([H, C.c_int64, dp, C.c_double, C.c_int32], C.c_int32),
        "phy_world_rigid_apply_torque": ([H, C.c_int64, dp, C.c_double, C.c_int32], C.c_int32),
        "phy_world_rigid_get_velocity": ([H, C.c_int64, dp], C.c_int32),
        "phy_world_rigid_get_angular_velocity": ([H, C.c_int64, dp], C.c_int32),
        "phy_world_save": ([H, C.c_char_p], C.c_int32),
        "phy_world_load": ([C.c_char_p], H),
        "phy_world_destroy": ([H], None),
        "phy_last_error": ([], C.c_char_p),
        "phy_sph_params_defaults": ([C.POINTER(PhySphParams)], None),
        "phy_world_create_fluid_with": ([C.POINTER(PhySphParams)], H),
        "phy_fluid_set_params": ([H, C.POINTER(PhySphParams)], C.c_int32),
        "phy_fluid_get_params": ([H, C.POINTER(PhySphParams)], C.c_int32),
        "phy_fluid_fill_box": ([H, dp, dp, C.c_double, C.c_double], C.c_int32),
        "phy_fluid_lattice_mass": ([H, C.c_double], C.c_double),
        "phy_fluid_set_particles": ([H, up, dp, dp, dp, u32p], C.c_int32),
        "phy_fluid_set_particles_f32": ([H, up, fp, fp, fp, u32p], C.c_int32),
        "phy_fluid_get_field": ([H, C.c_int32, dp, up], up),
        "phy_fluid_get_field_f32": ([H, C.c_int32, fp, up], up),
        "phy_fluid_set_keep_aux": ([H, C.c_int32], C.c_int32),
        "phy_fluid_set_body_acc": ([H, dp, up], C.c_int32),
        "phy_fluid_get_body_acc": ([H, dp, up], C.c_int32),
        "phy_fluid_couple_bodies": ([H, C.POINTER(PhyCoupleBody), C.c_uint32, C.c_double, C.c_double],
                                    C.c_int32),
        "phy_fluid_couple_points": ([H, C.POINTER(PhyCouplePoint), C.c_uint32, C.c_double, C.c_double,
                                     C.c_double], C.c_int32),
        "phy_fluid_couple_heat": ([H, C.POINTER(PhyHeatGrid), C.c_double, C.c_double, C.c_double, C.c_double],
                                  C.c_int32),
        "phy_fluid_sample_velocity": ([H, dp, C.c_uint32, dp], C.c_int32),
        "phy_fluid_diagnostics_f64": ([H, dp], C.c_int32),
        "phy_fluid_neighbor_stats": ([H, u32p, u64p, u64p, up], C.c_int32),
        "phy_fluid_gather_stats": ([H, C.c_int32, u32p, u64p, u64p, up], C.c_int32),
        "phy_fluid_sync": ([H], C.c_int32),
        "phy_fluid_profile_begin": ([H], C.c_int32),
        "phy_fluid_profile_end": ([H, dp, u32p], C.c_int32),
        "phy_fluid_tape_begin": ([H, C.c_uint32], C.c_int32),
        "phy_fluid_tape_steps": ([H], C.c_uint32),
        "phy_fluid_tape_backward": ([H, dp, dp, dp, dp, dp, dp], C.c_int32),
        "phy_fluid_tape_backward_full": ([H, dp, dp, dp, dp, dp, dp], C.c_int32),
        "phy_fluid_tape_end": ([H], C.c_int32),
        "phy_fluid_set_graph": ([H, C.c_int32], C.c_int32),
        "phy_fluid_batch_create": ([C.POINTER(PhySphParams), C.c_uint32], H),
        "phy_fluid_batch_destroy": ([H], None),
        "phy_fluid_batch_scenes": ([H], C.c_uint32),
        "phy_fluid_batch_set_params": ([H, C.c_uint32, C.POINTER(PhySphParams)], C.c_int32),
        "phy_fluid_batch_get_params": ([H, C.c_uint32, C.POINTER(PhySphParams)], C.c_int32),
        "phy_fluid_batch_set_particles_f32": ([H, u32p, fp, fp, fp, u32p], C.c_int32),
        "phy_fluid_batch_offsets": ([H, u32p], C.c_int32),
        "phy_fluid_batch_step": ([H, C.c_double, C.c_uint32], C.c_int32),
        "phy_fluid_batch_step_checked": ([H, C.c_double, C.POINTER(C.c_int32)], C.c_int32),
        "phy_fluid_batch_set_graph": ([H, C.c_int32], C.c_int32),
        "phy_fluid_batch_set_keep_aux": ([H, C.c_int32], C.c_int32),
        "phy_fluid_batch_sync": ([H], C.c_int32),
        "phy_fluid_batch_get_field_f32": ([H, C.c_int32, fp, up], up),
        "phy_fluid_batch_diagnostics_f64": ([H, dp], C.c_int32),
        "phy_fluid_batch_neighbor_stats": ([H, u32p, u64p, u64p, up], C.c_int32),
        "phy_fluid_batch_gather_stats": ([H, C.c_int32, u32p, u64p, u64p, up], C.c_int32),
        "phy_fluid_batch_tape_begin": ([H, C.c_uint32], C.c_int32),
        "phy_fluid_batch_tape_steps": ([H], C.c_uint32),
        "phy_fluid_batch_tape_backward": ([H, fp, fp, fp, fp, fp, dp], C.c_int32),
        "phy_fluid_batch_tape_backward_full": ([H, fp, fp, fp, fp, fp, dp], C.c_int32),
        "phy_fluid_batch_tape_end": ([H], C.c_int32),
        "phy_fluid_batch_profile_begin": ([H], C.c_int32),
        "phy_fluid_batch_profile_end": ([H, dp, u32p], C.c_int32),
        "phy_device_mem_info": ([u64p, u64p], C.c_int32),
        "phy_device_mem_stats": ([u64p, u64p], C.c_int32),
        "phy_fluid_diagnostics_f64_noalloc": ([C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p],
                                              C.c_int32),
        "phy_device_mem_reset_peak": ([], C.c_int32),
        "phy_fluid_get_tuning": ([H, C.POINTER(PhyTuning)], C.c_int32),
        "phy_fluid_set_tuning": ([H, C.POINTER(PhyTuning)], C.c_int32),
        "phy_fluid_autotune": ([H, C.c_double, C.c_uint32, C.c_char_p, C.POINTER(PhyTuning), dp],
                               C.c_int32),
        "phy_fluid_batch_get_tuning": ([H, C.POINTER(PhyTuning)], C.c_int32),
        "phy_fluid_batch_set_tuning": ([H, C.POINTER(PhyTuning)], C.c_int32),
        "phy_fluid_batch_autotune": ([H, C.c_double, C.c_uint32, C.c_char_p, C.POINTER(PhyTuning), dp],
                                     C.c_int32),
        "phy_fluid_flat_export": ([H, C.POINTER(PhyFlatInfo), fp, fp, fp, up, C.POINTER(C.c_int32), up,
                                   C.POINTER(C.c_int32)], C.c_int32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def last_error() -> str:
    return (lib().phy_last_error() or b"").decode(errors="replace")


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


class World:
    """A ``World<f64>`` handle (crates/phy-core/src/world.rs:78-318) from the C ABI."""

    def __init__(self, handle):
        if not handle:
            raise PhyError(f"world creation failed: {last_error()}")
        self._h = handle

    @classmethod
    def create_fluid(cls) -> "World":
        """``phy_world_create_fluid`` (crates/phy-ffi/src/lib.rs:113-126)."""
        return cls(lib().phy_world_create_fluid())

    @classmethod
    def load(cls, path: str) -> "World":
        """``phy_world_load``: a world from a phy-io JSON archive (lib.rs:532-542)."""
        return cls(lib().phy_world_load(os.fsencode(path)))

    def save(self, path: str) -> None:
        """``phy_world_save``: write the phy-io JSON archive (lib.rs:517-530)."""
        if lib().phy_world_save(self._h, os.fsencode(path)) != 0:
            raise PhyError(last_error())

    def sub_count(self) -> int:
        return int(lib().phy_world_sub_count(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().phy_world_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def step(self, dt: float, n: int = 1) -> None:
        rc = lib().phy_world_step(self._h, dt) if n == 1 else lib().phy_world_step_n(self._h, dt, n)
        if rc != 0:
            raise PhyError(f"phy_world_step rc={rc}: {last_error()}")

    def step_checked(self, dt: float) -> int:
        return int(lib().phy_world_step_checked(self._h, dt))

    def time(self) -> float:
        return lib().phy_world_time(self._h)

    def sync(self) -> None:
        if lib().phy_fluid_sync(self._h) != 0:
            raise PhyError(last_error())

    def fluid_count(self) -> int:
        return int(lib().phy_world_fluid_count(self._h))

    def positions(self) -> np.ndarray:
        n = self.fluid_count()
        out = np.zeros(3 * n)
        got = lib().phy_world_get_fluid_positions(self._h, _dp(out), out.size)
        assert got == n
        return out.reshape(n, 3)

    def velocities(self) -> np.ndarray:
        n = self.fluid_count()
        out = np.zeros(3 * n)
        got = lib().phy_world_get_fluid_velocities(self._h, _dp(out), out.size)
        assert got == n
        return out.reshape(n, 3)


class FluidWorld(World):
    """``FluidWorld<f64>`` in a one-subsystem World, computed on the GPU in FP32."""

    def __init__(self, params: PhySphParams | None = None):
        self.params = (params or PhySphParams.defaults()).copy()
        super().__init__(lib().phy_world_create_fluid_with(C.byref(self.params)))

    def set_params(self, params: PhySphParams) -> None:
        self.params = params.copy()
        if lib().phy_fluid_set_params(self._h, C.byref(self.params)) != 0:
            raise PhyError(last_error())

    def __len__(self) -> int:
        return self.fluid_count()

    def fill_box(self, mn, mx, spacing: float, margin: float) -> None:
        a = np.ascontiguousarray(mn, dtype=np.float64)
        b = np.ascontiguousarray(mx, dtype=np.float64)
        if lib().phy_fluid_fill_box(self._h, _dp(a), _dp(b), spacing, margin) != 0:
            raise PhyError(last_error())

    def lattice_mass(self, s: float) -> float:
        return lib().phy_fluid_lattice_mass(self._h, s)

    def set_particles(self, pos, vel=None, mass=None, material=None) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 3)
        mass = None if mass is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,)))
        mat = None if material is None else np.ascontiguousarray(material, dtype=np.uint32)
        rc = lib().phy_fluid_set_particles(
            self._h, n, _dp(pos), _dp(vel), _dp(mass),
            None if mat is None else mat.ctypes.data_as(C.POINTER(C.c_uint32)))
        if rc != 0:
            raise PhyError(last_error())

    def set_particles_f32(self, pos, vel=None, mass=None, material=None) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        mass = None if mass is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(mass, dtype=np.float32), (n,)))
        mat = None if material is None else np.ascontiguousarray(material, dtype=np.uint32)
        rc = lib().phy_fluid_set_particles_f32(
            self._h, n, _fp(pos), _fp(vel), _fp(mass),
            None if mat is None else mat.ctypes.data_as(C.POINTER(C.c_uint32)))
        if rc != 0:
            raise PhyError(last_error())

    def get(self, field: str, f32: bool = False) -> np.ndarray:
        n = len(self)
        comps = 3 if field in VEC_FIELDS else 1
        out = np.zeros(n * comps, dtype=np.float32 if f32 else np.float64)
        fn = lib().phy_fluid_get_field_f32 if f32 else lib().phy_fluid_get_field
        got = fn(self._h, FIELDS[field], _fp(out) if f32 else _dp(out), out.size)
        if got != n:
            raise PhyError(f"get({field}) wrote {got} of {n}: {last_error()}")
        return out.reshape(n, 3) if comps == 3 else out

    def set_keep_aux(self, on: bool = True) -> None:
        if lib().phy_fluid_set_keep_aux(self._h, 1 if on else 0) != 0:
            raise PhyError(last_error())

    def set_body_acc(self, b) -> None:
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
        if lib().phy_fluid_set_body_acc(self._h, _dp(b), b.size) != 0:
            raise PhyError(last_error())

    def body_acc(self) -> np.ndarray:
        """The pending body_acc (zero after a step)."""
        out = np.zeros((len(self), 3))
        if lib().phy_fluid_get_body_acc(self._h, _dp(out), out.size) != 0:
            raise PhyError(last_error())
        return out

    # ---- coupling (world.rs:400-773) ----

    def couple_bodies(self, pose7, vel3, inv_mass, shape, dims3, dt: float, friction: float = 0.0):
        """FluidWorld::couple_bodies; returns the bodies' new velocities."""
        pose = np.asarray(pose7, dtype=np.float64).reshape(-1, 7)
        nb = pose.shape[0]
        vel = np.asarray(vel3, dtype=np.float64).reshape(nb, 3)
        im = np.asarray(inv_mass, dtype=np.float64).reshape(nb)
        sh = np.asarray(shape, dtype=np.int32).reshape(nb)
        dm = np.asarray(dims3, dtype=np.float64).reshape(nb, 3)
        arr = (PhyCoupleBody * max(nb, 1))()
        for k in range(nb):
            arr[k].pose[:] = list(pose[k])
            arr[k].vel[:] = list(vel[k])
            arr[k].inv_mass = float(im[k])
            arr[k].shape = int(sh[k])
            arr[k].dims[:] = list(dm[k])
        if lib().phy_fluid_couple_bodies(self._h, arr, nb, dt, friction) != 0:
            raise PhyError(last_error())
        return np.array([list(arr[k].vel) for k in range(nb)]).reshape(nb, 3)

    def couple_points(self, pos3, vel3, mass, force3, dt: float, drag: float, soft_density: float):
        """FluidWorld::couple_points; returns the points' forces."""
        pos = np.asarray(pos3, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        vel = np.asarray(vel3, dtype=np.float64).reshape(n, 3)
        m = np.broadcast_to(np.asarray(mass, dtype=np.float64), (n,))
        f = np.asarray(force3, dtype=np.float64).reshape(n, 3)
        arr = (PhyCouplePoint * max(n, 1))()
        for k in range(n):
            arr[k].pos[:] = list(pos[k])
            arr[k].vel[:] = list(vel[k])
            arr[k].mass = float(m[k])
            arr[k].force[:] = list(f[k])
        if lib().phy_fluid_couple_points(self._h, arr, n, dt, drag, soft_density) != 0:
            raise PhyError(last_error())
        return np.array([list(arr[k].force) for k in range(n)]).reshape(n, 3)

    def couple_heat(self, u, origin, dx: float, src, dt: float, t_ref: float, beta: float,
                    heat_gain: float):
        """FluidWorld::couple_heat against u[z, y, x]; returns the source grid."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        nz, ny, nx = u.shape
        out = np.array(src, dtype=np.float64).reshape(nz, ny, nx)
        g = PhyHeatGrid()
        g.nx, g.ny, g.nz = nx, ny, nz
        g.origin[:] = [float(v) for v in origin]
        g.dx = dx
        g.u = _dp(u)
        g.src = _dp(out)
        if lib().phy_fluid_couple_heat(self._h, C.byref(g), dt, t_ref, beta, heat_gain) != 0:
            raise PhyError(last_error())
        return out

    def sample_velocity(self, q3) -> np.ndarray:
        """The nearest-particle velocity sampler couple_heat installs."""
        q = np.ascontiguousarray(q3, dtype=np.float64).reshape(-1, 3)
        out = np.zeros_like(q)
        if lib().phy_fluid_sample_velocity(self._h, _dp(q), q.shape[0], _dp(out)) != 0:
            raise PhyError(last_error())
        return out

    def diagnostics(self) -> np.ndarray:
        out = np.zeros(5)
        if lib().phy_fluid_diagnostics_f64(self._h, _dp(out)) != 0:
            raise PhyError(last_error())
        return out

    def profile_begin(self) -> None:
        if lib().phy_fluid_profile_begin(self._h) != 0:
            raise PhyError(last_error())

    def profile_end(self):
        """Summed ms of {sort, density, force, step} and the steps timed."""
        ms = np.zeros(4)
        steps = C.c_uint32(0)
        if lib().phy_fluid_profile_end(self._h, _dp(ms), C.byref(steps)) != 0:
            raise PhyError(last_error())
        return ms, int(steps.value)

    def set_graph(self, on: bool = True) -> None:
        """Replay the step from a captured hipGraph (bit-identical results)."""
        if lib().phy_fluid_set_graph(self._h, 1 if on else 0) != 0:
            raise PhyError(last_error())

    def tuning(self):
        """(density_list, force_list, split, cell_div, block)"""
        t = PhyTuning()
        if lib().phy_fluid_get_tuning(self._h, C.byref(t)) != 0:
            raise PhyError(last_error())
        return t.as_tuple()

    def set_tuning(self, density_list: int, force_list: int, split: int = 0, cell_div: int = 2,
                   block: int = 128) -> None:
        t = PhyTuning(density_list, force_list, split, cell_div, block)
        if lib().phy_fluid_set_tuning(self._h, C.byref(t)) != 0:
            raise PhyError(last_error())

    def autotune(self, dt: float, steps: int = 5, cache_path: str | None = None):
        """-> (cached?, (density_list, force_list, split, cell_div, block), ms[12]
        of the timed list candidates)."""
        return _autotune(lib().phy_fluid_autotune, self._h, dt, steps, cache_path)

    def flat_export(self):
        """-> dict of SphFlatData fields (pos, vel, scalar: (n, 4) float32;
        cell_start, sorted: int32; grid_min, nc; parameters)."""
        info = PhyFlatInfo()
        rc = lib().phy_fluid_flat_export(self._h, C.byref(info), None, None, None, 0, None, 0, None)
        if rc < 0:
            raise PhyError(last_error())
        n, nc = int(info.n), int(info.ncells)
        pos = np.zeros((n, 4), np.float32)
        vel = np.zeros((n, 4), np.float32)
        sca = np.zeros((n, 4), np.float32)
        cs = np.zeros(nc + 1, np.int32)
        so = np.zeros(max(n, 1), np.int32)
        i32 = C.POINTER(C.c_int32)
        rc = lib().phy_fluid_flat_export(self._h, C.byref(info), _fp(pos), _fp(vel), _fp(sca), max(n, 1),
                                         cs.ctypes.data_as(i32), cs.size, so.ctypes.data_as(i32))
        if rc != 0:
            raise PhyError(f"flat export: rc {rc}: {last_error()}")
        return {"n": n, "pos": pos, "vel": vel, "scalar": sca, "cell_start": cs, "sorted": so[:n],
                "grid_min": np.array(list(info.grid_min)), "nc": np.array(list(info.nc)),
                "h": info.h, "rest_density": info.rest_density, "stiffness": info.stiffness,
                "shear_min": info.shear_min, "gravity": np.array(list(info.gravity)),
                "visc_k": np.array(list(info.visc_k)[:info.n_visc]),
                "visc_n": np.array(list(info.visc_n)[:info.n_visc])}

    def tape_begin(self, checkpoint_every: int) -> None:
        """Start recording a rollout for reverse-mode gradients."""
        if lib().phy_fluid_tape_begin(self._h, int(checkpoint_every)) != 0:
            raise PhyError(last_error())

    def tape_steps(self) -> int:
        return int(lib().phy_fluid_tape_steps(self._h))

    def tape_backward(self, xbar=None, vbar=None, rhobar=None):
        """Cotangents of (final x, final v, last-step rho) -> (x0_bar, v0_bar,
        dL/d stiffness, dL/d visc_k[0])."""
        n = len(self)
        xb = None if xbar is None else np.ascontiguousarray(xbar, dtype=np.float64).reshape(n, 3)
        vb = None if vbar is None else np.ascontiguousarray(vbar, dtype=np.float64).reshape(n, 3)
        rb = None if rhobar is None else np.ascontiguousarray(rhobar, dtype=np.float64).reshape(n)
        x0 = np.zeros((n, 3))
        v0 = np.zeros((n, 3))
        dp = np.zeros(2)
        rc = lib().phy_fluid_tape_backward(self._h, _dp(xb), _dp(vb), _dp(rb), _dp(x0), _dp(v0), _dp(dp))
        if rc != 0:
            raise PhyError(last_error())
        return x0, v0, float(dp[0]), float(dp[1])

    def tape_backward_full(self, xbar=None, vbar=None, rhobar=None):
        """As tape_backward, with every parameter gradient: -> (x0_bar, v0_bar,
        dL/d stiffness, dL/d visc_k[0..16], dL/d visc_n[0..16])."""
        n = len(self)
        xb = None if xbar is None else np.ascontiguousarray(xbar, dtype=np.float64).reshape(n, 3)
        vb = None if vbar is None else np.ascontiguousarray(vbar, dtype=np.float64).reshape(n, 3)
        rb = None if rhobar is None else np.ascontiguousarray(rhobar, dtype=np.float64).reshape(n)
        x0 = np.zeros((n, 3))
        v0 = np.zeros((n, 3))
        dp = np.zeros(TAPE_PARAMS)
        rc = lib().phy_fluid_tape_backward_full(self._h, _dp(xb), _dp(vb), _dp(rb), _dp(x0), _dp(v0),
                                                _dp(dp))
        if rc != 0:
            raise PhyError(last_error())
        return x0, v0, float(dp[0]), dp[1:1 + MAX_MAT].copy(), dp[1 + MAX_MAT:].copy()

    def tape_end(self) -> None:
        if lib().phy_fluid_tape_end(self._h) != 0:
            raise PhyError(last_error())

    def neighbor_stats(self):
        n = len(self)
        c = np.zeros(n, dtype=np.uint32)
        s = np.zeros(n, dtype=np.uint64)
        x = np.zeros(n, dtype=np.uint64)
        rc = lib().phy_fluid_neighbor_stats(
            self._h, c.ctypes.data_as(C.POINTER(C.c_uint32)), s.ctypes.data_as(C.POINTER(C.c_uint64)),
            x.ctypes.data_as(C.POINTER(C.c_uint64)), n)
        if rc != 0:
            raise PhyError(last_error())
        return c, s, x

    def gather_stats(self, which: str):
        """Neighbour statistics taken by the step's own gather: ``which`` is
        "density" (r < h, self included) or "force" (j != i, 0 < r < h, NaN
        kept)."""
        n = len(self)
        c = np.zeros(n, dtype=np.uint32)
        s = np.zeros(n, dtype=np.uint64)
        x = np.zeros(n, dtype=np.uint64)
        rc = lib().phy_fluid_gather_stats(
            self._h, _PASS[which], c.ctypes.data_as(C.POINTER(C.c_uint32)),
            s.ctypes.data_as(C.POINTER(C.c_uint64)), x.ctypes.data_as(C.POINTER(C.c_uint64)), n)
        if rc != 0:
            raise PhyError(last_error())
        return c, s, x


_PASS = {"density": 0, "force": 1}


def _autotune(fn, h, dt, steps, cache_path):
    t = PhyTuning()
    ms = np.full(12, np.nan)
    path = None if cache_path is None else os.fsencode(cache_path)
    rc = fn(h, dt, int(steps), path, C.byref(t), _dp(ms))
    if rc < 0:
        raise PhyError(last_error())
    return bool(rc == 1), t.as_tuple(), ms


def device_mem_info():
    """(free, total) bytes of the current HIP device."""
    f, t = C.c_uint64(0), C.c_uint64(0)
    if lib().phy_device_mem_info(C.byref(f), C.byref(t)) != 0:
        raise PhyError(last_error())
    return int(f.value), int(t.value)


def device_mem_stats():
    """(in_use, peak) bytes of device memory held by the library's buffers."""
    a, b = C.c_uint64(0), C.c_uint64(0)
    if lib().phy_device_mem_stats(C.byref(a), C.byref(b)) != 0:
        raise PhyError(last_error())
    return int(a.value), int(b.value)


def device_mem_reset_peak() -> None:
    if lib().phy_device_mem_reset_peak() != 0:
        raise PhyError(last_error())


def _u32(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint32))


class FluidBatch:
    """Independent scenes (one FluidWorld each) stepped together on the GPU.

    ``params``: one PhySphParams per scene. Particles are passed and returned
    concatenated in scene order; ``offsets()`` gives each scene's range.
    """

    def __init__(self, params):
        params = [p.copy() for p in params]
        arr = (PhySphParams * len(params))(*params)
        h = lib().phy_fluid_batch_create(arr, len(params))
        if not h:
            raise PhyError(last_error())
        self._h = h
        self.nscenes = len(params)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().phy_fluid_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ok(self, rc) -> None:
        if rc != 0:
            raise PhyError(last_error())

    def set_params(self, scene: int, params: PhySphParams) -> None:
        self._ok(lib().phy_fluid_batch_set_params(self._h, scene, C.byref(params)))

    def get_params(self, scene: int) -> PhySphParams:
        p = PhySphParams()
        self._ok(lib().phy_fluid_batch_get_params(self._h, scene, C.byref(p)))
        return p

    def set_particles_f32(self, counts, pos, vel=None, mass=None, material=None) -> None:
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        n = int(counts.sum())
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(n, 3)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        mass = None if mass is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(mass, dtype=np.float32), (n,)))
        mat = None if material is None else np.ascontiguousarray(material, dtype=np.uint32)
        self._ok(lib().phy_fluid_batch_set_particles_f32(self._h, _u32(counts), _fp(pos), _fp(vel),
                                                          _fp(mass), _u32(mat)))

    def offsets(self) -> np.ndarray:
        o = np.zeros(self.nscenes + 1, dtype=np.uint32)
        self._ok(lib().phy_fluid_batch_offsets(self._h, _u32(o)))
        return o

    def __len__(self) -> int:
        return int(self.offsets()[-1])

    def step(self, dt: float, n: int = 1) -> None:
        self._ok(lib().phy_fluid_batch_step(self._h, dt, n))

    def step_checked(self, dt: float) -> np.ndarray:
        st = np.zeros(self.nscenes, dtype=np.int32)
        rc = lib().phy_fluid_batch_step_checked(self._h, dt, st.ctypes.data_as(C.POINTER(C.c_int32)))
        if rc < 0:
            raise PhyError(last_error())
        return st

    def set_graph(self, on: bool = True) -> None:
        self._ok(lib().phy_fluid_batch_set_graph(self._h, 1 if on else 0))

    def set_keep_aux(self, on: bool = True) -> None:
        self._ok(lib().phy_fluid_batch_set_keep_aux(self._h, 1 if on else 0))

    def tuning(self):
        t = PhyTuning()
        self._ok(lib().phy_fluid_batch_get_tuning(self._h, C.byref(t)))
        return t.as_tuple()

    def set_tuning(self, density_list: int, force_list: int, split: int = 0, cell_div: int = 2,
                   block: int = 128) -> None:
        t = PhyTuning(density_list, force_list, split, cell_div, block)
        self._ok(lib().phy_fluid_batch_set_tuning(self._h, C.byref(t)))

    def autotune(self, dt: float, steps: int = 5, cache_path: str | None = None):
        return _autotune(lib().phy_fluid_batch_autotune, self._h, dt, steps, cache_path)

    def sync(self) -> None:
        self._ok(lib().phy_fluid_batch_sync(self._h))

    def get(self, field: str) -> np.ndarray:
        n = len(self)
        comps = 3 if field in VEC_FIELDS else 1
        out = np.zeros(n * comps, dtype=np.float32)
        got = lib().phy_fluid_batch_get_field_f32(self._h, FIELDS[field], _fp(out), out.size)
        if got != n:
            raise PhyError(f"get({field}) wrote {got} of {n}: {last_error()}")
        return out.reshape(n, 3) if comps == 3 else out

    def diagnostics(self) -> np.ndarray:
        out = np.zeros((self.nscenes, 5))
        self._ok(lib().phy_fluid_batch_diagnostics_f64(self._h, _dp(out)))
        return out

    def neighbor_stats(self):
        n = len(self)
        c = np.zeros(n, dtype=np.uint32)
        s = np.zeros(n, dtype=np.uint64)
        x = np.zeros(n, dtype=np.uint64)
        self._ok(lib().phy_fluid_batch_neighbor_stats(
            self._h, _u32(c), s.ctypes.data_as(C.POINTER(C.c_uint64)),
            x.ctypes.data_as(C.POINTER(C.c_uint64)), n))
        return c, s, x

    def gather_stats(self, which: str):
        """As FluidWorld.gather_stats, concatenated; indices are scene-local."""
        n = len(self)
        c = np.zeros(n, dtype=np.uint32)
        s = np.zeros(n, dtype=np.uint64)
        x = np.zeros(n, dtype=np.uint64)
        self._ok(lib().phy_fluid_batch_gather_stats(
            self._h, _PASS[which], _u32(c), s.ctypes.data_as(C.POINTER(C.c_uint64)),
            x.ctypes.data_as(C.POINTER(C.c_uint64)), n))
        return c, s, x

    def tape_begin(self, checkpoint_every: int) -> None:
        self._ok(lib().phy_fluid_batch_tape_begin(self._h, int(checkpoint_every)))

    def tape_steps(self) -> int:
        return int(lib().phy_fluid_batch_tape_steps(self._h))

    def tape_backward(self, xbar=None, vbar=None, rhobar=None):
        """-> (x0_bar, v0_bar, dL/d stiffness per scene, dL/d visc_k[0] per scene)."""
        n = len(self)
        xb = None if xbar is None else np.ascontiguousarray(xbar, dtype=np.float32).reshape(n, 3)
        vb = None if vbar is None else np.ascontiguousarray(vbar, dtype=np.float32).reshape(n, 3)
        rb = None if rhobar is None else np.ascontiguousarray(rhobar, dtype=np.float32).reshape(n)
        x0 = np.zeros((n, 3), dtype=np.float32)
        v0 = np.zeros((n, 3), dtype=np.float32)
        dp = np.zeros((self.nscenes, 2))
        self._ok(lib().phy_fluid_batch_tape_backward(self._h, _fp(xb), _fp(vb), _fp(rb), _fp(x0),
                                                      _fp(v0), _dp(dp)))
        return x0, v0, dp[:, 0].copy(), dp[:, 1].copy()

    def tape_backward_full(self, xbar=None, vbar=None, rhobar=None):
        """-> (x0_bar, v0_bar, dL/d stiffness [S], dL/d visc_k [S, 16], dL/d visc_n [S, 16])."""
        n = len(self)
        xb = None if xbar is None else np.ascontiguousarray(xbar, dtype=np.float32).reshape(n, 3)
        vb = None if vbar is None else np.ascontiguousarray(vbar, dtype=np.float32).reshape(n, 3)
        rb = None if rhobar is None else np.ascontiguousarray(rhobar, dtype=np.float32).reshape(n)
        x0 = np.zeros((n, 3), dtype=np.float32)
        v0 = np.zeros((n, 3), dtype=np.float32)
        dp = np.zeros((self.nscenes, TAPE_PARAMS))
        self._ok(lib().phy_fluid_batch_tape_backward_full(self._h, _fp(xb), _fp(vb), _fp(rb),
                                                           _fp(x0), _fp(v0), _dp(dp)))
        return (x0, v0, dp[:, 0].copy(), dp[:, 1:1 + MAX_MAT].copy(),
                dp[:, 1 + MAX_MAT:].copy())

    def tape_end(self) -> None:
        self._ok(lib().phy_fluid_batch_tape_end(self._h))

    def profile_begin(self) -> None:
        self._ok(lib().phy_fluid_batch_profile_begin(self._h))

    def profile_end(self):
        ms = np.zeros(4)
        steps = C.c_uint32(0)
        self._ok(lib().phy_fluid_batch_profile_end(self._h, _dp(ms), C.byref(steps)))
        return ms, int(steps.value)
