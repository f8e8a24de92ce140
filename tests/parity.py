"""Shared helpers of the GPU parity tests: the same FP32 state fed to the HIP
path (through the C ABI) and to the f64 oracle, and the error metrics.

Error metric (norm-wise relative, per field): max_i |a_i - b_i| / max_i |b_i|
with |.| the Euclidean norm for vector fields. Interior pressure forces cancel,
so per-particle relative error is ill-posed; the norm-wise form is the one the
tolerances below are stated in.
"""
from __future__ import annotations

import numpy as np

import oracle as O
from physics_amd import ffi, scenes

TOL_STEP = 1e-5     # one-step parity, every field (FP32 + Kahan vs f64)
DT32 = float(np.float32(scenes.DT))   # dt as the device sees it


def normwise(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 2:
        num = np.sqrt(((a - b) ** 2).sum(1)).max(initial=0.0)
        den = np.sqrt((b ** 2).sum(1)).max(initial=0.0)
    else:
        num = np.abs(a - b).max(initial=0.0)
        den = np.abs(b).max(initial=0.0)
    if den == 0.0:
        return float(num)
    return float(num / den)


def params_pair(h=scenes.H32, bounds=((-5, -5, -5), (5, 5, 5)), gravity=-9.81, visc_k=(3.5,),
                visc_n=(1.0,), stiffness=250.0, rest_density=1000.0):
    gp = ffi.PhySphParams.defaults()
    op = O.Params.defaults()
    for p in (gp, op):
        p.h = h
        p.stiffness = stiffness
        p.rest_density = rest_density
        p.gravity[1] = gravity
        p.set_visc(visc_k, visc_n)
        for a in range(3):
            p.bounds_min[a] = bounds[0][a]
            p.bounds_max[a] = bounds[1][a]
    return gp, op


def make_pair(pos32, vel32=None, mass=1.0, material=None, **kw):
    """GPU world and oracle world holding the identical FP32 state."""
    gp, op = params_pair(**kw)
    pos32 = np.ascontiguousarray(pos32, dtype=np.float32)
    n = len(pos32)
    vel32 = np.zeros_like(pos32) if vel32 is None else np.ascontiguousarray(vel32, np.float32)
    m32 = np.ascontiguousarray(np.broadcast_to(np.float32(mass), (n,)) if np.isscalar(mass)
                               else np.asarray(mass, np.float32))
    g = ffi.FluidWorld(gp)
    g.set_particles_f32(pos32, vel32, m32, material)
    g.set_keep_aux(True)
    o = O.World(op)
    o.set_particles(pos32.astype(np.float64), vel32.astype(np.float64), m32.astype(np.float64),
                    None if material is None else np.asarray(material, np.uint64))
    return g, o


def step_both(g, o, dt=DT32):
    x0 = g.get("pos")
    g.step(dt)
    o.step(dt)
    return x0


def compare_step(g, o, x0):
    """Norm-wise errors of every stored per-particle field after one step."""
    err = {}
    for f in ("rho", "p", "acc", "xsph", "mu_eff", "vel"):
        err[f] = normwise(g.get(f), o.get(f))
    # positions: compare displacements, which carry the step's arithmetic. The
    # GPU stores x' in FP32, so up to one FP32 ulp of x' per component is storage
    # rounding, not arithmetic error; only the excess counts.
    xg, xo = g.get("pos"), o.get("pos")
    ulp = np.abs(np.spacing(xo.astype(np.float32)).astype(np.float64))
    excess = np.maximum(np.abs(xg - xo) - ulp, 0.0)
    dxo = xo - x0
    den = np.sqrt((dxo ** 2).sum(1)).max(initial=0.0)
    err["dx"] = float(np.sqrt((excess ** 2).sum(1)).max(initial=0.0) / den) if den > 0 else 0.0
    err["pos"] = normwise(xg, xo)
    return err


def sync_oracle_to_gpu(g, o, material=None):
    """Feed the oracle the GPU's current FP32 state (as f64)."""
    o.set_particles(g.get("pos"), g.get("vel"), g.get("mass"),
                    None if material is None else np.asarray(material, np.uint64))
