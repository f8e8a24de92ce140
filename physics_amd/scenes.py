"""Synthetic dam-break scenes of the benchmark configurations.

Randomness comes from the reference's portable SplitMix64 ``Rng``
(crates/phy-core/src/replay.rs:29-53), so every scene is a pure function of
its seed. Lattice blocks follow ``FluidWorld::fill_box``'s index order
(x outer, y, z inner; crates/phy-fluid/src/sph/world.rs:130-143) and carry the
mass of ``FluidWorld::lattice_mass`` (world.rs:148-169).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

H32 = float(np.float32(0.2))     # FP32-representable support radius (0.2f)
S32 = float(np.float32(0.1))     # lattice spacing h/2 (0.1f)
DT = 0.01                        # the README / perf_sweep step (no default in the reference)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """``Rng::next_u64`` for ``count`` draws from ``Rng::seed(seed)`` (replay.rs:40-46)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def rng_f64(seed: int, count: int) -> np.ndarray:
    """``Rng::next_f64`` = (u >> 11) / 2^53 (replay.rs:49-52)."""
    u = splitmix64(seed, count)
    return (u >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def kernel_coeffs(h: float):
    """(poly6, spiky-gradient, viscosity-laplacian) coefficients, kernels.rs:22-35."""
    pi = math.pi
    h2 = h * h
    h3 = h2 * h
    h6 = h3 * h3
    h9 = h6 * h3
    return 315.0 / (64.0 * pi * h9), -45.0 / (pi * h6), 45.0 / (pi * h6)


def lattice_mass(h: float, s: float, rest_density: float = 1000.0, fallback: float = 0.02) -> float:
    """``FluidWorld::lattice_mass`` (world.rs:148-169) in f64."""
    c6, _, _ = kernel_coeffs(h)
    h2 = h * h
    n = int(math.ceil(h / s)) + 1
    sum_w = 0.0
    for i in range(-n, n + 1):
        for j in range(-n, n + 1):
            for l in range(-n, n + 1):
                rr = math.sqrt(float(i * i + j * j + l * l))
                r = s * rr
                if r < h:
                    d = h2 - r * r
                    sum_w += c6 * d * d * d
    return rest_density / sum_w if sum_w > 0.0 else fallback


@dataclass
class Scene:
    pos: np.ndarray          # (n, 3) float32
    vel: np.ndarray          # (n, 3) float32
    mass: float
    h: float
    bounds_min: tuple
    bounds_max: tuple
    dt: float = DT


def lattice_block(nx: int, ny: int, nz: int, spacing: float, origin, jitter_frac: float,
                  seed: int) -> np.ndarray:
    """Lattice block (x outer, y, z inner) with each coordinate jittered by
    spacing * jitter_frac * (2u - 1), u drawn in particle order x, y, z."""
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1).astype(np.float64)
    pos = np.asarray(origin, dtype=np.float64) + (idx + 0.5) * spacing
    if jitter_frac:
        u = rng_f64(seed, pos.size).reshape(-1, 3)
        pos = pos + spacing * jitter_frac * (2.0 * u - 1.0)
    return pos.astype(np.float32)


def dam_break(n_side=(100, 100, 100), seed: int = 1, h: float = H32, spacing: float = S32,
              tank_factor: float = 4.0, jitter_frac: float = 0.01) -> Scene:
    """The 1M-particle FP32 dam break at its default size: a column of
    100 x 100 x 100 particles at spacing h/2 in the corner of a closed cubic
    tank 4x the column width, +-1 % seeded jitter, zero velocity."""
    nx, ny, nz = n_side
    pos = lattice_block(nx, ny, nz, spacing, (0.0, 0.0, 0.0), jitter_frac, seed)
    width = max(nx, nz) * spacing
    side = tank_factor * width
    m = lattice_mass(h, spacing)
    return Scene(pos=pos, vel=np.zeros_like(pos), mass=m, h=h, bounds_min=(0.0, 0.0, 0.0),
                 bounds_max=(side, side, side))


def apply_scene(params, scene: Scene):
    """Copy a scene's h and tank bounds into a PhySphParams-like structure."""
    params.h = scene.h
    for a in range(3):
        params.bounds_min[a] = scene.bounds_min[a]
        params.bounds_max[a] = scene.bounds_max[a]
    return params


def rollout_block(nside: int = 32, seed: int = 1, vel_sigma: float = 0.1,
                  mass_scale: float = 1.0) -> Scene:
    """The differentiable-rollout block: nside^3 lattice at spacing h/2 with
    +-1 % seeded jitter and seeded N(0, vel_sigma) velocities (Box-Muller over
    SplitMix64), placed in the default +-5 m tank high enough that a 100-step
    (1 s) free fall does not reach the floor."""
    width = nside * S32
    origin = (-0.5 * width, 4.7 - width, -0.5 * width)
    pos = lattice_block(nside, nside, nside, S32, origin, 0.01, seed)
    n = len(pos)
    u1 = rng_f64(seed + 1000, 3 * n)
    u2 = rng_f64(seed + 2000, 3 * n)
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    vel = (vel_sigma * z).reshape(n, 3).astype(np.float32)
    m = lattice_mass(H32, S32) * mass_scale
    return Scene(pos=pos, vel=vel, mass=m, h=H32, bounds_min=(-5.0, -5.0, -5.0),
                 bounds_max=(5.0, 5.0, 5.0))


def rollout_loss_cotangents(x: np.ndarray, rho: np.ndarray, rest_density: float = 1000.0):
    """L = |centroid(x)|^2 + mean (rho - rho0)^2 and its cotangents (dL/dx, dL/drho)."""
    n = len(x)
    c = x.mean(0)
    L = float(c @ c + ((rho - rest_density) ** 2).mean())
    return L, np.tile(2.0 * c / n, (n, 1)), 2.0 * (rho - rest_density) / n


def batch_counts(nscenes: int, seed: int, lo: int = 2000, hi: int = 64000) -> np.ndarray:
    """Particle counts round(exp(U[ln lo, ln hi])) of a batch, drawn from the seed."""
    u = rng_f64(seed, nscenes)
    return np.rint(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def batch_scene(n: int, seed: int, vel_sigma: float = 0.1) -> Scene:
    """One scene of a batched run: the first n particles (fill order x, y, z)
    of an a x a x a lattice block, a = ceil(cbrt(n)), spacing h/2, +-1 %
    jitter, N(0, vel_sigma) velocities, in the corner of its own tank of twice
    the block's width per axis."""
    a = max(1, int(math.ceil(round(n ** (1.0 / 3.0), 9))))
    while a ** 3 < n:
        a += 1
    pos = lattice_block(a, a, a, S32, (0.0, 0.0, 0.0), 0.01, seed)[:n]
    u1 = rng_f64(seed + 1000, 3 * n)
    u2 = rng_f64(seed + 2000, 3 * n)
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    vel = (vel_sigma * z).reshape(n, 3).astype(np.float32)
    side = 2.0 * a * S32
    return Scene(pos=pos, vel=vel, mass=lattice_mass(H32, S32), h=H32,
                 bounds_min=(0.0, 0.0, 0.0), bounds_max=(side, side, side))


def batch_scenes(nscenes: int = 256, seed: int = 7, lo: int = 2000, hi: int = 64000):
    """The batched-scenes workload: nscenes independent dam breaks with
    log-uniform particle counts in [lo, hi]."""
    return [batch_scene(int(n), seed * 100003 + s) for s, n in enumerate(batch_counts(nscenes, seed, lo, hi))]


def inverse_design_rollouts(nrollouts: int = 16, seed: int = 5, nx: int = 64, ny: int = 64,
                            nz: int = 32, base_visc: float = 3.5):
    """The batched inverse-design workload: nrollouts blocks of nx x ny x nz
    particles (spacing h/2, +-1 % jitter, N(0, 0.1) velocities) in free fall
    with the walls out of reach (+-100 m), each with its own visc_k[0] drawn
    uniform in [0.5, 2] x base_visc. Returns [(Scene, visc_k0)]."""
    u = rng_f64(seed, nrollouts)
    out = []
    for r in range(nrollouts):
        s = seed * 7919 + r
        origin = (-0.5 * nx * S32, -0.5 * ny * S32, -0.5 * nz * S32)
        pos = lattice_block(nx, ny, nz, S32, origin, 0.01, s)
        n = len(pos)
        u1 = rng_f64(s + 1000, 3 * n)
        u2 = rng_f64(s + 2000, 3 * n)
        z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
        vel = (0.1 * z).reshape(n, 3).astype(np.float32)
        sc = Scene(pos=pos, vel=vel, mass=lattice_mass(H32, S32), h=H32,
                   bounds_min=(-100.0, -100.0, -100.0), bounds_max=(100.0, 100.0, 100.0))
        out.append((sc, base_visc * (0.5 + 1.5 * float(u[r]))))
    return out
