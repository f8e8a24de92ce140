"""Reverse-mode gradients over the whole material table: several materials,
power-law (non-Newtonian) viscosity mu = visc_k[m] max(shear, shear_min)^(visc_n[m] - 1)
(crates/phy-fluid/src/sph/world.rs:313-327), against central finite
differences of the f64 oracle.

Loss and tolerance as in test_gpu_adjoint: L = |centroid(x_N)|^2 +
mean (rho - rho0)^2, 1e-3 relative per direction. Parameter and v0
directions use eps = 1e-5: the shear floor max(shear, shear_min) is a C0 kink
in v0 that a 1e-4 step crosses (on the power-law case the 1e-4 difference of
one v0 direction is 3 % off the 1e-5 and 1e-6 ones, which agree to 1e-9),
while 1e-5 differences of every direction here agree with 1e-6 ones to
6e-6. x0 directions use a short rollout with eps = 1e-8: the shear-rate sum
|v_j - v_i| / (r + 1e-4) (m_j / rho_j) does not vanish at r = h, so mu jumps
when a pair crosses the support edge, and a 1e-6 step in x0 crosses such an
edge on two of the four seeded directions (20 % off); 1e-7, 1e-8 and 1e-9
differences agree to 3e-7 there.
Material tags outside the table (here 5) select material 0, as in the
reference (world.rs:318-322), so those particles feed visc_k[0] / visc_n[0].
"""
import numpy as np
import pytest

import oracle as O
from physics_amd import ffi, scenes

pytestmark = pytest.mark.gpu

TOL = 1e-3
DT32 = float(np.float32(scenes.DT))
M = ffi.MAX_MAT


def _setup(nside=6, seed=3):
    pos = scenes.lattice_block(nside, nside, nside, scenes.S32, (0.2, 0.5, -0.3), 0.01, seed)
    n = len(pos)
    vel = (0.1 * np.sqrt(-2 * np.log(1 - scenes.rng_f64(seed + 10, 3 * n))) *
           np.cos(2 * np.pi * scenes.rng_f64(seed + 20, 3 * n))).reshape(n, 3).astype(np.float32)
    mass = np.float32(scenes.lattice_mass(scenes.H32, scenes.S32) * 1.3)
    u = scenes.rng_f64(seed + 30, n)
    mat = np.where(u < 0.45, 0, np.where(u < 0.9, 1, 5)).astype(np.uint32)
    return pos, vel, mass, mat


def _params(cls, k, vk, vn, shear_min=None):
    p = cls.defaults()
    p.h = scenes.H32
    p.stiffness = k
    p.set_visc(vk, vn)
    if shear_min is not None:
        p.shear_min = shear_min
    for a in range(3):
        p.bounds_min[a] = -100.0
        p.bounds_max[a] = 100.0
    return p


def _loss(x, rho):
    c = x.mean(0)
    return float(c @ c + ((rho - 1000.0) ** 2).mean())


def _cotangents(x, rho):
    n = len(x)
    c = x.mean(0)
    return np.tile(2.0 * c / n, (n, 1)), 2.0 * (rho - 1000.0) / n


def _gpu_grad(case, steps, K):
    pos, vel, mass, mat = case["state"]
    g = ffi.FluidWorld(_params(ffi.PhySphParams, case["k"], case["vk"], case["vn"], case.get("smin")))
    n = len(pos)
    g.set_particles_f32(pos, vel, np.full(n, mass, np.float32), mat)
    g.tape_begin(K)
    g.step(DT32, steps)
    x, rho = g.get("pos"), g.get("rho")
    xbar, rhobar = _cotangents(x, rho)
    x0b, v0b, dk, dvk, dvn = g.tape_backward_full(xbar, None, rhobar)
    return _loss(x, rho), x0b, v0b, dk, dvk, dvn, g


def _oracle_loss(case, steps, dx=0.0, dv=0.0, dk=0.0, dvk=None, dvn=None):
    pos, vel, mass, mat = case["state"]
    nm = len(case["vk"])
    vk = np.array(case["vk"], float) + (0.0 if dvk is None else dvk[:nm])
    vn = np.array(case["vn"], float) + (0.0 if dvn is None else dvn[:nm])
    w = O.World(_params(O.Params, case["k"] + dk, vk, vn, case.get("smin")))
    w.set_particles(pos.astype(np.float64) + dx, vel.astype(np.float64) + dv,
                    np.full(len(pos), float(mass)), mat)
    w.step(DT32, steps)
    return _loss(w.get("pos"), w.get("rho"))


def _param_directions(case, rng):
    """Seeded directions: each parameter alone, then v0 alone, then mixed."""
    nm = len(case["vk"])
    out = [("stiffness", 0.0, 10.0, np.zeros(M), np.zeros(M))]
    for m in range(nm):
        e = np.zeros(M)
        e[m] = 0.5
        out.append((f"visc_k[{m}]", 0.0, 0.0, e, np.zeros(M)))
    for m in range(nm):
        e = np.zeros(M)
        e[m] = 0.05
        out.append((f"visc_n[{m}]", 0.0, 0.0, np.zeros(M), e))
    out.append(("v0", 1.0, 0.0, np.zeros(M), np.zeros(M)))
    for t in range(2):
        vk = np.zeros(M)
        vn = np.zeros(M)
        vk[:nm] = rng.normal(size=nm) * 0.5
        vn[:nm] = rng.normal(size=nm) * 0.05
        out.append((f"mixed {t}", 1.0, rng.normal() * 10.0, vk, vn))
    return out


def _check_params_and_v0(case, steps=30, K=7, seed=11):
    L, x0b, v0b, dk, dvk, dvn, _ = _gpu_grad(case, steps, K)
    L64 = _oracle_loss(case, steps)
    assert abs(L - L64) / abs(L64) < 1e-5
    nm = len(case["vk"])
    # gradients of materials outside the table are zero
    assert np.all(dvk[nm:] == 0.0) and np.all(dvn[nm:] == 0.0)
    rng = np.random.default_rng(seed)
    vel = case["state"][1]
    worst = 0.0
    eps = 1e-5
    for name, cv, ck, vk, vn in _param_directions(case, rng):
        dv = rng.normal(size=vel.shape) * cv
        lp = _oracle_loss(case, steps, 0.0, eps * dv, eps * ck, eps * vk, eps * vn)
        lm = _oracle_loss(case, steps, 0.0, -eps * dv, -eps * ck, -eps * vk, -eps * vn)
        fd = (lp - lm) / (2 * eps)
        ad = float((v0b * dv).sum() + dk * ck + (dvk * vk).sum() + (dvn * vn).sum())
        rel = abs(ad - fd) / max(abs(fd), 1e-30)
        print(f"{name}: adjoint {ad:.8e} fd {fd:.8e} rel {rel:.2e}", flush=True)
        worst = max(worst, rel)
    assert worst < TOL


def _check_positions(case, steps=10, K=4, seed=12):
    _, x0b, _, _, _, _, _ = _gpu_grad(case, steps, K)
    rng = np.random.default_rng(seed)
    pos = case["state"][0]
    worst = 0.0
    eps = 1e-8
    for t in range(4):
        dx = rng.normal(size=pos.shape)
        fd = (_oracle_loss(case, steps, eps * dx) - _oracle_loss(case, steps, -eps * dx)) / (2 * eps)
        ad = float((x0b * dx).sum())
        rel = abs(ad - fd) / abs(fd)
        print(f"x0 direction {t}: adjoint {ad:.8e} fd {fd:.8e} rel {rel:.2e}", flush=True)
        worst = max(worst, rel)
    assert worst < TOL


def _newtonian_pair():
    return {"state": _setup(6, 3), "k": 250.0, "vk": [3.5, 9.0], "vn": [1.0, 1.0]}


def _power_law_pair(seed=3, nside=6):
    return {"state": _setup(nside, seed), "k": 250.0, "vk": [3.5, 1.5], "vn": [0.7, 1.3]}


def test_two_newtonian_materials_match_fd():
    _check_params_and_v0(_newtonian_pair())


def test_power_law_materials_match_fd():
    _check_params_and_v0(_power_law_pair())


def test_power_law_below_shear_floor_matches_fd():
    """shear_min above most particles' shear rate: mu sits on the floor, the
    shear-rate cotangent vanishes there and d mu / d n = mu ln(shear_min)."""
    case = _power_law_pair(seed=8)
    case["smin"] = 0.05
    _check_params_and_v0(case, seed=21)


def test_power_law_positions_short_rollout():
    _check_positions(_power_law_pair(seed=4, nside=5))


def test_full_and_short_forms_agree_and_batch_matches_worlds():
    """phy_fluid_tape_backward's two values are the first two of the full
    layout; a batch of one two-material Newtonian scene and one power-law
    scene gives each scene the single-world gradients bit for bit."""
    cases = [_newtonian_pair(), _power_law_pair(seed=9)]
    singles = []
    for case in cases:
        L, x0b, v0b, dk, dvk, dvn, g = _gpu_grad(case, 12, 5)
        x, rho = g.get("pos"), g.get("rho")
        xbar, rhobar = _cotangents(x, rho)
        s = g.tape_backward(xbar, None, rhobar)
        assert np.array_equal(s[0], x0b) and np.array_equal(s[1], v0b)
        assert s[2] == dk and s[3] == dvk[0]
        singles.append((x0b, v0b, dk, dvk, dvn, xbar, rhobar))
        g.tape_end()

    params = [_params(ffi.PhySphParams, c["k"], c["vk"], c["vn"]) for c in cases]
    b = ffi.FluidBatch(params)
    st = [c["state"] for c in cases]
    counts = [len(s[0]) for s in st]
    b.set_particles_f32(counts, np.concatenate([s[0] for s in st]), np.concatenate([s[1] for s in st]),
                        np.concatenate([np.full(len(s[0]), s[2], np.float32) for s in st]),
                        np.concatenate([s[3] for s in st]))
    b.tape_begin(5)
    b.step(DT32, 12)
    xbar = np.concatenate([s[5] for s in singles]).astype(np.float32)
    rhobar = np.concatenate([s[6] for s in singles]).astype(np.float32)
    x0b, v0b, dk, dvk, dvn = b.tape_backward_full(xbar, None, rhobar)
    short = b.tape_backward(xbar, None, rhobar)
    b.tape_end()
    assert np.array_equal(short[2], dk) and np.array_equal(short[3], dvk[:, 0])
    off = b.offsets()
    for s, (sx, sv, sk, svk, svn, _, _) in enumerate(singles):
        lo, hi = off[s], off[s + 1]
        assert np.array_equal(x0b[lo:hi], sx.astype(np.float32))
        assert np.array_equal(v0b[lo:hi], sv.astype(np.float32))
        assert dk[s] == sk
        assert np.array_equal(dvk[s], svk) and np.array_equal(dvn[s], svn)


def test_recording_leaves_a_newtonian_trajectory_unchanged():
    """Recorded steps use the force variant that also sums the shear rate; for
    visc_n = 1 every output is bit-identical to the plain variant."""
    pos, vel, mass, mat = _setup(7, 13)
    n = len(pos)
    out = []
    for rec in (False, True):
        g = ffi.FluidWorld(_params(ffi.PhySphParams, 250.0, [3.5, 9.0], [1.0, 1.0]))
        g.set_particles_f32(pos, vel, np.full(n, mass, np.float32), mat)
        if rec:
            g.tape_begin(4)
        g.step(DT32, 20)
        out.append((g.get("pos"), g.get("vel"), g.get("rho")))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
