"""Benchmark: SPH particle-steps/s of the 1M-particle FP32 dam break on one GPU.

Workload: a 100 x 100 x 100 particle column (spacing h/2, +-1 % seeded jitter,
zero velocity) in a closed 40 m tank, dt = 0.01, the FP32 forward step of
FluidWorld::step (crates/phy-fluid/src/sph/world.rs:183-192) on the GPU.
Default: 20 untimed warm-up steps, then the remaining 980 steps of the
1000-step run timed. Inputs are resident in HBM before the timed region.

Prints one JSON line. The timed steps run in chunks, each bracketed by a device
synchronisation; state snapshots for the CPU baseline are taken between chunks,
outside the timed windows. Per-stage kernel times come from HIP events recorded
on the world's own stream (phy_fluid_profile_*).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# algorithmic HBM bytes per particle-step (compulsory FP32 traffic): sort
# 60 B, density 20 B, fused force + integrate 56 B (read x, v, rho, p; write x', v')
ALG_BYTES = {"sort": 60, "density": 20, "force": 56}
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (scaled by a plain copy's counts) on the default 1000-step workload
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_force_1M_dambreak_hbm_traffic.md")


def measured_counter(key: str, kernel: str):
    """A per-launch figure of `kernel` from the committed counter summary, or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith(f"{key} {kernel}:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def measured_traffic(kernel: str):
    """Bytes per launch of `kernel` from the committed counter summary, or None."""
    v = measured_counter("traffic_bytes_per_launch", kernel)
    return None if v is None else int(v)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=980)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n-side", type=int, default=100)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--chunks", type=int, default=4)
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of oracle CPU work for the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--verbose", action="store_true", help="per-chunk timings on stderr")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adj-side", type=int, default=32)
    ap.add_argument("--adj-steps", type=int, default=100)
    ap.add_argument("--adj-k", type=int, default=10)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--no-inverse", action="store_true")
    ap.add_argument("--inv-rollouts", type=int, default=16)
    ap.add_argument("--inv-steps", type=int, default=200)
    ap.add_argument("--inv-k", type=int, default=20)
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}), flush=True)
        sys.exit(2)
    from physics_amd import ffi, scenes

    n_side = a.n_side
    sc = scenes.dam_break((n_side, n_side, n_side), seed=a.seed)
    n = len(sc.pos)
    gp = scenes.apply_scene(ffi.PhySphParams.defaults(), sc)
    g = ffi.FluidWorld(gp)
    g.set_particles_f32(sc.pos, None, np.full(n, sc.mass, dtype=np.float32))
    dt = sc.dt

    for _ in range(a.warmup):
        g.step(dt)
    g.sync()

    K = a.steps
    chunks = max(1, min(a.chunks, K))
    sizes = [K // chunks + (1 if i < K % chunks else 0) for i in range(chunks)]
    wall = 0.0
    stage = np.zeros(4)
    timed_steps = 0
    snaps = []
    for c, k in enumerate(sizes):
        if k == 0:
            continue
        g.sync()
        g.profile_begin()
        t0 = time.perf_counter()
        g.step(dt, k)
        g.sync()
        wall += time.perf_counter() - t0
        ms, s = g.profile_end()
        stage += ms
        timed_steps += s
        if a.verbose:
            print(f"[bench] chunk {c}: {k} steps, {1e3 * (time.perf_counter() - t0) / k:.3f} ms/step wall, "
                  f"sort {ms[0] / s:.3f} density {ms[1] / s:.3f} force {ms[2] / s:.3f} ms/step",
                  file=sys.stderr, flush=True)
        if not a.no_cpu:
            snaps.append((g.get("pos"), g.get("vel")))
    assert timed_steps == K

    value = n * K / wall
    per = stage / K  # ms per step per stage
    names = ["sort", "density", "force"]
    dom = names[int(np.argmax(per[:3]))]
    t_dom = per[names.index(dom)] * 1e-3
    achieved = ALG_BYTES[dom] * n / t_dom / 1e9
    x = g.get("pos")
    assert np.isfinite(x).all(), "non-finite state in the benchmark run"
    # the counter summary describes the default workload only
    default_run = n_side == 100 and a.seed == 1 and a.warmup + K == 1000
    traffic = measured_traffic(f"k_{dom}") if default_run and dom != "sort" else None
    traffic_src = ("rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE mean over the 1000 dispatches of this "
                   "workload, scaled by a 1 GiB copy's counts (profiles/k_force_1M_dambreak_hbm_traffic.md)"
                   if traffic else None)

    out = {
        "metric": "SPH particle-steps/s at 1M (FP32 fwd) + adjoint/forward time ratio, 1 GPU",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"dam-break column {n_side}^3 = {n} particles, spacing h/2 (0.1f), "
                        f"+-1% SplitMix64 jitter seed {a.seed}, zero velocity, closed "
                        f"{sc.bounds_max[0]:.0f} m tank, dt {dt}, steps {a.warmup + 1}.."
                        f"{a.warmup + K} of the run",
            "particles": n,
            "h": sc.h,
            "parallelism": "single GPU",
        },
        "stages_ms_per_step": {"sort": per[0], "density": per[1], "force": per[2]},
        "adjoint_forward_ratio": None,
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_{dom}" if dom != "sort" else "sort kernels",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "alg_bytes_per_particle": ALG_BYTES[dom],
            # the kernel is bound by vector-ALU issue, not by HBM: its VALUBusy
            # from the same counter summary
            "valu_busy_pct": measured_counter("valu_busy_pct", f"k_{dom}") if traffic else None,
        },
    }

    if not a.no_adjoint:
        out["adjoint"] = adjoint_leg(a)
        out["adjoint_forward_ratio"] = out["adjoint"]["ratio"]
    if not a.no_batch:
        out["batched_scenes"] = batch_leg(a)
    if not a.no_inverse:
        out["inverse_design"] = inverse_leg(a)

    if not a.no_cpu and snaps:
        import oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        O.set_threads(threads)
        op = O.Params.defaults()
        op.h = sc.h
        for ax in range(3):
            op.bounds_min[ax] = sc.bounds_min[ax]
            op.bounds_max[ax] = sc.bounds_max[ax]
        w = O.World(op)
        mass = np.full(n, float(np.float32(sc.mass)))
        t_cpu, done, used = 0.0, 0, []
        # one oracle step from each snapshot, spread over the timed trajectory
        for idx in range(len(snaps)):
            pos, vel = snaps[idx]
            w.set_particles(pos, vel, mass)
            t_cpu += w.time_steps(dt, 1)
            done += 1
            used.append(idx)
            if t_cpu >= a.cpu_budget:
                break
        out["cpu_baseline"] = {
            "value": n * done / t_cpu,
            "unit": "particle-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"f64 oracle (C++ restatement of FluidWorld::step, OpenMP) one step "
                      f"from each of {done} GPU states taken at the end of timed chunks "
                      f"{used} of {len(sizes)}; {t_cpu:.1f} s of CPU work",
        }
    print(json.dumps(out), flush=True)


def adjoint_leg(a):
    """Forward rollout vs its reverse-mode gradient on the differentiable-rollout
    block: ratio = t(tape_backward: checkpoint recomputation + adjoint steps) /
    t(forward rollout of the same steps)."""
    from physics_amd import ffi, scenes
    sc = scenes.rollout_block(a.adj_side, seed=1)
    n = len(sc.pos)
    gp = scenes.apply_scene(ffi.PhySphParams.defaults(), sc)
    m = np.full(n, sc.mass, dtype=np.float32)
    dt = float(np.float32(sc.dt))
    g = ffi.FluidWorld(gp)
    # warm-up (kernel loading, allocations)
    g.set_particles_f32(sc.pos, sc.vel, m)
    g.step(dt, 5)
    g.tape_begin(a.adj_k)
    g.step(dt, 5)
    g.tape_backward(np.zeros((n, 3)), None, None)
    g.tape_end()
    # plain forward rollout
    g.set_particles_f32(sc.pos, sc.vel, m)
    g.sync()
    t0 = time.perf_counter()
    g.step(dt, a.adj_steps)
    g.sync()
    t_fwd = time.perf_counter() - t0
    # recorded forward + backward
    g.set_particles_f32(sc.pos, sc.vel, m)
    g.sync()
    t0 = time.perf_counter()
    g.tape_begin(a.adj_k)
    g.step(dt, a.adj_steps)
    g.sync()
    t_rec = time.perf_counter() - t0
    L, xbar, rhobar = scenes.rollout_loss_cotangents(g.get("pos"), g.get("rho"))
    t0 = time.perf_counter()
    _, v0b, dk, dmu = g.tape_backward(xbar, None, rhobar)[0:4]
    t_bwd = time.perf_counter() - t0
    g.tape_end()
    return {
        "config": f"rollout block {a.adj_side}^3 = {n} particles, {a.adj_steps} FP32 steps, "
                  f"checkpoint every {a.adj_k}; loss |centroid|^2 + mean (rho - rho0)^2",
        "forward_ms": t_fwd * 1e3,
        "recorded_forward_ms": t_rec * 1e3,
        "backward_ms": t_bwd * 1e3,
        "ratio": t_bwd / t_fwd,
        "loss": L,
        "dL_dstiffness": dk,
        "dL_dviscosity": dmu,
    }


def _concat(scs):
    pos = np.concatenate([sc.pos for sc in scs])
    vel = np.concatenate([sc.vel for sc in scs])
    mass = np.concatenate([np.full(len(sc.pos), sc.mass, dtype=np.float32) for sc in scs])
    return [len(sc.pos) for sc in scs], pos, vel, mass


def batch_leg(a):
    """Independent dam breaks of log-uniform size in one batch: every stage is
    one launch over all scenes and each step is one hipGraph replay. Also a
    one-step f64-oracle check of a sample of scenes after the timed run."""
    from physics_amd import ffi, scenes
    scs = scenes.batch_scenes(a.batch_scenes, seed=7)
    counts, pos, vel, mass = _concat(scs)
    params = [scenes.apply_scene(ffi.PhySphParams.defaults(), sc) for sc in scs]
    dt = float(np.float32(scenes.DT))
    b = ffi.FluidBatch(params)
    b.set_particles_f32(counts, pos, vel, mass)
    b.sync()
    t0 = time.perf_counter()
    b.step(dt, a.batch_steps)
    b.sync()
    wall = time.perf_counter() - t0
    # stage split of a few eager (event-profiled) steps at the end state
    b.profile_begin()
    b.step(dt, 5)
    ms, k = b.profile_end()
    res = {
        "config": f"{len(scs)} scenes, {sum(counts)} particles (log-uniform 2k..64k per scene, "
                  f"min {min(counts)}, max {max(counts)}), {a.batch_steps} FP32 steps from rest "
                  f"shape + N(0, 0.1) velocities, dt {dt:.6g}, hipGraph replay",
        "scene_steps_per_s": len(scs) * a.batch_steps / wall,
        "particle_steps_per_s": sum(counts) * a.batch_steps / wall,
        "ms_per_step": wall / a.batch_steps * 1e3,
        "eager_stage_ms_per_step_at_end": {"sort": ms[0] / k, "density": ms[1] / k, "force": ms[2] / k},
    }
    if not a.no_cpu:
        import oracle as O
        off = b.offsets()
        x0, v0, m = b.get("pos"), b.get("vel"), b.get("mass")
        b.step(dt)
        x1, v1, r1 = b.get("pos"), b.get("vel"), b.get("rho")
        worst = {"rho": 0.0, "vel": 0.0, "dx": 0.0}
        sample = list(range(0, len(scs), max(1, len(scs) // 16)))
        for s in sample:
            lo, hi = off[s], off[s + 1]
            o = O.World(scenes.apply_scene(O.Params.defaults(), scs[s]))
            o.set_particles(x0[lo:hi].astype(np.float64), v0[lo:hi].astype(np.float64),
                            m[lo:hi].astype(np.float64))
            o.step(dt)
            xo = o.get("pos")
            ulp = np.abs(np.spacing(xo.astype(np.float32)).astype(np.float64))
            exc = np.maximum(np.abs(x1[lo:hi] - xo) - ulp, 0.0)
            disp = np.sqrt(((xo - x0[lo:hi]) ** 2).sum(1)).max()
            e = {"rho": np.abs(r1[lo:hi] - o.get("rho")).max() / np.abs(o.get("rho")).max(),
                 "vel": np.sqrt(((v1[lo:hi] - o.get("vel")) ** 2).sum(1)).max()
                 / np.sqrt((o.get("vel") ** 2).sum(1)).max(),
                 "dx": np.sqrt((exc ** 2).sum(1)).max() / disp}
            for f in worst:
                worst[f] = max(worst[f], float(e[f]))
        res["oracle_one_step_error_worst_of_sample"] = worst
        res["oracle_sample_scenes"] = len(sample)
    b.close()
    return res


def _rollout_cotangents(x, rho, off):
    from physics_amd import scenes
    losses = []
    xbar = np.zeros_like(x)
    rhobar = np.zeros_like(rho)
    for s in range(len(off) - 1):
        lo, hi = off[s], off[s + 1]
        L, xbar[lo:hi], rhobar[lo:hi] = scenes.rollout_loss_cotangents(x[lo:hi], rho[lo:hi])
        losses.append(L)
    return losses, xbar, rhobar


def inverse_leg(a):
    """Batched differentiable rollouts with per-rollout viscosity: recorded
    forward (checkpoint every K) + reverse-mode gradient of every rollout."""
    from physics_amd import ffi, scenes
    ro = scenes.inverse_design_rollouts(a.inv_rollouts, seed=5)
    scs = [sc for sc, _ in ro]
    params = []
    for sc, mu in ro:
        p = scenes.apply_scene(ffi.PhySphParams.defaults(), sc)
        p.set_visc([mu], [1.0])
        params.append(p)
    counts, pos, vel, mass = _concat(scs)
    dt = float(np.float32(scenes.DT))
    free0, total = ffi.device_mem_info()
    b = ffi.FluidBatch(params)
    b.set_particles_f32(counts, pos, vel, mass)
    off = b.offsets()
    # warm-up: kernels loaded, graph captured, tape buffers allocated
    b.tape_begin(a.inv_k)
    b.step(dt, 2)
    b.tape_backward(np.zeros((len(pos), 3), np.float32), None, None)
    b.tape_end()
    b.set_particles_f32(counts, pos, vel, mass)
    b.sync()
    t0 = time.perf_counter()
    b.step(dt, a.inv_steps)
    b.sync()
    t_plain = time.perf_counter() - t0
    b.set_particles_f32(counts, pos, vel, mass)
    b.sync()
    t0 = time.perf_counter()
    b.tape_begin(a.inv_k)
    b.step(dt, a.inv_steps)
    b.sync()
    t_fwd = time.perf_counter() - t0
    losses, xbar, rhobar = _rollout_cotangents(b.get("pos"), b.get("rho"), off)
    t0 = time.perf_counter()
    _, v0b, dk, dmu = b.tape_backward(xbar, None, rhobar)
    t_bwd = time.perf_counter() - t0
    free1, _ = ffi.device_mem_info()
    b.tape_end()
    b.close()
    return {
        "config": f"{len(scs)} rollouts x {counts[0]} particles (64x64x32 blocks, free fall), "
                  f"visc_k[0] ~ U[0.5, 2] x 3.5 per rollout, {a.inv_steps} FP32 steps, "
                  f"checkpoint every {a.inv_k}",
        "rollouts_per_s_forward_plus_adjoint": len(scs) / (t_fwd + t_bwd),
        "forward_ms": t_plain * 1e3,
        "recorded_forward_ms": t_fwd * 1e3,
        "backward_ms": t_bwd * 1e3,
        "adjoint_forward_ratio": t_bwd / t_plain,
        "hbm_in_use_after_backward_GB": (free0 - free1) / 1e9,
        "dL_dviscosity": [float(x) for x in dmu],
        "dL_dstiffness": [float(x) for x in dk],
        "loss": losses,
    }


if __name__ == "__main__":
    main()
