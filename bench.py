"""Benchmark: SPH particle-steps/s of the 1M-particle FP32 dam break on one GPU.

Workload: a 100 x 100 x 100 particle column (spacing h/2, +-1 % seeded jitter,
zero velocity) in a closed 40 m tank, dt = 0.01, the FP32 forward step of
FluidWorld::step (crates/phy-fluid/src/sph/world.rs:183-192) on the GPU, a
1000-step run. Inputs are resident in HBM before the timed region.

The headline `value` times steps warmup + 1 .. warmup + steps of the run
(default: 20 warm-up steps, then the remaining 980). The column collapses onto
the floor within ~150 steps into a sheet of ~1,000 neighbours per particle, so
a short window at the start of the run is the cheap phase; the output also
carries, under "windows":
  initial    the headline window;
  collapsed  the same run continued untimed to step 300, then the same number
             of steps timed (when they fit in the run);
  full_run   a fresh world's whole 1000-step run in one bracket (when the
             headline window does not already cover it).
Each window has its stage split, roofline and (with the CPU baseline) the f64
oracle's particle-steps/s sampled from that window's states.

Prints one JSON line. Timed steps run in chunks, each bracketed by a device
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
FP32_PEAK_TFLOPS = 157.3
# HBM bytes per launch and VALUBusy measured with rocprofv3 --pmc on the
# dispatches of each window of the 1M-particle run (scaled by a plain copy's
# counts); lines "<key> <kernel> [<window>]: <value>"
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_force_1M_dambreak_hbm_traffic.md")


def measured_counter(key: str, kernel: str, window):
    """A per-launch figure of `kernel` in `window` from the committed counter
    summary, or None (no summary line for exactly that window)."""
    if window is None:
        return None
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith(f"{key} {kernel} [{window}]:"):
                return float(line.split("]:")[1])
    except OSError:
        pass
    return None


def measured_traffic(kernel: str, window):
    """Bytes per launch of `kernel` in `window` from the counter summary, or None."""
    v = measured_counter("traffic_bytes_per_launch", kernel, window)
    return None if v is None else int(v)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=980)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n-side", type=int, default=100)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--chunks", type=int, default=4)
    ap.add_argument("--run-steps", type=int, default=1000,
                    help="steps of the configured run (the full-run leg and the collapsed window live in it)")
    ap.add_argument("--collapsed-start", type=int, default=300,
                    help="steps run before the collapsed-sheet window")
    ap.add_argument("--cpu-budget", type=float, default=10.0,
                    help="seconds of oracle CPU work per window for the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-full-run", action="store_true")
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


def cpu_info():
    """Host CPU model (lscpu's "Model name", from /proc/cpuinfo) and the
    number of CPUs this process may run on (nproc)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    return model, nproc


def oracle_threads():
    """Oracle threads for the multi-core baseline: the process's CPU share
    (OMP_NUM_THREADS where the host sets it, else nproc)."""
    _, nproc = cpu_info()
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return env if env > 0 else nproc


def time_window(g, dt, k_steps, chunks, tag, verbose, snaps=None):
    """Times k_steps steps of world g in chunks, each bracketed by a device
    synchronisation, with HIP-event stage times on the world's stream.
    Returns (wall seconds, stage ms summed over the steps {sort, density,
    force, step}). When snaps is a list, the state at the end of every chunk
    is appended to it (outside the timed windows)."""
    chunks = max(1, min(chunks, k_steps))
    sizes = [k_steps // chunks + (1 if i < k_steps % chunks else 0) for i in range(chunks)]
    wall = 0.0
    stage = np.zeros(4)
    timed = 0
    for c, k in enumerate(sizes):
        if k == 0:
            continue
        g.sync()
        g.profile_begin()
        t0 = time.perf_counter()
        g.step(dt, k)
        g.sync()
        w = time.perf_counter() - t0
        wall += w
        ms, n = g.profile_end()
        stage += ms
        timed += n
        if verbose:
            print(f"[bench] {tag} chunk {c}: {k} steps, {1e3 * w / k:.3f} ms/step wall, "
                  f"sort {ms[0] / n:.3f} density {ms[1] / n:.3f} force {ms[2] / n:.3f} ms/step",
                  file=sys.stderr, flush=True)
        if snaps is not None:
            snaps.append((g.get("pos"), g.get("vel")))
    assert timed == k_steps
    return wall, stage


def window_report(g, n, k_steps, wall, stage, first_step, traffic_key):
    """Throughput, stage split and roofline of one timed window. The
    dominant kernel's achieved bandwidth is its algorithmic bytes per
    launch over its live HIP-event average launch time; the pair-flop
    fraction counts 16 flop per density neighbour and 75 per force neighbour
    (the reference's per-pair arithmetic, neighbour counts of the window's
    end state from the gathers' own statistics) against the FP32 vector peak."""
    per = stage / k_steps
    names = ["sort", "density", "force"]
    dom = names[int(np.argmax(per[:3]))]
    t_dom = per[names.index(dom)] * 1e-3
    achieved = ALG_BYTES[dom] * n / t_dom / 1e9
    dc, _, _ = g.gather_stats("density")
    fc, _, _ = g.gather_stats("force")
    flops = 16.0 * float(dc.astype(np.float64).sum()) + 75.0 * float(fc.astype(np.float64).sum())
    pair_tflops = flops / (per[3] * 1e-3) / 1e12
    traffic = measured_traffic(f"k_{dom}", traffic_key) if dom != "sort" else None
    valu = measured_counter("valu_busy_pct", f"k_{dom}", traffic_key) if dom != "sort" else None
    return {
        "steps": f"{first_step}..{first_step + k_steps - 1}",
        "value": n * k_steps / wall,
        "unit": "particle-steps/s",
        "ms_per_step": wall / k_steps * 1e3,
        "stages_ms_per_step": {"sort": per[0], "density": per[1], "force": per[2]},
        "neighbours_per_particle": {"density": float(dc.mean()), "force": float(fc.mean())},
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_{dom}" if dom != "sort" else "sort kernels",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": (f"rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of this window's dispatches, "
                               f"scaled by a 1 GiB copy's counts ({os.path.relpath(TRAFFIC_PROFILE, ROOT)})"
                               if traffic else None),
            "alg_bytes_per_particle": ALG_BYTES[dom],
            "pair_flops_per_step": flops,
            "pair_tflops": pair_tflops,
            "pair_flop_frac_of_fp32_peak": pair_tflops / FP32_PEAK_TFLOPS,
            "valu_busy_pct": valu,
        },
    }


def cpu_sample(sc, n, dt, snaps, budget, threads):
    """f64 oracle steps from GPU states, on `threads` host threads, until
    `budget` seconds of CPU work: particle-steps/s and what was sampled."""
    import oracle as O
    O.set_threads(threads)
    op = O.Params.defaults()
    op.h = sc.h
    for ax in range(3):
        op.bounds_min[ax] = sc.bounds_min[ax]
        op.bounds_max[ax] = sc.bounds_max[ax]
    w = O.World(op)
    mass = np.full(n, float(np.float32(sc.mass)))
    # one step from each state in turn, cycling over the states, until the
    # budget is spent (every state at least once)
    t_cpu, done = 0.0, 0
    while done < len(snaps) or t_cpu < budget:
        pos, vel = snaps[done % len(snaps)]
        w.set_particles(pos, vel, mass)
        t_cpu += w.time_steps(dt, 1)
        done += 1
    return n * done / t_cpu, done, t_cpu


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1",
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

    # the headline window: steps warmup + 1 .. warmup + steps of the run
    K = a.steps
    snaps = []
    wall, stage = time_window(g, dt, K, a.chunks, "initial", a.verbose, None if a.no_cpu else snaps)
    done_steps = a.warmup + K
    x = g.get("pos")
    assert np.isfinite(x).all(), "non-finite state in the benchmark run"
    key = f"initial {a.warmup + 1}..{done_steps}" if n_side == 100 and a.seed == 1 else None
    head = window_report(g, n, K, wall, stage, a.warmup + 1, key)

    # the collapsed sheet: the same run continued (untimed) to step
    # collapsed_start, then the same number of steps timed
    coll = None
    snaps_c = []
    if a.collapsed_start >= done_steps and a.collapsed_start + K <= a.run_steps:
        g.step(dt, a.collapsed_start - done_steps)
        g.sync()
        wall_c, stage_c = time_window(g, dt, K, a.chunks, "collapsed", a.verbose,
                                      None if a.no_cpu else snaps_c)
        key_c = (f"collapsed {a.collapsed_start + 1}..{a.collapsed_start + K}"
                 if n_side == 100 and a.seed == 1 else None)
        coll = window_report(g, n, K, wall_c, stage_c, a.collapsed_start + 1, key_c)
    del g

    out = {
        "metric": "SPH particle-steps/s at 1M (FP32 fwd) + adjoint/forward time ratio, 1 GPU",
        "value": head["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"dam-break column {n_side}^3 = {n} particles, spacing h/2 (0.1f), "
                        f"+-1% SplitMix64 jitter seed {a.seed}, zero velocity, closed "
                        f"{sc.bounds_max[0]:.0f} m tank, dt {dt}, {a.run_steps}-step run; value: "
                        f"steps {a.warmup + 1}..{a.warmup + K} of the run",
            "particles": n,
            "h": sc.h,
            "parallelism": "single GPU",
        },
        "stages_ms_per_step": head["stages_ms_per_step"],
        "adjoint_forward_ratio": None,
        "roofline": head["roofline"],
        "windows": {"initial": head},
    }
    if coll is not None:
        out["windows"]["collapsed"] = coll

    if not a.no_full_run and a.warmup + K < a.run_steps:
        out["windows"]["full_run"] = full_run_leg(a, sc, n, dt)

    if not a.no_adjoint:
        out["adjoint"] = adjoint_leg(a)
        out["adjoint_forward_ratio"] = out["adjoint"]["ratio"]
    if not a.no_batch:
        out["batched_scenes"] = batch_leg(a)
    if not a.no_inverse:
        out["inverse_design"] = inverse_leg(a)

    if not a.no_cpu and snaps:
        model, nproc = cpu_info()
        threads = oracle_threads()
        v, done, t = cpu_sample(sc, n, dt, snaps, a.cpu_budget, threads)
        base = {
            "value": v,
            "unit": "particle-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"f64 oracle (C++ restatement of FluidWorld::step, OpenMP over particles, "
                      f"serial hash grid, integrate and bounds): {done} single steps, each from one of the "
                      f"{len(snaps)} GPU states at the ends of the headline window's chunks (steps "
                      f"{head['steps']}) in turn, {t:.1f} s of CPU work",
            "cpu_model": model,
            "nproc": nproc,
        }
        v1, d1, t1 = cpu_sample(sc, n, dt, snaps[-1:], 0.0, 1)
        base["value_1_thread"] = v1
        base["sample_1_thread"] = f"one oracle step on 1 thread from the last headline-window state ({t1:.1f} s)"
        if snaps_c:
            vc, dc_, tc = cpu_sample(sc, n, dt, snaps_c, a.cpu_budget, threads)
            base["collapsed"] = {
                "value": vc, "cores": threads,
                "sample": f"{dc_} single steps, each from one of the {len(snaps_c)} GPU states of the "
                          f"collapsed window (steps {coll['steps']}) in turn, {tc:.1f} s of CPU work"}
            coll["cpu_baseline"] = base["collapsed"]
            coll["gpu_over_cpu"] = coll["value"] / vc
        head["gpu_over_cpu"] = head["value"] / v
        out["cpu_baseline"] = base
    print(json.dumps(out), flush=True)


def full_run_leg(a, sc, n, dt):
    """The configured run from its first step: a fresh world (kernels already
    loaded), run_steps steps in one synchronised bracket, HIP-event stages."""
    from physics_amd import ffi, scenes
    gp = scenes.apply_scene(ffi.PhySphParams.defaults(), sc)
    g = ffi.FluidWorld(gp)
    g.set_particles_f32(sc.pos, None, np.full(n, sc.mass, dtype=np.float32))
    g.sync()
    g.profile_begin()
    t0 = time.perf_counter()
    g.step(dt, a.run_steps)
    g.sync()
    wall = time.perf_counter() - t0
    ms, k = g.profile_end()
    assert np.isfinite(g.get("pos")).all()
    per = ms / k
    return {
        "steps": f"1..{a.run_steps}",
        "value": n * a.run_steps / wall,
        "unit": "particle-steps/s",
        "ms_per_step": wall / a.run_steps * 1e3,
        "stages_ms_per_step": {"sort": per[0], "density": per[1], "force": per[2]},
    }


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
    # the library's own device-memory high-water mark over the recorded
    # rollout and its backward pass (every buffer of the batch included)
    in_use0, _ = ffi.device_mem_stats()
    ffi.device_mem_reset_peak()
    t0 = time.perf_counter()
    b.tape_begin(a.inv_k)
    b.step(dt, a.inv_steps)
    b.sync()
    t_fwd = time.perf_counter() - t0
    losses, xbar, rhobar = _rollout_cotangents(b.get("pos"), b.get("rho"), off)
    t0 = time.perf_counter()
    _, v0b, dk, dmu = b.tape_backward(xbar, None, rhobar)
    t_bwd = time.perf_counter() - t0
    _, peak = ffi.device_mem_stats()
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
        "peak_hbm_GB": peak / 1e9,
        "peak_hbm_source": "high-water mark of the library's own device allocations "
                           "(phy_device_mem_stats) over the recorded forward and backward",
        "hbm_before_tape_GB": in_use0 / 1e9,
        "dL_dviscosity": [float(x) for x in dmu],
        "dL_dstiffness": [float(x) for x in dk],
        "loss": losses,
    }


if __name__ == "__main__":
    main()
