"""Per-window kernel figures of a rocprofv3 run of bench.py.

bench.py steps one 1M-particle world through its warm-up, the headline window
(steps W + 1 .. W + K), an untimed stretch and the collapsed window (steps C + 1
.. C + K), then, unless switched off, a fresh world's full run (steps 1 .. R)
and the adjoint, batched-scene and inverse-design legs. This module splits the
dispatches of a `rocprofv3 --kernel-trace` CSV (`*_kernel_trace.csv`) or of a
`rocprofv3 --pmc` CSV (`*_counter_collection.csv`) of such a run into those
windows by time: the 1M world's force kernel (the last launch of every step)
marks the steps, and a window of steps a..b holds every dispatch that starts
after step a - 1's force kernel ends and ends by step b's. Per window and
kernel it prints the launches per step and the mean duration (kernel trace)
or the mean counter values per launch (counter collection); dispatches outside
the 1M world (the other legs) are summarised per kernel and grid size.

usage: python -m physics_amd.trace_windows CSV [CSV ...] [--warmup 5] [--steps 20]
           [--collapsed-start 300] [--run-steps 1000] [--particles 1000000]
           [--lines-for KERNEL [--fetch-scale B --write-scale B]]
Several CSVs (one per rocprofv3 run of the same command, e.g. one counter
each) are merged per window and kernel. --lines-for prints the counter-summary
lines bench.py reads, with HBM bytes per launch from FETCH_SIZE and WRITE_SIZE
scaled by a plain copy's bytes per counter unit.
"""
from __future__ import annotations

import argparse
import collections
import csv
import re
import sys

K_TILE = 256


def _short(name: str) -> str:
    name = name.split("(")[0]
    for pre in ("void ", "phy::"):
        name = name.replace(pre, "")
    return name


def _grid(row) -> int:
    if "Grid_Size" in row:
        return int(float(row["Grid_Size"]))
    return int(float(row.get("Grid_Size_X", 0))) * max(int(float(row.get("Grid_Size_Y", 1))), 1) * \
        max(int(float(row.get("Grid_Size_Z", 1))), 1)


def load(path: str):
    """Dispatches as dicts {name, grid, start, end, counters{}} in start order."""
    rows = list(csv.DictReader(open(path)))
    if not rows:
        return []
    disp = collections.OrderedDict()
    for r in rows:
        key = r.get("Dispatch_Id") or r.get("Correlation_Id")
        d = disp.get(key)
        if d is None:
            d = {"name": _short(r["Kernel_Name"]), "grid": _grid(r), "start": int(r["Start_Timestamp"]),
                 "end": int(r["End_Timestamp"]), "counters": {}}
            disp[key] = d
        if "Counter_Name" in r:
            d["counters"][r["Counter_Name"]] = d["counters"].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return sorted(disp.values(), key=lambda d: d["start"])


def world_force(disps, particles: int):
    """The 1M world's step-marking force launches (the step's form, not the
    statistics form), in order."""
    grid = -(-particles // K_TILE) * K_TILE
    pat = re.compile(r"^k_force<\d+, (true|false), \d, false, \d+(, (true|false))?>$")
    return [d for d in disps if pat.match(d["name"]) and d["grid"] == grid]


def windows(a):
    w = [(f"initial {a.warmup + 1}..{a.warmup + a.steps}", 0, a.warmup + 1, a.warmup + a.steps)]
    if a.collapsed_start >= a.warmup + a.steps and a.collapsed_start + a.steps <= a.run_steps:
        w.append((f"collapsed {a.collapsed_start + 1}..{a.collapsed_start + a.steps}", 0,
                  a.collapsed_start + 1, a.collapsed_start + a.steps))
    return w


# not part of a step: the gathers' statistics forms and the readbacks and
# copies bench.py makes between a window's chunks (outside its timed brackets)
NOT_STEP = re.compile(r"^(k_density<\d+, \d, true|k_force<\d+, (true|false), \d, true|k_unsort|__amd_rocclr)")


def split(disps, force, a):
    """{window: [dispatches]} for the first world's windows and the full run
    (dispatches that are not part of a step left out)."""
    first = a.collapsed_start + a.steps if len(windows(a)) > 1 else a.warmup + a.steps
    out = collections.OrderedDict()

    def take(tag, f0, s0, s1):
        # step s of the world whose first force launch is force[f0] ends with
        # force[f0 + s - 1]; the first step of a world starts with its first
        # AABB pass after the previous world's last step
        if s0 > 1:
            lo = force[f0 + s0 - 2]["end"]
        else:
            prev = force[f0 - 1]["end"] if f0 > 0 else -1
            lo = min(d["start"] for d in disps if d["start"] > prev and d["name"] == "k_aabb") - 1
        hi = force[f0 + s1 - 1]["end"]
        out[tag] = [d for d in disps if d["start"] > lo and d["end"] <= hi and not NOT_STEP.match(d["name"])]

    for tag, f0, s0, s1 in windows(a):
        take(tag, f0, s0, s1)
    if len(force) >= first + a.run_steps:
        take(f"full run 1..{a.run_steps}", first, 1, a.run_steps)
    return out


def summarise(group, nsteps):
    by = collections.OrderedDict()
    for d in group:
        by.setdefault(d["name"], []).append(d)
    res = []
    for name, ds in by.items():
        dur = [(d["end"] - d["start"]) * 1e-3 for d in ds]  # us
        ctr = collections.defaultdict(float)
        for d in ds:
            for c, v in d["counters"].items():
                ctr[c] += v
        res.append((name, len(ds), len(ds) / nsteps if nsteps else None, sum(dur) / len(dur),
                    sum(dur) / nsteps * 1e-3 if nsteps else None, {c: v / len(ds) for c, v in ctr.items()}))
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("csv", nargs="+", help="one CSV per rocprofv3 run of the same bench command")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--collapsed-start", type=int, default=300)
    ap.add_argument("--run-steps", type=int, default=1000)
    ap.add_argument("--particles", type=int, default=1000000)
    ap.add_argument("--lines-for", default=None,
                    help="print bench's counter-summary lines for this kernel (k_force, k_density)")
    ap.add_argument("--fetch-scale", type=float, default=None,
                    help="HBM bytes read per FETCH_SIZE KiB-unit, from a plain copy")
    ap.add_argument("--write-scale", type=float, default=None,
                    help="HBM bytes written per WRITE_SIZE KiB-unit, from a plain copy")
    a = ap.parse_args(argv)
    merged = collections.OrderedDict()  # window -> kernel -> [launches, per, us, ms, counters]
    others = collections.OrderedDict()
    for path in a.csv:
        disps = load(path)
        force = world_force(disps, a.particles)
        if len(force) < a.warmup + a.steps:
            sys.exit(f"{path}: {len(force)} step-marking force launches of the {a.particles}-particle "
                     f"world: fewer than the warm-up and headline window")
        inside = set()
        for tag, group in split(disps, force, a).items():
            s0, s1 = (int(x) for x in tag.split()[-1].split(".."))
            inside.update(id(d) for d in group)
            w = merged.setdefault(tag, collections.OrderedDict())
            for name, cnt, per, us, ms, ctr in summarise(group, s1 - s0 + 1):
                if name in w:
                    w[name][4].update(ctr)
                else:
                    w[name] = [cnt, per, us, ms, dict(ctr)]
        for d in disps:
            if id(d) not in inside:
                others.setdefault((d["name"], d["grid"]), []).append(d)
    if a.lines_for:
        for tag, w in merged.items():
            for name, (cnt, per, us, ms, ctr) in w.items():
                if not name.startswith(a.lines_for + "<"):
                    continue
                for c, v in sorted(ctr.items()):
                    print(f"{c} {a.lines_for} [{tag}]: {v:.1f}   ({cnt} launches of {name})")
                if "FETCH_SIZE" in ctr and "WRITE_SIZE" in ctr and a.fetch_scale and a.write_scale:
                    b = ctr["FETCH_SIZE"] * a.fetch_scale + ctr["WRITE_SIZE"] * a.write_scale
                    print(f"traffic_bytes_per_launch {a.lines_for} [{tag}]: {b:.0f}")
                if "VALUBusy" in ctr:
                    print(f"valu_busy_pct {a.lines_for} [{tag}]: {ctr['VALUBusy']:.1f}")
        return
    for tag, w in merged.items():
        s0, s1 = (int(x) for x in tag.split()[-1].split(".."))
        total = sum(r[3] for r in w.values())
        print(f"## window {tag}: {s1 - s0 + 1} steps, kernel time {total:.4f} ms/step")
        ctrs = sorted({c for r in w.values() for c in r[4]})
        print("| kernel | launches/step | mean us | ms/step |" + "".join(f" {c} |" for c in ctrs))
        print("|---|---|---|---|" + "---|" * len(ctrs))
        for name, (cnt, per, us, ms, ctr) in sorted(w.items(), key=lambda kv: -kv[1][3]):
            print(f"| {name} | {per:.2f} | {us:.2f} | {ms:.4f} |" + "".join(f" {ctr.get(c, 0.0):.1f} |" for c in ctrs))
        print()
    print("## other dispatches (warm-up, untimed steps, statistics passes, other legs)")
    print("| kernel | grid | launches | mean us | total ms |")
    print("|---|---|---|---|---|")
    for (name, grid), ds in sorted(others.items(), key=lambda kv: -sum(d["end"] - d["start"] for d in kv[1])):
        tot = sum(d["end"] - d["start"] for d in ds) * 1e-6
        print(f"| {name} | {grid} | {len(ds)} | {tot / len(ds) * 1e3:.2f} | {tot:.3f} |")


if __name__ == "__main__":
    main()
