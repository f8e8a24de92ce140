"""Builds physics_amd/libphy_ffi.so (gfx950) from physics_amd/csrc with hipcc.

Usage: python -m physics_amd.build   (or physics_amd.build.build())
       python -m physics_amd.build --hbm-copy   also builds the plain-copy
       program whose rocprofv3 --pmc counts calibrate FETCH_SIZE / WRITE_SIZE
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libphy_ffi.so")
# counter-calibration copy program (rocprofv3 --pmc reference point)
COPY_EXE = os.path.join(HERE, "hbm_copy")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["sph_kernels.hip", "sph_stats.hip", "sph_adjoint.hip", "sph_couple.hip", "fluid_device.cpp", "world.cpp",
           "archive.cpp", "phy_ffi.cpp"]
# the coupling's f64 arithmetic rounds operation by operation like the reference
EXTRA = {"sph_couple.hip": ["-ffp-contract=off"]}

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 generation (some of their forms
# have given wrong results on gfx950 without any error).
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]


def build(verbose: bool = False, hbm_copy: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC, *COMMON, *EXTRA.get(src, []), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the gfx-only feature: harmless
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            failed.append(src)
            print(f"[build] {src} FAILED:\n" + "\n".join(lines), file=sys.stderr)
        elif lines and verbose:
            print(f"[build] {src}:\n" + "\n".join(lines))
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if hbm_copy:
        build_hbm_copy()
    return LIB


def build_hbm_copy() -> str:
    """The counter-calibration copy program (profiling only, not the library)."""
    cmd = [HIPCC, *COMMON, os.path.join(CSRC, "hbm_copy.hip"), "-o", COPY_EXE + ".tmp"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for hbm_copy.hip:\n" + r.stdout.decode(errors="replace"))
    os.replace(COPY_EXE + ".tmp", COPY_EXE)
    return COPY_EXE


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, hbm_copy="--hbm-copy" in sys.argv))
