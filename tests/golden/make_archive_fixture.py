"""Writes tests/golden/fluid_archive_small.json: a world archive in the layout
the reference's phy-io writes (crates/phy-io/src/archive.rs:139-324 with the
serde field order of FluidSubsystem, FluidWorld, SphParams and Particle; f64
scalars as {"__f64": bits}, archive.rs:31-100; Vec3 as u64 bit arrays,
crates/phy-rigid/src/shape.rs:18-53). The state is a 3 x 3 x 3 lattice with f64
coordinates that FP32 does not represent exactly, two materials and a pending
body acceleration, at time 0.25.

Run: python tests/golden/make_archive_fixture.py
"""
import json
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))


def bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def f64(x: float) -> dict:
    return {"__f64": bits(x)}


def vec3(v) -> list:
    return [bits(float(c)) for c in v]


def archive() -> dict:
    params = {
        "rest_density": f64(1000.0), "stiffness": f64(250.0), "viscosity": f64(3.5),
        "visc_k": [f64(3.5), f64(10.0)], "visc_n": [f64(1.0), f64(1.0)],
        "shear_min": f64(0.01), "mass": f64(0.02), "h": f64(0.2),
        "gravity": vec3((0.0, -9.81, 0.0)), "bounds_min": vec3((-5.0, -5.0, -5.0)),
        "bounds_max": vec3((5.0, 5.0, 5.0)), "boundary_damp": f64(0.4), "xsph_eps": f64(0.5),
    }
    particles = []
    k = 0
    for i in range(3):
        for j in range(3):
            for l in range(3):
                pos = (0.1 * i + 1e-9 * k, 0.1 * j + 0.3333333333333333, -0.1 * l)
                vel = (0.01 * k, -0.02, 1.0 / 3.0)
                particles.append({
                    "pos": vec3(pos), "vel": vec3(vel), "acc": vec3((0.0, 0.0, 0.0)),
                    "body_acc": vec3((0.0, 1.5 if k == 4 else 0.0, 0.0)),
                    "rho": f64(0.0), "p": f64(0.0), "mass": f64(0.7 + 0.01 * k), "material": k % 2,
                })
                k += 1
    fluid = {"world": {"params": params, "particles": particles}, "thermal_expansion": f64(0.0),
             "heat_gain": f64(0.0), "t_ref": f64(0.0), "drag": f64(1.0), "friction": f64(0.1)}
    return {"version": 1, "time": f64(0.25), "subsystems": [{"Fluid": fluid}]}


if __name__ == "__main__":
    with open(os.path.join(HERE, "fluid_archive_small.json"), "w") as f:
        json.dump(archive(), f, indent=2)
        f.write("\n")
    print("wrote fluid_archive_small.json")
