// archive.h — the reference's JSON world archive (phy-io, crates/phy-io/src/
// archive.rs:139-324) for worlds of fluid subsystems.
//
// Layout (serde field order of WorldArchive, FluidSubsystem, FluidWorld's
// params + particles, SphParams, Particle):
//   {"version": 1, "time": F, "subsystems": [{"Fluid": {"world": {"params":
//    {...}, "particles": [...]}, "thermal_expansion": F, "heat_gain": F,
//    "t_ref": F, "drag": F, "friction": F}}]}
// where every f64 scalar F is written as {"__f64": <u64 bit pattern>}
// (archive.rs:31-100) and every Vec3 as an array of three u64 bit patterns
// (phy-rigid shape.rs:18-53, serde_geom). Integers (version, material) are
// plain numbers. Loading also accepts plain JSON numbers for scalars.
#pragma once
#include <memory>
#include <string>

#include "world.h"

namespace phy {

// Pretty-printed JSON of the world (two-space indent, like serde_json's
// to_string_pretty). Particle state is the device's FP32 state as f64.
std::string save_world_json(World& w);

// Rebuilds a world from an archive. Throws std::runtime_error on malformed
// input, on a non-finite {"__f64"} scalar (the reference's loader rejects
// those, archive.rs:77-81) and on subsystem kinds this library does not host.
std::unique_ptr<World> load_world_json(const std::string& json);

}  // namespace phy
