// sph_couple.h — device interface of the fluid coupling passes
// (crates/phy-fluid/src/sph/world.rs:400-773): rigid bodies, point masses,
// a heat grid, and the nearest-particle velocity sampler.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_types.h"

namespace phy {

enum CoupleShape : int32_t { kShapeSphere = 0, kShapeBox = 1, kShapeCapsule = 2 };

// f64 world parameters the coupling reads (SphParams bounds, gravity,
// rest density, the `mass` field and the grid's 1/h).
struct CoupleParams {
    double lo[3], hi[3], g[3];
    double rest, mass, inv_cell, h;
};

// The parts of phy-rigid's Body<f64> the coupling uses (shape.rs:433-463).
struct CoupleBody {
    double pos[3];
    double q[4];     // orientation, w i j k (body -> world)
    double vel[3];   // updated by the contact impulses and the buoyancy
    double inv_mass; // <= 0: static
    double dims[3];  // sphere: r; box: half extents; capsule: half_height, r
    int32_t shape;
    int32_t pad;
};

// CouplePoint<f64> (sph/couple.rs:10-19)
struct CouplePoint {
    double pos[3], vel[3];
    double mass;
    double force[3];
};

struct Contact {
    uint32_t slot;  // particle slot in the current state order
    uint32_t body;
};

// ---- bodies (world.rs:414-503) ----
// cnt[id[i]] = number of bodies containing particle slot i (original index order)
void launch_contact_count(hipStream_t st, int n, const float4* posm, const uint32_t* id,
                          const CoupleBody* bodies, int nb, uint32_t* cnt);
// contacts in (original index, body) order at off[id[i]]
void launch_contact_fill(hipStream_t st, int n, const float4* posm, const uint32_t* id,
                         const CoupleBody* bodies, int nb, const uint32_t* off, Contact* out);
// resolves the contacts in order (one wave), then applies each dynamic body's
// buoyancy; displaced: nb doubles, zero on entry
void launch_contact_resolve(hipStream_t st, const Contact* c, const uint32_t* total, float4* posm,
                            float4* vel, CoupleBody* bodies, int nb, double* displaced,
                            const CoupleParams& cp, double dt);

// out[i] = in[id[i]] (original index order -> state order)
void launch_gather_by_id4(hipStream_t st, int n, const uint32_t* id, const float4* in, float4* out);

// ---- points (world.rs:691-752) ----
// inv[id[i]] = i
void launch_inverse_perm(hipStream_t st, int n, const uint32_t* id, uint32_t* inv);
void launch_widen_vel(hipStream_t st, int n, const float4* vel, double* vel64);
void launch_narrow_vel(hipStream_t st, int n, const double* vel64, float4* vel);
// sequential over points, one workgroup; the cell table (posm_s, id_s, start,
// dom) bins the current positions
void launch_couple_points(hipStream_t st, CouplePoint* pts, int np, const CoupleParams& cp, double dt,
                          double drag, double soft_density, const float4* posm_s, const uint32_t* id_s,
                          const uint32_t* start, const Domain* dom, const uint32_t* inv, double* vel64);

// ---- heat grid (world.rs:555-678) ----
struct HeatGrid {
    uint32_t nx, ny, nz;
    double origin[3];
    double dx;
    const double* u;  // x fastest: u[x + nx (y + ny z)]
};
struct HeatScratch {
    double* q;                  // per particle slot: source term
    uint32_t* qcell;            // per particle slot: heat cell or 0xffffffff
    unsigned long long* qmax;   // 1: bit pattern of the largest finite |q|
    int* scale;                 // 1: fixed-point exponent
    unsigned long long* lo;     // per heat cell: 128-bit fixed-point sum of |q|
    unsigned long long* hi;
    uint32_t* flags;            // per heat cell: 1 touched, 2 NaN, 4 infinite
};
// body[i] += -g beta (T(x_i) - t_ref) (FP32 state, f64 arithmetic); the
// source terms go to scratch; heat_gain <= 0 skips them
void launch_heat_body(hipStream_t st, int n, const float4* posm, const float4* vel, float4* body,
                      const CoupleParams& cp, const HeatGrid& hg, double dt, double t_ref, double beta,
                      double heat_gain, const HeatScratch& hs);
// src[c] += sum of the cell's source terms (deterministic fixed-point sum)
void launch_heat_source(hipStream_t st, int n, double dt, const HeatScratch& hs, uint32_t ncell,
                        double* src);

// ---- velocity sampler (world.rs:614-643) ----
void launch_nearest_velocity(hipStream_t st, int nq, const double* q3, double* out3, const float4* posm_s,
                             const uint32_t* id_s, const uint32_t* start, const Domain* dom,
                             const uint32_t* inv, const float4* vel, const CoupleParams& cp);

// the scan of the cell table, reused for the contact offsets: counts[0..*total)
// -> start[0..*total] (exclusive, start[*total] = sum); counts are cleared
void launch_exclusive_scan(hipStream_t st, uint32_t* counts, const uint32_t* total, uint32_t* partial,
                           uint32_t* start);

}  // namespace phy
