/* phy_ffi.h — C ABI of the MI355X phy-fluid library (libphy_ffi.so).
 *
 * Drop-in for the fluid subset of the reference ABI, crates/phy-ffi/phy_ffi.h
 * (ABI version 2): same symbol names, signatures, ownership and error codes.
 * The world's particle state lives on the GPU in FP32; the f64 readback channel
 * returns the exact (double) of the FP32 value, so the f32 channel is always
 * the exact (float) cast of the f64 channel (crates/phy-ffi/src/lib.rs:978-1001).
 *
 * Symbols below the "additive" line have no counterpart in the reference; they
 * are a backwards-compatible minor-version extension (policy lib.rs:19-41).
 *
 * Handles are not thread-safe: serialise calls on one handle. Distinct handles
 * are independent (each owns its own HIP stream and buffers).
 */
#ifndef PHY_FFI_H
#define PHY_FFI_H

#include <stdbool.h>
#include <stdint.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* phy_ffi.h:15-24 */
#define PHY_FFI_ABI_VERSION 2
#define PHY_FFI_VERSION_MAJOR 0
#define PHY_FFI_VERSION_MINOR 3
#define PHY_FFI_VERSION_PATCH 0

/* Opaque world handle (phy_ffi.h:28-30). */
typedef struct PhyWorldHandle PhyWorldHandle;

/* ---- reference ABI v2 ------------------------------------------------------ */

/* Returns PHY_FFI_ABI_VERSION (2). Replaces lib.rs:106-109 / phy_ffi.h:36. */
uint32_t phy_ffi_abi_version(void);

/* Dam-break fluid world: SphParams::defaults() and
 * fill_box((-4,-4,-1), (0,4,1), 0.18, 0.05) = 10,648 particles.
 * NULL on failure (e.g. no GPU; see phy_last_error). Replaces lib.rs:113-126,
 * 209-214 / phy_ffi.h:55. */
PhyWorldHandle *phy_world_create_fluid(void);

/* Non-fluid factories: outside this library (rigid, granular, coupled, soft,
 * field, optics, solid worlds). Exported so ABI v2 stays link-compatible; they
 * return NULL, the documented failure value (phy_ffi.h:57-68, lib.rs:216-328). */
PhyWorldHandle *phy_world_create_rigid(void);
PhyWorldHandle *phy_world_create_rigid_empty(void);
PhyWorldHandle *phy_world_create_granular(void);
PhyWorldHandle *phy_world_create_coupled(void);
PhyWorldHandle *phy_world_create_soft_empty(void);
PhyWorldHandle *phy_world_create_field_empty(void);
PhyWorldHandle *phy_world_create_optics_empty(void);
PhyWorldHandle *phy_world_create_solid_empty(void);

/* Step dt seconds: 0 ok, -1 null handle or internal error.
 * Replaces lib.rs:330-338 / phy_ffi.h:71. */
int32_t phy_world_step(PhyWorldHandle *w, double dt);

/* Step with the NaN/Inf watchdog: 0 ok, -1 null/error, 2 non-finite value
 * (world stops at that frame), 3 stalled (never produced), 4 other.
 * Replaces lib.rs:350-363 / phy_ffi.h:83. */
int32_t phy_world_step_checked(PhyWorldHandle *w, double dt);

/* Simulation time (f64 accumulation of dt); NaN for a null handle.
 * Replaces lib.rs:365-372 / phy_ffi.h:87. */
double phy_world_time(PhyWorldHandle *w);

/* Number of subsystems; 0 for null. Replaces lib.rs:374-378 / phy_ffi.h:90. */
uintptr_t phy_world_sub_count(PhyWorldHandle *w);

/* Total particles over all fluid subsystems; 0 for null.
 * Replaces lib.rs:380-397 / phy_ffi.h:93. */
uintptr_t phy_world_fluid_count(PhyWorldHandle *w);

/* Positions / velocities (x,y,z interleaved, index order) into buf of len
 * doubles; returns particles written, truncating at len/3.
 * Replaces lib.rs:399-460 / phy_ffi.h:97, 102. */
uintptr_t phy_world_get_fluid_positions(PhyWorldHandle *w, double *buf, uintptr_t len);
uintptr_t phy_world_get_fluid_velocities(PhyWorldHandle *w, double *buf, uintptr_t len);

/* Rigid-body readbacks: no rigid subsystem exists in this library; 0 bodies.
 * Replaces lib.rs:462-515 / phy_ffi.h:105, 109. */
uintptr_t phy_world_rigid_count(PhyWorldHandle *w);
uintptr_t phy_world_get_rigid_transforms(PhyWorldHandle *w, double *buf, uintptr_t len);

/* JSON save / load of the phy-io world archive (crates/phy-io/src/archive.rs:
 * version, time, tagged subsystems; f64 scalars as {"__f64": bits}, Vec3 as
 * u64 bit patterns). Save writes every fluid subsystem (other kinds are not
 * hosted here); load rebuilds the world and returns NULL on malformed input, a
 * non-finite scalar or a non-fluid subsystem. Positions and velocities load
 * rounded to FP32. 0 / -1 and handle / NULL. Replaces lib.rs:517-542 /
 * phy_ffi.h:112, 115. */
int32_t phy_world_save(PhyWorldHandle *w, const char *path);
PhyWorldHandle *phy_world_load(const char *path);

/* Free a handle; null-safe. A second destroy of the same non-null handle is a
 * caller error (as in the reference). Replaces lib.rs:544-553 / phy_ffi.h:118. */
void phy_world_destroy(PhyWorldHandle *w);

/* f32 channel. Replaces lib.rs:582-653, 655-689 / phy_ffi.h:129-141. */
int32_t phy_world_step_f32(PhyWorldHandle *w, float dt);
uintptr_t phy_world_get_fluid_positions_f32(PhyWorldHandle *w, float *buf, uintptr_t len);
uintptr_t phy_world_get_fluid_velocities_f32(PhyWorldHandle *w, float *buf, uintptr_t len);
uintptr_t phy_world_get_rigid_transforms_f32(PhyWorldHandle *w, float *buf, uintptr_t len);

/* Rigid-body control: no rigid subsystem here; -1 (documented failure).
 * Replaces lib.rs:712-899 / phy_ffi.h:149-175. */
int64_t phy_world_rigid_add_body(PhyWorldHandle *w, int32_t shape_kind, double mass,
                                 const double *pos7, const double *inertia3);
int32_t phy_world_rigid_apply_force(PhyWorldHandle *w, int64_t id, const double *f3, double dt,
                                    int32_t _mode);
int32_t phy_world_rigid_apply_torque(PhyWorldHandle *w, int64_t id, const double *t3, double dt,
                                     int32_t _mode);
int32_t phy_world_rigid_get_velocity(PhyWorldHandle *w, int64_t id, double *out3);
int32_t phy_world_rigid_get_angular_velocity(PhyWorldHandle *w, int64_t id, double *out3);

/* ---- additive (no counterpart in the reference) ------------------------------ */

#define PHY_MAX_MATERIALS 16

/* SphParams<f64> as a POD (crates/phy-fluid/src/sph/params.rs:12-50); the
 * material tables hold n_visc_k / n_visc_n valid entries. */
typedef struct PhySphParams {
    double rest_density;
    double stiffness;
    double viscosity;
    double visc_k[PHY_MAX_MATERIALS];
    double visc_n[PHY_MAX_MATERIALS];
    uint32_t n_visc_k;
    uint32_t n_visc_n;
    double shear_min;
    double mass;
    double h;
    double gravity[3];
    double bounds_min[3];
    double bounds_max[3];
    double boundary_damp;
    double xsph_eps;
} PhySphParams;

/* Particle fields for phy_fluid_get_field*: vector fields give 3 values per
 * particle, scalar fields 1. ACC, XSPH and MU_EFF hold the last step's values
 * and need phy_fluid_set_keep_aux(w, 1) before that step. */
enum {
    PHY_FIELD_POS = 0,
    PHY_FIELD_VEL = 1,
    PHY_FIELD_ACC = 2,
    PHY_FIELD_XSPH = 3,
    PHY_FIELD_RHO = 4,
    PHY_FIELD_P = 5,
    PHY_FIELD_MASS = 6,
    PHY_FIELD_MU_EFF = 7
};

/* Message of the last failed call on this thread ("" if none). */
const char *phy_last_error(void);

/* Fill out with SphParams::defaults() (params.rs:54-85). */
void phy_sph_params_defaults(PhySphParams *out);

/* World with one empty fluid subsystem of the given parameters; NULL on error. */
PhyWorldHandle *phy_world_create_fluid_with(const PhySphParams *params);

/* Replace / read the parameters of the first fluid subsystem. 0 ok, -1 error. */
int32_t phy_fluid_set_params(PhyWorldHandle *w, const PhySphParams *params);
int32_t phy_fluid_get_params(PhyWorldHandle *w, PhySphParams *out);

/* FluidWorld::fill_box / lattice_mass (world.rs:124-169) on the first fluid
 * subsystem. fill_box returns 0 ok / -1; lattice_mass NaN on error. */
int32_t phy_fluid_fill_box(PhyWorldHandle *w, const double *min3, const double *max3,
                           double spacing, double margin);
double phy_fluid_lattice_mass(PhyWorldHandle *w, double spacing);

/* Replace all particles of the first fluid subsystem (index order). vel, mass
 * and material may be NULL (zero velocity, SphParams::mass, material 0).
 * Values are rounded to FP32. At most 2^27 particles. 0 ok, -1 error. */
int32_t phy_fluid_set_particles(PhyWorldHandle *w, uintptr_t n, const double *pos3,
                                const double *vel3, const double *mass,
                                const uint32_t *material);
int32_t phy_fluid_set_particles_f32(PhyWorldHandle *w, uintptr_t n, const float *pos3,
                                    const float *vel3, const float *mass,
                                    const uint32_t *material);

/* Read a PHY_FIELD_* field in index order into buf (len values); returns the
 * particles written (truncating like the position readback), 0 on error. */
uintptr_t phy_fluid_get_field(PhyWorldHandle *w, int32_t field, double *buf, uintptr_t len);
uintptr_t phy_fluid_get_field_f32(PhyWorldHandle *w, int32_t field, float *buf, uintptr_t len);

/* Keep acc / xsph / mu_eff of each step (costs 36 B per particle-step of HBM
 * writes). 0 ok, -1 error. */
int32_t phy_fluid_set_keep_aux(PhyWorldHandle *w, int32_t on);

/* body_acc accumulator (3 per particle, index order), consumed by the next step
 * (world.rs:356, 361). 0 ok, -1 error. */
int32_t phy_fluid_set_body_acc(PhyWorldHandle *w, const double *b3, uintptr_t len);
/* The pending body_acc (3 per particle, index order; zero after a step). */
int32_t phy_fluid_get_body_acc(PhyWorldHandle *w, double *out3, uintptr_t len);

/* ---- Coupling (crates/phy-fluid/src/sph/world.rs:400-773) ----
 * The reference's FluidWorld::couple_bodies / couple_points / couple_heat,
 * called on the current state (typically after a step), with the bodies,
 * points and heat grid handed over as plain arrays: this library hosts no
 * rigid or heat subsystem, so FluidSubsystem::couple inside a World stays a
 * no-op. Particles change in place; rho, p and aux keep the last step's
 * values; a recording tape is invalidated. 0 ok, -1 error. */
#define PHY_SHAPE_SPHERE 0  /* dims[0] = radius */
#define PHY_SHAPE_BOX 1     /* dims = half extents */
#define PHY_SHAPE_CAPSULE 2 /* dims[0] = half height (along local y), dims[1] = radius */

/* Body<f64> as couple_bodies reads it (phy-rigid shape.rs:433-463). */
typedef struct PhyCoupleBody {
    double pose[7];   /* position xyz + orientation quaternion w i j k (body -> world),
                         the layout of phy_world_rigid_add_body's pos7 */
    double vel[3];    /* linear velocity; updated in place */
    double inv_mass;  /* <= 0: static (impenetrable, no buoyancy) */
    int32_t shape;    /* PHY_SHAPE_* */
    int32_t _pad;
    double dims[3];
} PhyCoupleBody;

/* Particles inside a body are pushed to its support point along their local
 * direction (world.rs:755-773); a dynamic body takes the contact impulses and
 * rho0 V g dt of buoyancy when any particle entered it (world.rs:414-503). */
int32_t phy_fluid_couple_bodies(PhyWorldHandle *w, PhyCoupleBody *bodies, uint32_t n, double dt,
                                double friction);

/* CouplePoint<f64> (sph/couple.rs:10-19). */
typedef struct PhyCouplePoint {
    double pos[3];
    double vel[3];
    double mass;
    double force[3];  /* out: buoyancy + drag; untouched outside the fluid bounds */
} PhyCouplePoint;

/* Points in order: mass-weighted fluid velocity of the 27 cells around each
 * point, force, and the reaction -force dt spread over those particles by
 * mass fraction (world.rs:691-752). */
int32_t phy_fluid_couple_points(PhyWorldHandle *w, PhyCouplePoint *pts, uint32_t n, double dt, double drag,
                                double soft_density);

/* A temperature grid (phy-field ScalarField: x fastest, idx = x + nx (y + ny z)). */
typedef struct PhyHeatGrid {
    uint32_t nx, ny, nz, _pad;
    double origin[3];
    double dx;
    const double *u;  /* temperatures, nx*ny*nz */
    double *src;      /* heat sources, nx*ny*nz, accumulated in place (NULL when heat_gain <= 0) */
} PhyHeatGrid;

/* body_acc += -g beta (T(x) - t_ref) with T trilinear (t_ref outside the grid)
 * for particles inside the fluid bounds; with heat_gain > 0, src[cell] +=
 * heat_gain |v| dt (world.rs:555-678). */
int32_t phy_fluid_couple_heat(PhyWorldHandle *w, const PhyHeatGrid *grid, double dt, double t_ref, double beta,
                              double heat_gain);

/* The velocity sampler couple_heat installs on the heat field (world.rs:614-643):
 * for each query point the velocity of the nearest particle of the current
 * state (the first in index order on ties; particle 0 when no distance is
 * finite; zero without particles). */
int32_t phy_fluid_sample_velocity(PhyWorldHandle *w, const double *q3, uint32_t n, double *out3);

/* FP64 device reductions over the current state: out5 = total mass, momentum
 * x, y, z (sum m v) and kinetic energy (sum |v|^2 m / 2) (subsystem.rs:149-163).
 * 0 ok, -1 error. */
int32_t phy_fluid_diagnostics_f64(PhyWorldHandle *w, double *out5);

/* As phy_fluid_diagnostics_f64 for one scene held in caller-owned DEVICE
 * arrays, allocating nothing (no counterpart in the reference): for callers
 * that cannot allocate device memory, such as code under HIP graph capture
 * and tight serving loops. posm4: n x (x, y, z, mass) floats, vel4: n x (vx,
 * vy, vz, unused) floats, the library's particle layout; out5: 5 doubles in
 * device memory (mass, momentum x y z, kinetic energy), written
 * asynchronously on `stream` (a hipStream_t; NULL for the null stream). The
 * sums are bit-identical to phy_fluid_diagnostics_f64 of a world holding the
 * same particles in the same order. At most 2^24 particles per call (the
 * per-tile FP64 partials live in a fixed 2.5 MiB device array that each GPU
 * holds until the process ends); larger n returns -1 without launching.
 * Calls that overlap on one GPU share that array and race: serialise them
 * (one stream, or events). 0 ok (launched), -1 error. */
int32_t phy_fluid_diagnostics_f64_noalloc(const float *posm4, const float *vel4, uint64_t n, double *out5,
                                          void *stream);

/* Density-predicate neighbour statistics of the current state (index order):
 * neighbour count (self included), sum of neighbour indices and xor of their
 * SplitMix64 hashes. For exact neighbour-set comparison. 0 ok, -1 error. */
int32_t phy_fluid_neighbor_stats(PhyWorldHandle *w, uint32_t *counts, uint64_t *sums,
                                 uint64_t *xors, uintptr_t len);

/* Neighbour statistics, in the layout of phy_fluid_neighbor_stats, taken by
 * the step's own gathers on the current state with the current tuning: pass 0
 * the density pass (j with r < h, self included; world.rs:220-230), pass 1 the
 * force pass (j != i with 0 < r < h, a NaN distance kept; world.rs:292-299).
 * These are the sets the density and force sums run over, for exact
 * comparison with the reference's. rho / p / aux of the last step are kept.
 * 0 ok, -1 error. */
int32_t phy_fluid_gather_stats(PhyWorldHandle *w, int32_t pass, uint32_t *counts, uint64_t *sums,
                               uint64_t *xors, uintptr_t len);

/* n steps of dt in one call (same result as n phy_world_step calls). */
int32_t phy_world_step_n(PhyWorldHandle *w, double dt, uint32_t n);

/* Reverse-mode gradients of a rollout (adjoint of FluidWorld::step).
 * phy_fluid_tape_begin(w, K) marks the current state as the rollout start;
 * every following phy_world_step* is recorded and the state is checkpointed
 * every K steps (memory: one 40 B/particle checkpoint per K steps plus K step
 * slots of ~120 B/particle). phy_fluid_tape_backward takes the cotangents of
 * the final positions and velocities (3 per particle) and of the density of
 * the last step (1 per particle), in index order, any of them NULL for zero,
 * and writes the cotangents of the initial positions / velocities (3 per
 * particle, may be NULL) and dparams[0] = dL/d stiffness, dparams[1] =
 * dL/d visc_k[0] (FP64 device reductions). Segments are recomputed from their
 * checkpoints; the world state is unchanged. Any material table and power-law
 * exponents; no body_acc. Returns 0 ok, -1 error. Replacing the particles
 * while recording invalidates the tape. */
int32_t phy_fluid_tape_begin(PhyWorldHandle *w, uint32_t checkpoint_every);
uint32_t phy_fluid_tape_steps(PhyWorldHandle *w);
int32_t phy_fluid_tape_backward(PhyWorldHandle *w, const double *xbar_final, const double *vbar_final,
                                const double *rhobar_last, double *xbar0, double *vbar0,
                                double *dparams);

/* Parameter gradients of the full material table per scene. */
#define PHY_TAPE_PARAMS (1 + 2 * PHY_MAX_MATERIALS)
/* As phy_fluid_tape_backward, with dparams holding PHY_TAPE_PARAMS values:
 * [0] dL/d stiffness, [1 + m] dL/d visc_k[m], [1 + PHY_MAX_MATERIALS + m]
 * dL/d visc_n[m] for m < PHY_MAX_MATERIALS (zero for m >= n_visc_k). The
 * visc_n gradient is defined for Newtonian materials too (visc_n = 1):
 * dmu/dn = mu ln(max(shear, shear_min)). */
int32_t phy_fluid_tape_backward_full(PhyWorldHandle *w, const double *xbar_final,
                                     const double *vbar_final, const double *rhobar_last,
                                     double *xbar0, double *vbar0, double *dparams);
int32_t phy_fluid_tape_end(PhyWorldHandle *w);

/* Steps are asynchronous on the world's HIP stream; block until the GPU has
 * finished every step issued so far. 0 ok, -1 error. */
int32_t phy_fluid_sync(PhyWorldHandle *w);

/* Per-stage timing with HIP events on the world's stream. Between begin and
 * end every step records its stages; end synchronises and writes the summed
 * milliseconds {sort, density, force, whole step} to ms4 and the number of
 * timed steps to *steps (may be NULL). 0 ok, -1 error. */
int32_t phy_fluid_profile_begin(PhyWorldHandle *w);
int32_t phy_fluid_profile_end(PhyWorldHandle *w, double *ms4, uint32_t *steps);

/* Capture the fluid step of this world in a hipGraph and replay it on every
 * following step (recaptured when dt, the particles or the parameters change).
 * Results are bit-identical to eager steps. 0 ok, -1 error. */
int32_t phy_fluid_set_graph(PhyWorldHandle *w, int32_t on);

/* Launch choices of the step: per-lane LDS neighbour-list length of the
 * density and of the force pass (8, 16 or 32; default 16 / 32), the launch
 * form (split = 1: each pass is one launch for the waves whose lanes share a
 * grid cell plus one for the others; 0: one launch), the neighbour grid's
 * cell size (cell_div = 2: cells of h / 2, a 5 x 5 x 5-cell window, the
 * default; 1: cells of h, the reference's 27 cells) and the gathers'
 * workgroup size (block = 128 lanes, the default, or 256). The force pass's compensated sums are blocked per list
 * drain, and the cell size orders the candidates, so force_list and cell_div
 * are part of the summation order: keep them fixed (for example through the
 * autotuner's cache) when runs must repeat bit for bit. The block size
 * changes no result. */
typedef struct PhyTuning {
    int32_t density_list;
    int32_t force_list;
    int32_t split;
    int32_t cell_div;
    int32_t block;
} PhyTuning;

int32_t phy_fluid_get_tuning(PhyWorldHandle *w, PhyTuning *out);
int32_t phy_fluid_set_tuning(PhyWorldHandle *w, const PhyTuning *t);

/* Autotuner. With cache_path naming a file that holds a choice for this
 * device and workload class (architecture, log2 particle count, log2 scene
 * count, non-Newtonian flag), that choice is applied and 1 is returned.
 * Otherwise `steps` steps of dt are timed from the current state for each
 * candidate: the cell size (h / 2 and h, whole steps), then at that cell size
 * the list lengths 8, 16, 32 in one and in two launches, then the workgroup
 * size (256 and 128 lanes); the fastest choice is applied (and stored in
 * cache_path when given), the state is restored, and 0 is returned. rho / p /
 * aux fields are undefined until the next step. out (may be NULL) receives
 * the applied choice; ms12 (may be NULL) the timed milliseconds per step of
 * the list trials: density for lists 8, 16, 32 in one launch, then in two,
 * then force in the same order. -1 on error. */
int32_t phy_fluid_autotune(PhyWorldHandle *w, double dt, uint32_t steps, const char *cache_path,
                           PhyTuning *out, double *ms12);

/* Flat export: FluidWorld::to_gpu_flat over Grid::to_flat (world.rs:782-859,
 * grid.rs:62-175, SphFlatData in sph/gpu_flat.rs:14-30), the reference's
 * hand-off format for GPU consumers, with the grid built from the current
 * positions (as after build_grid). Cells: the occupied key box padded to
 * [min - 1, max + 2] per axis, x fastest (c = i + nc0 (j + nc1 k) relative to
 * grid_min); cell_start: ncells + 1 prefix sums; sorted: particle indices,
 * ascending inside a cell. pos4 / vel4: (x, y, z, 0); scalar4: (rho, p, mass,
 * material); 4 floats per particle. info is always filled; the arrays are
 * written and 0 returned when n_cap >= n and cell_cap >= ncells + 1, else 1
 * (call again with larger buffers); -1 on error. An empty world has one cell
 * and cell_start = {0, 0}. */
typedef struct PhyFlatInfo {
    uint64_t n;
    uint64_t ncells;
    int64_t grid_min[3];
    uint64_t nc[3];
    float h, rest_density, stiffness, shear_min;
    float gravity[3];
    uint32_t n_visc;
    float visc_k[PHY_MAX_MATERIALS];
    float visc_n[PHY_MAX_MATERIALS];
} PhyFlatInfo;

int32_t phy_fluid_flat_export(PhyWorldHandle *w, PhyFlatInfo *info, float *pos4, float *vel4,
                              float *scalar4, uintptr_t n_cap, int32_t *cell_start,
                              uintptr_t cell_cap, int32_t *sorted);

/* Device memory of the current HIP device (hipMemGetInfo): free and total
 * bytes. 0 ok, -1 error. */
int32_t phy_device_mem_info(uint64_t *free_bytes, uint64_t *total_bytes);

/* Device memory held by this library's own buffers over every world and
 * batch of the process: bytes in use now and the high-water mark since the
 * last phy_device_mem_reset_peak (which sets it to the bytes in use). Either
 * pointer may be NULL. 0 ok, -1 error. */
int32_t phy_device_mem_stats(uint64_t *in_use, uint64_t *peak);
int32_t phy_device_mem_reset_peak(void);

/* ---- Batches of independent scenes (no counterpart in the reference) ----
 * A batch holds n_scenes independent FluidWorlds of uneven particle counts in
 * one set of device arrays: scene s owns particles [offsets[s], offsets[s+1])
 * in concatenated index order and follows FluidWorld::step with its own
 * SphParams (bounds, viscosity, ...). All scenes share dt; every stage of a
 * step is one launch over all scenes, and the step is replayed from a captured
 * hipGraph (phy_fluid_batch_set_graph(b, 0) steps eagerly). Each scene's
 * result is bit-identical to the same scene stepped alone. */
typedef struct PhyFluidBatch PhyFluidBatch;

/* params: n_scenes entries. NULL on error (phy_last_error). */
PhyFluidBatch *phy_fluid_batch_create(const PhySphParams *params, uint32_t n_scenes);
void phy_fluid_batch_destroy(PhyFluidBatch *b);
uint32_t phy_fluid_batch_scenes(PhyFluidBatch *b);
int32_t phy_fluid_batch_set_params(PhyFluidBatch *b, uint32_t scene, const PhySphParams *params);
int32_t phy_fluid_batch_get_params(PhyFluidBatch *b, uint32_t scene, PhySphParams *out);

/* Replace every scene's particles: counts[n_scenes], then FP32 arrays
 * concatenated in scene order (pos3 / vel3: 3 per particle; mass, material: 1;
 * vel3, mass, material may be NULL: zero, SphParams::mass, 0). 0 ok, -1. */
int32_t phy_fluid_batch_set_particles_f32(PhyFluidBatch *b, const uint32_t *counts,
                                          const float *pos3, const float *vel3,
                                          const float *mass, const uint32_t *material);
/* offsets: n_scenes + 1 entries (first particle of each scene, then the total). */
int32_t phy_fluid_batch_offsets(PhyFluidBatch *b, uint32_t *offsets);

/* n steps of dt for every scene. 0 ok, -1 error. */
int32_t phy_fluid_batch_step(PhyFluidBatch *b, double dt, uint32_t n);
/* One step, then the watchdog of every scene (FluidSubsystem::validate,
 * subsystem.rs:74-90): status[s] = 0 ok, 1 non-finite position, 2 non-finite
 * velocity. Returns the number of failing scenes, -1 on error. */
int32_t phy_fluid_batch_step_checked(PhyFluidBatch *b, double dt, int32_t *status);
int32_t phy_fluid_batch_set_graph(PhyFluidBatch *b, int32_t on);
int32_t phy_fluid_batch_set_keep_aux(PhyFluidBatch *b, int32_t on);
int32_t phy_fluid_batch_sync(PhyFluidBatch *b);

/* PHY_FIELD_* of every particle in concatenated index order; returns the
 * particles written (whole particles that fit in len), 0 on error. */
uintptr_t phy_fluid_batch_get_field_f32(PhyFluidBatch *b, int32_t field, float *buf, uintptr_t len);
/* out: 5 per scene (mass, momentum x y z, kinetic energy), FP64 on device. */
int32_t phy_fluid_batch_diagnostics_f64(PhyFluidBatch *b, double *out);
/* As phy_fluid_neighbor_stats, concatenated; neighbour indices are scene-local. */
int32_t phy_fluid_batch_neighbor_stats(PhyFluidBatch *b, uint32_t *counts, uint64_t *sums,
                                       uint64_t *xors, uintptr_t len);

/* As phy_fluid_gather_stats, concatenated; neighbour indices are scene-local. */
int32_t phy_fluid_batch_gather_stats(PhyFluidBatch *b, int32_t pass, uint32_t *counts, uint64_t *sums,
                                     uint64_t *xors, uintptr_t len);

/* Adjoint tape over every scene (as phy_fluid_tape_*, FP32 cotangents in
 * concatenated order); dparams: 2 per scene (dL/d stiffness, dL/d visc_k[0]). */
int32_t phy_fluid_batch_tape_begin(PhyFluidBatch *b, uint32_t checkpoint_every);
uint32_t phy_fluid_batch_tape_steps(PhyFluidBatch *b);
int32_t phy_fluid_batch_tape_backward(PhyFluidBatch *b, const float *xbar_final,
                                      const float *vbar_final, const float *rhobar_last,
                                      float *xbar0, float *vbar0, double *dparams);
/* As phy_fluid_batch_tape_backward with PHY_TAPE_PARAMS values per scene
 * (layout of phy_fluid_tape_backward_full). */
int32_t phy_fluid_batch_tape_backward_full(PhyFluidBatch *b, const float *xbar_final,
                                           const float *vbar_final, const float *rhobar_last,
                                           float *xbar0, float *vbar0, double *dparams);
int32_t phy_fluid_batch_tape_end(PhyFluidBatch *b);

/* As phy_fluid_get_tuning / set_tuning / autotune, for a whole batch. */
int32_t phy_fluid_batch_get_tuning(PhyFluidBatch *b, PhyTuning *out);
int32_t phy_fluid_batch_set_tuning(PhyFluidBatch *b, const PhyTuning *t);
int32_t phy_fluid_batch_autotune(PhyFluidBatch *b, double dt, uint32_t steps, const char *cache_path,
                                 PhyTuning *out, double *ms12);

/* As phy_fluid_profile_* (profiled steps run eagerly, not from the graph). */
int32_t phy_fluid_batch_profile_begin(PhyFluidBatch *b);
int32_t phy_fluid_batch_profile_end(PhyFluidBatch *b, double *ms4, uint32_t *steps);

#ifdef __cplusplus
}
#endif

#endif /* PHY_FFI_H */
