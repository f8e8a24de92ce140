// sph_types.h — plain-old-data shared by the host runtime and the gfx950 kernels.
//
// Particle state lives on the device in structure-of-arrays form, in cell order
// (particles sorted by their uniform-grid cell; inside a dense cell by z, ties
// by original index):
//   posm[s] = float4(x, y, z, mass)        16 B
//   velm[s] = float4(vx, vy, vz, m/rho)    16 B  (w filled by the density pass)
//   mp[s]   = mass * p / rho^2             4 B   (filled by the density pass)
//   rho[s], prs[s]                         4 B each
//   id[s]   = original particle index      4 B
//   mat[s]  = material tag                 4 B   (only read with >1 material)
// The reference keeps an AoS Vec<Particle<f64>> of 160 B per particle
// (crates/phy-fluid/src/particle.rs:17-46).
#pragma once
#include <stdint.h>

namespace phy {

constexpr int kMaxMaterials = 16;
// Particles per tile: the per-particle kernels run one 256-lane workgroup per
// tile, and a tile never spans two scenes.
constexpr int kTile = 256;
// Largest dense cell table (cells) one world may allocate; particles whose cell
// falls outside the table go to the overflow segment (exact, but O(n) per visit).
constexpr uint64_t kMaxDenseCells = 1ull << 26;

// Step constants in device-ready form. The neighbour predicate and the cell
// keys use the exact f64 support radius; the per-pair arithmetic runs in FP32
// with coefficients computed in f64 and rounded once.
struct DevParams {
    double h64;          // SphParams::h
    double inv_cell;     // 1.0 / h in f64 (grid.rs:27: T::one() / self.cell)
    double s_thresh;     // smallest f64 s with sqrt(s) >= h: r < h  <=>  s < s_thresh
    double h2_64, c6_64;  // f64 h*h and poly6 coefficient (density pass)
    double rest64, stiff64;
    // Largest |dz| of a pair with r < h whose x / y cell columns are less than
    // cdiv apart in both (zw[0]), cdiv apart in one (zw[1]) or in both axes
    // (zw[2]); rounded up (all three equal h for cells of h). The uniform scan
    // trims each z-run to the lanes' z-range widened by these.
    double zw[3];
    double cell64;       // edge of a dense-table cell: h / cdiv
    int32_t cdiv;        // cells per support radius along each axis: 1 (cells of h) or 2 (h / 2)
    int32_t pad0;
    float s_lo, s_hi;    // FP32 filter: s32 < s_lo => inside; s32 > s_hi => outside
    float s_hq;          // s_thresh (1 + 2^-16) rounded up: bound of the column-local filter
    float q_band;        // s_thresh 2^-19 rounded up: |q32| > q_band decides h^2 - s in FP32
    float m_uni;         // the scene's common particle mass, NaN when the masses differ
    float h, h2;         // FP32 support radius and its square
    float c6, cs, cv;    // poly6, spiky-gradient, viscosity-laplacian coefficients
    float rest, stiff, shear_min, r_eps;
    float g[3], lo[3], hi[3];
    float damp, xsph_eps;
    int32_t n_mat;       // number of material entries used (visc_k.len())
    int32_t newtonian;   // 1 when every visc_n entry is exactly 1 (mu_eff = visc_k)
    float visc_k[kMaxMaterials];
    float visc_n[kMaxMaterials];
};

// One independent scene of a batch (a single FluidWorld is a batch of one).
// Scenes occupy consecutive particle ranges [off, off + n) of the concatenated
// device arrays and consecutive tile ranges [tile0, tile0 + ntiles).
struct SceneInfo {
    uint32_t off, n;
    uint32_t tile0, ntiles;
    unsigned long long cap;  // dense cell-table capacity of the scene (cells)
    long long bkey[6];       // cell keys of the scene's bounds box: lo xyz, hi xyz
};

// Dense grid domain of one scene for one step. Keys are cell keys
// floor(cdiv fl(x (1/h))) with cdiv = 1 (the reference's i64 cell keys,
// grid.rs:26-33) or 2 (half cells, an exact refinement: scaling by 2 is exact,
// so reference key = half key >> 1); the table covers kmin .. kmin + dims - 1
// per axis with z the fastest-varying axis, so the z-adjacent cells of a
// neighbour row are one contiguous run of sorted particles. Local cell index
// ncells is the scene's overflow segment. The scene's cells 0 .. ncells sit at
// toff in the concatenated cell table of the step.
// XCDs of the MI355X (workgroups are dealt to them round-robin)
constexpr int kXcds = 8;

// AABB copies in use: one per XCD for up to kXcds scenes, else one.
__host__ __device__ inline int aabb_copies(int nscenes) { return nscenes <= kXcds ? kXcds : 1; }

struct Domain {
    long long kmin[3];
    long long dims[3];
    uint32_t ncells;
    uint32_t toff;
    // cells per support radius (cdiv of the scene's parameters): a dense
    // particle's candidates lie within +-reach cells per axis
    int32_t reach;
    int32_t pad0;
    // AABB of the finite keys of the current positions (filled by k_aabb).
    // With at most kXcds scenes, one copy per XCD (workgroup b runs on XCD b
    // mod 8): a workgroup's atomics share their words with an eighth of the
    // grid only; larger batches spread theirs over their scenes and use copy 0.
    long long amin[kXcds][3];
    long long amax[kXcds][3];
    // Watchdog (subsystem.rs:74-90): smallest particle index (concatenated
    // original order) with a non-finite position / velocity after the step;
    // 0xffffffff when none.
    uint32_t bad_pos;
    uint32_t bad_vel;
};

// Device view of a batch, passed by value to every per-particle kernel.
struct SceneTable {
    const SceneInfo* info;        // [nscenes]
    const uint32_t* tile_scene;   // [ntiles] scene of each tile
    const DevParams* params;      // [nscenes]
    Domain* dom;                  // [nscenes], per step-buffer set
    uint32_t* table_total;        // cells of this step's concatenated table (incl. overflow cells)
    int nscenes;
    int ntiles;
    int n;                        // particles over all scenes
};

}  // namespace phy
