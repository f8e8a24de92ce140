// sph_kernels.h — host-side launch interface of sph_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_types.h"

namespace phy {

// spare entries after the sorted positions (read, never used, by the
// eight-candidate loads of the uniform scan)
constexpr int kScanPad = 8;
constexpr int kScanBlocks = 1024;  // fixed grid of the three-phase cell scan
constexpr int kForceList = 32;     // default per-lane LDS neighbour list length of k_force
constexpr int kDensityList = 16;   // default per-lane LDS neighbour list length of k_density
constexpr int kAdjList = 16;       // per-lane list length of the adjoint gathers

// Launch choices of the step, picked by the autotuner. The list length sets
// the LDS per workgroup (L x 4 block bytes, hence occupancy) against the drain
// count. The force pass's compensated sums are blocked per drain, so its list
// length is part of the summation order; the density pass's order does not
// depend on it. The cell size sets the neighbour search's candidate order
// (both passes' summation order) and its candidate count: cells of h / 2 give
// a 5 x 5 x 5-cell window (15.6 h^3 of candidate volume), cells of h the
// reference's 27 cells (27 h^3) in 9 rows instead of 25. The block size is
// the gathers' workgroup size; it changes no result.
struct Tuning {
    int density_list = kDensityList;  // 8, 16 or 32
    int force_list = kForceList;      // 8, 16 or 32
    // 1: each pass is two launches, one for waves whose lanes share a dense
    // cell (compiled without the per-lane scan, so with fewer registers and
    // more waves in flight) and one for the others; 0: one launch for all.
    int split = 0;
    int cell_div = 2;                 // cells per h along each axis: 2 (h / 2) or 1 (h)
    int block = 128;                  // gather workgroup: 128 or 256 lanes
};
bool tuning_valid(const Tuning& t);

struct SortBuffers {
    // inputs (current order)
    const float4* posm_in;
    const float4* vel_in;
    const uint32_t* id_in;
    const uint32_t* mat_in;   // may be null
    const float4* body_in;    // may be null
    // scratch
    uint32_t* cell;
    uint32_t* rank;
    uint32_t* counts;         // kept all-zero between steps
    uint32_t* partial;
    uint32_t* start;
    uint32_t* tmp;                 // particle of each scatter slot
    uint32_t* ctmp;                // its cell
    unsigned long long* keytmp;    // its within-cell order key (z bits, index)
    // outputs (cell order)
    float4* posm_out;
    float4* velm_out;
    uint32_t* id_out;
    uint32_t* mat_out;
    float4* body_out;
    uint32_t* cell_out;
    float4* posl_out;         // column-local coordinates of the uniform scan (may be null)
};

struct ForceArgs {
    const float4* posm;
    const float4* posl;       // column-local coordinates (k_reorder), kScanPad spare entries
    const float4* velm;
    const float* rho;
    const float* prs;
    const float* mp;
    // with every scene's masses equal (m_uni): the density pass's (x, y, z,
    // m p / rho^2) of each sorted slot, so a neighbour takes two loads (null:
    // x and m from posm, m p / rho^2 from mp)
    const float4* frec = nullptr;
    const uint32_t* cell;
    const uint32_t* start;
    const uint32_t* id;
    const uint32_t* mat;      // may be null when one material
    const float4* body;       // may be null
    float4* posm_out;         // new state, in this step's cell order
    float4* vel_out;
    uint32_t* id_out;         // original index of each new-state slot
    uint32_t* mat_out;        // material of each new-state slot (with mat)
    float4* acc_out;          // optional (null: not stored)
    float* mu_out;
    float4* xsph_out;         // w: wall-clamp flags (bit per axis)
    float4* fvisc_out;        // optional (adjoint tape): viscous force per unit mu, w: shear sum
};

// Outputs of the statistics form of the density / force gathers (original
// index order): per particle the count, the sum and the xor of SplitMix64
// hashes of the scene-local original indices of the neighbours the pass's
// exact test accepts.
struct GatherStats {
    const uint32_t* id;        // original index of each sorted slot
    uint32_t* cnt;
    unsigned long long* sum;
    unsigned long long* xr;
};

// One recorded step for the adjoint, in that step's cell order.
struct AdjSlot {
    const float4* posm;       // x, m
    const float4* velm;       // v, m / rho
    const float* rho;
    const float* prs;
    const uint32_t* cell;
    const uint32_t* start;
    const uint32_t* id;
    const float4* acc;
    const float4* xsph;       // w: wall-clamp flags
    const float4* fvisc;      // viscous force per unit mu, w: shear-rate sum
    const float* mu;          // effective viscosity of each particle
    const uint32_t* mat;      // material tags
    float dt;
};

// Parameter-gradient layout per scene: [0] dL/d stiffness, [1 + m] dL/d
// visc_k[m], [1 + kMaxMaterials + m] dL/d visc_n[m].
constexpr int kAdjParams = 1 + 2 * kMaxMaterials;

struct AdjBuffers {
    const float4* xbar_in;    // cotangents of the step's output x', v' (original order)
    const float4* vbar_in;
    float4* Fb;               // cell order: (F_bar, P), (s_bar, 1/rho), local terms
    float4* Sb;
    float4* xb_loc;
    float4* vb_loc;
    float* rhob_loc;
    float* rhob;
    float4* xb_acc;
    float4* xbar_out;         // cotangents of the step's input x, v (original order)
    float4* vbar_out;
    float2* mush;             // per particle: (mu, shear-rate cotangent)
    double* part;             // per-tile FP64 partials, kAdjParams per tile
    double* totals;           // running parameter gradients, kAdjParams per scene
};

// T.dom must be the domain array of the recorded step. shear: the force pass
// has non-Newtonian terms (the power law's shear-rate dependence).
void launch_adjoint_step(hipStream_t st, const SceneTable& T, const AdjSlot& S, const AdjBuffers& B,
                         const float* rhob_seed, bool shear);

// Every launch covers all scenes of T; T.dom / T.table_total belong to the
// step's buffer set. Diagnostics write 5 doubles per scene; partial holds
// 5 per tile.
void launch_domain_reset(hipStream_t st, Domain* dom, int nscenes);
void launch_domain(hipStream_t st, const SceneTable& T, const float4* posm);
void launch_sort(hipStream_t st, const SceneTable& T, const SortBuffers& b);
void launch_density(hipStream_t st, const SceneTable& T, const float4* posm, const float4* posl,
                    const uint32_t* cellv, const uint32_t* start, float* rho, float* prs,
                    float4* velm, float* mp, float4* frec, const Tuning& tu);
void launch_force(hipStream_t st, const SceneTable& T, float dt, const ForceArgs& a, bool shear,
                  const Tuning& tu);
// FP64 diagnostics of one scene held in caller-owned device arrays (posm:
// x, y, z, mass; vel: vx, vy, vz, unused), with no allocation: the per-tile
// partials live in a file-scope device array of kDiagNoAllocMaxTiles tiles.
// out5 (device) receives mass, momentum x y z, kinetic energy, bit-identical
// to launch_diag on a one-scene world holding the same arrays in the same
// order. Calls on one GPU share the array: overlapping calls race.
constexpr uint64_t kDiagNoAllocMaxParticles = 1ull << 24;
constexpr uint32_t kDiagNoAllocMaxTiles = (uint32_t)(kDiagNoAllocMaxParticles / 256);
constexpr int kMaxDevices = 64;
void launch_diag_noalloc(hipStream_t st, uint64_t n, const float4* posm, const float4* vel, double* out5);
void launch_diag(hipStream_t st, const SceneTable& T, const float4* posm, const float4* vel,
                 double* partial, double* out);
void launch_unsort4(hipStream_t st, int n, const uint32_t* id, const float4* in, float4* out);
void launch_unsort1(hipStream_t st, int n, const uint32_t* id, const float* in, float* out);
// The step's own density (pass 0) or force (pass 1) gather on a sorted state
// (a: the step's sorted buffers; rho / p / mp need not hold values: the
// accept tests read positions only), with statistics instead of outputs.
void launch_gather_stats(hipStream_t st, const SceneTable& T, int pass, const ForceArgs& a,
                         bool shear, const Tuning& tu, const GatherStats& gs);
void launch_neighbor_stats(hipStream_t st, const SceneTable& T, const float4* posm,
                           const uint32_t* cellv, const uint32_t* start, const uint32_t* id,
                           uint32_t* cnt, unsigned long long* sum, unsigned long long* xr);

}  // namespace phy
