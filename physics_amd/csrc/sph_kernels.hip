// sph_kernels.hip — gfx950 kernels of one FluidWorld::step
// (crates/phy-fluid/src/sph/world.rs:183-192):
//
//   k_aabb / k_domain_finalize   dense-grid domain from the finite cell keys
//   k_bin, k_scan_*, k_scatter   counting sort of particles into cells of size h / 2 (or h)
//   k_reorder                    deterministic within-cell rank + gather to cell order
//   k_density                    rho / p gather (world.rs:204-246)
//   k_force                      forces (world.rs:252-347) fused with integrate
//                                (352-363), enforce_bounds (366-397) and the
//                                validate watchdog (subsystem.rs:74-90)
//   k_diag_*                     FP64 mass / momentum / kinetic-energy reductions
//                                (subsystem.rs:149-163)
//
// Every float sum is a gather in a fixed order (a function of the state: the
// scan's row and candidate order over the cell order), compensated with
// Kahan's algorithm or summed in f64, so repeated runs are bit-identical. The only atomics are
// integer ones (cell counts, min/max of keys, watchdog indices) whose results do
// not depend on arrival order.
//
// Every kernel serves a batch of independent scenes in one launch: the
// per-particle kernels run one workgroup per 256-particle tile and look up the
// tile's scene (parameters, domain, cell-table offset) in a SceneTable; each
// scene has its own dense cell table inside one concatenated table, scanned in
// one pass. A single world is a batch of one scene.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "sph_types.h"
#include "sph_device.h"
#include "sph_couple.h"
#include "sph_kernels.h"
#include "sph_gathers.h"

namespace phy {

// ---------------------------------------------------------------------------
// domain

__global__ __launch_bounds__(256) void k_domain_reset(Domain* dom, int nscenes) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nscenes) return;
    Domain* d = dom + s;
    for (int x = 0; x < kXcds; ++x)
        for (int a = 0; a < 3; ++a) {
            d->amin[x][a] = 0x7fffffffffffffffll;
            d->amax[x][a] = (long long)0x8000000000000000ull;
        }
    d->bad_pos = 0xffffffffu;
    d->bad_vel = 0xffffffffu;
}

// Key AABB per scene. Each workgroup walks a contiguous range of tiles, four
// at a time with their loads issued together, and flushes its running min /
// max with one set of atomics per scene it touched.
__global__ __launch_bounds__(256) void k_aabb(SceneTable T, const float4* __restrict__ posm) {
    constexpr int U = 4;
    __shared__ long long sh[6][4];
    const int per = (T.ntiles + (int)gridDim.x - 1) / (int)gridDim.x;
    const int t0 = (int)blockIdx.x * per;
    const int t1 = min(T.ntiles, t0 + per);
    long long mn[3], mx[3];
    auto reset = [&]() {
        for (int a = 0; a < 3; ++a) {
            mn[a] = 0x7fffffffffffffffll;
            mx[a] = (long long)0x8000000000000000ull;
        }
    };
    auto flush = [&](int s) {  // block-uniform: every thread calls it
        for (int a = 0; a < 3; ++a) {
            mn[a] = wave_min_ll(mn[a]);
            mx[a] = wave_max_ll(mx[a]);
        }
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0)
            for (int a = 0; a < 3; ++a) {
                sh[a][w] = mn[a];
                sh[3 + a][w] = mx[a];
            }
        __syncthreads();
        if (threadIdx.x < 3) {
            const int a = threadIdx.x;
            long long lo = sh[a][0], hi = sh[3 + a][0];
            for (int k = 1; k < 4; ++k) {
                lo = sh[a][k] < lo ? sh[a][k] : lo;
                hi = sh[3 + a][k] > hi ? sh[3 + a][k] : hi;
            }
            if (lo <= hi) {
                const int c = (int)blockIdx.x % aabb_copies(T.nscenes);
                atomicMin(&T.dom[s].amin[c][a], lo);
                atomicMax(&T.dom[s].amax[c][a], hi);
            }
        }
        __syncthreads();
    };
    reset();
    int cur = -1;
    for (int t = t0; t < t1; t += U) {
        int sc[U];
        bool ok[U];
        float4 p[U];
        double inv[U], cdv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) sc[u] = t + u < t1 ? (int)T.tile_scene[t + u] : -1;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ok[u] = false;
            inv[u] = 0.0;
            cdv[u] = 1.0;
            p[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (sc[u] >= 0) {
                const SceneInfo* si = T.info + sc[u];
                const uint32_t local = (uint32_t)(t + u - (int)si->tile0) * (uint32_t)kTile + threadIdx.x;
                ok[u] = local < si->n;
                if (ok[u]) p[u] = posm[si->off + local];
                inv[u] = T.params[sc[u]].inv_cell;
                cdv[u] = (double)T.params[sc[u]].cdiv;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (sc[u] < 0) continue;
            if (sc[u] != cur) {
                if (cur >= 0) flush(cur);
                reset();
                cur = sc[u];
            }
            long long k[3];
            // cell keys; particles whose reference key fails go to the overflow segment
            if (ok[u] && (hkey((double)p[u].x * inv[u], cdv[u], k[0]) & hkey((double)p[u].y * inv[u], cdv[u], k[1]) &
                          hkey((double)p[u].z * inv[u], cdv[u], k[2]))) {
                for (int a = 0; a < 3; ++a) {
                    mn[a] = k[a] < mn[a] ? k[a] : mn[a];
                    mx[a] = k[a] > mx[a] ? k[a] : mx[a];
                }
            }
        }
    }
    if (cur >= 0) flush(cur);
}

// Turns one scene's key AABB into its dense table domain, clipped to the
// scene's table capacity. Particles outside the domain are binned to the
// scene's overflow cell.
__device__ void finalize_domain(Domain* d, unsigned long long cap, const long long* bk, int cdiv, int ncopies) {
    long long lo[3], hi[3];
    const long long* blo = bk;
    const long long* bhi = bk + 3;
    long long mn[3], mx[3];
    for (int a = 0; a < 3; ++a) {
        mn[a] = d->amin[0][a];
        mx[a] = d->amax[0][a];
        for (int x = 1; x < ncopies; ++x) {
            mn[a] = d->amin[x][a] < mn[a] ? d->amin[x][a] : mn[a];
            mx[a] = d->amax[x][a] > mx[a] ? d->amax[x][a] : mx[a];
        }
    }
    bool any = mn[0] <= mx[0];
    for (int a = 0; a < 3; ++a) {
        lo[a] = any ? mn[a] : blo[a];
        hi[a] = any ? mx[a] : blo[a];
    }
    // extents as unsigned differences: exact for any hi >= lo
    auto ext = [&](int a) { return ((unsigned long long)hi[a] - (unsigned long long)lo[a]) + 1ull; };
    auto vol = [&]() {
        double v = 1.0;
        for (int a = 0; a < 3; ++a) v *= (double)ext(a);
        return v;
    };
    if (vol() > (double)cap) {
        // keep the part inside the bounds box first
        for (int a = 0; a < 3; ++a) {
            lo[a] = lo[a] > blo[a] ? lo[a] : blo[a];
            hi[a] = hi[a] < bhi[a] ? hi[a] : bhi[a];
            if (hi[a] < lo[a]) hi[a] = lo[a];
        }
    }
    while (vol() > (double)cap) {
        // shrink the longest axis symmetrically
        int am = 0;
        for (int a = 1; a < 3; ++a)
            if (ext(a) > ext(am)) am = a;
        unsigned long long cut = ext(am) / 4 + 1;
        lo[am] = (long long)((unsigned long long)lo[am] + cut / 2);
        hi[am] = (long long)((unsigned long long)hi[am] - (cut - cut / 2));
        if (hi[am] < lo[am]) hi[am] = lo[am];
    }
    unsigned long long nc = 1;
    for (int a = 0; a < 3; ++a) {
        d->kmin[a] = lo[a];
        d->dims[a] = hi[a] - lo[a] + 1;
        nc *= (unsigned long long)d->dims[a];
    }
    d->ncells = (uint32_t)nc;
    d->reach = cdiv;
}

// One workgroup: every scene's domain, then the scenes' table offsets as an
// exclusive scan of (ncells + 1) in scene order, and the table's total size.
// It also resets what the next domain pass and this step's force pass
// accumulate into: the key AABB (consumed here) and the watchdog words (the
// host reads them after the step), so a step needs no reset launch.
__global__ __launch_bounds__(256) void k_domain_finalize(SceneTable T) {
    __shared__ uint32_t sh[4];
    uint32_t carry = 0;
    const int tid = threadIdx.x;
    for (int base = 0; base < T.nscenes; base += 256) {
        const int s = base + tid;
        uint32_t cells = 0;
        if (s < T.nscenes) {
            Domain* dm = T.dom + s;
            const int nc = aabb_copies(T.nscenes);
            finalize_domain(dm, T.info[s].cap, T.info[s].bkey, T.params[s].cdiv, nc);
            cells = dm->ncells + 1u;
            for (int x = 0; x < nc; ++x)
                for (int a = 0; a < 3; ++a) {
                    dm->amin[x][a] = 0x7fffffffffffffffll;
                    dm->amax[x][a] = (long long)0x8000000000000000ull;
                }
            dm->bad_pos = 0xffffffffu;
            dm->bad_vel = 0xffffffffu;
        }
        uint32_t x = cells;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if ((tid & 63) >= o) x += y;
        }
        const int w = tid >> 6;
        if ((tid & 63) == 63) sh[w] = x;
        __syncthreads();
        uint32_t before = 0;
        for (int k = 0; k < w; ++k) before += sh[k];
        if (s < T.nscenes) T.dom[s].toff = carry + before + x - cells;
        carry += sh[0] + sh[1] + sh[2] + sh[3];
        __syncthreads();
    }
    if (tid == 0) *T.table_total = carry;
}

// ---------------------------------------------------------------------------
// counting sort

// Particles arrive in the previous step's cell order, so equal cells form runs
// of lanes: one atomic per run (wave-aggregated), each lane's arrival rank is
// the run's base plus its position in the run. The arrival order is not
// deterministic; k_reorder replaces it by the (z, original index) order.
__global__ __launch_bounds__(256) void k_bin(SceneTable T, const float4* __restrict__ posm,
                                             uint32_t* cell, uint32_t* rank, uint32_t* counts) {
    const TileLoc L = tile_loc(T);
    const int s = L.i;
    const int lane = threadIdx.x & 63;
    const bool valid = L.act;
    const Domain d = T.dom[L.s];
    const double inv = T.params[L.s].inv_cell;
    uint32_t c = 0xffffffffu;
    if (valid) {
        float4 p = posm[s];
        c = d.toff + cell_of(p.x, p.y, p.z, inv, d);
    }
    const uint32_t prev = __shfl_up(c, 1, 64);
    const bool head = lane == 0 || prev != c;
    const unsigned long long heads = __ballot(head);
    const unsigned long long le = heads & (~0ull >> (63 - lane));     // heads at lanes <= lane
    const int first = 63 - __clzll((long long)le);                     // this run's first lane
    const unsigned long long above = heads & ~(~0ull >> (63 - first)); // heads after it
    const int next = above ? __ffsll((long long)above) - 1 : 64;
    uint32_t base = 0;
    if (head && valid) base = atomicAdd(&counts[c], (uint32_t)(next - first));
    base = __shfl(base, first, 64);
    if (valid) {
        cell[s] = c;
        rank[s] = base + (uint32_t)(lane - first);
    }
}

// Two-phase exclusive scan over the concatenated cell table (every scene's
// cells and overflow cell) with a fixed grid; the table size is read on the
// device. Phase 1 sums each chunk; phase 2 scans the chunks and also clears
// counts for the next step.
__device__ __forceinline__ void chunk_of(uint32_t total, int b, int nb, uint32_t& lo, uint32_t& hi) {
    uint32_t per = (total + nb - 1) / nb;
    lo = min(total, per * (uint32_t)b);
    hi = min(total, lo + per);
}

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    uint32_t t = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
    return t;
}

__global__ __launch_bounds__(256) void k_scan_partial(const uint32_t* __restrict__ counts,
                                                      const uint32_t* __restrict__ table_total,
                                                      uint32_t* partial) {
    __shared__ uint32_t sh[4];
    uint32_t total = *table_total;
    uint32_t lo, hi;
    chunk_of(total, blockIdx.x, gridDim.x, lo, hi);
    uint32_t v = 0;
    for (uint32_t c = lo + threadIdx.x; c < hi; c += blockDim.x) v += counts[c];
    v = block_sum_u32(v, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

// Phase 2: workgroup b scans chunk b from the sum of the chunk totals before
// it (at most kScanBlocks - 1 of them, summed by the workgroup itself), and
// the last workgroup writes the grand total after the table.
__global__ __launch_bounds__(256) void k_scan_final(uint32_t* counts,
                                                    const uint32_t* __restrict__ table_total,
                                                    const uint32_t* __restrict__ partial,
                                                    uint32_t* start) {
    __shared__ uint32_t sh[256];
    uint32_t total = *table_total;
    uint32_t lo, hi;
    chunk_of(total, blockIdx.x, gridDim.x, lo, hi);
    int t = threadIdx.x;
    uint32_t v0 = 0;
    for (int k = t; k < (int)blockIdx.x; k += 256) v0 += partial[k];
    uint32_t carry = block_sum_u32(v0, sh);
    for (uint32_t base = lo; base < hi; base += 256) {
        uint32_t c = base + t;
        uint32_t v = c < hi ? counts[c] : 0;
        // inclusive scan in the block
        uint32_t x = v;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o, 64);
            if ((t & 63) >= o) x += y;
        }
        int w = t >> 6;
        __syncthreads();
        if ((t & 63) == 63) sh[w] = x;
        __syncthreads();
        uint32_t wo = 0;
        for (int i = 0; i < w; ++i) wo += sh[i];
        uint32_t incl = x + wo;
        if (c < hi) {
            start[c] = carry + incl - v;
            counts[c] = 0;
        }
        uint32_t tot = sh[0] + sh[1] + sh[2] + sh[3];
        carry += tot;
        __syncthreads();
    }
    if (blockIdx.x == gridDim.x - 1 && t == 0) start[total] = carry;
}

// Within-cell order key of a dense-cell member: z as order-preserving bits
// (-0 folded onto +0) above the original index, so one unsigned compare is
// the float compare of z with index tie-break. Overflow members compare the
// low word (the index) only.
__device__ __forceinline__ unsigned long long zid_key(float z, uint32_t id) {
    uint32_t b = __float_as_uint(z);
    b = b == 0x80000000u ? 0u : b;
    b ^= (b & 0x80000000u) ? 0xffffffffu : 0x80000000u;
    return ((unsigned long long)b << 32) | id;
}

__global__ __launch_bounds__(256) void k_scatter(int n, const uint32_t* __restrict__ cell,
                                                 const uint32_t* __restrict__ rank,
                                                 const uint32_t* __restrict__ start,
                                                 const uint32_t* __restrict__ id,
                                                 const float4* __restrict__ posm, uint32_t* tmp,
                                                 uint32_t* ctmp, unsigned long long* keytmp) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t c = cell[s];
    const uint32_t q = start[c] + rank[s];
    tmp[q] = (uint32_t)s;
    ctmp[q] = c;
    keytmp[q] = zid_key(posm[s].z, id[s]);
}

// Slot q of the scatter holds some member of its cell in arrival order; its
// final slot is the cell start plus the number of members ordered before it.
// Dense cells order their members by z, ties by original index, so each z-run
// of cells is sorted by z (the uniform scan trims runs by z); the overflow
// segment keeps ascending original index (grid.rs:37-38). Either order is a
// function of the state alone. Slots of a scene stay inside the scene's
// particle range, so the tile names the scene; the stored cell index is the
// scene-local one. The cells the tile touches span the slot window
// [start of the first lane's cell, end of the last lane's cell); when that
// window holds at most kReorderWin slots its keys are staged in LDS and every
// lane counts from LDS. A longer window (a cell of more than ~1,800 members,
// such as a crowded overflow segment) is counted from global memory.
constexpr uint32_t kReorderWin = 2048;

__global__ __launch_bounds__(256) void k_reorder(
    SceneTable T, const uint32_t* __restrict__ tmp, const uint32_t* __restrict__ ctmp,
    const unsigned long long* __restrict__ keytmp, const uint32_t* __restrict__ start,
    const float4* __restrict__ posm_in, const float4* __restrict__ vel_in,
    const uint32_t* __restrict__ mat_in, const float4* __restrict__ body_in, float4* posm_out,
    float4* velm_out, uint32_t* id_out, uint32_t* mat_out, float4* body_out, uint32_t* cell_out,
    float4* posl_out) {
    __shared__ unsigned long long ksh[kReorderWin];
    __shared__ uint32_t wsh[2];
    const TileLoc L = tile_loc(T);
    const int tid = threadIdx.x;
    const SceneInfo* si = T.info + L.s;
    const uint32_t te = si->off + si->n;  // end of the scene's slots
    const uint32_t toff = T.dom[L.s].toff;
    const uint32_t ovf = toff + T.dom[L.s].ncells;
    uint32_t s = 0, c = 0, lo = 0, hi = 0;
    if (L.act) {
        s = tmp[L.i];
        c = ctmp[L.i];
        lo = start[c];
        hi = start[c + 1];
    }
    // lane 0 is active in every tile
    if (tid == 0) wsh[0] = lo;
    if (L.act && ((uint32_t)L.i + 1u == te || tid == kTile - 1)) wsh[1] = hi;
    __syncthreads();
    const uint32_t wlo = wsh[0], whi = wsh[1];
    const bool win = whi - wlo <= kReorderWin;
    if (win)
        for (uint32_t t = wlo + (uint32_t)tid; t < whi; t += kTile) ksh[t - wlo] = keytmp[t];
    __syncthreads();
    if (!L.act) return;
    uint32_t r = 0;
    unsigned long long mk;
    if (win) {
        mk = ksh[(uint32_t)L.i - wlo];
        if (c == ovf) {
            for (uint32_t t = lo; t < hi; ++t) r += (uint32_t)ksh[t - wlo] < (uint32_t)mk ? 1u : 0u;
        } else {
            for (uint32_t t = lo; t < hi; ++t) r += ksh[t - wlo] < mk ? 1u : 0u;
        }
    } else {
        mk = keytmp[L.i];
        if (c == ovf) {
            for (uint32_t t = lo; t < hi; ++t) r += (uint32_t)keytmp[t] < (uint32_t)mk ? 1u : 0u;
        } else {
            for (uint32_t t = lo; t < hi; ++t) r += keytmp[t] < mk ? 1u : 0u;
        }
    }
    const uint32_t my = (uint32_t)mk;
    uint32_t dst = lo + r;
    const float4 p = posm_in[s];
    posm_out[dst] = p;
    if (posl_out)
        posl_out[dst] = c == ovf ? make_float4(0.f, 0.f, 0.f, 0.f)
                                 : col_local(p, c - toff, T.dom[L.s], T.params[L.s].cell64);
    float4 v = vel_in[s];
    v.w = 0.f;
    velm_out[dst] = v;
    id_out[dst] = my;
    cell_out[dst] = c - toff;
    if (mat_in) mat_out[dst] = mat_in[s];
    if (body_in) body_out[dst] = body_in[s];
}

// ---------------------------------------------------------------------------
// FP64 diagnostics (subsystem.rs:149-163): mass, momentum, kinetic energy per
// scene. Tile partials by a fixed tree, then per scene a fixed-shape sum of its
// tiles' partials, so the result does not depend on scheduling.

__device__ __forceinline__ double block_tree_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// The five sums of one tile (256 particles, one per lane) into partial[t].
__device__ __forceinline__ void diag_tile(bool act, const float4& pm, const float4& v, size_t t,
                                          double* partial, double* sh) {
    double acc[5] = {0, 0, 0, 0, 0};
    if (act) {
        const double m = pm.w;
        const double vx = v.x, vy = v.y, vz = v.z;
        acc[0] = m;
        acc[1] = vx * m;
        acc[2] = vy * m;
        acc[3] = vz * m;
        acc[4] = (((vx * vx + vy * vy) + vz * vz) * m) / 2.0;
    }
    for (int k = 0; k < 5; ++k) {
        const double r = block_tree_sum(acc[k], sh);
        if (threadIdx.x == 0) partial[t * 5 + k] = r;
    }
}

// Tiles tile0 .. tile0 + ntiles - 1 of partial summed in a fixed shape into out[0..5).
__device__ __forceinline__ void diag_sum_tiles(const double* __restrict__ partial, uint32_t tile0,
                                               uint32_t ntiles, double* out, double* sh) {
    double acc[5] = {0, 0, 0, 0, 0};
    for (uint32_t t = threadIdx.x; t < ntiles; t += 256)
        for (int k = 0; k < 5; ++k) acc[k] += partial[(size_t)(tile0 + t) * 5 + k];
    for (int k = 0; k < 5; ++k) {
        const double r = block_tree_sum(acc[k], sh);
        if (threadIdx.x == 0) out[k] = r;
    }
}

__global__ __launch_bounds__(256) void k_diag_partial(SceneTable T, const float4* __restrict__ posm,
                                                      const float4* __restrict__ vel,
                                                      double* partial) {
    __shared__ double sh[256];
    const TileLoc L = tile_loc(T);
    float4 pm = make_float4(0.f, 0.f, 0.f, 0.f), v = pm;
    if (L.act) {
        pm = posm[L.i];
        v = vel[L.i];
    }
    diag_tile(L.act, pm, v, (size_t)L.t, partial, sh);
}

// one workgroup per scene
__global__ __launch_bounds__(256) void k_diag_final(SceneTable T, const double* __restrict__ partial,
                                                    double* out) {
    __shared__ double sh[256];
    const SceneInfo si = T.info[blockIdx.x];
    diag_sum_tiles(partial, si.tile0, si.ntiles, out + (size_t)blockIdx.x * 5, sh);
}

// The same two passes over caller-owned arrays of one scene (tile t = particles
// 256 t .. 256 t + 255), so a scene's sums are bit-identical to k_diag_*'s.
__global__ __launch_bounds__(256) void k_diag_partial_flat(uint32_t n, const float4* __restrict__ posm,
                                                           const float4* __restrict__ vel,
                                                           double* partial) {
    __shared__ double sh[256];
    const size_t t = blockIdx.x;
    const size_t i = t * 256 + threadIdx.x;
    const bool act = i < n;
    float4 pm = make_float4(0.f, 0.f, 0.f, 0.f), v = pm;
    if (act) {
        pm = posm[i];
        v = vel[i];
    }
    diag_tile(act, pm, v, t, partial, sh);
}

__global__ __launch_bounds__(256) void k_diag_final_flat(uint32_t ntiles, const double* __restrict__ partial,
                                                         double* out) {
    __shared__ double sh[256];
    diag_sum_tiles(partial, 0u, ntiles, out, sh);
}

// ---------------------------------------------------------------------------
// original-order gather for readback: out[id[s]] = in[s]

__global__ __launch_bounds__(256) void k_unsort4(int n, const uint32_t* __restrict__ id,
                                                 const float4* __restrict__ in, float4* out) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) out[id[s]] = in[s];
}
__global__ __launch_bounds__(256) void k_unsort1(int n, const uint32_t* __restrict__ id,
                                                 const float* __restrict__ in, float* out) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) out[id[s]] = in[s];
}

// neighbour statistics of the density predicate, for exact set comparison:
// count, sum of original indices, xor of SplitMix64-hashed original indices.
// Indices are scene-local (original index minus the scene's offset); the
// outputs are written in concatenated original order.
__global__ __launch_bounds__(256) void k_neighbor_stats(SceneTable T, const float4* __restrict__ posm,
                                                        const uint32_t* __restrict__ cellv,
                                                        const uint32_t* __restrict__ start_all,
                                                        const uint32_t* __restrict__ idv,
                                                        uint32_t* cnt_out,
                                                        unsigned long long* sum_out,
                                                        unsigned long long* xor_out) {
    const TileLoc L = tile_loc(T);
    if (!L.act) return;
    const int i = L.i;
    const Domain d = T.dom[L.s];
    const DevParams P = T.params[L.s];
    const uint32_t* __restrict__ start = start_all + d.toff;
    const uint32_t off = T.info[L.s].off;
    float4 pi = posm[i];
    uint32_t c = 0;
    unsigned long long s1 = 0, x1 = 0;
    for_each_candidate(cellv[i], pi.x, pi.y, pi.z, d, P.inv_cell, start, posm, [&](uint32_t j) {
        float4 pj = posm[j];
        float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        float s = dx * dx + dy * dy + dz * dz;
        if (classify(P, s, pi.x, pi.y, pi.z, pj.x, pj.y, pj.z) == 1) {
            unsigned long long id = idv[j] - off;
            ++c;
            s1 += id;
            unsigned long long z = id + 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            x1 ^= z ^ (z >> 31);
        }
    });
    uint32_t o = idv[i];
    cnt_out[o] = c;
    sum_out[o] = s1;
    xor_out[o] = x1;
}

// ---------------------------------------------------------------------------
// launch wrappers

static inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

void launch_domain_reset(hipStream_t st, Domain* dom, int nscenes) {
    if (nscenes > 0) hipLaunchKernelGGL(k_domain_reset, dim3(nblk(nscenes, 256)), dim3(256), 0, st, dom, nscenes);
}

// dom must hold reset AABB words (a new array through launch_domain_reset;
// afterwards k_domain_finalize resets them for the next pass)
void launch_domain(hipStream_t st, const SceneTable& T, const float4* posm) {
    if (T.ntiles > 0) {
        // about eight workgroups per scene, at least 512: more workgroups per
        // scene only add atomic traffic on that scene's domain
        int nb = max(512, 8 * T.nscenes);
        nb = T.ntiles < nb ? T.ntiles : nb;
        hipLaunchKernelGGL(k_aabb, dim3(nb), dim3(256), 0, st, T, posm);
    }
    hipLaunchKernelGGL(k_domain_finalize, dim3(1), dim3(256), 0, st, T);
}

void launch_sort(hipStream_t st, const SceneTable& T, const SortBuffers& b) {
    if (T.ntiles > 0)
        hipLaunchKernelGGL(k_bin, dim3(T.ntiles), dim3(256), 0, st, T, b.posm_in, b.cell, b.rank,
                           b.counts);
    hipLaunchKernelGGL(k_scan_partial, dim3(kScanBlocks), dim3(256), 0, st, b.counts, T.table_total,
                       b.partial);
    hipLaunchKernelGGL(k_scan_final, dim3(kScanBlocks), dim3(256), 0, st, b.counts, T.table_total,
                       b.partial, b.start);
    if (T.n > 0)
        hipLaunchKernelGGL(k_scatter, dim3(nblk(T.n, 256)), dim3(256), 0, st, T.n, b.cell, b.rank,
                           b.start, b.id_in, b.posm_in, b.tmp, b.ctmp, b.keytmp);
    if (T.ntiles > 0)
        hipLaunchKernelGGL(k_reorder, dim3(T.ntiles), dim3(256), 0, st, T, b.tmp, b.ctmp, b.keytmp,
                           b.start, b.posm_in, b.vel_in, b.mat_in, b.body_in, b.posm_out, b.velm_out,
                           b.id_out, b.mat_out, b.body_out, b.cell_out, b.posl_out);
}

void launch_exclusive_scan(hipStream_t st, uint32_t* counts, const uint32_t* total, uint32_t* partial,
                           uint32_t* start) {
    hipLaunchKernelGGL(k_scan_partial, dim3(kScanBlocks), dim3(256), 0, st, counts, total, partial);
    hipLaunchKernelGGL(k_scan_final, dim3(kScanBlocks), dim3(256), 0, st, counts, total, partial, start);
}

bool tuning_valid(const Tuning& t) {
    auto ok = [](int L) { return L == 8 || L == 16 || L == 32; };
    return ok(t.density_list) && ok(t.force_list) && (t.split == 0 || t.split == 1) &&
           (t.cell_div == 1 || t.cell_div == 2) && (t.block == 128 || t.block == 256);
}

void launch_density(hipStream_t st, const SceneTable& T, const float4* posm, const float4* posl,
                    const uint32_t* cellv, const uint32_t* start, float* rho, float* prs,
                    float4* velm, float* mp, float4* frec, const Tuning& tu) {
    density_any<false>(st, T, posm, posl, cellv, start, rho, prs, velm, mp, frec, tu, GatherStats{});
}

void launch_force(hipStream_t st, const SceneTable& T, float dt, const ForceArgs& a, bool shear,
                  const Tuning& tu) {
    force_any<false>(st, T, dt, a, shear, tu, GatherStats{});
}

void launch_diag(hipStream_t st, const SceneTable& T, const float4* posm, const float4* vel,
                 double* partial, double* out) {
    if (T.ntiles > 0)
        hipLaunchKernelGGL(k_diag_partial, dim3(T.ntiles), dim3(256), 0, st, T, posm, vel, partial);
    hipLaunchKernelGGL(k_diag_final, dim3(T.nscenes), dim3(256), 0, st, T, partial, out);
}

// ---------------------------------------------------------------------------
// FP64 diagnostics without allocation: the tile partials live in a file-scope
// device array instead of a buffer of the world. Each GPU holds its own copy;
// host code reaches it only through the address hipGetSymbolAddress returns
// for that device (the symbol's name is a host-side stand-in).

static __device__ __align__(256) double scratch[(size_t)kDiagNoAllocMaxTiles * 5];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}

// The scratch array's address on the current device, looked up once per
// device with that device current.
static double* scratch_on_current_device() {
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("diagnostics_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw std::runtime_error("diagnostics_noalloc: no address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    return static_cast<double*>(g_scratch_ptr[dev]);
}

void launch_diag_noalloc(hipStream_t st, uint64_t n, const float4* posm, const float4* vel, double* out5) {
    if (n > kDiagNoAllocMaxParticles)
        throw std::invalid_argument("diagnostics_noalloc: at most 2^24 particles per call");
    double* part = scratch_on_current_device();
    const uint32_t ntiles = (uint32_t)((n + 255) / 256);
    if (ntiles > 0)
        hipLaunchKernelGGL(k_diag_partial_flat, dim3(ntiles), dim3(256), 0, st, (uint32_t)n, posm, vel, part);
    hipLaunchKernelGGL(k_diag_final_flat, dim3(1), dim3(256), 0, st, ntiles, (const double*)part, out5);
    check(hipGetLastError(), "diagnostics_noalloc launch");
}

void launch_unsort4(hipStream_t st, int n, const uint32_t* id, const float4* in, float4* out) {
    hipLaunchKernelGGL(k_unsort4, dim3(nblk(n, 256)), dim3(256), 0, st, n, id, in, out);
}
void launch_unsort1(hipStream_t st, int n, const uint32_t* id, const float* in, float* out) {
    hipLaunchKernelGGL(k_unsort1, dim3(nblk(n, 256)), dim3(256), 0, st, n, id, in, out);
}

void launch_neighbor_stats(hipStream_t st, const SceneTable& T, const float4* posm,
                           const uint32_t* cellv, const uint32_t* start, const uint32_t* id,
                           uint32_t* cnt, unsigned long long* sum, unsigned long long* xr) {
    if (T.ntiles == 0) return;
    hipLaunchKernelGGL(k_neighbor_stats, dim3(T.ntiles), dim3(256), 0, st, T, posm, cellv, start, id,
                       cnt, sum, xr);
}

}  // namespace phy
