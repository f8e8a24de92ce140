// sph_device.h — device helpers shared by the forward and adjoint kernels:
// reference cell keys, the exact neighbour predicate, the cell-table walk and
// the per-lane neighbour compaction (gather). The forward and adjoint passes
// use the same accept tests, so they visit identical neighbour sets in the
// identical order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include <type_traits>

#include "sph_types.h"

namespace phy {

// ---------------------------------------------------------------------------
// helpers

// num-traits ToPrimitive::to_i64 on the floored f64 key (grid.rs:29-31):
// NaN and values outside [-2^63, 2^63) become 0 (unwrap_or(0)).
__device__ __forceinline__ long long key1(float x, double inv) {
    double f = floor((double)x * inv);
    if (f >= -9223372036854775808.0 && f < 9223372036854775808.0) return (long long)f;
    return 0;
}

// The dense table bins particles into cells of size h / cdiv (cdiv 1 or 2):
// the cell key floor(cdiv fl(x (1/h))) is the reference key for cdiv = 1 and
// an exact refinement of it for cdiv = 2 (scaling by 2 is exact, so key1 =
// hkey >> 1). Returns false where the reference key fails (NaN, out of the i64
// range): such particles live in the overflow segment, which is matched by
// reference-key adjacency.
__device__ __forceinline__ bool hkey(double v, double cdiv, long long& k) {
    const double f = floor(v * cdiv);
    if (f >= -9223372036854775808.0 && f < 9223372036854775808.0 && v >= -9223372036854775808.0 &&
        v < 9223372036854775808.0) {
        k = (long long)f;
        return true;
    }
    k = 0;
    return false;
}

// A dense particle's candidates lie within +-reach cells per axis (reach =
// cdiv): a pair with r < h has |fl(x_j/h) - fl(x_i/h)| < 1 up to rounding,
// i.e. reference keys at most 1 apart and half keys at most 2 apart (see the
// neighbour-search notes of the design). The window has span = 2 reach + 1
// cells per row and span^2 rows (x, y) of z-runs.
__device__ __forceinline__ int win_span(const Domain& d) { return 2 * d.reach + 1; }

__device__ __forceinline__ bool finite3(float x, float y, float z) {
    return isfinite(x) && isfinite(y) && isfinite(z);
}

// Exact f64 distance predicate of the reference, r = sqrt((dx*dx + dy*dy) + dz*dz)
// in f64 on the FP32 coordinates (world.rs:226, 295-296; nalgebra norm order).
// Contraction is switched off so the f64 sum rounds exactly like the oracle's.
__device__ __attribute__((noinline)) double exact_s64(float xi, float yi, float zi, float xj,
                                                      float yj, float zj) {
#pragma clang fp contract(off)
    double dx = (double)xi - (double)xj;
    double dy = (double)yi - (double)yj;
    double dz = (double)zi - (double)zj;
    return (dx * dx + dy * dy) + dz * dz;
}

// Classifies a candidate pair from its FP32 squared distance s32 (any rounding,
// FMA allowed): returns 1 inside (r < h), 0 outside, 2 for NaN distance.
// |s32 - s64| <= 6u * s, so the filter margins (2^-19 relative) are safe; the
// band in between is decided by the exact f64 expression.
__device__ __forceinline__ int classify(const DevParams& P, float s32, float xi, float yi, float zi,
                                        float xj, float yj, float zj) {
    if (s32 < P.s_lo) return 1;
    if (s32 > P.s_hi) return 0;
    double s = exact_s64(xi, yi, zi, xj, yj, zj);
    if (s < P.s_thresh) return 1;
    if (s != s) return 2;
    return 0;
}

__device__ __forceinline__ void kahan_add(float& sum, float& c, float x) {
    float y = x - c;
    float t = sum + y;
    c = (t - sum) - y;
    sum = t;
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap so consecutive
// logical blocks (spatially adjacent cells) share one XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    static_assert(kXcds == 8, "the remap divides by shifts");
    int per = nb >> 3, rem = nb & 7;
    int x = b & 7, k = b >> 3;
    return x * per + min(x, rem) + k;
}

// This workgroup's tile, its scene and this lane's particle slot. The tile
// index is XCD-remapped (xcd_block), so consecutive tiles of a scene share an
// L2. Lanes past the end of the scene are inactive and must not touch i.
struct TileLoc {
    int t;       // tile
    uint32_t s;  // scene
    int i;       // particle slot in the concatenated arrays
    bool act;
};

__device__ __forceinline__ TileLoc tile_loc(const SceneTable& T) {
    TileLoc L;
    L.t = xcd_block(blockIdx.x, gridDim.x);
    L.s = T.tile_scene[L.t];
    const SceneInfo* si = T.info + L.s;
    const uint32_t local = (uint32_t)(L.t - (int)si->tile0) * (uint32_t)kTile + threadIdx.x;
    L.act = local < si->n;
    L.i = (int)(si->off + local);
    return L;
}

// tile_loc for workgroups of BS lanes (kTile / BS workgroups per tile; the
// XCD remap keeps a tile's workgroups, and consecutive tiles, on one XCD).
template <int BS>
__device__ __forceinline__ TileLoc tile_loc_bs(const SceneTable& T) {
    static_assert(BS == 128 || BS == 256, "workgroups of 128 or 256 lanes");
    if (BS == kTile) return tile_loc(T);
    constexpr int sub = kTile / BS;
    TileLoc L;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    L.t = b / sub;
    L.s = T.tile_scene[L.t];
    const SceneInfo* si = T.info + L.s;
    const uint32_t local = (uint32_t)(L.t - (int)si->tile0) * (uint32_t)kTile + (uint32_t)(b % sub) * BS +
                           threadIdx.x;
    L.act = local < si->n;
    L.i = (int)(si->off + local);
    return L;
}

__device__ __forceinline__ long long wave_min_ll(long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        long long w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ long long wave_max_ll(long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

__device__ __forceinline__ uint32_t cell_of(float x, float y, float z, double inv, const Domain& d) {
    long long h0, h1, h2;
    const double cd = (double)d.reach;
    const bool ok = hkey((double)x * inv, cd, h0) & hkey((double)y * inv, cd, h1) & hkey((double)z * inv, cd, h2);
    // wrapping differences: an out-of-domain key maps to a huge unsigned value
    const unsigned long long k0 = (unsigned long long)h0 - (unsigned long long)d.kmin[0];
    const unsigned long long k1 = (unsigned long long)h1 - (unsigned long long)d.kmin[1];
    const unsigned long long k2 = (unsigned long long)h2 - (unsigned long long)d.kmin[2];
    if (!ok || k0 >= (unsigned long long)d.dims[0] || k1 >= (unsigned long long)d.dims[1] ||
        k2 >= (unsigned long long)d.dims[2])
        return d.ncells;
    return (uint32_t)((k0 * d.dims[1] + k1) * d.dims[2] + k2);
}

// Cell-key interval (table-relative, per axis) of the 27 reference cells
// around reference key k: the reference keys k - 1, k, k + 1 (wrapping, like
// grid.rs:44-48) that fall inside the table, as cells of the table (a
// reference cell is 1 or 2 table cells per axis). Empty: lo > hi.
__device__ __forceinline__ void ref_block(long long k, const Domain& d, int a, long long& lo,
                                          long long& hi) {
    const int sh = d.reach - 1;  // cell key >> sh = reference key
    const long long hlo = d.kmin[a] >> sh, hhi = (d.kmin[a] + d.dims[a] - 1) >> sh;
    const unsigned long long span = (unsigned long long)(hhi - hlo);
    long long kmn = 0, kmx = -1;
    bool any = false;
    for (int dk = -1; dk <= 1; ++dk) {
        const long long K = (long long)((unsigned long long)k + (unsigned long long)(long long)dk);
        if ((unsigned long long)K - (unsigned long long)hlo <= span) {
            kmn = any ? (K < kmn ? K : kmn) : K;
            kmx = any ? (K > kmx ? K : kmx) : K;
            any = true;
        }
    }
    if (!any) {
        lo = 0;
        hi = -1;
        return;
    }
    const long long c0 = (long long)((unsigned long long)kmn << sh),
                    c1 = (long long)((unsigned long long)(kmx + 1) << sh) - 1;
    lo = (c0 > d.kmin[a] ? c0 : d.kmin[a]) - d.kmin[a];
    hi = (c1 < d.kmin[a] + d.dims[a] - 1 ? c1 : d.kmin[a] + d.dims[a] - 1) - d.kmin[a];
}

// ---------------------------------------------------------------------------
// neighbour visiting
//
// Rows (dx, dy) of the cell window with the z-adjacent cells of each row as
// one contiguous run [start[cz - reach], start[cz + reach + 1]).

struct CellCoord {
    long long cx, cy, cz;
};

struct Quad4 {
    float4 p[4];
};

// Coordinates of dense cell c < ncells. The table holds fewer than 2^32 cells,
// so every extent fits 32 bits and 32-bit division suffices.
__device__ __forceinline__ CellCoord decode(uint32_t c, const Domain& d) {
    CellCoord r;
    const uint32_t dz = (uint32_t)d.dims[2], dy = (uint32_t)d.dims[1];
    const uint32_t t = c / dz;
    const uint32_t x = t / dy;
    r.cz = (long long)(c - t * dz);
    r.cy = (long long)(t - x * dy);
    r.cx = (long long)x;
    return r;
}

// Run of row k (0..span^2-1) of the window of an in-domain particle: cells
// (cx + dx, cy + dy, cz - reach .. cz + reach), contiguous in z; empty when
// the row leaves the domain (those cells hold no dense particles).
__device__ __forceinline__ void row_run(int k, const CellCoord& cc, const Domain& d,
                                        const uint32_t* __restrict__ start, uint32_t& a,
                                        uint32_t& b) {
    const int R = d.reach, span = 2 * R + 1;
    long long x = cc.cx + (k / span) - R;
    long long y = cc.cy + (k % span) - R;
    if (x < 0 || x >= d.dims[0] || y < 0 || y >= d.dims[1]) {
        a = b = 0;
        return;
    }
    long long z0 = cc.cz > R ? cc.cz - R : 0;
    long long z1 = cc.cz + R < d.dims[2] ? cc.cz + R : d.dims[2] - 1;
    long long base = (x * d.dims[1] + y) * d.dims[2];
    a = start[base + z0];
    b = start[base + z1 + 1];
}

// Cells per axis of the reference block of an out-of-domain particle.
__device__ __forceinline__ int block_span(const Domain& d) { return 3 * d.reach; }

// Run of row k (0..block_span^2-1) of the reference block of an out-of-domain
// particle (intervals from ref_block); empty rows have a == b.
__device__ __forceinline__ void block_run(int k, const long long* lo, const long long* hi,
                                          const Domain& d, const uint32_t* __restrict__ start,
                                          uint32_t& a, uint32_t& b) {
    const int bs = block_span(d);
    const long long x = lo[0] + k / bs, y = lo[1] + k % bs;
    if (x > hi[0] || y > hi[1] || lo[2] > hi[2]) {
        a = b = 0;
        return;
    }
    const long long base = (x * d.dims[1] + y) * d.dims[2];
    a = start[base + lo[2]];
    b = start[base + hi[2] + 1];
}

// Reference cell adjacency on full i64 keys with wrapping offsets (grid.rs:44-48).
__device__ __forceinline__ bool keys_adjacent(float xi, float yi, float zi, float xj, float yj,
                                              float zj, double inv) {
    unsigned long long a = (unsigned long long)key1(xj, inv) - (unsigned long long)key1(xi, inv) + 1ull;
    unsigned long long b = (unsigned long long)key1(yj, inv) - (unsigned long long)key1(yi, inv) + 1ull;
    unsigned long long c = (unsigned long long)key1(zj, inv) - (unsigned long long)key1(zi, inv) + 1ull;
    return a <= 2ull && b <= 2ull && c <= 2ull;
}

// Visits every candidate j of particle (xi, yi, zi) in the reference order and
// calls f(j). Dense cells come from the table; the overflow segment (particles
// binned outside the dense domain, e.g. NaN keys) is checked by key adjacency.
template <class F>
__device__ __forceinline__ void for_each_candidate(uint32_t ci, float xi, float yi, float zi,
                                                   const Domain& d, double inv,
                                                   const uint32_t* __restrict__ start,
                                                   const float4* __restrict__ posm, F&& f) {
    if (ci < d.ncells) {
        CellCoord cc = decode(ci, d);
        const int rows = win_span(d) * win_span(d);
        for (int k = 0; k < rows; ++k) {
            uint32_t a, b;
            row_run(k, cc, d, start, a, b);
            for (uint32_t j = a; j < b; ++j) f(j);
        }
    } else {
        // out-of-domain particle: the in-domain cells of its 27 reference cells
        long long lo[3], hi[3];
        ref_block(key1(xi, inv), d, 0, lo[0], hi[0]);
        ref_block(key1(yi, inv), d, 1, lo[1], hi[1]);
        ref_block(key1(zi, inv), d, 2, lo[2], hi[2]);
        const int rows = block_span(d) * block_span(d);
        for (int k = 0; k < rows; ++k) {
            uint32_t a, b;
            block_run(k, lo, hi, d, start, a, b);
            for (uint32_t j = a; j < b; ++j) f(j);
        }
    }
    uint32_t o0 = start[d.ncells], o1 = start[d.ncells + 1];
    for (uint32_t j = o0; j < o1; ++j) {
        float4 q = posm[j];
        if (keys_adjacent(xi, yi, zi, q.x, q.y, q.z, inv)) f(j);
    }
}

// ---------------------------------------------------------------------------
// neighbour compaction
//
// Candidates are tested with the cheap FP32 filter (exact f64 fallback in the
// band) and accepted ones are appended to a per-lane list in LDS; the wave
// drains the lists whenever one is full, so the heavy per-neighbour work runs
// on dense lists instead of on every candidate under divergence (about 25 % of
// the half-cell window's candidates are neighbours at rest, ~50 % in a
// collapsed sheet).
//
// Two scans produce the same lists in the same order:
//  * uniform: every active lane of the wave sits in the same dense cell (the
//    common case once the column has collapsed, ~100 particles per half cell). The
//    candidate sequence is then wave-uniform: loop control runs on the scalar
//    unit and candidate loads are uniform.
//  * per-lane: each lane has its own window of rows; the wave walks the rows
//    together, each lane loading its own z-run of the row, six candidates of
//    it per round.


__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// true when the condition holds on any active lane; the lane mask stays in
// scalar registers (no vector-register round trip)
__device__ __forceinline__ bool wave_any(bool b) { return __builtin_amdgcn_ballot_w64(b) != 0ull; }

// Element idx of an array with a wave-uniform base: a 32-bit byte offset from
// the base (arrays here hold < 2^27 elements of <= 16 B), so the load can use
// the scalar-base + 32-bit-offset address form instead of a 64-bit per-lane
// address.
template <class T>
__device__ __forceinline__ T ldg(const T* __restrict__ base, uint32_t idx) {
    return *(const T*)((const char*)base + (idx * (uint32_t)sizeof(T)));
}

// Element o / 16 of an array, o the byte offset of a 16-byte record (the
// neighbour lists hold these offsets, so 16-byte loads need no shift).
template <class T>
__device__ __forceinline__ T at(const T* __restrict__ base, uint32_t o) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "4, 8 or 16-byte elements");
    constexpr uint32_t sh = sizeof(T) == 16 ? 0u : (sizeof(T) == 8 ? 1u : 2u);
    return *(const T*)((const char*)base + (o >> sh));
}

// test(j, pj) -> bool, a conservative FP32 filter (every true neighbour
// passes; the exact decision is proc's); load(o) -> T and proc(o, T, listed, scan)
// take the byte offset o = 16 j of candidate j (proc accumulates; a slot that
// is not listed, or that proc's exact test rejects, must contribute exactly
// nothing); flush(scan) after every drain; scan is std::true_type in the
// uniform scan's drains and std::false_type in the per-lane scan's, so a proc
// can keep state that only one of them uses. self_j is a valid index: every list
// slot starts as its offset, and every slot written later holds a candidate
// or an entry at most seven past a run's end (inside the array's kScanPad
// spare entries), so drain slots beyond a lane's count read in-bounds records. With
// PAD those slots hold self_j itself: a proc that multiplies its zero weights
// into the neighbour's data needs finite data there (a stale slot may name a
// NaN particle); a proc that selects its terms does not. Overflow-segment
// candidates additionally need reference key adjacency. Appends are
// branch-free: every candidate writes the list head and the head advances
// only when the filter passes (with PAD a drain first writes self_j's offset
// at the head of a list that is not full);
// with PAD each drained slot is reset to self_j. again() -> bool, called once
// a uniform scan has drained its last slot (wave-uniform result): true scans
// the shared window again from the start (the caller resets its sums).
// Which waves a gather launch serves: every wave, only waves whose active
// lanes all sit in one dense cell (uniform scan), or only the others (per-lane
// scan). The two-launch tuning launches the last two separately, so the
// uniform kernel does not carry the per-lane scan's registers.
enum GatherMode { kAllWaves = 0, kUniformWaves = 1, kOtherWaves = 2 };

// The wave's active lanes all sit in dense half cells of one (x, y) column,
// within kColSpan consecutive z cells [c0, c1]: their candidates are then
// covered by one shared window (rows of z-runs cz(c0) - reach .. cz(c1) +
// reach), scanned with wave-uniform loads. The window is larger than a
// lane's own by at most kColSpan - 1 cells per row; every extra candidate is
// still tested with the exact predicate.
constexpr uint32_t kColSpan = 3;

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

struct WaveCol {
    bool ok;
    uint32_t c0, c1;  // lowest and highest cell of the wave's lanes
};

__device__ __forceinline__ WaveCol wave_column(bool act, uint32_t ci, const Domain& d) {
    WaveCol w;
    w.c0 = uni(wave_min_u32(act ? ci : 0xffffffffu));
    w.c1 = uni(wave_max_u32(act ? ci : 0u));
    const uint32_t dz = (uint32_t)d.dims[2];
    w.ok = w.c0 <= w.c1 && w.c1 < d.ncells && w.c1 - w.c0 < kColSpan && w.c0 / dz == w.c1 / dz;
    return w;
}

__device__ __forceinline__ bool wave_uniform(bool act, uint32_t ci, const Domain& d) {
    return wave_column(act, ci, d).ok;
}

// z half-widths of the trimmed rows of a uniform scan (DevParams::zw). Off
// for domains with keys of magnitude 2^40 and more: the column gap that zw[1]
// and zw[2] assume holds for smaller coordinates.
struct ZWin {
    double w0, w1, w2;
    bool on;
};

__device__ __forceinline__ ZWin zwin(const DevParams& P, const Domain& d) {
    ZWin z{P.zw[0], P.zw[1], P.zw[2], true};
    for (int a = 0; a < 3; ++a)
        z.on = z.on && d.kmin[a] > -(1ll << 40) && d.kmin[a] < (1ll << 40) &&
               d.dims[a] < (1ll << 40) && d.kmin[a] + d.dims[a] < (1ll << 40);
    return z;
}

__device__ __forceinline__ float wave_min_f(float v) {
    for (int o = 32; o > 0; o >>= 1) {
        const float w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

__device__ __forceinline__ float wave_max_f(float v) {
    for (int o = 32; o > 0; o >>= 1) {
        const float w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// FP32 bounds that contain [lo64, hi64] (rounded outward)
__device__ __forceinline__ float f32_down(double v) {
    float f = (float)v;
    return (double)f > v ? nextafterf(f, -INFINITY) : f;
}
__device__ __forceinline__ float f32_up(double v) {
    float f = (float)v;
    return (double)f < v ? nextafterf(f, INFINITY) : f;
}

// Trims the z-run [a, b) of a uniform scan to the candidates with z in
// [zlo, zhi]. Inside a dense cell the members are ordered by z, and the cells
// of a run follow in z order (keys are monotone in z), so the run is sorted by
// z: the candidates below zlo form a prefix and those above zhi a suffix,
// found 64 at a time with one coalesced load and a ballot. ZW: z is the w
// component of src (column-local coordinates), else its z component.
template <bool ZW>
__device__ __forceinline__ void trim_run(const float4* __restrict__ src, float zlo, float zhi,
                                         uint32_t& a, uint32_t& b) {
    const uint32_t lane = threadIdx.x & 63u;
    auto zof = [&](uint32_t j) { return ZW ? src[j].w : src[j].z; };
    for (;;) {
        const uint32_t j = a + lane;
        const bool below = j < b && zof(j) < zlo;
        const uint32_t nb = (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(below));
        a += nb;
        if (nb < 64u) break;
    }
    for (;;) {
        const uint32_t j = b - 1u - lane;
        const bool above = lane < b - a && zof(j) > zhi;
        const uint32_t na = (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(above));
        b -= na;
        if (na < 64u) break;
    }
}

// Column-local coordinates of a particle in dense cell c: (x - X, y - Y,
// (x - X)^2 + (y - Y)^2, z), X and Y the lower corner of its cell column.
// The uniform scan filters with them: for a candidate of row (ex, ey) and a
// lane with local (u, v), dx = x_j - (u - ex cell) and likewise dy, so
// dx^2 + dy^2 = n_j - 2 xr x_j - 2 yr y_j + (xr^2 + yr^2): two FMAs per
// candidate against per-row lane constants, all terms of size h^2 (no
// cancellation of large coordinates).
__device__ __forceinline__ float4 col_local(const float4& p, uint32_t c, const Domain& d, double cell) {
    const CellCoord cc = decode(c, d);
    const double X = (double)(d.kmin[0] + cc.cx) * cell, Y = (double)(d.kmin[1] + cc.cy) * cell;
    const float xl = (float)((double)p.x - X), yl = (float)((double)p.y - Y);
    return make_float4(xl, yl, fmaf(yl, yl, xl * xl), p.z);
}

// The column-local filter of a uniform scan (posl: column-local coordinates
// of the sorted particles, kScanPad spare entries; li: the lane's own; hh:
// the cell edge h / cdiv; s_hq: the bound, NaN to pass everything).
// Conservative: its FP32 rounding is below 1e-5 h^2 while s_hq sits s_thresh
// 2^-16 above s_thresh.
struct ColLocal {
    const float4* __restrict__ posl;
    float4 li;
    float hh, s_hq;
};

// Run of row (ex, ey) of a column wave's shared window: z cells cz0 - reach
// .. cz1 + reach of column (cx + ex, cy + ey).
__device__ __forceinline__ void col_run(int ex, int ey, const CellCoord& cc, long long cz1, const Domain& d,
                                        const uint32_t* __restrict__ start, uint32_t& a,
                                        uint32_t& b) {
    const int R = d.reach;
    long long x = cc.cx + ex;
    long long y = cc.cy + ey;
    if (x < 0 || x >= d.dims[0] || y < 0 || y >= d.dims[1]) {
        a = b = 0;
        return;
    }
    long long z0 = cc.cz > R ? cc.cz - R : 0;
    long long z1 = cz1 + R < d.dims[2] ? cz1 + R : d.dims[2] - 1;
    long long base = (x * d.dims[1] + y) * d.dims[2];
    a = start[base + z0];
    b = start[base + z1 + 1];
}


// Slots per drain round as a type, so one drain can be instantiated at two
// widths.
template <int N>
struct DrainWidth {
    static constexpr int value = N;
};

// S: list slots drained per round (1, 2 or 4: more keeps more neighbour
// loads in flight, fewer needs fewer registers) in the uniform scan; the
// per-lane scan, whose row state stays live across its drains, drains at most
// two.
// zt: the z half-widths of the uniform scan's trimmed rows. With LOC the
// uniform scan filters with the column-local coordinates of cl instead of
// test (which still filters the overflow segment and the per-lane scan).
template <int L, int MODE, int S, bool LOC, bool PAD, int R, int BS = 256, class Test, class Load,
          class Proc, class Flush, class Again>
__device__ __forceinline__ void gather(bool act, uint32_t ci, float xi, float yi, float zi,
                                       uint32_t self_j, const Domain& d, double inv,
                                       const ZWin zt, const ColLocal& cl,
                                       const uint32_t* __restrict__ start,
                                       const float4* __restrict__ posm, uint32_t (*lst)[BS],
                                       int tid, Test&& test, Load&& load, Proc&& proc,
                                       Flush&& flush, Again&& again) {
    static_assert(S == 1 || S == 2 || S == 4, "drain width");
    static_assert(L % S == 0, "whole drain rounds per list");
    static_assert(R >= 1 && R <= 8 && R <= L, "per-lane round width");
    static_assert(BS == 128 || BS == 256, "workgroups of 128 or 256 lanes");
    // the lane's list head as a 32-bit byte offset into the list array: it
    // advances one row (4 BS bytes) per append, so an append is one select
    // and one add; the count is derived from it where needed
    constexpr uint32_t kRowB = 4u * BS;
    constexpr uint32_t kRowShift = BS == 256 ? 10u : 9u;
    char* const lbase = reinterpret_cast<char*>(&lst[0][0]);
    const uint32_t lp0 = (uint32_t)tid * 4u;
    uint32_t lp = lp0;
    auto count = [&]() { return (int)((lp - lp0) >> kRowShift); };
    const uint32_t self_o = self_j << 4;
    for (int k = 0; k < L; ++k) lst[k][tid] = self_o;
    // list-full tests: at least L - 8 + 1, L - R + 1 or L entries
    const uint32_t lp8 = lp0 + (L - 8) * kRowB;
    const uint32_t lpR = lp0 + (L - R) * kRowB;
    const uint32_t lpL = lp0 + L * kRowB;
    auto drain = [&](auto width, auto scan) {
        constexpr int W = decltype(width)::value;
        // the head slot may hold the last failing candidate: with PAD it
        // names the lane again (a full list has no head slot)
        if (PAD && lp < lpL) *reinterpret_cast<uint32_t*>(lbase + lp) = self_o;
        const int cnt = count();
        // do-while: the loop test sits at the bottom, so the accumulators
        // stay in place across iterations
        if (!wave_any(0 < cnt)) {
            flush(scan);
            return;
        }
        int k = 0;
        do {
            if (W == 4) {
                const bool v0 = k < cnt, v1 = k + 1 < cnt, v2 = k + 2 < cnt, v3 = k + 3 < cnt;
                const uint32_t o0 = lst[k][tid], o1 = lst[k + 1][tid], o2 = lst[k + 2][tid],
                               o3 = lst[k + 3][tid];
                if (PAD) {
                    lst[k][tid] = self_o;
                    lst[k + 1][tid] = self_o;
                    lst[k + 2][tid] = self_o;
                    lst[k + 3][tid] = self_o;
                }
                auto t0 = load(o0);
                auto t1 = load(o1);
                auto t2 = load(o2);
                auto t3 = load(o3);
                proc(o0, t0, v0, scan);
                proc(o1, t1, v1, scan);
                proc(o2, t2, v2, scan);
                proc(o3, t3, v3, scan);
            } else if (W == 2) {
                const bool v0 = k < cnt, v1 = k + 1 < cnt;
                const uint32_t o0 = lst[k][tid], o1 = lst[k + 1][tid];
                if (PAD) {
                    lst[k][tid] = self_o;
                    lst[k + 1][tid] = self_o;
                }
                auto t0 = load(o0);
                auto t1 = load(o1);
                proc(o0, t0, v0, scan);
                proc(o1, t1, v1, scan);
            } else {
                const bool v0 = k < cnt;
                const uint32_t o0 = lst[k][tid];
                if (PAD) lst[k][tid] = self_o;
                auto t0 = load(o0);
                proc(o0, t0, v0, scan);
            }
            k += W;
        } while (wave_any(k < cnt));
        lp = lp0;
        flush(scan);
    };
    // The slot at the head gets j whatever ok says, and the head advances
    // only when ok (no branch); j must be a valid index: it stays in the list
    // until overwritten, and drains read every slot. Only the head slot can
    // hold a failing candidate. With PAD a drain first writes the lane's own
    // offset there and puts it back into every slot it reads, so a slot at or
    // past the count always names the lane.
    auto push = [&](uint32_t j, bool ok) {
        *reinterpret_cast<uint32_t*>(lbase + lp) = j << 4;
        lp += ok ? kRowB : 0u;
    };
    auto ov_ok = [&](const float4& pj) {
        return keys_adjacent(xi, yi, zi, pj.x, pj.y, pj.z, inv);
    };
    const WaveCol wc = MODE == kOtherWaves ? WaveCol{false, 0u, 0u} : wave_column(act, ci, d);
    const bool uniform = MODE == kUniformWaves ? true : MODE == kOtherWaves ? false : wc.ok;
    if (MODE != kOtherWaves && uniform) {
        const CellCoord cc = decode(wc.c0, d);
        const long long cz1 = cc.cz + (long long)(wc.c1 - wc.c0);
        // the lanes' z-range: a candidate farther than the row's half-width
        // below or above it is no lane's neighbour
        const double zmn = (double)wave_min_f(act ? zi : INFINITY);
        const double zmx = (double)wave_max_f(act ? zi : -INFINITY);
        // the widths as plain values: a select between fields of a struct in
        // memory becomes a load from a selected address, kept on the stack
        const double zw0 = zt.w0, zw1 = zt.w1, zw2 = zt.w2;
        // rows in mirror order (0, 24, 1, 23, ..., 11, 13, 12 for 5 x 5 rows):
        // a row off to one side fills the lists of the lanes on that side
        // first, and its mirror row right after evens the counts out
        const int reach = (int)uni((uint32_t)d.reach), span = 2 * reach + 1, nrows = span * span;
        do {
            for (int kr = 0; kr < nrows; ++kr) {
                const int row = (kr & 1) ? nrows - 1 - (kr >> 1) : (kr >> 1);
                const int rx = row / span;
                const int ex = rx - reach, ey = row - rx * span - reach;
                uint32_t a, b;
                col_run(ex, ey, cc, cz1, d, start, a, b);
                a = uni(a);
                b = uni(b);
                if (zt.on && a < b) {
                    const int cls = (ex == reach || ex == -reach) + (ey == reach || ey == -reach);
                    const double w = cls == 0 ? zw0 : (cls == 1 ? zw1 : zw2);
                    trim_run<LOC>(LOC ? cl.posl : posm, f32_down(zmn - w), f32_up(zmx + w), a, b);
                    a = uni(a);
                    b = uni(b);
                }
                // the lane's position relative to the row's column corner
                const float xr = cl.li.x - (float)ex * cl.hh, yr = cl.li.y - (float)ey * cl.hh;
                const float mx = -2.f * xr, my = -2.f * yr;
                // an inactive lane's bound is -inf: it takes no candidate (the
                // scan's data are finite), with no per-candidate lane mask
                const float thr = act ? cl.s_hq - fmaf(yr, yr, xr * xr) : -INFINITY;
                auto loc_maybe = [&](const float4& q) {
                    const float dz = q.w - zi;
                    const float t = fmaf(q.y, my, fmaf(q.x, mx, fmaf(dz, dz, q.z)));
                    return !(t >= thr);  // a NaN bound passes everything
                };
                const float4* __restrict__ src = LOC ? cl.posl : posm;
                const bool la = LOC || act;
                auto maybe = [&](uint32_t j, const float4& q) { return LOC ? loc_maybe(q) : test(j, q); };
                // Both ends of the z-sorted run at once: four candidates from the
                // front and four from the back per round. Lanes sit at different z,
                // so a front-to-back walk would fill the low lanes' lists first and
                // the high lanes' last; taking both ends keeps the counts closer and
                // the drains' lanes busier. Each round is two 64-byte wave-uniform
                // loads; cnt <= L - 8 before it, so its eight writes stay inside
                // the list (non-short-circuit flags: no branches).
                uint32_t lo = a, hi = b;
                while (hi - lo >= 8u) {
                    const Quad4 qa = *reinterpret_cast<const Quad4*>(src + lo);
                    const Quad4 qb = *reinterpret_cast<const Quad4*>(src + hi - 4u);
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u) push(lo + u, maybe(lo + u, qa.p[u]) & la);
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u) push(hi - 4u + u, maybe(hi - 4u + u, qb.p[u]) & la);
                    lo += 4u;
                    hi -= 4u;
                    if (wave_any(lp > lp8)) drain(DrainWidth<S>{}, std::true_type{});
                }
                if (lo < hi) {
                    // the last 1..7 candidates: one round from lo (the arrays have
                    // kScanPad spare entries, so every index of the round is in
                    // bounds); a candidate past the run gets a false flag
                    const uint32_t last = hi - 1u;
                    const Quad4 qa = *reinterpret_cast<const Quad4*>(src + lo);
                    const Quad4 qb = *reinterpret_cast<const Quad4*>(src + lo + 4u);
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u)
                        push(lo + u, maybe(lo + u, qa.p[u]) & (la & (lo + u <= last)));
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u)
                        push(lo + 4u + u, maybe(lo + 4u + u, qb.p[u]) & (la & (lo + 4u + u <= last)));
                    if (wave_any(lp > lp8)) drain(DrainWidth<S>{}, std::true_type{});
                }
            }
            const uint32_t o0 = uni(start[d.ncells]), o1 = uni(start[d.ncells + 1]);
            for (uint32_t j = o0; j < o1; ++j) {
                const float4 pj = posm[j];
                push(j, act && ov_ok(pj) && test(j, pj));
                if (wave_any(lp == lpL)) drain(DrainWidth<S>{}, std::true_type{});
            }
            drain(DrainWidth<S>{}, std::true_type{});
        } while (again());
    } else if (MODE != kUniformWaves) {
        // The per-lane scan. Each lane's window is a rectangle of table columns
        // [x0, x1] x [y0, y1] with z cells z0..z1 (empty when z0 > z1), as
        // 32-bit table-relative values with wrapping arithmetic (a row left of
        // the table wraps to a huge x and fails x <= x1): for a dense particle
        // its 5 x 5 half-cell (3 x 3 cell) window, for an out-of-domain one
        // the in-domain cells of its 27 reference cells (6 x 6 half cells).
        // The lanes walk the rows together: a wave-uniform loop over span x
        // span rows (x-major), span the widest window of the wave, each lane
        // loading the bounds of its own z-run of the row (a row outside its
        // rectangle is empty); then the overflow segment. A lane's rows come
        // in its own x-major order either way. The bounds of the next row are
        // loaded while the current row is tested.
        uint32_t x0 = 1u, x1 = 0u, y0 = 1u, y1 = 0u, z0 = 1u, z1 = 0u, step = 0u;
        if (act) {
            const uint32_t r = (uint32_t)d.reach;
            if (ci < d.ncells) {
                const CellCoord cc = decode(ci, d);
                const uint32_t cx = (uint32_t)cc.cx, cy = (uint32_t)cc.cy, cz = (uint32_t)cc.cz;
                x0 = cx - r;
                y0 = cy - r;
                x1 = min(cx + r, (uint32_t)d.dims[0] - 1u);
                y1 = min(cy + r, (uint32_t)d.dims[1] - 1u);
                z0 = cz > r ? cz - r : 0u;
                z1 = min(cz + r, (uint32_t)d.dims[2] - 1u);
                step = 2u * r + 1u;
            } else {
                long long lx, hx, ly, hy, lz, hz;
                ref_block(key1(xi, inv), d, 0, lx, hx);
                ref_block(key1(yi, inv), d, 1, ly, hy);
                ref_block(key1(zi, inv), d, 2, lz, hz);
                if (lx <= hx && ly <= hy && lz <= hz) {
                    x0 = (uint32_t)lx;
                    x1 = (uint32_t)hx;
                    y0 = (uint32_t)ly;
                    y1 = (uint32_t)hy;
                    z0 = (uint32_t)lz;
                    z1 = (uint32_t)hz;
                }
                step = 3u * r;
            }
        }
        const bool zok = act && z0 <= z1;
        const uint32_t D1 = (uint32_t)d.dims[1], D2 = (uint32_t)d.dims[2];
        const uint32_t span = uni(wave_max_u32(step)), nrows = span * span;
        auto bounds = [&](uint32_t k, uint32_t& ja, uint32_t& jb) {
            const uint32_t ra = k / span, rb = k - ra * span;  // wave-uniform
            const uint32_t x = x0 + ra, y = y0 + rb;
            const bool ok = zok && x <= x1 && y <= y1;
            const uint32_t base = ok ? (x * D1 + y) * D2 : 0u;
            ja = start[base + (ok ? z0 : 0u)];
            jb = start[base + (ok ? z1 + 1u : 0u)];
        };
        // R consecutive entries of the lane's run from one base offset (j + R
        // - 1 <= last + 7 stays inside the array's kScanPad spare entries; an
        // entry past the run gets a false flag); OV: the overflow segment's
        // reference key adjacency. The count is at most L - R before a round.
        auto round = [&](uint32_t& j, uint32_t jend, auto ov) {
            const uint32_t last = jend - 1u;
            const char* pb = reinterpret_cast<const char*>(posm) + (size_t)(j << 4);
            float4 pq[R];
#pragma unroll
            for (int u = 0; u < R; ++u) pq[u] = *reinterpret_cast<const float4*>(pb + 16 * u);
#pragma unroll
            for (int u = 0; u < R; ++u) {
                bool t = (j + (uint32_t)u <= last) & test(j + u, pq[u]);
                if (decltype(ov)::value) t = t & ov_ok(pq[u]);
                push(j + u, t);
            }
            j = min(j + (uint32_t)R, jend);
        };
        uint32_t pa = 0u, pb = 0u;
        if (nrows > 0u) bounds(0u, pa, pb);
        for (uint32_t k = 0; k < nrows; ++k) {
            uint32_t j = pa;
            const uint32_t jend = pb;
            if (k + 1u < nrows) bounds(k + 1u, pa, pb);
            while (wave_any(j < jend)) {
                if (j < jend) round(j, jend, std::false_type{});
                if (wave_any(lp > lpR)) drain(DrainWidth<(S < 2 ? S : 2)>{}, std::false_type{});
            }
        }
        {
            const uint32_t o0 = uni(start[d.ncells]), o1 = uni(start[d.ncells + 1]);
            uint32_t j = act ? o0 : 0u;
            const uint32_t jend = act ? o1 : 0u;
            while (wave_any(j < jend)) {
                if (j < jend) round(j, jend, std::true_type{});
                if (wave_any(lp > lpR)) drain(DrainWidth<(S < 2 ? S : 2)>{}, std::false_type{});
            }
        }
        drain(DrainWidth<(S < 2 ? S : 2)>{}, std::false_type{});
    }
}

// Candidate filters of the compaction. Both are conservative: |s32 - s64| <=
// 6u s for any FP32 evaluation order, so every pair with r < h has
// s32 < s_thresh (1 + 6u) < s_hi. The exact decision is made per listed slot.
//
// Density pass (world.rs:225-229): r < h, self included; a NaN distance fails.
__device__ __forceinline__ bool density_maybe(const DevParams& P, float xi, float yi, float zi,
                                              const float4& pj) {
    const float dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
    return dx * dx + dy * dy + dz * dz < P.s_hi;
}

// Force pass (world.rs:292-299): a NaN distance passes (it contributes NaN).
__device__ __forceinline__ bool force_maybe(const DevParams& P, float xi, float yi, float zi,
                                            const float4& pj) {
    const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
    return !(dx * dx + dy * dy + dz * dz >= P.s_hi);
}

// The reference's f64 squared distance from f64 coordinate differences,
// rounded exactly like the oracle's (no contraction).
__device__ __forceinline__ double sq_exact(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    return (dx * dx + dy * dy) + dz * dz;
}

// Exact density-pass decision (r < h, self included) for a listed slot with
// FP32 squared distance s32; the f64 band and NaN take the exact predicate.
// The vote is one compare (an unlisted slot's key is 0, inside), and the f64
// call runs for every lane behind it: the common path changes no execution
// mask.
__device__ __forceinline__ bool density_valid(const DevParams& P, bool listed, float xi, float yi,
                                              float zi, const float4& pj, float s32) {
    const float key = listed ? s32 : 0.f;
    const bool fast = key < P.s_lo;
    bool ex = false;
    if (wave_any(!(key < P.s_lo))) ex = !fast && exact_s64(xi, yi, zi, pj.x, pj.y, pj.z) < P.s_thresh;
    return listed && (fast || ex);
}

// The FP32 side of the force-pass decision in integer form. s32 is +0,
// positive or NaN, so with t = bits(s32) - 1 (+0 wraps to the top) 0 < s32 <
// s_lo is t < fast_bound (0 when s_lo is not positive: nothing is clear). A
// slot with t >= fast_bound needs the exact predicate unless it names the
// particle itself; its key is t, the particle's own slots' key 0, so the
// wave's vote is one compare.
__device__ __forceinline__ uint32_t fast_bound(const DevParams& P) {
    return P.s_lo > 0.f ? __float_as_uint(P.s_lo) - 1u : 0u;
}

// Exact force-pass decision for a listed slot j of particle i (world.rs:
// 292-299): skip j == i, r >= h and r <= 0; a NaN distance is kept. i and j
// are compared for equality only (indices or list offsets alike). s32 is the
// slot's FP32 squared distance. Clear cases are decided in FP32; the f64 band,
// NaN and coincident distinct points take the exact f64 predicate, which runs
// for every lane when some lane needs it (the common path changes no
// execution mask).
__device__ __forceinline__ bool force_valid(const DevParams& P, bool listed, uint32_t i, uint32_t j,
                                            float xi, float yi, float zi, const float4& pj,
                                            float s32) {
    const uint32_t tb = fast_bound(P), t = __float_as_uint(s32) - 1u;
    const bool fast = listed && t < tb;
    const uint32_t key = listed && j != i ? t : 0u;
    bool ex = false;
    if (wave_any(key >= tb)) {
        const double e = exact_s64(xi, yi, zi, pj.x, pj.y, pj.z);
        const bool nz = xi != pj.x || yi != pj.y || zi != pj.z;
        ex = key >= tb && ((e < P.s_thresh && nz) || e != e);
    }
    return fast || ex;
}

// force_valid for a proc that multiplies its zero weights into the
// neighbour's velocity: a listed pair the exact test rejects (a band pair
// beyond h, a coincident point) gets the particle's own velocity vi in vj, so
// a NaN velocity there cannot reach the sums. The replacement sits inside the
// rarely taken exact branch.
__device__ __forceinline__ bool force_valid_v(const DevParams& P, bool listed, uint32_t i,
                                              uint32_t j, float xi, float yi, float zi,
                                              const float4& pj, float s32, const float4& vi,
                                              float4& vj) {
    const uint32_t tb = fast_bound(P), t = __float_as_uint(s32) - 1u;
    const bool fast = listed && t < tb;
    const uint32_t key = listed && j != i ? t : 0u;
    bool ex = false;
    if (wave_any(key >= tb)) {
        const double e = exact_s64(xi, yi, zi, pj.x, pj.y, pj.z);
        const bool nz = xi != pj.x || yi != pj.y || zi != pj.z;
        const bool need = key >= tb;
        ex = need && ((e < P.s_thresh && nz) || e != e);
        if (need && !ex) vj = vi;
    }
    return fast || ex;
}

}  // namespace phy
