// sph_gathers.h — the density and force gather kernels (world.rs:204-347 with
// the fused integrate / bounds / watchdog of world.rs:352-397 and
// subsystem.rs:74-90) and their launch templates, included by sph_kernels.hip
// (the step's instantiations) and sph_stats.hip (the statistics forms, which
// run the same gathers and record the neighbour sets instead of outputs).
// Template parameters: L the per-lane list length, MODE the waves a launch
// serves, STATS the statistics form, BS the workgroup size (128 or 256 lanes:
// 2 or 1 workgroups per 256-particle tile).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include <type_traits>

#include "sph_types.h"
#include "sph_device.h"
#include "sph_kernels.h"

namespace phy {

// ---------------------------------------------------------------------------
// neighbour-set statistics of the gathers themselves
//
// With STATS the density and force kernels run their own gather (the same
// scan, filters, lists and exact tests as the step) and, instead of writing
// the step's outputs, record for each particle the count, the sum and the xor
// of SplitMix64-hashed scene-local original indices of every listed slot their
// exact test accepts: the neighbour sets the physics sums over, for exact
// comparison with the reference's sets.

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct SetAcc {
    uint32_t c = 0;
    unsigned long long s = 0, x = 0;
    __device__ __forceinline__ void add(bool ok, uint32_t id) {
        c += ok ? 1u : 0u;
        s += ok ? (unsigned long long)id : 0ull;
        x ^= ok ? mix64(id) : 0ull;
    }
    __device__ __forceinline__ void store(const GatherStats& g, uint32_t o) const {
        g.cnt[o] = c;
        g.sum[o] = s;
        g.xr[o] = x;
    }
};

// ---------------------------------------------------------------------------
// density / EOS (world.rs:204-246)
//
// Near the rest density the pair geometry and the sum run in f64: p = k (rho -
// rho0) cancels there and would amplify an FP32 density error of ~2e-6 by
// rho / (rho - rho0). MI355X executes f64 FMA at the unpacked FP32 rate, and the
// compaction limits it to true neighbours.
//
// Compressed waves take FP32 pair terms, candidate-major. A column wave (64
// particles within three half cells of one column: about 20x the rest density
// or more) walks its shared window (the uniform scan's rows, trimmed to the
// lanes' z-range, and the overflow segment) four candidates at a time with
// wave-uniform loads, and every lane evaluates every candidate: q = h^2 - s by
// an FMA chain over FP32 differences of the sorted positions, decided by its
// sign outside a band of +-2^-19 h^2 and by the exact f64 predicate inside it
// (the same sets as the list path), and sums T = sum m_j q^3 in FP32 partials
// of 16 candidates folded into an f64 total. No lists, no drains: with about
// half of the window's candidates being neighbours, this costs fewer
// operations than filtering, listing and draining them. Error, with u =
// 2^-24, for a neighbour (s <= h^2, so every step of the chain lies in [0,
// h^2]): h^2 rounded (u h^2), the differences squared (2u s) and three
// roundings (2u h^2 + u q) give |q32 - q| <= 5u h^2; a term's error is at most
// 15u h^2 m q^2 + 3u m q^3, and a partial of at most 16 positive terms adds
// 15u. With E = sum m q^2 summed alongside (in FP32: an error below 2^-11
// relative for up to 2^12 terms), the relative error of T is at most (15 R +
// 18) u, R = h^2 E / T (1.5 for neighbours spread evenly over the ball). A lane
// accepts its FP32 density when R <= 1.99 (48u), rho - rho0 >= 0.75 rho (p =
// k (rho - rho0) then within 64u = 4e-6 relative) and 2^-100 <= T <= 2^100 (no
// FP32 overflow, underflow negligible); when any lane of the wave does not,
// the wave sums again in f64 through the list path. A neighbour inside the
// band whose q32 rounds to 0 or below adds 0 (its true term is below m (5u
// h^2)^3).
constexpr double kLowpShare = 0.75;
constexpr double kLowpRatio = 1.99;

// The FP32 candidate-major density sums of a column wave (see above) into sum
// (T, f64) and esum (E); with STATS the accepted sets into sa as well.
template <bool STATS>
__device__ __forceinline__ void density_column(const WaveCol& wc, bool act, const float4& pi,
                                               const DevParams& P, const Domain& d, const ZWin& zt,
                                               const uint32_t* __restrict__ start,
                                               const float4* __restrict__ posm, const GatherStats& gs,
                                               uint32_t soff, double& sum, float& esum, SetAcc& sa) {
    float part = 0.f;
    // q = h^2 - s in FP32: an FMA chain over the sorted positions' differences
    auto qfp = [&](const float4& q) {
        const float dx = q.x - pi.x, dy = q.y - pi.y, dz = q.z - pi.z;
        return fmaf(-dx, dx, fmaf(-dy, dy, fmaf(-dz, dz, P.h2)));
    };
    // the reference's own f64 distance, inline (a call here would cost the
    // kernel registers); NaN fails
    auto exact_in = [&](const float4& q) {
        const double ex = (double)pi.x - (double)q.x, ey = (double)pi.y - (double)q.y,
                     ez = (double)pi.z - (double)q.z;
        return sq_exact(ex, ey, ez) < P.s_thresh;
    };
    const float qb = P.q_band;
    // a candidate's decision: |q| > q_band decides by the sign of q, the band
    // between by the exact predicate (NaN fails both)
    auto decide = [&](float qv, const float4& q) { return fabsf(qv) > qb ? qv > 0.f : exact_in(q); };
    const CellCoord cc = decode(wc.c0, d);
    const long long cz1 = cc.cz + (long long)(wc.c1 - wc.c0);
    const double zmn = (double)wave_min_f(act ? pi.z : INFINITY);
    const double zmx = (double)wave_max_f(act ? pi.z : -INFINITY);
    const double zw0 = zt.w0, zw1 = zt.w1, zw2 = zt.w2;
    const int reach = (int)uni((uint32_t)d.reach);
    // UM: every particle of the scene has the mass m_uni, factored out of the
    // rows' terms (T and E scaled once after the rows)
    auto rows = [&](auto um) {
        constexpr bool UM = decltype(um)::value;
        for (int ex = -reach; ex <= reach; ++ex)
            for (int ey = -reach; ey <= reach; ++ey) {
                uint32_t a, b;
                col_run(ex, ey, cc, cz1, d, start, a, b);
                a = uni(a);
                b = uni(b);
                if (zt.on && a < b) {
                    const int cls = (ex == reach || ex == -reach) + (ey == reach || ey == -reach);
                    const double w = cls == 0 ? zw0 : (cls == 1 ? zw1 : zw2);
                    trim_run<false>(posm, f32_down(zmn - w), f32_up(zmx + w), a, b);
                    a = uni(a);
                    b = uni(b);
                }
                // four candidates per round from one 64-byte wave-uniform load
                // (the array has kScanPad spare entries). Each lane takes max(q,
                // 0), the term of every candidate outside the band; a round that
                // finds one of its lanes inside the band (|q| <= q_band, from a
                // min over the four |q|) lets those lanes zero the candidates the
                // exact predicate rejects. Full rounds first, then the run's last
                // partial round with the candidates past its end zeroed. A
                // 16-candidate partial is folded into the f64 total.
                int round = 0;
                auto group = [&](uint32_t j, auto tail) {
                    constexpr bool TAIL = decltype(tail)::value;
                    const Quad4 qm = *reinterpret_cast<const Quad4*>(posm + j);
                    float qe[4];
                    float band = INFINITY;
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u) {
                        const float qv = qfp(qm.p[u]);
                        band = fminf(band, fabsf(qv));
                        qe[u] = fmaxf(qv, 0.f);  // NaN: 0
                    }
                    bool vb[4];
                    if (STATS) {
#pragma unroll
                        for (uint32_t u = 0; u < 4; ++u)
                            vb[u] = (!TAIL || j + u < b) && decide(qfp(qm.p[u]), qm.p[u]);
                    }
                    if (wave_any(band <= qb)) {
#pragma unroll
                        for (uint32_t u = 0; u < 4; ++u)
                            if (fabsf(qfp(qm.p[u])) <= qb && !exact_in(qm.p[u])) qe[u] = 0.f;
                    }
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u) {
                        const bool in = !TAIL || j + u < b;
                        float mq2;
                        if (UM) {
                            const float q = in ? qe[u] : 0.f;
                            mq2 = q * q;
                            qe[u] = q;
                        } else {
                            mq2 = (in ? qm.p[u].w : 0.f) * (qe[u] * qe[u]);
                        }
                        part = fmaf(mq2, qe[u], part);
                        esum += mq2;
                        if (STATS) sa.add(vb[u], gs.id[j + u] - soff);
                    }
                    if (++round == 4) {
                        sum += (double)part;
                        part = 0.f;
                        round = 0;
                    }
                };
                uint32_t j = a;
                for (; j + 4u <= b; j += 4u) group(j, std::false_type{});
                if (j < b) group(j, std::true_type{});
                sum += (double)part;
                part = 0.f;
            }
    };
    const float mu = P.m_uni;
    if (mu == mu) {
        rows(std::true_type{});
        sum *= (double)mu;
        esum *= mu;
    } else {
        rows(std::false_type{});
    }
    const uint32_t o0 = uni(start[d.ncells]), o1 = uni(start[d.ncells + 1]);
    for (uint32_t j = o0; j < o1; ++j) {
        // the overflow segment: per-lane membership (key adjacency), and data
        // that may be non-finite, so both factors are selected
        const float4 q = posm[j];
        const float qv = qfp(q);
        const bool valid = keys_adjacent(pi.x, pi.y, pi.z, q.x, q.y, q.z, P.inv_cell) && decide(qv, q);
        const float qe = valid ? fmaxf(qv, 0.f) : 0.f;
        const float mq2 = valid ? q.w * (qe * qe) : 0.f;
        part = fmaf(mq2, qe, part);
        esum += mq2;
        if (STATS) sa.add(valid, gs.id[j] - soff);
        sum += (double)part;
        part = 0.f;
    }
}

template <int L, int MODE, bool STATS, int BS>
// Seven waves per SIMD (at most 72 VGPRs; the kernel needs 78 unhinted and
// fits in 71 without spilling): 1M collapsed sheet 0.506 -> 0.490 ms per
// density pass; eight (63 VGPRs, a few spilled dwords) is slower.
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(7))) void k_density(SceneTable T, const float4* __restrict__ posm,
                                                 const float4* __restrict__ posl,
                                                 const uint32_t* __restrict__ cellv,
                                                 const uint32_t* __restrict__ start_all,
                                                 float* rho_out, float* p_out, float4* velm,
                                                 float* mp_out, float4* frec, GatherStats gs) {
    __shared__ uint32_t lst[L][BS];
    const int tid = threadIdx.x;
    const TileLoc TL = tile_loc_bs<BS>(T);
    const int i = TL.i;
    const bool act = TL.act;
    const Domain d = T.dom[TL.s];
    uint32_t ci = act ? cellv[i] : d.ncells;
    if (MODE != kAllWaves && wave_uniform(act, ci, d) != (MODE == kUniformWaves)) return;
    const DevParams P = T.params[TL.s];
    const uint32_t* __restrict__ start = start_all + d.toff;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    constexpr bool LOC = MODE != kOtherWaves;
    ColLocal cl{posl, make_float4(0.f, 0.f, 0.f, 0.f), (float)P.cell64, P.s_hq};
    if (act) {
        pi = posm[i];
        if (LOC) cl.li = posl[i];
    }
    double sum = 0.0;
    SetAcc sa;
    const uint32_t soff = STATS ? T.info[TL.s].off : 0u;
    bool done = false;
    if (MODE != kOtherWaves) {
        const WaveCol wc = wave_column(act, ci, d);
        if (wc.ok) {
            float esum = 0.f;
            density_column<STATS>(wc, act, pi, P, d, zwin(P, d), start, posm, gs, soff, sum, esum, sa);
            const double rho = P.c6_64 * sum;
            const bool keep = !act || (sum >= 0x1p-100 && sum <= 0x1p100 &&
                                       P.h2_64 * (double)esum <= kLowpRatio * sum && rho - P.rest64 >= kLowpShare * rho);
            done = !wave_any(!keep);
            if (!done) {
                sum = 0.0;
                sa = SetAcc{};
            }
        }
    }
    if (!done) {
        const double xi = pi.x, yi = pi.y, zi = pi.z;
        // four neighbour loads in flight per drain round (force keeps two: four
        // would cost it a wave per SIMD); the per-lane scan takes six candidates
        // per round: a particle's run of a row holds about five at rest density
        // with half cells, so most runs take one round instead of two (1M column
        // at rest: 0.168 -> 0.136 ms per density pass, for five waves per SIMD
        // instead of six)
        gather<L, MODE, (MODE == kUniformWaves ? 1 : 4), LOC, false, 6, BS>(
            act, ci, pi.x, pi.y, pi.z, act ? (uint32_t)i : 0u, d, P.inv_cell, zwin(P, d), cl, start, posm, lst,
            tid, [&](uint32_t j, const float4& pj) { return density_maybe(P, pi.x, pi.y, pi.z, pj); },
            [&](uint32_t o) { return at(posm, o); },
            [&](uint32_t j, const float4& pj, bool listed, auto) {
                // the reference's own f64 distance decides r < h exactly (NaN fails)
                const double dx = xi - (double)pj.x, dy = yi - (double)pj.y, dz = zi - (double)pj.z;
                const double s = sq_exact(dx, dy, dz);
                const bool valid = listed && s < P.s_thresh;
                const double q = P.h2_64 - s;
                const double t = (double)pj.w * (q * q * q);
                sum += valid ? t : 0.0;
                if (STATS) sa.add(valid, at(gs.id, j) - soff);
            },
            [&](auto) {}, [] { return false; });
    }
    if (!act) return;
    if (STATS) {
        sa.store(gs, gs.id[i]);
        return;
    }
    double rho = P.c6_64 * sum;
    double over = rho - P.rest64;
    double p = over > 0.0 ? P.stiff64 * over : 0.0;
    rho_out[i] = (float)rho;
    p_out[i] = (float)p;
    velm[i].w = (float)((double)pi.w / rho);  // m_j / rho_j of the force pass
    const float mpj = (float)((double)pi.w * (p / (rho * rho)));  // m_j * p_j / rho_j^2
    mp_out[i] = mpj;
    if (frec) frec[i] = make_float4(pi.x, pi.y, pi.z, mpj);
}

// ---------------------------------------------------------------------------
// forces + integrate + bounds + watchdog (world.rs:252-397, subsystem.rs:74-90)
//
// Summation over blocks: the contributions of one drained batch (at most L
// neighbours of this lane) are summed with plain FMAs, then each batch partial
// is added to the running total, with Kahan's algorithm for the pressure
// force (its terms point every way and cancel in the interior) and plainly for
// the viscosity, XSPH and shear sums (a 1M collapsed sheet: acc, xsph and v
// within 1.4e-7, 3.4e-7 and 9e-8 of the oracle; full Kahan 1.8e-7, 3.4e-7,
// 7.5e-8). Ten fewer registers than Kahan on all nine sums, and with the
// occupancy hint of five waves per SIMD (a few values spill outside the
// loops) the kernel runs five waves instead of four: 1M at rest 0.393 ->
// 0.379 ms per step, collapsed 2.83 -> 2.78 ms.

template <bool SHEAR>
struct ForceAcc {
    float b[10];   // batch partials: pressure xyz, viscosity xyz, XSPH xyz, shear
    float s[10];   // totals
    float c[3];    // compensation of the pressure totals
    __device__ __forceinline__ void zero() {
        for (int k = 0; k < 10; ++k) b[k] = s[k] = 0.f;
        for (int k = 0; k < 3; ++k) c[k] = 0.f;
    }
    __device__ __forceinline__ void merge() {
        for (int k = 0; k < (SHEAR ? 10 : 9); ++k) {
            if (k < 3) kahan_add(s[k], c[k], b[k]);
            else s[k] += b[k];
            b[k] = 0.f;
        }
    }
};

struct NbrData {
    float4 p, v;
    float mp;
};

// Returns the exact force-pass decision for the slot (the pair's terms are
// added only when it holds).
// UM: the scene's masses are all m_uni, and nb.p carries m p / rho^2 in w
// (the density pass's record) instead of the mass.
template <bool SHEAR, bool UM = false>
__device__ __forceinline__ bool pair_force(ForceAcc<SHEAR>& A, const DevParams& P, uint32_t i,
                                           uint32_t j, float4 pi, float4 vi, float Pi,
                                           const NbrData& nb, bool listed) {
    const float4 pj = nb.p;
    float4 vj = nb.v;
    float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
    float s = dx * dx + dy * dy + dz * dz;
    const bool valid = force_valid_v(P, listed, i, j, pi.x, pi.y, pi.z, pj, s, vi, vj);
    // distinct FP32 points closer than ~1e-19 underflow s; keep r tiny but finite.
    // (the IEEE maximum: a NaN distance stays NaN, as in the reference)
    s = __builtin_elementwise_maximum(s, 1.17549435e-38f);
    float inv_r = __builtin_amdgcn_rsqf(s);
    float r = s * inv_r;
    float hr = P.h - r;
    float dvx = vj.x - vi.x, dvy = vj.y - vi.y, dvz = vj.z - vi.z;
    // m_j / rho_j, zero for a rejected or padding slot (the particle itself, a
    // band pair outside r < h, a coincident point): its viscosity, XSPH and
    // shear weights vanish with it (h - r, q and the velocity differences are
    // finite there), so they need no selects of their own
    const float mr = valid ? vj.w : 0.f;
    float q = P.h2 - s;
    // pressure: dir * (m_i m_j (P_i + P_j) * cs (h - r)^2), dir = d / r; selected
    // away for a rejected slot, whose m p / rho^2 may be NaN (rho = 0)
    float a = (UM ? fmaf(P.m_uni, Pi, pj.w) : fmaf(pj.w, Pi, nb.mp)) * (hr * hr) * inv_r;
    a = valid ? a : 0.f;
    // viscosity: (v_j - v_i) * (m_i m_j / rho_j * cv (h - r))
    float bb = mr * hr;
    // XSPH: (v_j - v_i) * (m_j / rho_j * poly6(r))
    float w = mr * (q * q * q);
    A.b[0] = fmaf(a, dx, A.b[0]);
    A.b[1] = fmaf(a, dy, A.b[1]);
    A.b[2] = fmaf(a, dz, A.b[2]);
    A.b[3] = fmaf(bb, dvx, A.b[3]);
    A.b[4] = fmaf(bb, dvy, A.b[4]);
    A.b[5] = fmaf(bb, dvz, A.b[5]);
    A.b[6] = fmaf(w, dvx, A.b[6]);
    A.b[7] = fmaf(w, dvy, A.b[7]);
    A.b[8] = fmaf(w, dvz, A.b[8]);
    if (SHEAR) {
        // shear-rate proxy: |v_j - v_i| / (r + 1e-4) * (m_j / rho_j); only a
        // non-Newtonian material (visc_n != 1) reads it
        float dv = sqrtf(dvx * dvx + dvy * dvy + dvz * dvz);
        float t = dv * __builtin_amdgcn_rcpf(r + P.r_eps) * mr;
        A.b[9] += t;
    }
    return valid;
}

template <int L, bool SHEAR, int MODE, bool STATS, int BS, bool UM>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(5))) void k_force(
    SceneTable T, float dt, const float4* __restrict__ posm, const float4* __restrict__ posl,
    const float4* __restrict__ velm,
    const float* __restrict__ rho, const float* __restrict__ prs, const float* __restrict__ mp,
    const float4* __restrict__ frec,
    const uint32_t* __restrict__ cellv, const uint32_t* __restrict__ start_all,
    const uint32_t* __restrict__ idv, const uint32_t* __restrict__ matv,
    const float4* __restrict__ body, float4* posm_out, float4* vel_out, uint32_t* id_out,
    uint32_t* mat_out, float4* acc_out, float* mu_out, float4* xsph_out, float4* fvisc_out,
    GatherStats gs) {
    __shared__ uint32_t lst[L][BS];
    const int tid = threadIdx.x;
    const TileLoc TL = tile_loc_bs<BS>(T);
    const int i = TL.i;
    const bool act = TL.act;
    const Domain d = T.dom[TL.s];
    uint32_t ci = act ? cellv[i] : d.ncells;
    if (MODE != kAllWaves && wave_uniform(act, ci, d) != (MODE == kUniformWaves)) return;
    const DevParams P = T.params[TL.s];
    const uint32_t* __restrict__ start = start_all + d.toff;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f), vi = pi;
    float rho_i = 1.f, Pi = 0.f;
    if (act) {
        pi = posm[i];
        vi = velm[i];
        rho_i = rho[i];
        Pi = prs[i] / (rho_i * rho_i);
    }
    constexpr bool LOC = MODE != kOtherWaves;
    ColLocal cl{posl, make_float4(0.f, 0.f, 0.f, 0.f), (float)P.cell64, P.s_hq};
    if (LOC && act) cl.li = posl[i];
    ForceAcc<SHEAR> A;
    A.zero();
    SetAcc sa;
    const uint32_t soff = STATS ? T.info[TL.s].off : 0u;
    // per-lane rounds of six candidates: a row of five half cells at rest
    // spacing holds about five, so a row takes one round instead of two
    // the slot as a byte offset of 16- and 4-byte records: the stores after
    // the gather use it with the arrays' scalar bases (no 64-bit per-lane
    // address stays live across the gather)
    const uint32_t self_j = act ? (uint32_t)i : 0u;
    // UM: the per-lane scan and the overflow segment read candidates' x, y, z
    // from the records the drains load too, so a drain finds the lines its
    // scan brought into L1
    gather<L, MODE, (MODE == kUniformWaves ? 1 : 2), LOC, true, 6, BS>(
        act, ci, pi.x, pi.y, pi.z, self_j, d, P.inv_cell, zwin(P, d), cl, start, UM ? frec : posm, lst, tid,
        [&](uint32_t j, const float4& pj) { return force_maybe(P, pi.x, pi.y, pi.z, pj); },
        // UM: two loads per neighbour (the density pass's record and velm)
        [&](uint32_t o) {
            if (UM) return NbrData{at(frec, o), at(velm, o), 0.f};
            return NbrData{at(posm, o), at(velm, o), at(mp, o)};
        },
        // every slot at or past the lane's count names the lane itself, which
        // no validity test accepts, so the listed flag is not needed
        [&](uint32_t o, const NbrData& nb, bool, auto) {
            const bool ok = pair_force<SHEAR, UM>(A, P, self_j << 4, o, pi, vi, Pi, nb, true);
            if (STATS) sa.add(ok, at(gs.id, o) - soff);
        },
        [&](auto) { A.merge(); }, [] { return false; });
    if (!act) return;
    const uint32_t o16 = self_j << 4;
    if (STATS) {
        sa.store(gs, at(gs.id, o16));
        return;
    }
    auto put = [&](auto* base, auto v) {
        using E = std::remove_reference_t<decltype(*base)>;
        *reinterpret_cast<E*>(reinterpret_cast<char*>(base) + (o16 >> (sizeof(E) == 16 ? 0 : 2))) = v;
    };

    // power-law effective viscosity (world.rs:319-327)
    int m = 0;
    if (matv) {
        const uint32_t mt = at(matv, o16);
        put(mat_out, mt);  // the new state keeps this step's cell order
        if (P.n_mat > 1) m = mt < (uint32_t)P.n_mat ? (int)mt : 0;
    }
    // indexed through the table (a dynamic index into the register copy P
    // would move that copy to scratch memory)
    const DevParams* __restrict__ Pm = T.params + TL.s;
    float kc = Pm->visc_k[m];
    float mu;
    if (!SHEAR) {
        mu = kc;  // powf(x, 0) == 1 for every x
    } else {
        float sh = A.s[9];
        float sreg = sh > P.shear_min ? sh : P.shear_min;
        mu = kc * powf(sreg, Pm->visc_n[m] - 1.f);
    }
    // rho_i again from memory: kept in a register across the gather it is
    // the one value the five-wave occupancy hint would spill (an opaque copy
    // of the offset keeps the compiler from reusing the first load)
    uint32_t o_rho = o16;
    asm volatile("" : "+v"(o_rho));
    rho_i = at(rho, o_rho);
    float acc[3] = {0.f, 0.f, 0.f}, fvis[3] = {0.f, 0.f, 0.f};
    if (rho_i > 0.f) {
        float mcs = pi.w * P.cs, mcv = pi.w * P.cv * mu;
        for (int a = 0; a < 3; ++a) {
            fvis[a] = mcv * A.s[3 + a];
            acc[a] = (mcs * A.s[a] + fvis[a]) / rho_i;
        }
    }
    float xs[3] = {P.c6 * A.s[6], P.c6 * A.s[7], P.c6 * A.s[8]};

    // integrate (world.rs:352-363) and enforce_bounds (366-397)
    float bacc[3] = {0.f, 0.f, 0.f};
    if (body) {
        float4 b = at(body, o16);
        bacc[0] = b.x;
        bacc[1] = b.y;
        bacc[2] = b.z;
    }
    float x[3] = {pi.x, pi.y, pi.z}, v[3] = {vi.x, vi.y, vi.z};
    uint32_t clamp = 0;  // bit a: axis a clamped to a wall (the adjoint's bounds mask)
    for (int a = 0; a < 3; ++a) {
        float aa = (acc[a] + P.g[a]) + bacc[a];
        v[a] = v[a] + aa * dt;
        v[a] = v[a] + xs[a] * P.xsph_eps;
        x[a] = x[a] + v[a] * dt;
        if (x[a] < P.lo[a]) {
            x[a] = P.lo[a];
            v[a] = -v[a] * P.damp;
            clamp |= 1u << a;
        } else if (x[a] > P.hi[a]) {
            x[a] = P.hi[a];
            v[a] = -v[a] * P.damp;
            clamp |= 1u << a;
        }
    }
    put(posm_out, make_float4(x[0], x[1], x[2], pi.w));
    put(vel_out, make_float4(v[0], v[1], v[2], 0.f));
    if (acc_out) {
        put(acc_out, make_float4(acc[0], acc[1], acc[2], 0.f));
        put(mu_out, mu);
        put(xsph_out, make_float4(xs[0], xs[1], xs[2], __uint_as_float(clamp)));
    }
    // adjoint tape: viscous force per unit mu (m cv sum ...) and the raw
    // shear-rate sum of the power law (0 for a Newtonian pass)
    if (fvisc_out) {
        const float u = rho_i > 0.f ? pi.w * P.cv : 0.f;
        put(fvisc_out, make_float4(u * A.s[3], u * A.s[4], u * A.s[5], SHEAR ? A.s[9] : 0.f));
    }

    // watchdog: smallest original index with a non-finite position / velocity
    uint32_t id = at(idv, o16);
    put(id_out, id);
    uint32_t bp = finite3(x[0], x[1], x[2]) ? 0xffffffffu : id;
    uint32_t bv = finite3(v[0], v[1], v[2]) ? 0xffffffffu : id;
    if (bp != 0xffffffffu) atomicMin(&T.dom[TL.s].bad_pos, bp);
    if (bv != 0xffffffffu) atomicMin(&T.dom[TL.s].bad_vel, bv);
}


// ---------------------------------------------------------------------------
// launch templates

template <int L, bool STATS, int BS>
static void density_l(hipStream_t st, const SceneTable& T, bool split, const float4* posm,
                      const float4* posl, const uint32_t* cellv, const uint32_t* start, float* rho,
                      float* prs, float4* velm, float* mp, float4* frec, const GatherStats& gs) {
    const dim3 grid(T.ntiles * (kTile / BS)), block(BS);
    if (split) {
        hipLaunchKernelGGL((k_density<L, kUniformWaves, STATS, BS>), grid, block, 0, st, T,
                           posm, posl, cellv, start, rho, prs, velm, mp, frec, gs);
        hipLaunchKernelGGL((k_density<L, kOtherWaves, STATS, BS>), grid, block, 0, st, T,
                           posm, posl, cellv, start, rho, prs, velm, mp, frec, gs);
    } else {
        hipLaunchKernelGGL((k_density<L, kAllWaves, STATS, BS>), grid, block, 0, st, T, posm,
                           posl, cellv, start, rho, prs, velm, mp, frec, gs);
    }
}

template <bool STATS>
static void density_any(hipStream_t st, const SceneTable& T, const float4* posm, const float4* posl,
                        const uint32_t* cellv, const uint32_t* start, float* rho, float* prs,
                        float4* velm, float* mp, float4* frec, const Tuning& tu, const GatherStats& gs) {
    if (T.ntiles == 0) return;
    const bool sp = tu.split != 0;
    switch (tu.density_list * 1000 + tu.block) {
        case 8128: density_l<8, STATS, 128>(st, T, sp, posm, posl, cellv, start, rho, prs, velm, mp, frec, gs); break;
        case 16128: density_l<16, STATS, 128>(st, T, sp, posm, posl, cellv, start, rho, prs, velm, mp, frec, gs); break;
        case 32128: density_l<32, STATS, 128>(st, T, sp, posm, posl, cellv, start, rho, prs, velm, mp, frec, gs); break;
        case 8256: density_l<8, STATS, 256>(st, T, sp, posm, posl, cellv, start, rho, prs, velm, mp, frec, gs); break;
        case 16256: density_l<16, STATS, 256>(st, T, sp, posm, posl, cellv, start, rho, prs, velm, mp, frec, gs); break;
        default: density_l<32, STATS, 256>(st, T, sp, posm, posl, cellv, start, rho, prs, velm, mp, frec, gs); break;
    }
}


template <int L, bool SHEAR, int MODE, bool STATS, int BS, bool UM>
static void force_one(hipStream_t st, const SceneTable& T, float dt, const ForceArgs& a,
                      const GatherStats& gs) {
    hipLaunchKernelGGL((k_force<L, SHEAR, MODE, STATS, BS, UM>), dim3(T.ntiles * (kTile / BS)), dim3(BS), 0,
                       st, T, dt, a.posm, a.posl, a.velm, a.rho, a.prs, a.mp, a.frec, a.cell, a.start, a.id,
                       a.mat, a.body,
                       a.posm_out, a.vel_out, a.id_out, a.mat_out, a.acc_out, a.mu_out, a.xsph_out,
                       a.fvisc_out, gs);
}

template <int L, bool SHEAR, bool STATS, int BS, bool UM>
static void force_l(hipStream_t st, const SceneTable& T, bool split, float dt, const ForceArgs& a,
                    const GatherStats& gs) {
    if (split) {
        force_one<L, SHEAR, kUniformWaves, STATS, BS, UM>(st, T, dt, a, gs);
        force_one<L, SHEAR, kOtherWaves, STATS, BS, UM>(st, T, dt, a, gs);
    } else {
        force_one<L, SHEAR, kAllWaves, STATS, BS, UM>(st, T, dt, a, gs);
    }
}

template <bool SHEAR, bool STATS, int BS>
static void force_list(hipStream_t st, const SceneTable& T, bool split, float dt, const ForceArgs& a,
                       int list, const GatherStats& gs) {
    // the statistics forms read no neighbour data beyond positions: one form
    if (!STATS && a.frec) {
        switch (list) {
            case 8: force_l<8, SHEAR, STATS, BS, !STATS>(st, T, split, dt, a, gs); break;
            case 16: force_l<16, SHEAR, STATS, BS, !STATS>(st, T, split, dt, a, gs); break;
            default: force_l<32, SHEAR, STATS, BS, !STATS>(st, T, split, dt, a, gs); break;
        }
        return;
    }
    switch (list) {
        case 8: force_l<8, SHEAR, STATS, BS, false>(st, T, split, dt, a, gs); break;
        case 16: force_l<16, SHEAR, STATS, BS, false>(st, T, split, dt, a, gs); break;
        default: force_l<32, SHEAR, STATS, BS, false>(st, T, split, dt, a, gs); break;
    }
}

template <bool STATS>
static void force_any(hipStream_t st, const SceneTable& T, float dt, const ForceArgs& a, bool shear,
                      const Tuning& tu, const GatherStats& gs) {
    if (T.ntiles == 0) return;
    const bool sp = tu.split != 0;
    if (tu.block == 128) {
        if (shear) force_list<true, STATS, 128>(st, T, sp, dt, a, tu.force_list, gs);
        else force_list<false, STATS, 128>(st, T, sp, dt, a, tu.force_list, gs);
    } else {
        if (shear) force_list<true, STATS, 256>(st, T, sp, dt, a, tu.force_list, gs);
        else force_list<false, STATS, 256>(st, T, sp, dt, a, tu.force_list, gs);
    }
}

}  // namespace phy
