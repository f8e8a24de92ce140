// sph_adjoint.hip — reverse-mode (vector-Jacobian product) of one SPH step.
//
// The forward step (crates/phy-fluid/src/sph/world.rs:183-397) maps
// (x, v; k = stiffness, visc_k[m], visc_n[m]) to (x', v') and the density rho.
// Given the cotangents of (x', v') (and of rho on the last step), three
// passes produce the cotangents of (x, v) and accumulate the parameter
// gradients:
//
//   k_adj_local    per particle: wall clamp and semi-implicit Euler backwards;
//                  F_bar = a_bar / rho (acceleration = force / rho), the XSPH
//                  cotangent s_bar, the local rho_bar term, and the power-law
//                  viscosity backwards (dL/d visc_k, dL/d visc_n per material,
//                  the shear-rate cotangent).
//   k_adj_force_newton / k_adj_force_shear (one body, without and with the
//                  shear-rate terms of power-law materials)
//                  gather over the force neighbours: every pair term of the
//                  pressure, viscosity (each particle's own mu), XSPH and
//                  shear-rate sums, differentiated with respect to the
//                  particle's own position, velocity, P = p / rho^2 and rho,
//                  collected from both its own sums and its
//                  neighbours' sums. The neighbour relation is symmetric, so
//                  the transpose is a gather, not a scatter; then the EOS
//                  p = k max(0, rho - rho0) backwards (dL/dk).
//   k_adj_density  gather over the density neighbours: rho = sum m W(r) backwards.
//
// Neighbour sets and visiting order are the forward's (shared accept tests in
// sph_device.h) and all sums are gathers, so the adjoint is deterministic.
// Sums are plain FP32 within a particle; parameter gradients are reduced in
// FP64 over fixed per-tile partials.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "sph_types.h"
#include "sph_device.h"
#include "sph_kernels.h"

namespace phy {

__device__ __forceinline__ double block_sum_f64(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
    return t;
}

// Integrate + bounds backwards (world.rs:352-397), the acc = F / rho split and
// the effective viscosity backwards: mu = visc_k[m] max(shear, shear_min)^(n-1)
// (world.rs:319-327) gives dL/d visc_k[m], dL/d visc_n[m] (FP64 tile partials
// per material) and the cotangent of the shear-rate sum.
__global__ __launch_bounds__(256) void k_adj_local(SceneTable T, AdjSlot S,
                                                   const float4* __restrict__ xbar_in,
                                                   const float4* __restrict__ vbar_in, float4* Fb,
                                                   float4* Sb, float4* xb_loc, float4* vb_loc,
                                                   float* rhob_loc, float2* mush, double* part) {
    __shared__ double sh[4];
    const TileLoc TL = tile_loc(T);
    const int i = TL.i;
    const DevParams P = T.params[TL.s];
    double kb = 0.0, nb = 0.0;  // dL/d visc_k[m], dL/d visc_n[m] of this particle's material
    int m = 0;
    if (TL.act) {
        const uint32_t id = S.id[i];
        const float4 xo = xbar_in[id], vo = vbar_in[id];
        const float4 xs = S.xsph[i];
        const uint32_t clamp = __float_as_uint(xs.w);
        float xb[3] = {xo.x, xo.y, xo.z}, vb[3] = {vo.x, vo.y, vo.z};
        for (int a = 0; a < 3; ++a) {
            if (clamp & (1u << a)) {  // x' = wall, v' = -damp v'_pre
                xb[a] = 0.f;
                vb[a] = -P.damp * vb[a];
            }
        }
        float vt[3], ab[3], sb[3];
        for (int a = 0; a < 3; ++a) {
            vt[a] = vb[a] + xb[a] * S.dt;  // x' = x + v' dt
            ab[a] = vt[a] * S.dt;          // v' = v + (acc + g) dt + eps xsph
            sb[a] = vt[a] * P.xsph_eps;
        }
        const float rho = S.rho[i];
        const float4 acc = S.acc[i];
        const float mu = S.mu[i];
        float F[3] = {0.f, 0.f, 0.f}, rb = 0.f, shb = 0.f;
        if (P.n_mat > 1) {
            const uint32_t mt = S.mat[i];
            m = mt < (uint32_t)P.n_mat ? (int)mt : 0;
        }
        if (rho > 0.f) {  // acc = (f_press + f_visc) / rho (world.rs:332-336)
            for (int a = 0; a < 3; ++a) F[a] = ab[a] / rho;
            rb = -(ab[0] * acc.x + ab[1] * acc.y + ab[2] * acc.z) / rho;
            const float4 fv = S.fvisc[i];  // viscous force per unit mu; w: shear sum
            const double mub = (double)F[0] * fv.x + (double)F[1] * fv.y + (double)F[2] * fv.z;
            const DevParams* Pm = T.params + TL.s;  // dynamic material index: read through the table
            const double n = Pm->visc_n[m];
            const double sreg = fv.w > P.shear_min ? (double)fv.w : (double)P.shear_min;
            kb = mub * pow(sreg, n - 1.0);                   // d mu / d k = sreg^(n-1)
            nb = mub * (double)mu * log(sreg);                // d mu / d n = mu ln sreg
            if (fv.w > P.shear_min) shb = (float)(mub * (double)mu * (n - 1.0) / (double)fv.w);
        }
        const float prs = S.prs[i];
        Fb[i] = make_float4(F[0], F[1], F[2], prs / (rho * rho));
        Sb[i] = make_float4(sb[0], sb[1], sb[2], 1.f / rho);
        xb_loc[i] = make_float4(xb[0], xb[1], xb[2], 0.f);
        vb_loc[i] = make_float4(vt[0], vt[1], vt[2], 0.f);
        rhob_loc[i] = rb;
        mush[i] = make_float2(mu, shb);
    }
    // per-material tile partials (materials are uniform per scene table, so
    // every thread runs the same loop)
    const int nm = P.n_mat;
    for (int q = 0; q < nm; ++q) {
        const double tk = block_sum_f64(m == q ? kb : 0.0, sh);
        const double tn = block_sum_f64(m == q ? nb : 0.0, sh);
        if (threadIdx.x == 0) {
            part[(size_t)TL.t * kAdjParams + 1 + q] = tk;
            part[(size_t)TL.t * kAdjParams + 1 + kMaxMaterials + q] = tn;
        }
    }
}

struct AdjNbr {
    float4 p, v, F, s;
    float2 ms;  // (mu, shear cotangent)
};

template <bool SHEAR>
struct ForceAdjAcc {
    float A1;                     // sum m_b G (F_a - F_b) . D
    float A2[3];                  // sum m_b (P_a + P_b) [G dF + (G'/r)(dF . D) D]
    float A3[3], A4;              // sum mu_b m_b V F_b ; sum (m_b / rho_b) V
    float A5;                     // sum mu_b m_b V (F_b . dv)
    float A6[3];                  // sum m_b [mu_a (F_a . dv) / rho_b - mu_b (F_b . dv) / rho_a] D / r
    float A7, A8[3], A9;          // sum (m_b / rho_b) W ; sum W s_b ; sum W (s_b . dv)
    float A10[3];                 // sum (W'/r) [(m_b/rho_b)(s_a . dv) - (m_a/rho_a)(s_b . dv)] D
    // shear-rate sum sh_a = sum |dv| (m_b / rho_b) / (r + eps), backwards:
    float C1[3];                  // sum (shb_a mr_b + shb_b mr_a) / (r + eps) dv / |dv|
    float C2[3];                  // sum (shb_a mr_b + shb_b mr_a) |dv| D / (r (r + eps)^2)
    float C3;                     // sum shb_b |dv| / (r + eps)
};

template <int L, bool SHEAR>
__device__ __forceinline__ void adj_force_body(
    SceneTable T, AdjSlot S, const float4* __restrict__ Fb, const float4* __restrict__ Sb,
    const float4* __restrict__ xb_loc, const float4* __restrict__ vb_loc, const float* __restrict__ rhob_loc,
    const float2* __restrict__ mush, const float* __restrict__ rhob_seed, float* rhob_out, float4* xb_acc,
    float4* vbar_out, double* part) {
    __shared__ uint32_t lst[L][256];
    __shared__ double sh[4];
    const int tid = threadIdx.x;
    const TileLoc TL = tile_loc(T);
    const int a = TL.i;
    const bool act = TL.act;
    const Domain d = T.dom[TL.s];
    const DevParams P = T.params[TL.s];
    const uint32_t* __restrict__ start = S.start + d.toff;
    float4 pa = make_float4(0.f, 0.f, 0.f, 0.f), va = pa, Fa = pa, sa = pa;
    float2 msa = make_float2(0.f, 0.f);
    uint32_t ca = d.ncells;
    if (act) {
        pa = S.posm[a];
        va = S.velm[a];
        Fa = Fb[a];
        sa = Sb[a];
        msa = mush[a];
        ca = S.cell[a];
    }
    const float ma = pa.w, ira = sa.w, mra = va.w;  // m_a, 1/rho_a, m_a/rho_a
    const float mua = msa.x, shba = msa.y;
    const float h = P.h, h2 = P.h2;
    ForceAdjAcc<SHEAR> A = {};
    // per-lane rounds of six candidates, as in the forward force pass
    gather<L, kAllWaves, 2, false, true, 6>(
        act, ca, pa.x, pa.y, pa.z, act ? (uint32_t)a : 0u, d, P.inv_cell, zwin(P, d), ColLocal{}, start, S.posm, lst, tid,
        [&](uint32_t j, const float4& pj) { return force_maybe(P, pa.x, pa.y, pa.z, pj); },
        [&](uint32_t o) {
            return AdjNbr{at(S.posm, o), at(S.velm, o), at(Fb, o), at(Sb, o), at(mush, o)};
        },
        // a slot at or past the lane's count names the lane itself, which the
        // validity test rejects, so the listed flag is not needed
        [&](uint32_t o, const AdjNbr& nb, bool, auto) {
            const float Dx = nb.p.x - pa.x, Dy = nb.p.y - pa.y, Dz = nb.p.z - pa.z;
            float s = Dx * Dx + Dy * Dy + Dz * Dz;
            const bool valid = force_valid(P, true, (uint32_t)a << 4, o, pa.x, pa.y, pa.z, nb.p, s);
            s = __builtin_elementwise_maximum(s, 1.17549435e-38f);
            // padding slot (the particle itself, D = 0): evaluate the kernels at
            // r = h/2 so every term stays finite before its zero weight applies
            s = valid ? s : 0.25f * h2;
            const float ir = __builtin_amdgcn_rsqf(s);
            const float r = s * ir;
            const float hr = h - r;
            const float mb = nb.p.w, mrb = nb.v.w, irb = nb.s.w, mub = nb.ms.x;
            const float dvx = nb.v.x - va.x, dvy = nb.v.y - va.y, dvz = nb.v.z - va.z;
            const float dFx = nb.F.x - Fa.x, dFy = nb.F.y - Fa.y, dFz = nb.F.z - Fa.z;
            const float dFD = dFx * Dx + dFy * Dy + dFz * Dz;
            // kernels: G = cs (h-r)^2 / r, G'/r = -cs (h-r)(h+r) / r^3,
            // V = cv (h - r), W = c6 (h^2-r^2)^3, W'/r = -6 c6 (h^2-r^2)^2
            const float G = P.cs * hr * hr * ir;
            const float Gpr = -P.cs * hr * (h + r) * ir * ir * ir;
            const float V = P.cv * hr;
            const float q = h2 - s;
            const float W = P.c6 * q * q * q;
            const float Wpr = -6.f * P.c6 * q * q;
            const float Fb_dv = nb.F.x * dvx + nb.F.y * dvy + nb.F.z * dvz;
            const float Fa_dv = Fa.x * dvx + Fa.y * dvy + Fa.z * dvz;
            const float sb_dv = nb.s.x * dvx + nb.s.y * dvy + nb.s.z * dvz;
            const float sa_dv = sa.x * dvx + sa.y * dvy + sa.z * dvz;
            const float w = valid ? 1.f : 0.f;  // padding slot: the particle itself, finite
            const float mG = w * mb * G;
            const float PP = Fa.w + nb.F.w;
            A.A1 += -mG * dFD;
            const float c2a = w * mb * PP * G, c2b = w * mb * PP * Gpr * dFD;
            A.A2[0] += c2a * dFx + c2b * Dx;
            A.A2[1] += c2a * dFy + c2b * Dy;
            A.A2[2] += c2a * dFz + c2b * Dz;
            const float mV = w * mb * V;
            const float muV = mub * mV;
            A.A3[0] += muV * nb.F.x;
            A.A3[1] += muV * nb.F.y;
            A.A3[2] += muV * nb.F.z;
            A.A4 += w * mrb * V;
            A.A5 += muV * Fb_dv;
            const float c6v = w * mb * (mua * Fa_dv * irb - mub * Fb_dv * ira) * ir;
            A.A6[0] += c6v * Dx;
            A.A6[1] += c6v * Dy;
            A.A6[2] += c6v * Dz;
            const float wW = w * W;
            A.A7 += wW * mrb;
            A.A8[0] += wW * nb.s.x;
            A.A8[1] += wW * nb.s.y;
            A.A8[2] += wW * nb.s.z;
            A.A9 += wW * sb_dv;
            const float c10 = w * Wpr * (mrb * sa_dv - mra * sb_dv);
            A.A10[0] += c10 * Dx;
            A.A10[1] += c10 * Dy;
            A.A10[2] += c10 * Dz;
            if (SHEAR) {
                // |dv| is not differentiable at 0: its subgradient 0 there
                const float dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
                const float idv = dv2 > 0.f ? __builtin_amdgcn_rsqf(dv2) : 0.f;
                const float adv = dv2 * idv;
                const float ire = 1.f / (r + P.r_eps);
                const float cs1 = w * (shba * mrb + nb.ms.y * mra) * ire;
                A.C1[0] += cs1 * idv * dvx;
                A.C1[1] += cs1 * idv * dvy;
                A.C1[2] += cs1 * idv * dvz;
                const float cs2 = cs1 * adv * ire * ir;
                A.C2[0] += cs2 * Dx;
                A.C2[1] += cs2 * Dy;
                A.C2[2] += cs2 * Dz;
                A.C3 += w * nb.ms.y * adv * ire;
            }
        },
        [&](auto) {}, [] { return false; });
    double kb = 0.0;
    if (act) {
        const float4 xl = xb_loc[a], vl = vb_loc[a];
        const float rho = S.rho[a];
        float xb[3], vb[3];
        const float mcv = ma * P.cv;
        xb[0] = xl.x + ma * A.A2[0] + mcv * A.A6[0] - A.A10[0];
        xb[1] = xl.y + ma * A.A2[1] + mcv * A.A6[1] - A.A10[1];
        xb[2] = xl.z + ma * A.A2[2] + mcv * A.A6[2] - A.A10[2];
        vb[0] = vl.x + ma * (A.A3[0] * ira - mua * Fa.x * A.A4) - sa.x * A.A7 + mra * A.A8[0];
        vb[1] = vl.y + ma * (A.A3[1] * ira - mua * Fa.y * A.A4) - sa.y * A.A7 + mra * A.A8[1];
        vb[2] = vl.z + ma * (A.A3[2] * ira - mua * Fa.z * A.A4) - sa.z * A.A7 + mra * A.A8[2];
        float rb = rhob_loc[a] + (ma * ira * ira) * A.A5 + (ma * ira * ira) * A.A9;
        if (SHEAR) {
            for (int c = 0; c < 3; ++c) {
                xb[c] += A.C2[c];
                vb[c] -= A.C1[c];
            }
            rb -= (ma * ira * ira) * A.C3;
        }
        // EOS backwards: P = p / rho^2, p = k max(0, rho - rho0)
        const float Pbar = ma * A.A1;
        const float pbar = Pbar * ira * ira;
        const float prs = S.prs[a];
        rb += -2.f * Pbar * prs * ira * ira * ira;
        const float over = rho - P.rest;
        if (over > 0.f) {
            rb += P.stiff * pbar;
            kb = (double)pbar * (double)over;
        }
        const uint32_t id = S.id[a];
        if (rhob_seed) rb += rhob_seed[id];
        rhob_out[a] = rb;
        xb_acc[a] = make_float4(xb[0], xb[1], xb[2], 0.f);
        vbar_out[id] = make_float4(vb[0], vb[1], vb[2], 0.f);
    }
    kb = block_sum_f64(kb, sh);
    if (tid == 0) part[(size_t)TL.t * kAdjParams] = kb;
}

// Newtonian materials: held to four waves per SIMD (128 VGPRs, one value
// spilled), which the unconstrained build misses by one register; the
// shear-rate terms of power-law materials would spill far more, so that
// kernel keeps three waves.
template <int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_adj_force_newton(
    SceneTable T, AdjSlot S, const float4* __restrict__ Fb, const float4* __restrict__ Sb,
    const float4* __restrict__ xb_loc, const float4* __restrict__ vb_loc, const float* __restrict__ rhob_loc,
    const float2* __restrict__ mush, const float* __restrict__ rhob_seed, float* rhob_out, float4* xb_acc,
    float4* vbar_out, double* part) {
    adj_force_body<L, false>(T, S, Fb, Sb, xb_loc, vb_loc, rhob_loc, mush, rhob_seed, rhob_out, xb_acc, vbar_out, part);
}

template <int L>
__global__ __launch_bounds__(256) void k_adj_force_shear(
    SceneTable T, AdjSlot S, const float4* __restrict__ Fb, const float4* __restrict__ Sb,
    const float4* __restrict__ xb_loc, const float4* __restrict__ vb_loc, const float* __restrict__ rhob_loc,
    const float2* __restrict__ mush, const float* __restrict__ rhob_seed, float* rhob_out, float4* xb_acc,
    float4* vbar_out, double* part) {
    adj_force_body<L, true>(T, S, Fb, Sb, xb_loc, vb_loc, rhob_loc, mush, rhob_seed, rhob_out, xb_acc, vbar_out, part);
}

template <int L>
__global__ __launch_bounds__(256) void k_adj_density(SceneTable T, AdjSlot S,
                                                     const float* __restrict__ rhob,
                                                     const float4* __restrict__ xb_acc,
                                                     float4* xbar_out) {
    __shared__ uint32_t lst[L][256];
    const int tid = threadIdx.x;
    const TileLoc TL = tile_loc(T);
    const int a = TL.i;
    const bool act = TL.act;
    const Domain d = T.dom[TL.s];
    const DevParams P = T.params[TL.s];
    const uint32_t* __restrict__ start = S.start + d.toff;
    float4 pa = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t ca = d.ncells;
    if (act) {
        pa = S.posm[a];
        ca = S.cell[a];
    }
    float B1[3] = {0.f, 0.f, 0.f}, B2[3] = {0.f, 0.f, 0.f};
    struct DN {
        float4 p;
        float rb;
    };
    gather<L, kAllWaves, 2, false, true, 6>(
        act, ca, pa.x, pa.y, pa.z, act ? (uint32_t)a : 0u, d, P.inv_cell, zwin(P, d), ColLocal{}, start, S.posm, lst, tid,
        [&](uint32_t j, const float4& pj) { return density_maybe(P, pa.x, pa.y, pa.z, pj); },
        [&](uint32_t o) { return DN{at(S.posm, o), at(rhob, o)}; },
        [&](uint32_t o, const DN& nb, bool listed, auto) {
            const float Dx = nb.p.x - pa.x, Dy = nb.p.y - pa.y, Dz = nb.p.z - pa.z;
            const float s = Dx * Dx + Dy * Dy + Dz * Dz;
            const bool valid = density_valid(P, listed, pa.x, pa.y, pa.z, nb.p, s);
            const float q = P.h2 - s;
            const float Wpr = valid ? -6.f * P.c6 * q * q : 0.f;  // W'(r) / r
            const float c1 = Wpr * nb.p.w, c2 = Wpr * nb.rb;
            B1[0] += c1 * Dx;
            B1[1] += c1 * Dy;
            B1[2] += c1 * Dz;
            B2[0] += c2 * Dx;
            B2[1] += c2 * Dy;
            B2[2] += c2 * Dz;
        },
        [&](auto) {}, [] { return false; });
    if (!act) return;
    // rho_a = sum_b m_b W(r_ab): x_bar_a -= W'/r (rho_bar_a m_b + rho_bar_b m_a) D
    const float ra = rhob[a], ma = pa.w;
    const float4 xa = xb_acc[a];
    xbar_out[S.id[a]] = make_float4(xa.x - (ra * B1[0] + ma * B2[0]), xa.y - (ra * B1[1] + ma * B2[1]),
                                    xa.z - (ra * B1[2] + ma * B2[2]), 0.f);
}

// Adds one step's tile partials to the running FP64 totals of each scene
// (one workgroup per scene, fixed summation shape). Materials beyond the
// scene's table have no partials and stay zero.
__global__ __launch_bounds__(256) void k_adj_reduce(SceneTable T, const double* __restrict__ part,
                                                    double* totals) {
    __shared__ double sh[256];
    const SceneInfo si = T.info[blockIdx.x];
    const int nm = T.params[blockIdx.x].n_mat;
    for (int q = 0; q < kAdjParams; ++q) {
        const int mq = q == 0 ? -1 : (q - 1) % kMaxMaterials;
        if (mq >= nm) continue;  // uniform over the workgroup
        double v = 0.0;
        for (uint32_t t = threadIdx.x; t < si.ntiles; t += 256) v += part[(size_t)(si.tile0 + t) * kAdjParams + q];
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) totals[(size_t)blockIdx.x * kAdjParams + q] += sh[0];
        __syncthreads();
    }
}

void launch_adjoint_step(hipStream_t st, const SceneTable& T, const AdjSlot& S, const AdjBuffers& B,
                         const float* rhob_seed, bool shear) {
    if (T.ntiles > 0) {
        const int nb = T.ntiles;
        hipLaunchKernelGGL(k_adj_local, dim3(nb), dim3(256), 0, st, T, S, B.xbar_in, B.vbar_in, B.Fb,
                           B.Sb, B.xb_loc, B.vb_loc, B.rhob_loc, B.mush, B.part);
        if (shear)
            hipLaunchKernelGGL((k_adj_force_shear<kAdjList>), dim3(nb), dim3(256), 0, st, T, S, B.Fb,
                               B.Sb, B.xb_loc, B.vb_loc, B.rhob_loc, B.mush, rhob_seed, B.rhob, B.xb_acc,
                               B.vbar_out, B.part);
        else
            hipLaunchKernelGGL((k_adj_force_newton<kAdjList>), dim3(nb), dim3(256), 0, st, T, S, B.Fb,
                               B.Sb, B.xb_loc, B.vb_loc, B.rhob_loc, B.mush, rhob_seed, B.rhob, B.xb_acc,
                               B.vbar_out, B.part);
        hipLaunchKernelGGL(k_adj_density<kAdjList>, dim3(nb), dim3(256), 0, st, T, S, B.rhob,
                           B.xb_acc, B.xbar_out);
    }
    hipLaunchKernelGGL(k_adj_reduce, dim3(T.nscenes), dim3(256), 0, st, T, B.part, B.totals);
}

}  // namespace phy
