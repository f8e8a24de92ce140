// sph_couple.hip — fluid coupling on the device (crates/phy-fluid/src/sph/
// world.rs:400-773): rigid bodies (couple_bodies, push_out, body_volume),
// point masses (couple_points), a heat grid (couple_heat) and the
// nearest-particle velocity sampler that couple_heat installs.
//
// The particle state is FP32; the coupling arithmetic is f64 on it and this
// file is compiled with -ffp-contract=off, so every decision (containment,
// the sign of a normal velocity) and every f64 value matches the reference's
// operation-by-operation rounding on the same state. Results are rounded to
// FP32 once, when stored.
//
// Rigid-body geometry follows phy-rigid shape.rs:198-330 (support point,
// containment) and 433-463 (pose transforms, linear impulse). Quaternion
// rotation is nalgebra's UnitQuaternion * Vector3: t = 2 (q_v x v),
// v' = (t w + q_v x t) + v.

#include "sph_couple.h"
#include "sph_device.h"
#include "sph_kernels.h"

namespace phy {

namespace {

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 operator*(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ D3 operator/(D3 a, double s) { return d3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ D3 neg(D3 a) { return d3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ double dot(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ double norm(D3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ D3 normalize(D3 a) { return a / norm(a); }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
    return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ D3 ld3(const double* p) { return d3(p[0], p[1], p[2]); }
__device__ __forceinline__ void st3(double* p, D3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

// q = (w, i, j, k); conj flips the vector part (UnitQuaternion::inverse)
__device__ __forceinline__ D3 qrot(const double* q, D3 v, bool conj) {
    const double s = conj ? -1.0 : 1.0;
    const D3 u = d3(q[1] * s, q[2] * s, q[3] * s);
    const D3 t = cross(u, v) * 2.0;
    const D3 c = cross(u, t);
    return (t * q[0] + c) + v;
}

__device__ __forceinline__ D3 to_local(const CoupleBody& b, D3 p) { return qrot(b.q, p - ld3(b.pos), true); }
__device__ __forceinline__ D3 to_world(const CoupleBody& b, D3 p) { return qrot(b.q, p, false) + ld3(b.pos); }

__device__ bool contains_local(const CoupleBody& b, D3 p) {
    switch (b.shape) {
        case kShapeSphere: return norm(p) <= b.dims[0];
        case kShapeBox: return fabs(p.x) <= b.dims[0] && fabs(p.y) <= b.dims[1] && fabs(p.z) <= b.dims[2];
        case kShapeCapsule: {
            const double hh = b.dims[0], r = b.dims[1];
            double y = p.y;
            if (y < -hh) y = -hh;
            else if (y > hh) y = hh;
            const double cx = (p.x * p.x + (p.y - y) * (p.y - y)) + p.z * p.z;
            return cx <= r * r;
        }
        default: return false;
    }
}

__device__ D3 support_local(const CoupleBody& b, D3 dir) {
    switch (b.shape) {
        case kShapeSphere: return normalize(dir) * b.dims[0];
        case kShapeBox:
            return d3(dir.x >= 0.0 ? b.dims[0] : -b.dims[0], dir.y >= 0.0 ? b.dims[1] : -b.dims[1],
                      dir.z >= 0.0 ? b.dims[2] : -b.dims[2]);
        case kShapeCapsule: {
            const D3 n = normalize(dir);
            D3 end = d3(0.0, 0.0, 0.0);
            if (dir.y > 0.0) end = d3(0.0, b.dims[0], 0.0);
            else if (dir.y < 0.0) end = d3(0.0, -b.dims[0], 0.0);
            return end + n * b.dims[1];
        }
        default: return d3(0.0, 0.0, 0.0);
    }
}

// world.rs:507-540
__device__ double body_volume(const CoupleBody& b) {
    const double pi = 3.141592653589793;
    switch (b.shape) {
        case kShapeSphere: {
            const double r = b.dims[0];
            return 4.0 / 3.0 * pi * r * r * r;
        }
        case kShapeBox: return b.dims[0] * b.dims[1] * b.dims[2] * 8.0;
        case kShapeCapsule: {
            const double hh = b.dims[0], r = b.dims[1];
            return pi * r * r * 2.0 * hh + 4.0 / 3.0 * pi * r * r * r;
        }
        default: return 0.0;
    }
}

// world.rs:755-773 on a particle (pos, vel) and the body's current velocity
__device__ void push_out(D3& pos, D3& vel, const CoupleBody& b, D3 bvel) {
    const D3 local = to_local(b, pos);
    const D3 dir = norm(local) > 1e-9 ? normalize(local) : d3(0.0, 1.0, 0.0);
    const D3 sd = support_local(b, dir);
    pos = to_world(b, sd);
    const D3 v_rel = vel - bvel;
    const D3 n = normalize(to_world(b, normalize(sd)) - ld3(b.pos));
    const double vn = dot(v_rel, n);
    if (vn < 0.0) vel = vel - n * vn;
}

__device__ __forceinline__ bool body_contains(const CoupleBody& b, D3 pw) {
    return contains_local(b, to_local(b, pw));
}

inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------------------
// bodies

__global__ __launch_bounds__(256) void k_contact_count(int n, const float4* __restrict__ posm,
                                                       const uint32_t* __restrict__ id,
                                                       const CoupleBody* __restrict__ bodies, int nb,
                                                       uint32_t* cnt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 p = posm[i];
    const D3 pw = d3(p.x, p.y, p.z);
    uint32_t c = 0;
    for (int b = 0; b < nb; ++b) c += body_contains(bodies[b], pw) ? 1u : 0u;
    cnt[id[i]] = c;
}

__global__ __launch_bounds__(256) void k_contact_fill(int n, const float4* __restrict__ posm,
                                                      const uint32_t* __restrict__ id,
                                                      const CoupleBody* __restrict__ bodies, int nb,
                                                      const uint32_t* __restrict__ off, Contact* out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 p = posm[i];
    const D3 pw = d3(p.x, p.y, p.z);
    uint32_t k = off[id[i]];
    for (int b = 0; b < nb; ++b)
        if (body_contains(bodies[b], pw)) out[k++] = Contact{(uint32_t)i, (uint32_t)b};
}

// The contacts of world.rs:427-491 in the reference's order: particles by
// original index, bodies by index. Each contact depends on the particle's and
// the body's state left by the contacts before it, so one lane resolves them
// in order while the wave prefetches the next 64 contacts' particles. A
// particle's contacts are consecutive and its state stays in registers
// between them. The contact impulse on the particle is dropped, as in the
// reference: push_out's velocity replaces the particle's working copy
// (world.rs:488) before it is stored.
__global__ __launch_bounds__(64) void k_contact_resolve(const Contact* __restrict__ contacts,
                                                        const uint32_t* __restrict__ total, float4* posm,
                                                        float4* vel, CoupleBody* bodies, int nb,
                                                        double* displaced, CoupleParams cp, double dt) {
    __shared__ Contact cs[64];
    __shared__ float4 ps[64], vs[64];
    const int lane = threadIdx.x;
    const uint32_t K = *total;
    const double particle_vol = cp.mass / cp.rest;
    uint32_t cur = 0xffffffffu;
    D3 pos = d3(0.0, 0.0, 0.0), v = pos, p_world = pos;
    double m_p = 0.0;
    for (uint32_t base = 0; base < K; base += 64) {
        const uint32_t k = base + lane;
        if (k < K) {
            const Contact c = contacts[k];
            cs[lane] = c;
            ps[lane] = posm[c.slot];
            vs[lane] = vel[c.slot];
        }
        __syncthreads();
        if (lane == 0) {
            const uint32_t m = K - base < 64u ? K - base : 64u;
            for (uint32_t t = 0; t < m; ++t) {
                const Contact c = cs[t];
                if (c.slot != cur) {
                    if (cur != 0xffffffffu) {
                        posm[cur] = make_float4((float)pos.x, (float)pos.y, (float)pos.z, (float)m_p);
                        vel[cur] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.f);
                    }
                    cur = c.slot;
                    pos = d3(ps[t].x, ps[t].y, ps[t].z);
                    v = d3(vs[t].x, vs[t].y, vs[t].z);
                    m_p = ps[t].w;
                    p_world = pos;
                }
                const CoupleBody& b = bodies[c.body];
                const D3 local = to_local(b, p_world);
                const D3 n = norm(local) > 1e-9 ? normalize(local) : d3(0.0, 1.0, 0.0);
                D3 bvel = ld3(b.vel);
                if (b.inv_mass <= 0.0) {
                    push_out(pos, v, b, bvel);
                    continue;
                }
                const double v_rel_n = dot(v - bvel, n);
                if (v_rel_n < 0.0) {
                    const double inv_sum = (1.0 / m_p) + b.inv_mass;
                    const double j_mag = -(1.0 + 0.0) * v_rel_n / inv_sum;  // restitution 0
                    const D3 j = n * j_mag;
                    bvel = bvel + neg(j) * b.inv_mass;  // apply_impulse(-j)
                    st3(bodies[c.body].vel, bvel);
                }
                displaced[c.body] += particle_vol;
                push_out(pos, v, b, bvel);
            }
        }
        __syncthreads();
    }
    if (lane == 0 && cur != 0xffffffffu) {
        posm[cur] = make_float4((float)pos.x, (float)pos.y, (float)pos.z, (float)m_p);
        vel[cur] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.f);
    }
    __syncthreads();
    // buoyancy of each dynamic body (world.rs:493-502)
    const D3 g = ld3(cp.g);
    for (int b = lane; b < nb; b += 64) {
        CoupleBody& B = bodies[b];
        if (B.inv_mass <= 0.0) continue;
        const double v_sub = displaced[b] > 0.0 ? body_volume(B) : 0.0;
        const D3 buoy = neg(g) * (cp.rest * v_sub);
        st3(B.vel, ld3(B.vel) + (buoy * dt) * B.inv_mass);
    }
}

void launch_contact_count(hipStream_t st, int n, const float4* posm, const uint32_t* id,
                          const CoupleBody* bodies, int nb, uint32_t* cnt) {
    if (n > 0)
        hipLaunchKernelGGL(k_contact_count, dim3(nblk(n, 256)), dim3(256), 0, st, n, posm, id, bodies, nb,
                           cnt);
}

void launch_contact_fill(hipStream_t st, int n, const float4* posm, const uint32_t* id,
                         const CoupleBody* bodies, int nb, const uint32_t* off, Contact* out) {
    if (n > 0)
        hipLaunchKernelGGL(k_contact_fill, dim3(nblk(n, 256)), dim3(256), 0, st, n, posm, id, bodies, nb,
                           off, out);
}

void launch_contact_resolve(hipStream_t st, const Contact* c, const uint32_t* total, float4* posm,
                            float4* vel, CoupleBody* bodies, int nb, double* displaced,
                            const CoupleParams& cp, double dt) {
    hipLaunchKernelGGL(k_contact_resolve, dim3(1), dim3(64), 0, st, c, total, posm, vel, bodies, nb,
                       displaced, cp, dt);
}

// ---------------------------------------------------------------------------
// points

__global__ __launch_bounds__(256) void k_inverse_perm(int n, const uint32_t* __restrict__ id,
                                                      uint32_t* inv) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) inv[id[i]] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_widen_vel(int n, const float4* __restrict__ vel, double* v64) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 v = vel[i];
    v64[3 * i] = v.x;
    v64[3 * i + 1] = v.y;
    v64[3 * i + 2] = v.z;
}

__global__ __launch_bounds__(256) void k_narrow_vel(int n, const double* __restrict__ v64, float4* vel) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    vel[i] = make_float4((float)v64[3 * i], (float)v64[3 * i + 1], (float)v64[3 * i + 2], 0.f);
}

// the reference key of an f64 coordinate (grid.rs:26-33)
__device__ __forceinline__ long long key_d(double x, double inv) {
    const double f = floor(x * inv);
    if (f >= -9223372036854775808.0 && f < 9223372036854775808.0) return (long long)f;
    return 0;
}

// The dense runs of a point's 27 reference cells (as cells of the table, up
// to 36 z-runs for half cells, 9 for cells of h) plus the overflow segment
// (checked by key adjacency), flattened for the workgroup.
struct Runs {
    uint32_t a[37], len[37];
    uint32_t total;
    int nr;
};

__device__ void point_runs(const D3& q, const CoupleParams& cp, const uint32_t* __restrict__ start,
                           const Domain& d, Runs& R) {
    long long lo[3], hi[3];
    ref_block(key_d(q.x, cp.inv_cell), d, 0, lo[0], hi[0]);
    ref_block(key_d(q.y, cp.inv_cell), d, 1, lo[1], hi[1]);
    ref_block(key_d(q.z, cp.inv_cell), d, 2, lo[2], hi[2]);
    R.nr = 0;
    R.total = 0;
    const int rows = block_span(d) * block_span(d);
    for (int k = 0; k < rows; ++k) {
        uint32_t a, b;
        block_run(k, lo, hi, d, start, a, b);
        if (b > a) {
            R.a[R.nr] = a;
            R.len[R.nr] = b - a;
            R.total += b - a;
            ++R.nr;
        }
    }
    const uint32_t o0 = start[d.ncells], o1 = start[d.ncells + 1];
    R.a[36] = o0;
    R.len[36] = o1 - o0;
}

__device__ __forceinline__ bool point_adjacent(const D3& q, const float4& p, double inv) {
    const unsigned long long a =
        (unsigned long long)key1(p.x, inv) - (unsigned long long)key_d(q.x, inv) + 1ull;
    const unsigned long long b =
        (unsigned long long)key1(p.y, inv) - (unsigned long long)key_d(q.y, inv) + 1ull;
    const unsigned long long c =
        (unsigned long long)key1(p.z, inv) - (unsigned long long)key_d(q.z, inv) + 1ull;
    return a <= 2ull && b <= 2ull && c <= 2ull;
}

// Calls f(sorted slot) for this thread's share of the point's candidates.
template <class F>
__device__ __forceinline__ void for_point_candidates(const Runs& R, const D3& q, double inv,
                                                     const float4* __restrict__ posm_s, F&& f) {
    for (uint32_t t = threadIdx.x; t < R.total; t += blockDim.x) {
        uint32_t r = 0, u = t;
        while (u >= R.len[r]) u -= R.len[r++];
        f(R.a[r] + u);
    }
    for (uint32_t t = threadIdx.x; t < R.len[36]; t += blockDim.x) {
        const uint32_t j = R.a[36] + t;
        if (point_adjacent(q, posm_s[j], inv)) f(j);
    }
}

// fixed-shape tree over the workgroup (deterministic)
__device__ double block_sum_d(double v, double* sh) {
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

// world.rs:703-751, one workgroup walking the points in order (a point sees
// the velocity changes of the points before it). The neighbourhood sums use
// a fixed-shape tree instead of the reference's sequential order (same set,
// rounding differs at the ulp level); the reference's two passes over the
// neighbourhood give the same mass sum, so one is kept.
__global__ __launch_bounds__(256) void k_couple_points(CouplePoint* pts, int np, CoupleParams cp, double dt,
                                                       double drag, double soft_density,
                                                       const float4* __restrict__ posm_s,
                                                       const uint32_t* __restrict__ id_s,
                                                       const uint32_t* __restrict__ start,
                                                       const Domain* __restrict__ dom,
                                                       const uint32_t* __restrict__ inv, double* vel64) {
    __shared__ double sh[256];
    __shared__ Runs R;
    __shared__ double react[4];
    const Domain d = dom[0];
    start += d.toff;
    const D3 g = ld3(cp.g);
    for (int k = 0; k < np; ++k) {
        const D3 q = ld3(pts[k].pos);
        if (q.x < cp.lo[0] || q.x > cp.hi[0] || q.y < cp.lo[1] || q.y > cp.hi[1] || q.z < cp.lo[2] ||
            q.z > cp.hi[2])
            continue;
        if (threadIdx.x == 0) point_runs(q, cp, start, d, R);
        __syncthreads();
        double vx = 0.0, vy = 0.0, vz = 0.0, ws = 0.0;
        for_point_candidates(R, q, cp.inv_cell, posm_s, [&](uint32_t j) {
            const double w = posm_s[j].w;
            const double* v = vel64 + 3 * (size_t)inv[id_s[j]];
            vx += v[0] * w;
            vy += v[1] * w;
            vz += v[2] * w;
            ws += w;
        });
        vx = block_sum_d(vx, sh);
        vy = block_sum_d(vy, sh);
        vz = block_sum_d(vz, sh);
        ws = block_sum_d(ws, sh);
        if (threadIdx.x == 0) {
            const D3 v_f = ws > 0.0 ? d3(vx, vy, vz) / ws : d3(0.0, 0.0, 0.0);
            const double vol = soft_density > 0.0 ? pts[k].mass / soft_density : 0.0;
            const D3 f_b = neg(g) * (cp.rest * vol);
            const D3 f_d = (v_f - ld3(pts[k].vel)) * (drag * pts[k].mass);
            const D3 f = f_b + f_d;
            st3(pts[k].force, f);
            const D3 r = neg(f) * dt;
            react[0] = r.x;
            react[1] = r.y;
            react[2] = r.z;
        }
        __syncthreads();
        if (ws > 0.0) {
            const D3 r = d3(react[0], react[1], react[2]);
            for_point_candidates(R, q, cp.inv_cell, posm_s, [&](uint32_t j) {
                const double frac = (double)posm_s[j].w / ws;
                double* v = vel64 + 3 * (size_t)inv[id_s[j]];
                st3(v, ld3(v) + r * frac);
            });
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_gather_by_id4(int n, const uint32_t* __restrict__ id,
                                                       const float4* __restrict__ in, float4* out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = in[id[i]];
}

void launch_gather_by_id4(hipStream_t st, int n, const uint32_t* id, const float4* in, float4* out) {
    if (n > 0) hipLaunchKernelGGL(k_gather_by_id4, dim3(nblk(n, 256)), dim3(256), 0, st, n, id, in, out);
}

void launch_inverse_perm(hipStream_t st, int n, const uint32_t* id, uint32_t* inv) {
    if (n > 0) hipLaunchKernelGGL(k_inverse_perm, dim3(nblk(n, 256)), dim3(256), 0, st, n, id, inv);
}
void launch_widen_vel(hipStream_t st, int n, const float4* vel, double* vel64) {
    if (n > 0) hipLaunchKernelGGL(k_widen_vel, dim3(nblk(n, 256)), dim3(256), 0, st, n, vel, vel64);
}
void launch_narrow_vel(hipStream_t st, int n, const double* vel64, float4* vel) {
    if (n > 0) hipLaunchKernelGGL(k_narrow_vel, dim3(nblk(n, 256)), dim3(256), 0, st, n, vel64, vel);
}
void launch_couple_points(hipStream_t st, CouplePoint* pts, int np, const CoupleParams& cp, double dt,
                          double drag, double soft_density, const float4* posm_s, const uint32_t* id_s,
                          const uint32_t* start, const Domain* dom, const uint32_t* inv, double* vel64) {
    if (np > 0)
        hipLaunchKernelGGL(k_couple_points, dim3(1), dim3(256), 0, st, pts, np, cp, dt, drag, soft_density,
                           posm_s, id_s, start, dom, inv, vel64);
}

// ---------------------------------------------------------------------------
// heat grid

// usize::to_usize of a floored f64 (None -> 0)
__device__ __forceinline__ uint64_t to_usize0(double f) {
    return (f >= 0.0 && f < 18446744073709551616.0) ? (uint64_t)f : 0ull;
}

__device__ __forceinline__ double heat_at(const HeatGrid& hg, uint64_t x, uint64_t y, uint64_t z) {
    const uint64_t cx = x < hg.nx - 1 ? x : hg.nx - 1;
    const uint64_t cy = y < hg.ny - 1 ? y : hg.ny - 1;
    const uint64_t cz = z < hg.nz - 1 ? z : hg.nz - 1;
    return hg.u[cx + hg.nx * (cy + hg.ny * cz)];
}

// heat.rs:130-175
__device__ double trilinear(const HeatGrid& hg, uint64_t cx, uint64_t cy, uint64_t cz, double fx,
                            double fy, double fz) {
    const uint64_t x1 = cx + 1 < hg.nx - 1 ? cx + 1 : hg.nx - 1;
    const uint64_t y1 = cy + 1 < hg.ny - 1 ? cy + 1 : hg.ny - 1;
    const uint64_t z1 = cz + 1 < hg.nz - 1 ? cz + 1 : hg.nz - 1;
    const double c000 = heat_at(hg, cx, cy, cz), c100 = heat_at(hg, x1, cy, cz);
    const double c010 = heat_at(hg, cx, y1, cz), c110 = heat_at(hg, x1, y1, cz);
    const double c001 = heat_at(hg, cx, cy, z1), c101 = heat_at(hg, x1, cy, z1);
    const double c011 = heat_at(hg, cx, y1, z1), c111 = heat_at(hg, x1, y1, z1);
    const double x00 = c000 + (c100 - c000) * fx;
    const double x10 = c010 + (c110 - c010) * fx;
    const double x01 = c001 + (c101 - c001) * fx;
    const double x11 = c011 + (c111 - c011) * fx;
    const double y0 = x00 + (x10 - x00) * fy;
    const double y1v = x01 + (x11 - x01) * fy;
    return y0 + (y1v - y0) * fz;
}

__global__ __launch_bounds__(256) void k_heat_body(int n, const float4* __restrict__ posm,
                                                   const float4* __restrict__ vel, float4* body,
                                                   CoupleParams cp, HeatGrid hg, double dt, double t_ref,
                                                   double beta, double heat_gain, HeatScratch hs) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 p4 = posm[i];
    const D3 p = d3(p4.x, p4.y, p4.z);
    uint32_t qc = 0xffffffffu;
    if (!(p.x < cp.lo[0] || p.x > cp.hi[0] || p.y < cp.lo[1] || p.y > cp.hi[1] || p.z < cp.lo[2] ||
          p.z > cp.hi[2])) {
        const double dx = hg.dx;
        const double nx1 = (double)hg.nx - 1.0, ny1 = (double)hg.ny - 1.0, nz1 = (double)hg.nz - 1.0;
        const D3 o = ld3(hg.origin);
        const bool in = !(p.x < o.x || p.x > o.x + dx * nx1 || p.y < o.y || p.y > o.y + dx * ny1 ||
                          p.z < o.z || p.z > o.z + dx * nz1);
        double temp = t_ref;
        uint64_t cx = 0, cy = 0, cz = 0;
        if (in) {
            const double fx = floor((p.x - o.x) / dx), fy = floor((p.y - o.y) / dx),
                         fz = floor((p.z - o.z) / dx);
            cx = to_usize0(fx);
            cy = to_usize0(fy);
            cz = to_usize0(fz);
            temp = trilinear(hg, cx, cy, cz, (p.x - o.x) / dx - fx, (p.y - o.y) / dx - fy,
                             (p.z - o.z) / dx - fz);
        }
        const D3 corr = neg(ld3(cp.g)) * beta * (temp - t_ref);
        const float4 b = body[i];
        body[i] = make_float4((float)((double)b.x + corr.x), (float)((double)b.y + corr.y),
                              (float)((double)b.z + corr.z), 0.f);
        if (heat_gain > 0.0 && in && cx < hg.nx && cy < hg.ny && cz < hg.nz) {
            const float4 v4 = vel[i];
            const double speed = norm(d3(v4.x, v4.y, v4.z));
            const double q = heat_gain * speed * dt;
            qc = (uint32_t)(cx + hg.nx * (cy + hg.ny * cz));
            hs.q[i] = q;
            const double aq = fabs(q);
            if (aq <= 1.79769313486231570e308) atomicMax(hs.qmax, (unsigned long long)__double_as_longlong(aq));
        }
    }
    hs.qcell[i] = qc;
}

// Fixed-point exponent: every |q| 2^E is below 2^(125 - ceil(log2(n + 1))),
// so n of them sum below 2^125 in the 128-bit accumulators.
__global__ void k_heat_scale(int n, HeatScratch hs) {
    const double m = __longlong_as_double((long long)*hs.qmax);
    int L = 0;
    while ((1ll << L) < (long long)n + 1) ++L;
    *hs.scale = m > 0.0 ? 125 - (ilogb(m) + 1) - L : 0;
}

__global__ __launch_bounds__(256) void k_heat_accumulate(int n, HeatScratch hs) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = hs.qcell[i];
    if (c == 0xffffffffu) return;
    const double aq = fabs(hs.q[i]);
    uint32_t f = 1u;
    if (aq != aq) f |= 2u;
    else if (aq > 1.79769313486231570e308) f |= 4u;
    else {
        const double x = ldexp(aq, *hs.scale);                    // exact, < 2^125
        const double hd = floor(ldexp(x, -64));
        const unsigned long long hi = (unsigned long long)hd;
        const unsigned long long lo = (unsigned long long)(x - ldexp(hd, 64));
        const unsigned long long old = atomicAdd(&hs.lo[c], lo);
        const unsigned long long carry = old + lo < old ? 1ull : 0ull;
        if (hi + carry) atomicAdd(&hs.hi[c], hi + carry);
    }
    atomicOr(&hs.flags[c], f);
}

// q shares the sign of dt (heat_gain > 0, speed >= 0)
__global__ __launch_bounds__(256) void k_heat_finish(uint32_t ncell, double dt, HeatScratch hs, double* src) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncell) return;
    const uint32_t f = hs.flags[c];
    if (!(f & 1u)) return;
    double add;
    if (f & 2u) {
        add = __longlong_as_double(0x7ff8000000000000ll);
    } else if (f & 4u) {
        add = dt < 0.0 ? -__longlong_as_double(0x7ff0000000000000ll) : __longlong_as_double(0x7ff0000000000000ll);
    } else {
        const double s = ldexp((double)hs.hi[c], 64) + (double)hs.lo[c];
        add = ldexp(s, -*hs.scale);
        if (dt < 0.0) add = -add;
    }
    src[c] += add;
}

void launch_heat_body(hipStream_t st, int n, const float4* posm, const float4* vel, float4* body,
                      const CoupleParams& cp, const HeatGrid& hg, double dt, double t_ref, double beta,
                      double heat_gain, const HeatScratch& hs) {
    if (n > 0)
        hipLaunchKernelGGL(k_heat_body, dim3(nblk(n, 256)), dim3(256), 0, st, n, posm, vel, body, cp, hg, dt,
                           t_ref, beta, heat_gain, hs);
}

void launch_heat_source(hipStream_t st, int n, double dt, const HeatScratch& hs, uint32_t ncell,
                        double* src) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_heat_scale, dim3(1), dim3(1), 0, st, n, hs);
    hipLaunchKernelGGL(k_heat_accumulate, dim3(nblk(n, 256)), dim3(256), 0, st, n, hs);
    hipLaunchKernelGGL(k_heat_finish, dim3(nblk(ncell, 256)), dim3(256), 0, st, ncell, dt, hs, src);
}

// ---------------------------------------------------------------------------
// velocity sampler: exact nearest particle (first in index order on ties) by
// rings of table cells (edge h / cdiv) around the query's cell; rings are
// visited until the best squared distance is below ((r - 1) h / cdiv)^2, a
// full cell of margin over the smallest distance an unvisited ring can hold.
// The overflow segment is always scanned.

__device__ __forceinline__ void nearest_consider(const D3& q, const float4& p, uint32_t o, double& bd,
                                                 uint32_t& bo) {
    const double dxq = q.x - (double)p.x, dyq = q.y - (double)p.y, dzq = q.z - (double)p.z;
    const double dd = dxq * dxq + dyq * dyq + dzq * dzq;
    if (dd < bd || (dd == bd && o < bo)) {
        bd = dd;
        bo = o;
    }
}

__global__ __launch_bounds__(256) void k_nearest_velocity(int nq, const double* __restrict__ q3, double* out3,
                                                          const float4* __restrict__ posm_s,
                                                          const uint32_t* __restrict__ id_s,
                                                          const uint32_t* __restrict__ start,
                                                          const Domain* __restrict__ dom,
                                                          const uint32_t* __restrict__ inv,
                                                          const float4* __restrict__ vel, CoupleParams cp) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nq) return;
    const D3 q = ld3(q3 + 3 * (size_t)t);
    const Domain d = dom[0];
    start += d.toff;
    double bd = __longlong_as_double(0x7ff0000000000000ll);  // +inf
    uint32_t bo = 0xffffffffu;
    long long kq[3];
    const double cd = (double)d.reach;
    bool keyed = hkey(q.x * cp.inv_cell, cd, kq[0]) & hkey(q.y * cp.inv_cell, cd, kq[1]) &
                 hkey(q.z * cp.inv_cell, cd, kq[2]);
    for (int a = 0; a < 3; ++a)  // far from the table: no ring walk (and no key overflow)
        keyed = keyed && fabs((double)kq[a] - (double)d.kmin[a]) < 1099511627776.0;
    if (isfinite(q.x) && isfinite(q.y) && isfinite(q.z) && !keyed) {
        // a finite query beyond the key range or far from the table: every particle
        for (uint32_t j = start[0]; j < start[d.ncells + 1]; ++j)
            nearest_consider(q, posm_s[j], id_s[j], bd, bo);
    } else if (keyed && isfinite(q.x) && isfinite(q.y) && isfinite(q.z)) {
        // query key relative to the table, and the ring range that meets it
        long long rel[3], rmin = 0, rmax = 0;
        for (int a = 0; a < 3; ++a) {
            rel[a] = kq[a] - d.kmin[a];
            const long long below = -rel[a], above = rel[a] - (d.dims[a] - 1);
            const long long gap = below > 0 ? below : (above > 0 ? above : 0);
            rmin = gap > rmin ? gap : rmin;
            const long long far = rel[a] > d.dims[a] - 1 - rel[a] ? rel[a] : d.dims[a] - 1 - rel[a];
            rmax = far > rmax ? far : rmax;
        }
        if (d.ncells == 0) rmax = -1;
        for (long long r = rmin; r <= rmax; ++r) {
            const long long x0 = rel[0] - r > 0 ? rel[0] - r : 0;
            const long long x1 = rel[0] + r < d.dims[0] - 1 ? rel[0] + r : d.dims[0] - 1;
            const long long y0 = rel[1] - r > 0 ? rel[1] - r : 0;
            const long long y1 = rel[1] + r < d.dims[1] - 1 ? rel[1] + r : d.dims[1] - 1;
            for (long long x = x0; x <= x1; ++x)
                for (long long y = y0; y <= y1; ++y) {
                    // a row on the ring's x / y boundary takes every z of the
                    // ring; an inner row only its two z faces
                    const bool edge = (x == rel[0] - r || x == rel[0] + r || y == rel[1] - r || y == rel[1] + r);
                    long long za = rel[2] - r, zb = rel[2] + r, step = 2 * r > 0 ? 2 * r : 1;
                    if (edge) {
                        step = 1;
                        za = za > 0 ? za : 0;
                        zb = zb < d.dims[2] - 1 ? zb : d.dims[2] - 1;
                    }
                    for (long long z = za; z <= zb; z += step) {
                        if (z < 0 || z >= d.dims[2]) continue;
                        const uint32_t c = (uint32_t)((x * d.dims[1] + y) * d.dims[2] + z);
                        for (uint32_t j = start[c]; j < start[c + 1]; ++j)
                            nearest_consider(q, posm_s[j], id_s[j], bd, bo);
                    }
                }
            if (r >= 1) {
                const double m = (double)(r - 1) * (cp.h / cd);  // whole cells
                if (bd < m * m) break;
            }
        }
        for (uint32_t j = start[d.ncells]; j < start[d.ncells + 1]; ++j)
            nearest_consider(q, posm_s[j], id_s[j], bd, bo);
    }
    if (bo == 0xffffffffu) bo = 0;  // no distance below infinity: particle 0
    const float4 v = vel[inv[bo]];
    out3[3 * (size_t)t] = v.x;
    out3[3 * (size_t)t + 1] = v.y;
    out3[3 * (size_t)t + 2] = v.z;
}

void launch_nearest_velocity(hipStream_t st, int nq, const double* q3, double* out3, const float4* posm_s,
                             const uint32_t* id_s, const uint32_t* start, const Domain* dom,
                             const uint32_t* inv, const float4* vel, const CoupleParams& cp) {
    if (nq > 0)
        hipLaunchKernelGGL(k_nearest_velocity, dim3(nblk(nq, 256)), dim3(256), 0, st, nq, q3, out3, posm_s,
                           id_s, start, dom, inv, vel, cp);
}

}  // namespace phy
