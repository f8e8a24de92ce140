// sph_oracle.cpp — f64 CPU oracle for the phy-fluid SPH step.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker, never the product: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
// The shipped library (physics_amd/csrc) never links or calls it.
//
// It restates, in plain C++ f64, the reference solver of jojo240607/physics:
//   crates/phy-fluid/src/sph/world.rs   FluidWorld::{new, fill_box, lattice_mass,
//                                        step, compute_density_pressure,
//                                        compute_forces, integrate, enforce_bounds}
//   crates/phy-fluid/src/sph/grid.rs    Grid::{key_of, build, for_each_neighbor}
//   crates/phy-fluid/src/kernels.rs     Kernels::{new, poly6, spiky_grad_mag, visc_lap}
//   crates/phy-fluid/src/sph/params.rs  SphParams::defaults
//   crates/phy-fluid/src/sph/world.rs   couple_bodies, push_out, body_volume,
//                                        couple_heat (+ its velocity sampler),
//                                        couple_points (world.rs:400-773)
//   crates/phy-fluid/src/particle.rs    Particle::{new, with_material}
//   crates/phy-fluid/src/subsystem.rs   FluidSubsystem::{validate, total_momentum,
//                                        kinetic_energy}
// with the reference's operation order (left-associative products, nalgebra's
// (x*x + y*y) + z*z norm) and no FMA contraction (built with -ffp-contract=off).
// The two parallel passes use OpenMP over particles where the reference uses
// rayon par_iter + an order-preserving collect: every particle writes only its
// own output slot from snapshot inputs, so results do not depend on the thread
// count.
//
// Parity of the restatement itself is pinned by tests/test_oracle_*.py against
// closed forms, brute-force neighbour sets, symmetry and the reference's own
// test assertions; the ulp-level agreement with the Rust binary (nalgebra 0.33.3
// norm order, libm powf) cannot be executed here (no Rust toolchain).
//
// A run-time "mutation" switch plants known mistakes so the tests can show that
// each check actually fails on a plausible bug.

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>
#include <algorithm>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3 {
    double x, y, z;
};
inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 operator/(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
inline V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
// nalgebra Vector3::norm_squared / norm: (x*x + y*y) + z*z, then sqrt.
inline double norm_sq(V3 a) { return (a.x * a.x + a.y * a.y) + a.z * a.z; }
inline double norm(V3 a) { return std::sqrt(norm_sq(a)); }

// Planted-mistake switch used only by the oracle's own sensitivity tests.
enum Mutation : int {
    MUT_NONE = 0,
    MUT_DROP_XSPH = 1,        // integrate without the XSPH velocity correction
    MUT_VISC_SIGN = 2,        // (v_i - v_j) instead of (v_j - v_i) in viscosity
    MUT_NO_PRESSURE_CLAMP = 3,// p = k (rho - rho0) without max(0, .)
    MUT_SELF_IN_FORCE = 4,    // forget the j == i skip (adds nothing: r = 0 is
                              // skipped anyway — kept to show a benign change)
    MUT_DENSITY_NO_SELF = 5,  // drop the self term from the density sum
    MUT_BOUNDS_NO_DAMP = 6,   // reflect without the damping factor
    MUT_POLY6_H2 = 7,         // poly6 with (h - r) instead of (h^2 - r^2)
    MUT_DT_IN_XSPH = 8,       // v += xsph * eps * dt instead of v += xsph * eps
    // the power-law branch (world.rs:313-327)
    MUT_SHEAR_NO_REPS = 9,    // |dv| / r instead of |dv| / (r + 1e-4)
    MUT_SHEAR_MASS_I = 10,    // m_i / rho_i instead of m_j / rho_j
    MUT_POWER_EXP_N = 11,     // max'(shear, shear_min)^n instead of ^(n - 1)
    MUT_SHEAR_NO_FLOOR = 12,  // shear^(n - 1) without the shear_min floor
    MUT_SHEAR_DV_SQ = 13,     // |dv|^2 instead of |dv|
    MUT_MAT_OOR_LAST = 14,    // an out-of-table material uses the last entry, not 0
};
int g_mutation = MUT_NONE;

// crates/phy-fluid/src/sph/params.rs:12-50
constexpr int kMaxMat = 16;
struct Params {
    double rest_density, stiffness, viscosity;
    double visc_k[kMaxMat];
    double visc_n[kMaxMat];
    int32_t n_visc_k, n_visc_n;
    double shear_min, mass, h;
    double gravity[3], bounds_min[3], bounds_max[3];
    double boundary_damp, xsph_eps;
};

// crates/phy-fluid/src/particle.rs:17-46
struct Particle {
    V3 pos, vel, acc, body_acc;
    double rho, p, mass;
    uint64_t material;
    double mu_eff;
    V3 xsph;
};

// crates/phy-fluid/src/kernels.rs:22-65
struct Kernels {
    double h, h2, poly6_coeff, spiky_grad_coeff, visc_lap_coeff;
    explicit Kernels(double h_) {
        const double pi = 3.141592653589793;  // std::f64::consts::PI
        h = h_;
        h2 = h * h;
        double h3 = h2 * h;
        double h6 = h3 * h3;
        double h9 = h6 * h3;
        poly6_coeff = 315.0 / (64.0 * pi * h9);
        spiky_grad_coeff = -45.0 / (pi * h6);
        visc_lap_coeff = 45.0 / (pi * h6);
    }
    double poly6(double r) const {
        if (r >= h) return 0.0;
        double diff = (g_mutation == MUT_POLY6_H2) ? (h - r) : (h2 - r * r);
        return poly6_coeff * diff * diff * diff;
    }
    double spiky_grad_mag(double r) const {
        if (r >= h || r <= 0.0) return 0.0;
        double diff = h - r;
        return spiky_grad_coeff * diff * diff;
    }
    double visc_lap(double r) const {
        if (r >= h) return 0.0;
        return visc_lap_coeff * (h - r);
    }
};

// num-traits ToPrimitive::to_i64 for f64: None (-> unwrap_or(0)) for NaN and
// for values outside [-2^63, 2^63).
inline int64_t to_i64_or0(double v) {
    if (v >= -9223372036854775808.0 && v < 9223372036854775808.0) return (int64_t)v;
    return 0;
}

struct Key {
    int64_t i, j, k;
    bool operator==(const Key& o) const { return i == o.i && j == o.j && k == o.k; }
};
struct KeyHash {
    size_t operator()(const Key& k) const {
        uint64_t h = (uint64_t)k.i * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)k.j + 0xBF58476D1CE4E5B9ull + (h << 6) + (h >> 2);
        h ^= (uint64_t)k.k + 0x94D049BB133111EBull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

// crates/phy-fluid/src/sph/grid.rs:10-56
struct Grid {
    double cell;
    std::unordered_map<Key, std::vector<size_t>, KeyHash> map;
    explicit Grid(double c) : cell(c) {}
    Key key_of(V3 p) const {
        double inv = 1.0 / cell;  // T::one() / self.cell
        return Key{to_i64_or0(std::floor(p.x * inv)), to_i64_or0(std::floor(p.y * inv)),
                   to_i64_or0(std::floor(p.z * inv))};
    }
    void build(const std::vector<Particle>& ps) {
        map.clear();
        for (size_t i = 0; i < ps.size(); ++i) map[key_of(ps[i].pos)].push_back(i);
    }
    // di outermost, dk innermost; ascending index inside a cell.
    template <class F>
    void for_each_neighbor(V3 p, F&& f) const {
        Key c = key_of(p);
        for (int64_t di = -1; di <= 1; ++di)
            for (int64_t dj = -1; dj <= 1; ++dj)
                for (int64_t dk = -1; dk <= 1; ++dk) {
                    Key q{(int64_t)((uint64_t)c.i + (uint64_t)di), (int64_t)((uint64_t)c.j + (uint64_t)dj),
                          (int64_t)((uint64_t)c.k + (uint64_t)dk)};
                    auto it = map.find(q);
                    if (it == map.end()) continue;
                    for (size_t id : it->second) f(id);
                }
    }
};

struct World {
    Params params;
    std::vector<Particle> particles;
    Kernels kernels;
    Grid grid;
    explicit World(const Params& p) : params(p), kernels(p.h), grid(p.h) {}

    // world.rs:148-169
    double lattice_mass(double s) const {
        double h = params.h;
        int32_t n_shell = (int32_t)std::ceil(h / s) + 1;
        double sum_w = 0.0;
        for (int32_t i = -n_shell; i <= n_shell; ++i)
            for (int32_t j = -n_shell; j <= n_shell; ++j)
                for (int32_t l = -n_shell; l <= n_shell; ++l) {
                    double rr = std::sqrt((double)(i * i + j * j + l * l));
                    double r = s * rr;
                    if (r < h) sum_w += kernels.poly6(r);
                }
        if (sum_w > 0.0) return params.rest_density / sum_w;
        return params.mass;
    }

    // world.rs:124-144
    void fill_box(V3 mn, V3 mx, double spacing, double margin) {
        V3 lo = mn + v3(margin, margin, margin);
        V3 hi = mx - v3(margin, margin, margin);
        double d = spacing;
        double mass = lattice_mass(d);
        double x = lo.x;
        while (x <= hi.x) {
            double y = lo.y;
            while (y <= hi.y) {
                double z = lo.z;
                while (z <= hi.z) {
                    Particle pt{};
                    pt.pos = v3(x, y, z);
                    pt.mass = mass;
                    particles.push_back(pt);
                    z += d;
                }
                y += d;
            }
            x += d;
        }
    }

    // world.rs:183-192
    void step(double dt) {
        if (particles.empty()) return;
        grid.build(particles);
        compute_density_pressure();
        compute_forces();
        integrate(dt);
        enforce_bounds();
    }

    // world.rs:204-246
    void compute_density_pressure() {
        const size_t n = particles.size();
        const double h = params.h, rest = params.rest_density, stiff = params.stiffness;
        std::vector<V3> pos(n);
        std::vector<double> mass(n), rho_out(n), p_out(n);
        for (size_t i = 0; i < n; ++i) { pos[i] = particles[i].pos; mass[i] = particles[i].mass; }
#pragma omp parallel for schedule(dynamic, 256)
        for (long long ii = 0; ii < (long long)n; ++ii) {
            size_t i = (size_t)ii;
            V3 p_i = pos[i];
            double rho = 0.0;
            grid.for_each_neighbor(p_i, [&](size_t j) {
                if (g_mutation == MUT_DENSITY_NO_SELF && j == i) return;
                double r = norm(p_i - pos[j]);  // na_distance(&p_i, &pos[j])
                if (r < h) rho += mass[j] * kernels.poly6(r);
            });
            double over = rho - rest;
            double p = (over > 0.0) ? stiff * over : 0.0;
            if (g_mutation == MUT_NO_PRESSURE_CLAMP) p = stiff * over;
            rho_out[i] = rho;
            p_out[i] = p;
        }
        for (size_t i = 0; i < n; ++i) { particles[i].rho = rho_out[i]; particles[i].p = p_out[i]; }
    }

    // world.rs:252-347
    void compute_forces() {
        const size_t n = particles.size();
        const double h = params.h;
        const double r_eps = 1e-4;
        const double shear_min = params.shear_min;
        std::vector<V3> pos(n), vel(n);
        std::vector<double> rho(n), pr(n), mass(n);
        std::vector<uint64_t> mat(n);
        for (size_t i = 0; i < n; ++i) {
            pos[i] = particles[i].pos; vel[i] = particles[i].vel; rho[i] = particles[i].rho;
            pr[i] = particles[i].p; mass[i] = particles[i].mass; mat[i] = particles[i].material;
        }
        std::vector<V3> acc_out(n), xsph_out(n);
        std::vector<double> mu_out(n);
#pragma omp parallel for schedule(dynamic, 256)
        for (long long ii = 0; ii < (long long)n; ++ii) {
            size_t i = (size_t)ii;
            V3 p_i = pos[i], v_i = vel[i];
            double rho_i = rho[i], p_i_p = pr[i], m_i = mass[i];
            V3 f_press{0, 0, 0}, f_visc_raw{0, 0, 0}, xsph{0, 0, 0};
            double shear = 0.0;
            grid.for_each_neighbor(p_i, [&](size_t j) {
                if (j == i && g_mutation != MUT_SELF_IN_FORCE) return;
                V3 d = pos[j] - p_i;
                double r = norm(d);
                if (r >= h || r <= 0.0) return;  // NaN r falls through, as in the reference
                double rho_j = rho[j], p_j = pr[j];
                double grad = kernels.spiky_grad_mag(r);
                V3 dir = d / r;
                double coef = m_i * mass[j] * (p_i_p / (rho_i * rho_i) + p_j / (rho_j * rho_j));
                f_press += dir * (coef * grad);
                double lap = kernels.visc_lap(r);
                V3 dvel = (g_mutation == MUT_VISC_SIGN) ? (v_i - vel[j]) : (vel[j] - v_i);
                f_visc_raw += dvel * (m_i * mass[j] / rho_j * lap);
                double w = kernels.poly6(r);
                xsph += (vel[j] - v_i) * (mass[j] / rho_j * w);
                double dv = norm(vel[j] - v_i);
                if (g_mutation == MUT_SHEAR_DV_SQ) dv = norm_sq(vel[j] - v_i);
                const double rr = (g_mutation == MUT_SHEAR_NO_REPS) ? r : r + r_eps;
                const double mr = (g_mutation == MUT_SHEAR_MASS_I) ? m_i / rho_i : mass[j] / rho_j;
                shear += dv / rr * mr;
            });
            int n_k = params.n_visc_k;
            size_t m_idx = (mat[i] < (uint64_t)n_k) ? (size_t)mat[i] : 0;
            if (g_mutation == MUT_MAT_OOR_LAST && mat[i] >= (uint64_t)n_k) m_idx = (size_t)(n_k - 1);
            double kc = params.visc_k[m_idx];
            double nn = params.visc_n[m_idx];
            double sreg = (shear > shear_min) ? shear : shear_min;
            if (g_mutation == MUT_SHEAR_NO_FLOOR) sreg = shear;
            const double ex = (g_mutation == MUT_POWER_EXP_N) ? nn : nn - 1.0;
            double mu_eff = kc * std::pow(sreg, ex);
            V3 f_visc = f_visc_raw * mu_eff;
            V3 acc = (rho_i > 0.0) ? (f_press + f_visc) / rho_i : v3(0, 0, 0);
            acc_out[i] = acc;
            mu_out[i] = mu_eff;
            xsph_out[i] = xsph;
        }
        for (size_t i = 0; i < n; ++i) {
            particles[i].acc = acc_out[i];
            particles[i].mu_eff = mu_out[i];
            particles[i].xsph = xsph_out[i];
        }
    }

    // world.rs:352-363
    void integrate(double dt) {
        V3 g = v3(params.gravity[0], params.gravity[1], params.gravity[2]);
        double eps = params.xsph_eps;
        for (auto& pt : particles) {
            V3 a = pt.acc + g + pt.body_acc;
            pt.vel += a * dt;
            if (g_mutation == MUT_DT_IN_XSPH) pt.vel += pt.xsph * eps * dt;
            else if (g_mutation != MUT_DROP_XSPH) pt.vel += pt.xsph * eps;
            pt.pos += pt.vel * dt;
            pt.body_acc = v3(0, 0, 0);
        }
    }

    // world.rs:366-397
    void enforce_bounds() {
        const double* lo = params.bounds_min;
        const double* hi = params.bounds_max;
        double damp = (g_mutation == MUT_BOUNDS_NO_DAMP) ? 1.0 : params.boundary_damp;
        for (auto& pt : particles) {
            double* p = &pt.pos.x;
            double* v = &pt.vel.x;
            for (int a = 0; a < 3; ++a) {
                if (p[a] < lo[a]) { p[a] = lo[a]; v[a] = -v[a] * damp; }
                else if (p[a] > hi[a]) { p[a] = hi[a]; v[a] = -v[a] * damp; }
            }
        }
    }
};

// params.rs:54-85
// ---------------------------------------------------------------------------
// Coupling (world.rs:400-773): fluid <-> rigid bodies, fluid <-> heat grid,
// fluid <-> point masses, and the heat field's nearest-particle velocity
// sampler. Rigid-body geometry follows phy-rigid shape.rs:198-330 (support,
// containment) and 433-463 (pose transforms, linear impulse); the heat grid
// follows phy-field heat.rs:129-180 and grid.rs:80-82, 125-130, 203-208.
//
// Quaternion rotation restates nalgebra 0.33.3's UnitQuaternion * Vector3
// (t = 2 (q_v x v); v' = (t w + q_v x t) + v), a third-party routine that is
// not vendored in the reference: parity unpinned at the ulp level for rotated
// bodies (pinned against rotation matrices to 1e-15 by the oracle tests).

struct Quat {
    double w, i, j, k;
};
inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline V3 normalize(V3 a) { return a / norm(a); }
inline V3 qrot(Quat q, V3 v) {
    V3 u = v3(q.i, q.j, q.k);
    V3 t = cross(u, v) * 2.0;
    V3 c = cross(u, t);
    return (t * q.w + c) + v;
}
inline Quat qconj(Quat q) { return Quat{q.w, -q.i, -q.j, -q.k}; }

enum ShapeKind : int32_t { SHAPE_SPHERE = 0, SHAPE_BOX = 1, SHAPE_CAPSULE = 2 };

// Body<f64> restricted to what the coupling reads and writes.
struct CBody {
    V3 pos;
    Quat rot;
    V3 vel;
    double inv_mass;
    int32_t shape;
    V3 dims;  // sphere: r; box: half extents; capsule: half_height, r
};

inline V3 body_to_local(const CBody& b, V3 p) { return qrot(qconj(b.rot), p - b.pos); }
inline V3 body_to_world(const CBody& b, V3 p) { return qrot(b.rot, p) + b.pos; }

// f64::clamp (min <= max assumed; NaN stays NaN)
inline double clamp_f(double x, double lo, double hi) {
    if (x < lo) return lo;
    if (x > hi) return hi;
    return x;
}

bool contains_local(const CBody& b, V3 p) {
    switch (b.shape) {
        case SHAPE_SPHERE: return norm(p) <= b.dims.x;
        case SHAPE_BOX:
            return std::fabs(p.x) <= b.dims.x && std::fabs(p.y) <= b.dims.y && std::fabs(p.z) <= b.dims.z;
        case SHAPE_CAPSULE: {
            double hh = b.dims.x, r = b.dims.y;
            double y = clamp_f(p.y, -hh, hh);
            double cx = (p.x * p.x + (p.y - y) * (p.y - y)) + p.z * p.z;
            return cx <= r * r;
        }
        default: return false;
    }
}

V3 support_local(const CBody& b, V3 dir) {
    switch (b.shape) {
        case SHAPE_SPHERE: return normalize(dir) * b.dims.x;
        case SHAPE_BOX:
            return v3(dir.x >= 0.0 ? b.dims.x : -b.dims.x, dir.y >= 0.0 ? b.dims.y : -b.dims.y,
                      dir.z >= 0.0 ? b.dims.z : -b.dims.z);
        case SHAPE_CAPSULE: {
            V3 n = normalize(dir);
            V3 end = v3(0.0, 0.0, 0.0);
            if (dir.y > 0.0) end = v3(0.0, b.dims.x, 0.0);
            else if (dir.y < 0.0) end = v3(0.0, -b.dims.x, 0.0);
            return end + n * b.dims.y;
        }
        default: return v3(0.0, 0.0, 0.0);
    }
}

// world.rs:507-540 (the three closed forms this library hosts)
double body_volume(const CBody& b) {
    const double pi = 3.141592653589793;
    switch (b.shape) {
        case SHAPE_SPHERE: { double r = b.dims.x; return 4.0 / 3.0 * pi * r * r * r; }
        case SHAPE_BOX: return b.dims.x * b.dims.y * b.dims.z * 8.0;
        case SHAPE_CAPSULE: {
            double hh = b.dims.x, r = b.dims.y;
            return pi * r * r * 2.0 * hh + 4.0 / 3.0 * pi * r * r * r;
        }
        default: return 0.0;
    }
}

// world.rs:755-773
void push_out(Particle& pt, const CBody& b) {
    V3 local = body_to_local(b, pt.pos);
    V3 dir = norm(local) > 1e-9 ? normalize(local) : v3(0.0, 1.0, 0.0);
    V3 sd = support_local(b, dir);
    pt.pos = body_to_world(b, sd);
    V3 v_rel = pt.vel - b.vel;
    V3 n = normalize(body_to_world(b, normalize(sd)) - b.pos);
    double vn = dot(v_rel, n);
    if (vn < 0.0) pt.vel = pt.vel - n * vn;
}

// world.rs:414-503
void couple_bodies(World& W, CBody* bodies, size_t nb, double dt) {
    if (W.particles.empty() || nb == 0) return;
    const double rho_f = W.params.rest_density;
    const V3 g = v3(W.params.gravity[0], W.params.gravity[1], W.params.gravity[2]);
    const double particle_vol = W.params.mass / rho_f;
    const double restitution = 0.0;
    std::vector<double> displaced(nb, 0.0);
    for (size_t pi = 0; pi < W.particles.size(); ++pi) {
        Particle& pt = W.particles[pi];
        const V3 p_world = pt.pos;
        V3 v_p = pt.vel;
        const double m_p = pt.mass;
        for (size_t bi = 0; bi < nb; ++bi) {
            CBody& b = bodies[bi];
            V3 local = body_to_local(b, p_world);
            if (!contains_local(b, local)) continue;
            V3 n = norm(local) > 1e-9 ? normalize(local) : v3(0.0, 1.0, 0.0);
            V3 v_b = b.vel;
            if (b.inv_mass <= 0.0) {
                push_out(pt, b);
                v_p = pt.vel;
                continue;
            }
            double v_rel_n = dot(v_p - v_b, n);
            if (v_rel_n < 0.0) {
                double inv_sum = (1.0 / m_p) + b.inv_mass;
                double j_mag = -(1.0 + restitution) * v_rel_n / inv_sum;
                V3 j = n * j_mag;
                v_p += j / m_p;
                b.vel += neg(j) * b.inv_mass;  // apply_impulse(-j)
            }
            displaced[bi] += particle_vol;
            push_out(pt, b);
            v_p = pt.vel;  // the particle keeps push_out's velocity (world.rs:488)
        }
        pt.vel = v_p;
    }
    for (size_t bi = 0; bi < nb; ++bi) {
        CBody& b = bodies[bi];
        if (b.inv_mass <= 0.0) continue;
        double vol_b = body_volume(b);
        double v_sub = displaced[bi] > 0.0 ? vol_b : 0.0;
        V3 buoy = neg(g) * (rho_f * v_sub);
        b.vel += (buoy * dt) * b.inv_mass;
    }
}

// world.rs:691-752
struct CPoint {
    V3 pos, vel;
    double mass;
    V3 force;
};

void couple_points(World& W, CPoint* pts, size_t np, double dt, double drag, double soft_density) {
    if (np == 0 || W.particles.empty()) return;
    const double rho_f = W.params.rest_density;
    const V3 g = v3(W.params.gravity[0], W.params.gravity[1], W.params.gravity[2]);
    const double* lo = W.params.bounds_min;
    const double* hi = W.params.bounds_max;
    W.grid.build(W.particles);
    for (size_t k = 0; k < np; ++k) {
        CPoint& pt = pts[k];
        if (pt.pos.x < lo[0] || pt.pos.x > hi[0] || pt.pos.y < lo[1] || pt.pos.y > hi[1] ||
            pt.pos.z < lo[2] || pt.pos.z > hi[2])
            continue;
        V3 v_sum = v3(0.0, 0.0, 0.0);
        double w_sum = 0.0;
        W.grid.for_each_neighbor(pt.pos, [&](size_t j) {
            double w = W.particles[j].mass;
            v_sum += W.particles[j].vel * w;
            w_sum += w;
        });
        V3 v_f = w_sum > 0.0 ? v_sum / w_sum : v3(0.0, 0.0, 0.0);
        double vol = soft_density > 0.0 ? pt.mass / soft_density : 0.0;
        V3 f_b = neg(g) * (rho_f * vol);
        V3 f_d = (v_f - pt.vel) * (drag * pt.mass);
        V3 f = f_b + f_d;
        pt.force = f;
        double m_sum = 0.0;
        std::vector<size_t> idxs;
        W.grid.for_each_neighbor(pt.pos, [&](size_t j) {
            m_sum += W.particles[j].mass;
            idxs.push_back(j);
        });
        if (m_sum > 0.0) {
            V3 reaction = neg(f) * dt;
            for (size_t j : idxs) {
                double frac = W.particles[j].mass / m_sum;
                W.particles[j].vel += reaction * frac;
            }
        }
    }
}

// world.rs:555-678 (heat grid: x fastest, idx = x + nx (y + ny z))
struct CHeat {
    size_t nx, ny, nz;
    V3 origin;
    double dx;
    const double* u;
    double* src;
    double at(size_t x, size_t y, size_t z) const {
        size_t cx = std::min(x, nx - 1), cy = std::min(y, ny - 1), cz = std::min(z, nz - 1);
        return u[cx + nx * (cy + ny * cz)];
    }
    double trilinear(size_t cx, size_t cy, size_t cz, double fx, double fy, double fz) const {
        size_t x1 = std::min(cx + 1, nx - 1), y1 = std::min(cy + 1, ny - 1), z1 = std::min(cz + 1, nz - 1);
        double c000 = at(cx, cy, cz), c100 = at(x1, cy, cz), c010 = at(cx, y1, cz), c110 = at(x1, y1, cz);
        double c001 = at(cx, cy, z1), c101 = at(x1, cy, z1), c011 = at(cx, y1, z1), c111 = at(x1, y1, z1);
        double x00 = c000 + (c100 - c000) * fx;
        double x10 = c010 + (c110 - c010) * fx;
        double x01 = c001 + (c101 - c001) * fx;
        double x11 = c011 + (c111 - c011) * fx;
        double y0 = x00 + (x10 - x00) * fy;
        double y1v = x01 + (x11 - x01) * fy;
        return y0 + (y1v - y0) * fz;
    }
};

// usize::to_usize of a floored f64 (None -> unwrap_or(0))
inline size_t to_usize_or0(double f) {
    if (f >= 0.0 && f < 18446744073709551616.0) return (size_t)f;
    return 0;
}

void couple_heat(World& W, CHeat& H, double dt, double t_ref, double beta, double heat_gain) {
    if (W.particles.empty()) return;
    const V3 g = v3(W.params.gravity[0], W.params.gravity[1], W.params.gravity[2]);
    const double* lo = W.params.bounds_min;
    const double* hi = W.params.bounds_max;
    const double dx = H.dx;
    const double nx_m1 = (double)(H.nx ? H.nx - 1 : 0);
    const double ny_m1 = (double)(H.ny ? H.ny - 1 : 0);
    const double nz_m1 = (double)(H.nz ? H.nz - 1 : 0);
    const V3 hmin = H.origin;
    auto to_idx = [&](double x, double y, double z, size_t* c, double* t) {
        if (x < hmin.x || x > hmin.x + dx * nx_m1 || y < hmin.y || y > hmin.y + dx * ny_m1 ||
            z < hmin.z || z > hmin.z + dx * nz_m1)
            return false;
        double fx = std::floor((x - hmin.x) / dx);
        double fy = std::floor((y - hmin.y) / dx);
        double fz = std::floor((z - hmin.z) / dx);
        c[0] = to_usize_or0(fx);
        c[1] = to_usize_or0(fy);
        c[2] = to_usize_or0(fz);
        t[0] = (x - hmin.x) / dx - fx;
        t[1] = (y - hmin.y) / dx - fy;
        t[2] = (z - hmin.z) / dx - fz;
        return true;
    };
    for (Particle& pt : W.particles) {
        if (pt.pos.x < lo[0] || pt.pos.x > hi[0] || pt.pos.y < lo[1] || pt.pos.y > hi[1] ||
            pt.pos.z < lo[2] || pt.pos.z > hi[2])
            continue;
        size_t c[3];
        double t[3];
        bool in = to_idx(pt.pos.x, pt.pos.y, pt.pos.z, c, t);
        double temp = in ? H.trilinear(c[0], c[1], c[2], t[0], t[1], t[2]) : t_ref;
        V3 buoy_corr = neg(g) * beta * (temp - t_ref);
        pt.body_acc += buoy_corr;
        if (heat_gain > 0.0) {
            double speed = norm(pt.vel);
            if (in && c[0] < H.nx && c[1] < H.ny && c[2] < H.nz)
                H.src[c[0] + H.nx * (c[1] + H.ny * c[2])] += heat_gain * speed * dt;
        }
    }
}

// The velocity sampler couple_heat installs (world.rs:614-643): the velocity of
// the particle nearest to q, the first one in index order on ties; particle 0
// when no distance compares smaller than infinity; zero without particles.
V3 nearest_velocity(const World& W, V3 q) {
    if (W.particles.empty()) return v3(0.0, 0.0, 0.0);
    size_t bi = 0;
    double bd = INFINITY;
    for (size_t i = 0; i < W.particles.size(); ++i) {
        const V3 s = W.particles[i].pos;
        double dxq = q.x - s.x, dyq = q.y - s.y, dzq = q.z - s.z;
        double d = dxq * dxq + dyq * dyq + dzq * dzq;
        if (d < bd) {
            bd = d;
            bi = i;
        }
    }
    return W.particles[bi].vel;
}

void params_defaults(Params* p) {
    std::memset(p, 0, sizeof(*p));
    p->rest_density = 1000.0;
    p->stiffness = 250.0;
    p->viscosity = 3.5;
    p->visc_k[0] = 3.5;
    p->visc_n[0] = 1.0;
    p->n_visc_k = 1;
    p->n_visc_n = 1;
    p->shear_min = 0.01;
    p->mass = 0.02;
    p->h = 0.2;
    p->gravity[0] = 0.0; p->gravity[1] = -9.81; p->gravity[2] = 0.0;
    for (int a = 0; a < 3; ++a) { p->bounds_min[a] = -5.0; p->bounds_max[a] = 5.0; }
    p->boundary_damp = 0.4;
    p->xsph_eps = 0.5;
}

}  // namespace

extern "C" {

typedef Params OrParams;

int or_params_size(void) { return (int)sizeof(Params); }
void or_params_defaults(OrParams* p) { params_defaults(p); }
void or_set_mutation(int m) { g_mutation = m; }
int or_get_mutation(void) { return g_mutation; }

void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int or_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_kernels(double h, double* out3, double r, double* vals3) {
    Kernels k(h);
    out3[0] = k.poly6_coeff;
    out3[1] = k.spiky_grad_coeff;
    out3[2] = k.visc_lap_coeff;
    if (vals3) {
        vals3[0] = k.poly6(r);
        vals3[1] = k.spiky_grad_mag(r);
        vals3[2] = k.visc_lap(r);
    }
}

void* or_world_new(const OrParams* p) { return new World(*p); }
void or_world_free(void* w) { delete (World*)w; }
void* or_world_clone(void* w) { return new World(*(World*)w); }

void or_world_set_params(void* w, const OrParams* p) {
    World* W = (World*)w;
    W->params = *p;
    W->kernels = Kernels(p->h);
    W->grid = Grid(p->h);
}

size_t or_world_len(void* w) { return ((World*)w)->particles.size(); }
double or_world_lattice_mass(void* w, double s) { return ((World*)w)->lattice_mass(s); }
void or_world_fill_box(void* w, const double* mn, const double* mx, double spacing, double margin) {
    ((World*)w)->fill_box(v3(mn[0], mn[1], mn[2]), v3(mx[0], mx[1], mx[2]), spacing, margin);
}

// Replace the particle list. Null arrays leave the field at Particle::new's
// default (zero velocity, zero body_acc, material 0).
void or_world_set_particles(void* w, size_t n, const double* pos, const double* vel,
                            const double* mass, const uint64_t* material, const double* body_acc) {
    World* W = (World*)w;
    W->particles.assign(n, Particle{});
    for (size_t i = 0; i < n; ++i) {
        Particle& p = W->particles[i];
        p.pos = v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        if (vel) p.vel = v3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        p.mass = mass ? mass[i] : W->params.mass;
        if (material) p.material = material[i];
        if (body_acc) p.body_acc = v3(body_acc[3 * i], body_acc[3 * i + 1], body_acc[3 * i + 2]);
    }
}

// field: 0 pos, 1 vel, 2 acc, 3 body_acc, 4 xsph (3 per particle);
//        5 rho, 6 p, 7 mass, 8 mu_eff (1 per particle)
void or_world_get(void* w, int field, double* out) {
    World* W = (World*)w;
    size_t n = W->particles.size();
    for (size_t i = 0; i < n; ++i) {
        const Particle& p = W->particles[i];
        const V3* v = nullptr;
        switch (field) {
            case 0: v = &p.pos; break;
            case 1: v = &p.vel; break;
            case 2: v = &p.acc; break;
            case 3: v = &p.body_acc; break;
            case 4: v = &p.xsph; break;
            case 5: out[i] = p.rho; break;
            case 6: out[i] = p.p; break;
            case 7: out[i] = p.mass; break;
            case 8: out[i] = p.mu_eff; break;
            default: break;
        }
        if (v) { out[3 * i] = v->x; out[3 * i + 1] = v->y; out[3 * i + 2] = v->z; }
    }
}

void or_world_step(void* w, double dt) { ((World*)w)->step(dt); }

void or_world_step_n(void* w, double dt, int n) {
    for (int i = 0; i < n; ++i) ((World*)w)->step(dt);
}

// Individual stages, for stage-level parity (the grid is rebuilt by density).
void or_world_density(void* w) {
    World* W = (World*)w;
    W->grid.build(W->particles);
    W->compute_density_pressure();
}
void or_world_forces(void* w) { ((World*)w)->compute_forces(); }

// subsystem.rs:74-90 — 0 ok, 1 NonFinite{pos}, 2 NonFinite{vel}
int or_world_validate(void* w) {
    for (const auto& p : ((World*)w)->particles) {
        if (!std::isfinite(p.pos.x) || !std::isfinite(p.pos.y) || !std::isfinite(p.pos.z)) return 1;
        if (!std::isfinite(p.vel.x) || !std::isfinite(p.vel.y) || !std::isfinite(p.vel.z)) return 2;
    }
    return 0;
}

// subsystem.rs:149-155
void or_world_momentum(void* w, double* out3) {
    V3 s{0, 0, 0};
    for (const auto& p : ((World*)w)->particles) s += p.vel * p.mass;
    out3[0] = s.x; out3[1] = s.y; out3[2] = s.z;
}
// subsystem.rs:157-163
double or_world_kinetic_energy(void* w) {
    double e = 0.0;
    for (const auto& p : ((World*)w)->particles) e += (norm_sq(p.vel) * p.mass) / (1.0 + 1.0);
    return e;
}

// Density-pass neighbour set of particle i (j with r < h, self included), in the
// reference's visiting order. Returns the count; writes at most cap indices.
size_t or_world_neighbors(void* w, size_t i, uint64_t* out, size_t cap) {
    World* W = (World*)w;
    W->grid.build(W->particles);
    size_t cnt = 0;
    V3 p_i = W->particles[i].pos;
    W->grid.for_each_neighbor(p_i, [&](size_t j) {
        double r = norm(p_i - W->particles[j].pos);
        if (r < W->params.h) {
            if (cnt < cap) out[cnt] = j;
            ++cnt;
        }
    });
    return cnt;
}

// Per-particle neighbour statistics over the whole world (density predicate
// r < h, self included): count and two order-independent checksums of the
// neighbours' indices, for exact set comparison without storing the sets.
void or_world_neighbor_stats(void* w, uint32_t* counts, uint64_t* sum_ids, uint64_t* xor_hash) {
    World* W = (World*)w;
    W->grid.build(W->particles);
    const size_t n = W->particles.size();
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)n; ++ii) {
        size_t i = (size_t)ii;
        V3 p_i = W->particles[i].pos;
        uint32_t c = 0;
        uint64_t s = 0, x = 0;
        W->grid.for_each_neighbor(p_i, [&](size_t j) {
            double r = norm(p_i - W->particles[j].pos);
            if (r < W->params.h) {
                ++c;
                s += (uint64_t)j;
                uint64_t z = (uint64_t)j + 0x9E3779B97F4A7C15ull;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                x ^= z ^ (z >> 31);
            }
        });
        counts[i] = c;
        sum_ids[i] = s;
        xor_hash[i] = x;
    }
}

// The same statistics for the force pass's neighbour set (world.rs:292-299):
// j != i whose distance r = |p_j - p_i| fails both r >= h and r <= 0, so
// 0 < r < h or r NaN (a NaN distance falls through both tests and poisons the
// sums, as in the reference).
void or_world_force_neighbor_stats(void* w, uint32_t* counts, uint64_t* sum_ids, uint64_t* xor_hash) {
    World* W = (World*)w;
    W->grid.build(W->particles);
    const size_t n = W->particles.size();
    const double h = W->params.h;
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)n; ++ii) {
        size_t i = (size_t)ii;
        V3 p_i = W->particles[i].pos;
        uint32_t c = 0;
        uint64_t s = 0, x = 0;
        W->grid.for_each_neighbor(p_i, [&](size_t j) {
            if (j == i) return;
            double r = norm(W->particles[j].pos - p_i);
            if (r >= h || r <= 0.0) return;
            ++c;
            s += (uint64_t)j;
            uint64_t z = (uint64_t)j + 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            x ^= z ^ (z >> 31);
        });
        counts[i] = c;
        sum_ids[i] = s;
        xor_hash[i] = x;
    }
}

// Grid::build on the current positions, then Grid::to_flat
// (crates/phy-fluid/src/sph/grid.rs:62-175): occupied-key box padded to
// [min - 1, max + 2] per axis, x-fastest flat index, prefix-sum cell starts,
// particle ids of each cell in ascending order. Fills min3 / nc3 always;
// fills the arrays and returns 0 when the capacities suffice, else returns 1.
int or_world_flat_grid(void* wv, int64_t* min3, uint64_t* nc3, int32_t* cell_start, uint64_t cell_cap,
                       int32_t* sorted, uint64_t n_cap) {
    World* w = (World*)wv;
    Grid g(w->params.h);
    g.build(w->particles);
    int64_t mn[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, mx[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
    for (const auto& kv : g.map) {
        if (kv.second.empty()) continue;
        const int64_t k[3] = {kv.first.i, kv.first.j, kv.first.k};
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], k[a]);
            mx[a] = std::max(mx[a], k[a]);
        }
    }
    if (mn[0] > mx[0]) {  // empty grid
        for (int a = 0; a < 3; ++a) {
            min3[a] = 0;
            nc3[a] = 1;
        }
        if (cell_cap < 2) return 1;
        cell_start[0] = 0;
        cell_start[1] = 0;  // (the reference's vec![0] has one entry; +1 sentinel here)
        return 0;
    }
    uint64_t nc[3];
    for (int a = 0; a < 3; ++a) {
        mn[a] = mn[a] - 1;
        nc[a] = (uint64_t)(mx[a] - mn[a] + 3);
        min3[a] = mn[a];
        nc3[a] = nc[a];
    }
    const uint64_t ncell = nc[0] * nc[1] * nc[2];
    const uint64_t n = w->particles.size();
    if (cell_cap < ncell + 1 || n_cap < n) return 1;
    auto flat = [&](const Key& k) {
        return (uint64_t)(k.i - mn[0]) + nc[0] * ((uint64_t)(k.j - mn[1]) + nc[1] * (uint64_t)(k.k - mn[2]));
    };
    std::vector<uint32_t> counts(ncell, 0);
    for (const auto& kv : g.map)
        if (!kv.second.empty()) counts[flat(kv.first)] += (uint32_t)kv.second.size();
    int32_t acc = 0;
    for (uint64_t c = 0; c < ncell; ++c) {
        cell_start[c] = acc;
        acc += (int32_t)counts[c];
    }
    cell_start[ncell] = acc;
    std::vector<int32_t> cursor(cell_start, cell_start + ncell);
    for (const auto& kv : g.map) {
        if (kv.second.empty()) continue;
        const uint64_t c = flat(kv.first);
        for (size_t id : kv.second) sorted[cursor[c]++] = (int32_t)id;
    }
    return 0;
}

// Wall-clock of n steps (seconds), for the CPU baseline.
// Bodies: pose7 = pos xyz + quaternion w i j k (phy_ffi.h:147), vel3 in/out,
// inv_mass, shape (0 sphere, 1 box, 2 capsule), dims3.
void or_world_couple_bodies(void* w, size_t nb, const double* pose7, double* vel3,
                            const double* inv_mass, const int32_t* shape, const double* dims3,
                            double dt) {
    std::vector<CBody> b(nb);
    for (size_t k = 0; k < nb; ++k) {
        const double* q = pose7 + 7 * k;
        b[k].pos = v3(q[0], q[1], q[2]);
        b[k].rot = Quat{q[3], q[4], q[5], q[6]};
        b[k].vel = v3(vel3[3 * k], vel3[3 * k + 1], vel3[3 * k + 2]);
        b[k].inv_mass = inv_mass[k];
        b[k].shape = shape[k];
        b[k].dims = v3(dims3[3 * k], dims3[3 * k + 1], dims3[3 * k + 2]);
    }
    couple_bodies(*(World*)w, b.data(), nb, dt);
    for (size_t k = 0; k < nb; ++k) {
        vel3[3 * k] = b[k].vel.x;
        vel3[3 * k + 1] = b[k].vel.y;
        vel3[3 * k + 2] = b[k].vel.z;
    }
}

// Points: pos3, vel3, mass in; force3 in/out (untouched outside the bounds).
void or_world_couple_points(void* w, size_t np, const double* pos3, const double* vel3,
                            const double* mass, double* force3, double dt, double drag,
                            double soft_density) {
    std::vector<CPoint> p(np);
    for (size_t k = 0; k < np; ++k) {
        p[k].pos = v3(pos3[3 * k], pos3[3 * k + 1], pos3[3 * k + 2]);
        p[k].vel = v3(vel3[3 * k], vel3[3 * k + 1], vel3[3 * k + 2]);
        p[k].mass = mass[k];
        p[k].force = v3(force3[3 * k], force3[3 * k + 1], force3[3 * k + 2]);
    }
    couple_points(*(World*)w, p.data(), np, dt, drag, soft_density);
    for (size_t k = 0; k < np; ++k) {
        force3[3 * k] = p[k].force.x;
        force3[3 * k + 1] = p[k].force.y;
        force3[3 * k + 2] = p[k].force.z;
    }
}

void or_world_couple_heat(void* w, size_t nx, size_t ny, size_t nz, const double* origin3, double dx,
                          const double* u, double* src, double dt, double t_ref, double beta,
                          double heat_gain) {
    CHeat H{nx, ny, nz, v3(origin3[0], origin3[1], origin3[2]), dx, u, src};
    couple_heat(*(World*)w, H, dt, t_ref, beta, heat_gain);
}

void or_world_nearest_velocity(void* w, size_t nq, const double* q3, double* out3) {
    const World& W = *(World*)w;
    for (size_t k = 0; k < nq; ++k) {
        V3 v = nearest_velocity(W, v3(q3[3 * k], q3[3 * k + 1], q3[3 * k + 2]));
        out3[3 * k] = v.x;
        out3[3 * k + 1] = v.y;
        out3[3 * k + 2] = v.z;
    }
}

// Body-frame helpers for the oracle's own tests.
void or_quat_rotate(const double* q4, const double* v3in, double* out3) {
    V3 r = qrot(Quat{q4[0], q4[1], q4[2], q4[3]}, v3(v3in[0], v3in[1], v3in[2]));
    out3[0] = r.x;
    out3[1] = r.y;
    out3[2] = r.z;
}

double or_world_time_steps(void* w, double dt, int n) {
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n; ++i) ((World*)w)->step(dt);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
