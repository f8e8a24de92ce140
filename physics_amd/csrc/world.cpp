// world.cpp — host mirror implementation (see world.h for the reference map).
#include "world.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <stdexcept>

#include <cmath>
#include <stdexcept>

namespace phy {

// params.rs:54-85
SphParams SphParams::defaults() {
    SphParams p;
    p.rest_density = 1000.0;
    p.stiffness = 250.0;
    p.viscosity = 3.5;
    p.visc_k = {3.5};
    p.visc_n = {1.0};
    p.shear_min = 0.01;
    p.mass = 0.02;
    p.h = 0.2;
    p.gravity = Vec3{0.0, -9.81, 0.0};
    p.bounds_min = Vec3{-5.0, -5.0, -5.0};
    p.bounds_max = Vec3{5.0, 5.0, 5.0};
    p.boundary_damp = 0.4;
    p.xsph_eps = 0.5;
    return p;
}

HostParams SphParams::to_host() const {
    HostParams h;
    h.rest_density = rest_density;
    h.stiffness = stiffness;
    h.viscosity = viscosity;
    h.visc_k = visc_k;
    h.visc_n = visc_n;
    h.shear_min = shear_min;
    h.mass = mass;
    h.h = this->h;
    const Vec3* v[3] = {&gravity, &bounds_min, &bounds_max};
    double* o[3] = {h.gravity, h.bounds_min, h.bounds_max};
    for (int k = 0; k < 3; ++k) {
        o[k][0] = v[k]->x;
        o[k][1] = v[k]->y;
        o[k][2] = v[k]->z;
    }
    h.boundary_damp = boundary_damp;
    h.xsph_eps = xsph_eps;
    return h;
}

// particle.rs:50-70
Particle Particle::make(Vec3 pos, double mass) {
    Particle p;
    p.pos = pos;
    p.mass = mass;
    return p;
}
Particle Particle::with_material(Vec3 pos, double mass, size_t material) {
    Particle p = make(pos, mass);
    p.material = material;
    return p;
}

// ---------------------------------------------------------------------------

FluidWorld::FluidWorld(const SphParams& params) : params_(params), dev_(params.to_host()) {}

void FluidWorld::set_params(const SphParams& params) {
    params_ = params;
    dev_.set_params(params.to_host());
}

size_t FluidWorld::len() const { return host_dirty_ ? host_.size() : dev_.size(); }

// kernels.rs:39-45 in f64
static double poly6_f64(double h, double r) {
    const double pi = 3.141592653589793;
    double h2 = h * h, h3 = h2 * h, h6 = h3 * h3, h9 = h6 * h3;
    double c = 315.0 / (64.0 * pi * h9);
    if (r >= h) return 0.0;
    double diff = h2 - r * r;
    return c * diff * diff * diff;
}

// world.rs:148-169
double FluidWorld::lattice_mass(double s) const {
    double h = params_.h;
    double q = h / s;
    int32_t n_shell = (int32_t)(std::isfinite(q) ? std::ceil(q) : 1.0) + 1;
    double sum_w = 0.0;
    for (int32_t i = -n_shell; i <= n_shell; ++i)
        for (int32_t j = -n_shell; j <= n_shell; ++j)
            for (int32_t l = -n_shell; l <= n_shell; ++l) {
                double rr = std::sqrt((double)(i * i + j * j + l * l));
                double r = s * rr;
                if (r < h) sum_w += poly6_f64(h, r);
            }
    if (sum_w > 0.0) return params_.rest_density / sum_w;
    return params_.mass;
}

void FluidWorld::flush_host() {
    if (!host_dirty_) return;
    size_t n = host_.size();
    std::vector<float> pos(3 * n), vel(3 * n), mass(n), body;
    std::vector<uint32_t> mat(n);
    bool any_body = false;
    for (size_t i = 0; i < n; ++i) {
        const Particle& p = host_[i];
        pos[3 * i] = (float)p.pos.x;
        pos[3 * i + 1] = (float)p.pos.y;
        pos[3 * i + 2] = (float)p.pos.z;
        vel[3 * i] = (float)p.vel.x;
        vel[3 * i + 1] = (float)p.vel.y;
        vel[3 * i + 2] = (float)p.vel.z;
        mass[i] = (float)p.mass;
        if (p.material > 0xffffffffull) throw std::invalid_argument("material tag too large");
        mat[i] = (uint32_t)p.material;
        if (p.body_acc.x != 0 || p.body_acc.y != 0 || p.body_acc.z != 0) any_body = true;
    }
    if (any_body) {
        body.resize(3 * n);
        for (size_t i = 0; i < n; ++i) {
            body[3 * i] = (float)host_[i].body_acc.x;
            body[3 * i + 1] = (float)host_[i].body_acc.y;
            body[3 * i + 2] = (float)host_[i].body_acc.z;
        }
    }
    dev_.upload(n, pos.data(), vel.data(), mass.data(), mat.data(), any_body ? body.data() : nullptr);
    host_.clear();
    host_dirty_ = false;
    last_was_step_ = false;
}

std::vector<Particle> FluidWorld::particles() {
    if (host_dirty_) return host_;
    size_t n = dev_.size();
    std::vector<Particle> out(n);
    if (!n) return out;
    std::vector<float> a(3 * n), s(n);
    auto put3 = [&](Vec3 Particle::*f) {
        for (size_t i = 0; i < n; ++i) out[i].*f = Vec3{a[3 * i], a[3 * i + 1], a[3 * i + 2]};
    };
    dev_.download(F_POS, a.data());
    put3(&Particle::pos);
    dev_.download(F_VEL, a.data());
    put3(&Particle::vel);
    dev_.download(F_MASS, s.data());
    for (size_t i = 0; i < n; ++i) out[i].mass = s[i];
    {
        std::vector<uint32_t> m(n);
        dev_.download_material(m.data());
        for (size_t i = 0; i < n; ++i) out[i].material = m[i];
    }
    if (dev_.has_body_acc()) {
        dev_.get_body_acc(a.data());
        put3(&Particle::body_acc);
    }
    if (dev_.density_valid()) {
        dev_.download(F_RHO, s.data());
        for (size_t i = 0; i < n; ++i) out[i].rho = s[i];
        dev_.download(F_P, s.data());
        for (size_t i = 0; i < n; ++i) out[i].p = s[i];
    }
    if (dev_.aux_valid()) {
        dev_.download(F_ACC, a.data());
        put3(&Particle::acc);
        dev_.download(F_XSPH, a.data());
        put3(&Particle::xsph);
        dev_.download(F_MU_EFF, s.data());
        for (size_t i = 0; i < n; ++i) out[i].mu_eff = s[i];
    }
    return out;
}

void FluidWorld::set_particles(const std::vector<Particle>& ps) {
    host_ = ps;
    host_dirty_ = true;
    last_was_step_ = false;
}

void FluidWorld::add_particle(const Particle& p) {
    if (!host_dirty_) {
        host_ = particles();
        host_dirty_ = true;
    }
    host_.push_back(p);
    last_was_step_ = false;
}

// world.rs:124-144
void FluidWorld::fill_box(Vec3 mn, Vec3 mx, double spacing, double margin) {
    if (!host_dirty_) {
        host_ = particles();
        host_dirty_ = true;
    }
    Vec3 lo{mn.x + margin, mn.y + margin, mn.z + margin};
    Vec3 hi{mx.x - margin, mx.y - margin, mx.z - margin};
    double d = spacing;
    double mass = lattice_mass(d);
    double x = lo.x;
    while (x <= hi.x) {
        double y = lo.y;
        while (y <= hi.y) {
            double z = lo.z;
            while (z <= hi.z) {
                host_.push_back(Particle::make(Vec3{x, y, z}, mass));
                z += d;
            }
            y += d;
        }
        x += d;
    }
    last_was_step_ = false;
}

double FluidWorld::effective_viscosity(size_t i) {
    flush_host();
    if (i >= dev_.size()) throw std::out_of_range("particle index");
    if (!dev_.aux_valid()) throw std::logic_error("effective_viscosity needs keep_aux before step");
    std::vector<float> mu(dev_.size());
    dev_.download(F_MU_EFF, mu.data());
    return mu[i];
}

void FluidWorld::step(double dt) {
    flush_host();
    dev_.step((float)dt);
    last_was_step_ = true;
}

void FluidWorld::read_vec(Field f, double* out3) {
    flush_host();
    size_t n = dev_.size();
    std::vector<float> a(3 * n);
    dev_.download(f, a.data());
    for (size_t i = 0; i < 3 * n; ++i) out3[i] = (double)a[i];
}
void FluidWorld::read_vec_f32(Field f, float* out3) {
    flush_host();
    dev_.download(f, out3);
}
void FluidWorld::read_scalar(Field f, double* out) {
    flush_host();
    size_t n = dev_.size();
    std::vector<float> a(n);
    dev_.download(f, a.data());
    for (size_t i = 0; i < n; ++i) out[i] = (double)a[i];
}
void FluidWorld::read_scalar_f32(Field f, float* out) {
    flush_host();
    dev_.download(f, out);
}

void FluidWorld::upload(size_t n, const double* pos3, const double* vel3, const double* mass,
                        const uint32_t* material) {
    std::vector<float> p(3 * n), v, m;
    for (size_t i = 0; i < 3 * n; ++i) p[i] = (float)pos3[i];
    if (vel3) {
        v.resize(3 * n);
        for (size_t i = 0; i < 3 * n; ++i) v[i] = (float)vel3[i];
    }
    if (mass) {
        m.resize(n);
        for (size_t i = 0; i < n; ++i) m[i] = (float)mass[i];
    }
    upload_f32(n, p.data(), vel3 ? v.data() : nullptr, mass ? m.data() : nullptr, material);
}

void FluidWorld::upload_f32(size_t n, const float* pos3, const float* vel3, const float* mass,
                            const uint32_t* material) {
    host_.clear();
    host_dirty_ = false;
    dev_.upload(n, pos3, vel3, mass, material, nullptr);
    last_was_step_ = false;
}

void FluidWorld::set_body_acc(const double* b3) {
    flush_host();
    size_t n = dev_.size();
    std::vector<float> b(3 * n);
    for (size_t i = 0; i < 3 * n; ++i) b[i] = (float)b3[i];
    dev_.set_body_acc(b.data());
}

void FluidWorld::get_body_acc(double* out3) {
    flush_host();
    const size_t n = dev_.size();
    std::vector<float> b(3 * n);
    dev_.get_body_acc(b.data());
    for (size_t i = 0; i < 3 * n; ++i) out3[i] = b[i];
}

void FluidWorld::couple_bodies(CoupleBody* bodies, size_t n, double dt, double /*friction*/) {
    flush_host();
    if (n > 0xffffffffull) throw std::invalid_argument("couple_bodies: too many bodies");
    dev_.couple_bodies(bodies, (uint32_t)n, dt);
    last_was_step_ = false;
}

void FluidWorld::couple_points(CouplePoint* pts, size_t n, double dt, double drag, double soft_density) {
    flush_host();
    if (n > 0xffffffffull) throw std::invalid_argument("couple_points: too many points");
    dev_.couple_points(pts, (uint32_t)n, dt, drag, soft_density);
    last_was_step_ = false;
}

void FluidWorld::couple_heat(const HeatGrid& grid, double* src, double dt, double t_ref, double beta,
                             double heat_gain) {
    flush_host();
    // positions and velocities are untouched: the fused watchdog still holds
    dev_.couple_heat(grid, src, dt, t_ref, beta, heat_gain);
}

void FluidWorld::sample_velocity(const double* q3, size_t n, double* out3) {
    flush_host();
    if (n > 0xffffffffull) throw std::invalid_argument("sample_velocity: too many queries");
    dev_.sample_velocity(q3, (uint32_t)n, out3);
}

void FluidWorld::set_keep_aux(bool on) { dev_.set_keep_aux(on); }

void FluidWorld::tape_begin(uint32_t K) {
    flush_host();
    if (dev_.has_body_acc()) throw std::invalid_argument("tape: body_acc is not differentiated");
    dev_.tape_begin(K);
}

void FluidWorld::tape_backward(const double* xbar_T, const double* vbar_T, const double* rhobar_T,
                               double* xbar0, double* vbar0, double* dparams) {
    const size_t n = dev_.size();
    std::vector<float> xb, vb, rb, xo(3 * n), vo(3 * n);
    if (xbar_T) xb.assign(xbar_T, xbar_T + 3 * n);
    if (vbar_T) vb.assign(vbar_T, vbar_T + 3 * n);
    if (rhobar_T) rb.assign(rhobar_T, rhobar_T + n);
    dev_.tape_backward(xbar_T ? xb.data() : nullptr, vbar_T ? vb.data() : nullptr,
                       rhobar_T ? rb.data() : nullptr, xo.data(), vo.data(), dparams);
    for (size_t k = 0; k < 3 * n; ++k) {
        if (xbar0) xbar0[k] = xo[k];
        if (vbar0) vbar0[k] = vo[k];
    }
}

void FluidWorld::tape_end() { dev_.tape_end(); }
int FluidWorld::last_validate() { return dev_.last_validate(); }
void FluidWorld::diagnostics(double out5[5]) {
    flush_host();
    dev_.diagnostics(out5);
}
void FluidWorld::neighbor_stats(uint32_t* cnt, uint64_t* sum, uint64_t* xr) {
    flush_host();
    dev_.neighbor_stats(cnt, sum, xr);
}
void FluidWorld::gather_stats(int pass, uint32_t* cnt, uint64_t* sum, uint64_t* xr) {
    flush_host();
    dev_.gather_stats(pass, cnt, sum, xr);
}

// ---------------------------------------------------------------------------

FluidSubsystem::FluidSubsystem(std::unique_ptr<FluidWorld> w) : world_(std::move(w)) {}

void FluidSubsystem::step(double dt) { world_->step(dt); }

void FluidSubsystem::couple(World& world, double dt) {
    // subsystem.rs:92-143 looks up a HeatField (only when thermal_expansion or
    // heat_gain > 0) and a RigidSubsystem. This library hosts fluid subsystems only,
    // so neither lookup can succeed: the couple phase leaves the fluid untouched.
    (void)world;
    (void)dt;
}

// subsystem.rs:74-90. After a step the fused device watchdog already holds the
// first non-finite particle; otherwise scan a readback.
std::optional<WorldError> FluidSubsystem::validate() {
    int r = 0;
    if (world_->watchdog_current()) {
        r = world_->last_validate();
    } else {
        for (const Particle& p : world_->particles()) {
            if (!std::isfinite(p.pos.x) || !std::isfinite(p.pos.y) || !std::isfinite(p.pos.z)) {
                r = 1;
                break;
            }
            if (!std::isfinite(p.vel.x) || !std::isfinite(p.vel.y) || !std::isfinite(p.vel.z)) {
                r = 2;
                break;
            }
        }
    }
    if (r == 1) return WorldError{WorldError::NonFinite, "fluid", "pos"};
    if (r == 2) return WorldError{WorldError::NonFinite, "fluid", "vel"};
    return std::nullopt;
}

// subsystem.rs:149-163 via the FP64 device reduction
Vec3 FluidSubsystem::total_momentum() {
    double d[5];
    world_->diagnostics(d);
    return Vec3{d[1], d[2], d[3]};
}
double FluidSubsystem::kinetic_energy() {
    double d[5];
    world_->diagnostics(d);
    return d[4];
}

// ---------------------------------------------------------------------------

// to_i64 of the floored key (grid.rs:26-33): NaN / out of range -> 0
static long long flat_key(double x, double inv) {
    double f = std::floor(x * inv);
    if (f >= -9223372036854775808.0 && f < 9223372036854775808.0) return (long long)f;
    return 0;
}

bool FluidWorld::flat_export(FlatShape& sh, float* pos4, float* vel4, float* scalar4, uint64_t n_cap,
                             int32_t* cell_start, uint64_t cell_cap, int32_t* sorted) {
    const std::vector<Particle> ps = particles();
    const uint64_t n = ps.size();
    const double inv = 1.0 / params_.h;  // T::one() / self.cell
    std::vector<long long> keys(3 * n);
    long long mn[3] = {LLONG_MAX, LLONG_MAX, LLONG_MAX}, mx[3] = {LLONG_MIN, LLONG_MIN, LLONG_MIN};
    for (uint64_t i = 0; i < n; ++i) {
        const double c[3] = {ps[i].pos.x, ps[i].pos.y, ps[i].pos.z};
        for (int a = 0; a < 3; ++a) {
            const long long k = flat_key(c[a], inv);
            keys[3 * i + a] = k;
            mn[a] = std::min(mn[a], k);
            mx[a] = std::max(mx[a], k);
        }
    }
    sh.n = n;
    if (n == 0) {  // empty grid: one cell, origin 0
        for (int a = 0; a < 3; ++a) {
            sh.grid_min[a] = 0;
            sh.nc[a] = 1;
        }
        sh.ncells = 1;
        if (cell_cap < 2) return false;
        cell_start[0] = cell_start[1] = 0;
        return true;
    }
    // occupied-key box padded to [min - 1, max + 2] (grid.rs:107-114)
    double vol = 1.0;
    for (int a = 0; a < 3; ++a) {
        sh.grid_min[a] = mn[a] - 1;
        sh.nc[a] = (uint64_t)(mx[a] - sh.grid_min[a] + 3);
        vol *= (double)sh.nc[a];
    }
    if (!(vol < 2147483647.0)) throw std::runtime_error("flat export: more than 2^31 cells in the key box");
    sh.ncells = sh.nc[0] * sh.nc[1] * sh.nc[2];
    if (n_cap < n || cell_cap < sh.ncells + 1) return false;
    // x fastest: c = (i - min_i) + ncx ((j - min_j) + ncy (k - min_k))
    std::vector<uint32_t> cell(n);
    std::vector<int32_t> counts(sh.ncells + 1, 0);
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t ci = (uint64_t)(keys[3 * i] - sh.grid_min[0]);
        const uint64_t cj = (uint64_t)(keys[3 * i + 1] - sh.grid_min[1]);
        const uint64_t ck = (uint64_t)(keys[3 * i + 2] - sh.grid_min[2]);
        cell[i] = (uint32_t)(ci + sh.nc[0] * (cj + sh.nc[1] * ck));
        ++counts[cell[i]];
    }
    int32_t acc = 0;
    for (uint64_t c = 0; c < sh.ncells; ++c) {
        cell_start[c] = acc;
        acc += counts[c];
    }
    cell_start[sh.ncells] = acc;
    std::vector<int32_t> cursor(cell_start, cell_start + sh.ncells);
    for (uint64_t i = 0; i < n; ++i) sorted[cursor[cell[i]]++] = (int32_t)i;  // ascending ids per cell
    for (uint64_t i = 0; i < n; ++i) {
        const Particle& q = ps[i];
        const float p4[4] = {(float)q.pos.x, (float)q.pos.y, (float)q.pos.z, 0.f};
        const float v4[4] = {(float)q.vel.x, (float)q.vel.y, (float)q.vel.z, 0.f};
        const float s4[4] = {(float)q.rho, (float)q.p, (float)q.mass, (float)q.material};
        for (int a = 0; a < 4; ++a) {
            pos4[4 * i + a] = p4[a];
            vel4[4 * i + a] = v4[a];
            scalar4[4 * i + a] = s4[a];
        }
    }
    return true;
}

// ---------------------------------------------------------------------------

static std::vector<HostParams> hosts_of(const std::vector<SphParams>& ps) {
    std::vector<HostParams> h;
    for (const SphParams& p : ps) h.push_back(p.to_host());
    return h;
}

FluidBatch::FluidBatch(const std::vector<SphParams>& params)
    : params_(params), dev_(hosts_of(params)) {
    dev_.set_graph(true);
}

void FluidBatch::set_params(size_t s, const SphParams& p) {
    dev_.set_scene_params(s, p.to_host());
    params_.at(s) = p;
}

void FluidBatch::set_particles_f32(const uint32_t* counts, const float* pos3, const float* vel3,
                                   const float* mass, const uint32_t* material) {
    dev_.upload_scenes(counts, pos3, vel3, mass, material);
}

void FluidBatch::step_checked(double dt, int32_t* status) {
    dev_.step((float)dt);
    dev_.last_validate_all(status);
}

// ---------------------------------------------------------------------------

void World::step_skipping_seeded(double dt, const std::vector<bool>& skip, uint64_t seed) {
    seed_ = seed;
    started_ = true;
    for (size_t i = 0; i < subs_.size(); ++i) {
        if (i < skip.size() && skip[i]) continue;
        subs_[i]->step(dt);
    }
    size_t n = subs_.size();
    for (size_t i = 0; i < n; ++i) {
        std::unique_ptr<Subsystem> me = remove(i);
        me->couple(*this, dt);
        insert(i, std::move(me));
    }
    t_ += dt;
}

std::optional<WorldError> World::step_skipping_checked(double dt, const std::vector<bool>& skip) {
    step_skipping(dt, skip);
    for (auto& s : subs_) {
        auto e = s->validate();
        if (e) return e;
    }
    return std::nullopt;
}

Vec3 World::total_momentum() {
    Vec3 p;
    for (auto& s : subs_) {
        Vec3 q = s->total_momentum();
        p.x += q.x;
        p.y += q.y;
        p.z += q.z;
    }
    return p;
}

double World::kinetic_energy() {
    double e = 0.0;
    for (auto& s : subs_) e += s->kinetic_energy();
    return e;
}

std::unique_ptr<Subsystem> World::remove(size_t i) {
    std::unique_ptr<Subsystem> s = std::move(subs_.at(i));
    subs_.erase(subs_.begin() + (long)i);
    return s;
}

void World::insert(size_t i, std::unique_ptr<Subsystem> s) {
    subs_.insert(subs_.begin() + (long)i, std::move(s));
}

}  // namespace phy
