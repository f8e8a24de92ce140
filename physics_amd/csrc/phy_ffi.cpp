// phy_ffi.cpp — extern "C" boundary (include/phy_ffi.h).
//
// Every entry point catches all C++ exceptions: none may cross the C boundary,
// as the reference's guard() = catch_unwind (crates/phy-ffi/src/lib.rs:73-75).
// The reason of the last failure is kept per thread for phy_last_error().
#include "../../include/phy_ffi.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "archive.h"
#include "world.h"

static_assert(PHY_TAPE_PARAMS == phy::kAdjParams, "tape gradient layout");

using phy::FluidBatch;
using phy::FluidSubsystem;
using phy::FluidWorld;
using phy::World;

namespace {

thread_local std::string g_err;

void set_err(const char* what) { g_err = what ? what : "unknown error"; }

template <class R, class F>
R guard(R fail, F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        set_err(e.what());
    } catch (...) {
        set_err("unknown exception");
    }
    return fail;
}

World* as_world(PhyWorldHandle* w) { return reinterpret_cast<World*>(w); }
FluidBatch* as_batch(PhyFluidBatch* b) { return reinterpret_cast<FluidBatch*>(b); }
FluidBatch* need_batch(PhyFluidBatch* b) {
    if (!b) throw std::invalid_argument("null batch handle");
    return as_batch(b);
}
PhyWorldHandle* box_world(World* w) { return reinterpret_cast<PhyWorldHandle*>(w); }

// First fluid subsystem of the world, or null.
FluidSubsystem* first_fluid(World* w) {
    for (size_t i = 0; i < w->subsystem_count(); ++i)
        if (auto* f = dynamic_cast<FluidSubsystem*>(w->get(i))) return f;
    return nullptr;
}

FluidSubsystem* need_fluid(PhyWorldHandle* h) {
    if (!h) throw std::invalid_argument("null world handle");
    FluidSubsystem* f = first_fluid(as_world(h));
    if (!f) throw std::invalid_argument("world has no fluid subsystem");
    return f;
}

phy::SphParams from_c(const PhySphParams& c) {
    phy::SphParams p;
    p.rest_density = c.rest_density;
    p.stiffness = c.stiffness;
    p.viscosity = c.viscosity;
    if (c.n_visc_k == 0 || c.n_visc_k > PHY_MAX_MATERIALS || c.n_visc_n == 0 ||
        c.n_visc_n > PHY_MAX_MATERIALS)
        throw std::invalid_argument("PhySphParams: material table sizes must be in 1..16");
    p.visc_k.assign(c.visc_k, c.visc_k + c.n_visc_k);
    p.visc_n.assign(c.visc_n, c.visc_n + c.n_visc_n);
    p.shear_min = c.shear_min;
    p.mass = c.mass;
    p.h = c.h;
    p.gravity = phy::Vec3{c.gravity[0], c.gravity[1], c.gravity[2]};
    p.bounds_min = phy::Vec3{c.bounds_min[0], c.bounds_min[1], c.bounds_min[2]};
    p.bounds_max = phy::Vec3{c.bounds_max[0], c.bounds_max[1], c.bounds_max[2]};
    p.boundary_damp = c.boundary_damp;
    p.xsph_eps = c.xsph_eps;
    return p;
}

void to_c(const phy::SphParams& p, PhySphParams* c) {
    std::memset(c, 0, sizeof(*c));
    c->rest_density = p.rest_density;
    c->stiffness = p.stiffness;
    c->viscosity = p.viscosity;
    c->n_visc_k = (uint32_t)p.visc_k.size();
    c->n_visc_n = (uint32_t)p.visc_n.size();
    for (size_t i = 0; i < p.visc_k.size() && i < PHY_MAX_MATERIALS; ++i) c->visc_k[i] = p.visc_k[i];
    for (size_t i = 0; i < p.visc_n.size() && i < PHY_MAX_MATERIALS; ++i) c->visc_n[i] = p.visc_n[i];
    c->shear_min = p.shear_min;
    c->mass = p.mass;
    c->h = p.h;
    const phy::Vec3* v[3] = {&p.gravity, &p.bounds_min, &p.bounds_max};
    double* o[3] = {c->gravity, c->bounds_min, c->bounds_max};
    for (int k = 0; k < 3; ++k) {
        o[k][0] = v[k]->x;
        o[k][1] = v[k]->y;
        o[k][2] = v[k]->z;
    }
    c->boundary_damp = p.boundary_damp;
    c->xsph_eps = p.xsph_eps;
}

// lib.rs:113-126
World* build_fluid_world() {
    auto fw = std::make_unique<FluidWorld>(phy::SphParams::defaults());
    fw->fill_box(phy::Vec3{-4.0, -4.0, -1.0}, phy::Vec3{0.0, 4.0, 1.0}, 0.18, 0.05);
    auto* w = new World();
    w->add_subsystem(std::make_unique<FluidSubsystem>(std::move(fw)));
    return w;
}

// Concatenated readback over every fluid subsystem (lib.rs:399-429): writes
// whole particles while 3 more values fit; returns particles written.
template <class T, class Read>
uintptr_t read_all(PhyWorldHandle* h, T* buf, uintptr_t len, int comps, Read&& rd) {
    if (!h || !buf) return 0;
    World* w = as_world(h);
    uintptr_t written = 0;  // values
    for (size_t i = 0; i < w->subsystem_count(); ++i) {
        auto* f = dynamic_cast<FluidSubsystem*>(w->get(i));
        if (!f) continue;
        size_t n = f->world().len();
        if (!n) continue;
        std::vector<T> tmp(n * comps);
        rd(f->world(), tmp.data());
        for (size_t p = 0; p < n; ++p) {
            if (written + comps > len) return written / comps;
            std::memcpy(buf + written, tmp.data() + p * comps, comps * sizeof(T));
            written += comps;
        }
    }
    return written / comps;
}

int field_comps(int32_t f) {
    switch (f) {
        case PHY_FIELD_POS:
        case PHY_FIELD_VEL:
        case PHY_FIELD_ACC:
        case PHY_FIELD_XSPH:
            return 3;
        case PHY_FIELD_RHO:
        case PHY_FIELD_P:
        case PHY_FIELD_MASS:
        case PHY_FIELD_MU_EFF:
            return 1;
        default:
            return 0;
    }
}

}  // namespace

extern "C" {

uint32_t phy_ffi_abi_version(void) { return PHY_FFI_ABI_VERSION; }

const char* phy_last_error(void) { return g_err.c_str(); }

PhyWorldHandle* phy_world_create_fluid(void) {
    return guard<PhyWorldHandle*>(nullptr, [] { return box_world(build_fluid_world()); });
}

#define PHY_UNSUPPORTED_FACTORY(name, what)                                              \
    PhyWorldHandle* name(void) {                                                         \
        set_err(what " worlds are not provided by this library (fluid path only)");      \
        return nullptr;                                                                  \
    }
PHY_UNSUPPORTED_FACTORY(phy_world_create_rigid, "rigid")
PHY_UNSUPPORTED_FACTORY(phy_world_create_rigid_empty, "rigid")
PHY_UNSUPPORTED_FACTORY(phy_world_create_granular, "granular")
PHY_UNSUPPORTED_FACTORY(phy_world_create_coupled, "coupled fluid+rigid")
PHY_UNSUPPORTED_FACTORY(phy_world_create_soft_empty, "soft-body")
PHY_UNSUPPORTED_FACTORY(phy_world_create_field_empty, "field")
PHY_UNSUPPORTED_FACTORY(phy_world_create_optics_empty, "optics")
PHY_UNSUPPORTED_FACTORY(phy_world_create_solid_empty, "solid")
#undef PHY_UNSUPPORTED_FACTORY

int32_t phy_world_step(PhyWorldHandle* w, double dt) {
    if (!w) return -1;
    return guard<int32_t>(-1, [&] {
        as_world(w)->step(dt);
        return 0;
    });
}

int32_t phy_world_step_n(PhyWorldHandle* w, double dt, uint32_t n) {
    if (!w) return -1;
    return guard<int32_t>(-1, [&] {
        for (uint32_t i = 0; i < n; ++i) as_world(w)->step(dt);
        return 0;
    });
}

int32_t phy_world_step_checked(PhyWorldHandle* w, double dt) {
    if (!w) return -1;
    return guard<int32_t>(-1, [&] {
        auto e = as_world(w)->step_checked(dt);
        if (!e) return 0;
        if (e->kind == phy::WorldError::NonFinite) return 2;
        if (e->kind == phy::WorldError::Stalled) return 3;
        return 4;
    });
}

int32_t phy_world_step_f32(PhyWorldHandle* w, float dt) { return phy_world_step(w, (double)dt); }

double phy_world_time(PhyWorldHandle* w) {
    if (!w) return NAN;
    return as_world(w)->time();
}

uintptr_t phy_world_sub_count(PhyWorldHandle* w) {
    if (!w) return 0;
    return as_world(w)->subsystem_count();
}

uintptr_t phy_world_fluid_count(PhyWorldHandle* w) {
    if (!w) return 0;
    return guard<uintptr_t>(0, [&] {
        World* W = as_world(w);
        uintptr_t total = 0;
        for (size_t i = 0; i < W->subsystem_count(); ++i)
            if (auto* f = dynamic_cast<FluidSubsystem*>(W->get(i))) total += f->world().len();
        return total;
    });
}

uintptr_t phy_world_get_fluid_positions(PhyWorldHandle* w, double* buf, uintptr_t len) {
    return guard<uintptr_t>(0, [&] {
        return read_all<double>(w, buf, len, 3, [](FluidWorld& f, double* o) { f.read_vec(phy::F_POS, o); });
    });
}
uintptr_t phy_world_get_fluid_velocities(PhyWorldHandle* w, double* buf, uintptr_t len) {
    return guard<uintptr_t>(0, [&] {
        return read_all<double>(w, buf, len, 3, [](FluidWorld& f, double* o) { f.read_vec(phy::F_VEL, o); });
    });
}
uintptr_t phy_world_get_fluid_positions_f32(PhyWorldHandle* w, float* buf, uintptr_t len) {
    return guard<uintptr_t>(0, [&] {
        return read_all<float>(w, buf, len, 3, [](FluidWorld& f, float* o) { f.read_vec_f32(phy::F_POS, o); });
    });
}
uintptr_t phy_world_get_fluid_velocities_f32(PhyWorldHandle* w, float* buf, uintptr_t len) {
    return guard<uintptr_t>(0, [&] {
        return read_all<float>(w, buf, len, 3, [](FluidWorld& f, float* o) { f.read_vec_f32(phy::F_VEL, o); });
    });
}

uintptr_t phy_world_rigid_count(PhyWorldHandle* w) {
    (void)w;
    return 0;
}
uintptr_t phy_world_get_rigid_transforms(PhyWorldHandle* w, double* buf, uintptr_t len) {
    (void)w; (void)buf; (void)len;
    return 0;
}
uintptr_t phy_world_get_rigid_transforms_f32(PhyWorldHandle* w, float* buf, uintptr_t len) {
    (void)w; (void)buf; (void)len;
    return 0;
}
int64_t phy_world_rigid_add_body(PhyWorldHandle* w, int32_t shape_kind, double mass,
                                 const double* pos7, const double* inertia3) {
    (void)w; (void)shape_kind; (void)mass; (void)pos7; (void)inertia3;
    set_err("no rigid subsystem in this library");
    return -1;
}
int32_t phy_world_rigid_apply_force(PhyWorldHandle* w, int64_t id, const double* f3, double dt,
                                    int32_t mode) {
    (void)w; (void)id; (void)f3; (void)dt; (void)mode;
    set_err("no rigid subsystem in this library");
    return -1;
}
int32_t phy_world_rigid_apply_torque(PhyWorldHandle* w, int64_t id, const double* t3, double dt,
                                     int32_t mode) {
    (void)w; (void)id; (void)t3; (void)dt; (void)mode;
    set_err("no rigid subsystem in this library");
    return -1;
}
int32_t phy_world_rigid_get_velocity(PhyWorldHandle* w, int64_t id, double* out3) {
    (void)w; (void)id; (void)out3;
    set_err("no rigid subsystem in this library");
    return -1;
}
int32_t phy_world_rigid_get_angular_velocity(PhyWorldHandle* w, int64_t id, double* out3) {
    (void)w; (void)id; (void)out3;
    set_err("no rigid subsystem in this library");
    return -1;
}

// lib.rs:517-530: the phy-io JSON archive of the world, written to path.
int32_t phy_world_save(PhyWorldHandle* w, const char* path) {
    return guard<int32_t>(-1, [&] {
        if (!w) throw std::invalid_argument("null world handle");
        if (!path) throw std::invalid_argument("null path");
        const std::string json = phy::save_world_json(*as_world(w));
        std::FILE* f = std::fopen(path, "wb");
        if (!f) throw std::runtime_error(std::string("phy_world_save: cannot open ") + path);
        const size_t wr = std::fwrite(json.data(), 1, json.size(), f);
        const int rc = std::fclose(f);
        if (wr != json.size() || rc != 0) throw std::runtime_error(std::string("phy_world_save: write failed: ") + path);
        return 0;
    });
}

// lib.rs:532-542: a new world from an archive file; NULL on any failure.
PhyWorldHandle* phy_world_load(const char* path) {
    return guard<PhyWorldHandle*>(nullptr, [&] {
        if (!path) throw std::invalid_argument("null path");
        std::FILE* f = std::fopen(path, "rb");
        if (!f) throw std::runtime_error(std::string("phy_world_load: cannot open ") + path);
        std::string json;
        char buf[1 << 16];
        size_t r;
        while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) json.append(buf, r);
        std::fclose(f);
        return box_world(phy::load_world_json(json).release());
    });
}

void phy_world_destroy(PhyWorldHandle* w) {
    if (!w) return;
    guard<int>(0, [&] {
        delete as_world(w);
        return 0;
    });
}

// ---- additive ----------------------------------------------------------------

void phy_sph_params_defaults(PhySphParams* out) {
    if (out) to_c(phy::SphParams::defaults(), out);
}

PhyWorldHandle* phy_world_create_fluid_with(const PhySphParams* params) {
    if (!params) {
        set_err("null params");
        return nullptr;
    }
    return guard<PhyWorldHandle*>(nullptr, [&] {
        auto fw = std::make_unique<FluidWorld>(from_c(*params));
        auto* w = new World();
        w->add_subsystem(std::make_unique<FluidSubsystem>(std::move(fw)));
        return box_world(w);
    });
}

int32_t phy_fluid_set_params(PhyWorldHandle* w, const PhySphParams* params) {
    return guard<int32_t>(-1, [&] {
        if (!params) throw std::invalid_argument("null params");
        need_fluid(w)->world().set_params(from_c(*params));
        return 0;
    });
}

int32_t phy_fluid_get_params(PhyWorldHandle* w, PhySphParams* out) {
    return guard<int32_t>(-1, [&] {
        if (!out) throw std::invalid_argument("null out");
        to_c(need_fluid(w)->world().params(), out);
        return 0;
    });
}

int32_t phy_fluid_fill_box(PhyWorldHandle* w, const double* mn, const double* mx, double spacing,
                           double margin) {
    return guard<int32_t>(-1, [&] {
        if (!mn || !mx) throw std::invalid_argument("null box corner");
        need_fluid(w)->world().fill_box(phy::Vec3{mn[0], mn[1], mn[2]}, phy::Vec3{mx[0], mx[1], mx[2]},
                                        spacing, margin);
        return 0;
    });
}

double phy_fluid_lattice_mass(PhyWorldHandle* w, double spacing) {
    return guard<double>(NAN, [&] { return need_fluid(w)->world().lattice_mass(spacing); });
}

int32_t phy_fluid_set_particles(PhyWorldHandle* w, uintptr_t n, const double* pos3, const double* vel3,
                                const double* mass, const uint32_t* material) {
    return guard<int32_t>(-1, [&] {
        if (n && !pos3) throw std::invalid_argument("null positions");
        need_fluid(w)->world().upload(n, pos3, vel3, mass, material);
        return 0;
    });
}

int32_t phy_fluid_set_particles_f32(PhyWorldHandle* w, uintptr_t n, const float* pos3, const float* vel3,
                                    const float* mass, const uint32_t* material) {
    return guard<int32_t>(-1, [&] {
        if (n && !pos3) throw std::invalid_argument("null positions");
        need_fluid(w)->world().upload_f32(n, pos3, vel3, mass, material);
        return 0;
    });
}

uintptr_t phy_fluid_get_field(PhyWorldHandle* w, int32_t field, double* buf, uintptr_t len) {
    return guard<uintptr_t>(0, [&] {
        int c = field_comps(field);
        if (!c) throw std::invalid_argument("unknown field");
        FluidWorld& f = need_fluid(w)->world();
        size_t n = f.len();
        if (!buf || !n) return (uintptr_t)0;
        std::vector<double> tmp(n * c);
        if (c == 3) f.read_vec((phy::Field)field, tmp.data());
        else f.read_scalar((phy::Field)field, tmp.data());
        uintptr_t np = std::min<uintptr_t>(n, len / c);
        std::memcpy(buf, tmp.data(), np * c * sizeof(double));
        return np;
    });
}

uintptr_t phy_fluid_get_field_f32(PhyWorldHandle* w, int32_t field, float* buf, uintptr_t len) {
    return guard<uintptr_t>(0, [&] {
        int c = field_comps(field);
        if (!c) throw std::invalid_argument("unknown field");
        FluidWorld& f = need_fluid(w)->world();
        size_t n = f.len();
        if (!buf || !n) return (uintptr_t)0;
        std::vector<float> tmp(n * c);
        if (c == 3) f.read_vec_f32((phy::Field)field, tmp.data());
        else f.read_scalar_f32((phy::Field)field, tmp.data());
        uintptr_t np = std::min<uintptr_t>(n, len / c);
        std::memcpy(buf, tmp.data(), np * c * sizeof(float));
        return np;
    });
}

int32_t phy_fluid_set_keep_aux(PhyWorldHandle* w, int32_t on) {
    return guard<int32_t>(-1, [&] {
        need_fluid(w)->world().set_keep_aux(on != 0);
        return 0;
    });
}

int32_t phy_fluid_set_body_acc(PhyWorldHandle* w, const double* b3, uintptr_t len) {
    return guard<int32_t>(-1, [&] {
        FluidWorld& f = need_fluid(w)->world();
        if (!b3 || len < 3 * f.len()) throw std::invalid_argument("body_acc buffer too short");
        f.set_body_acc(b3);
        return 0;
    });
}

int32_t phy_fluid_get_body_acc(PhyWorldHandle* w, double* out3, uintptr_t len) {
    return guard<int32_t>(-1, [&] {
        FluidWorld& f = need_fluid(w)->world();
        if (!out3 || len < 3 * f.len()) throw std::invalid_argument("body_acc buffer too short");
        f.get_body_acc(out3);
        return 0;
    });
}

// ---- coupling (world.rs:400-773) ----

int32_t phy_fluid_couple_bodies(PhyWorldHandle* w, PhyCoupleBody* bodies, uint32_t n, double dt,
                                double friction) {
    return guard<int32_t>(-1, [&] {
        FluidWorld& f = need_fluid(w)->world();
        if (n && !bodies) throw std::invalid_argument("couple_bodies: null bodies");
        std::vector<phy::CoupleBody> b(n);
        for (uint32_t k = 0; k < n; ++k) {
            const PhyCoupleBody& s = bodies[k];
            if (s.shape < PHY_SHAPE_SPHERE || s.shape > PHY_SHAPE_CAPSULE)
                throw std::invalid_argument("couple_bodies: unknown shape kind");
            for (int a = 0; a < 3; ++a) {
                b[k].pos[a] = s.pose[a];
                b[k].vel[a] = s.vel[a];
                b[k].dims[a] = s.dims[a];
            }
            for (int a = 0; a < 4; ++a) b[k].q[a] = s.pose[3 + a];
            b[k].inv_mass = s.inv_mass;
            b[k].shape = s.shape;
            b[k].pad = 0;
        }
        f.couple_bodies(b.data(), n, dt, friction);
        for (uint32_t k = 0; k < n; ++k)
            for (int a = 0; a < 3; ++a) bodies[k].vel[a] = b[k].vel[a];
        return 0;
    });
}

int32_t phy_fluid_couple_points(PhyWorldHandle* w, PhyCouplePoint* pts, uint32_t n, double dt, double drag,
                                double soft_density) {
    return guard<int32_t>(-1, [&] {
        FluidWorld& f = need_fluid(w)->world();
        if (n && !pts) throw std::invalid_argument("couple_points: null points");
        std::vector<phy::CouplePoint> p(n);
        for (uint32_t k = 0; k < n; ++k) {
            for (int a = 0; a < 3; ++a) {
                p[k].pos[a] = pts[k].pos[a];
                p[k].vel[a] = pts[k].vel[a];
                p[k].force[a] = pts[k].force[a];
            }
            p[k].mass = pts[k].mass;
        }
        f.couple_points(p.data(), n, dt, drag, soft_density);
        for (uint32_t k = 0; k < n; ++k)
            for (int a = 0; a < 3; ++a) pts[k].force[a] = p[k].force[a];
        return 0;
    });
}

int32_t phy_fluid_couple_heat(PhyWorldHandle* w, const PhyHeatGrid* grid, double dt, double t_ref, double beta,
                              double heat_gain) {
    return guard<int32_t>(-1, [&] {
        FluidWorld& f = need_fluid(w)->world();
        if (!grid || !grid->u) throw std::invalid_argument("couple_heat: null grid");
        phy::HeatGrid g;
        g.nx = grid->nx;
        g.ny = grid->ny;
        g.nz = grid->nz;
        for (int a = 0; a < 3; ++a) g.origin[a] = grid->origin[a];
        g.dx = grid->dx;
        g.u = grid->u;
        f.couple_heat(g, grid->src, dt, t_ref, beta, heat_gain);
        return 0;
    });
}

int32_t phy_fluid_sample_velocity(PhyWorldHandle* w, const double* q3, uint32_t n, double* out3) {
    return guard<int32_t>(-1, [&] {
        FluidWorld& f = need_fluid(w)->world();
        if (n && (!q3 || !out3)) throw std::invalid_argument("sample_velocity: null buffer");
        f.sample_velocity(q3, n, out3);
        return 0;
    });
}

int32_t phy_fluid_diagnostics_f64(PhyWorldHandle* w, double* out5) {
    return guard<int32_t>(-1, [&] {
        if (!out5) throw std::invalid_argument("null out");
        need_fluid(w)->world().diagnostics(out5);
        return 0;
    });
}

int32_t phy_fluid_neighbor_stats(PhyWorldHandle* w, uint32_t* counts, uint64_t* sums, uint64_t* xors,
                                 uintptr_t len) {
    return guard<int32_t>(-1, [&] {
        FluidWorld& f = need_fluid(w)->world();
        if (!counts || !sums || !xors || len < f.len()) throw std::invalid_argument("buffers too short");
        f.neighbor_stats(counts, sums, xors);
        return 0;
    });
}

int32_t phy_fluid_gather_stats(PhyWorldHandle* w, int32_t pass, uint32_t* counts, uint64_t* sums,
                               uint64_t* xors, uintptr_t len) {
    return guard<int32_t>(-1, [&] {
        FluidWorld& f = need_fluid(w)->world();
        if (!counts || !sums || !xors || len < f.len()) throw std::invalid_argument("buffers too short");
        f.gather_stats(pass, counts, sums, xors);
        return 0;
    });
}

int32_t phy_fluid_tape_begin(PhyWorldHandle* w, uint32_t checkpoint_every) {
    return guard<int32_t>(-1, [&] {
        need_fluid(w)->world().tape_begin(checkpoint_every);
        return 0;
    });
}

uint32_t phy_fluid_tape_steps(PhyWorldHandle* w) {
    return guard<uint32_t>(0, [&] { return (uint32_t)need_fluid(w)->world().tape_steps(); });
}

int32_t phy_fluid_tape_backward(PhyWorldHandle* w, const double* xbar_T, const double* vbar_T,
                                const double* rhobar_T, double* xbar0, double* vbar0,
                                double* dparams) {
    return guard<int32_t>(-1, [&] {
        double g[phy::kAdjParams];
        need_fluid(w)->world().tape_backward(xbar_T, vbar_T, rhobar_T, xbar0, vbar0, g);
        if (dparams) {
            dparams[0] = g[0];
            dparams[1] = g[1];
        }
        return 0;
    });
}

int32_t phy_fluid_tape_backward_full(PhyWorldHandle* w, const double* xbar_T, const double* vbar_T,
                                     const double* rhobar_T, double* xbar0, double* vbar0,
                                     double* dparams) {
    return guard<int32_t>(-1, [&] {
        double g[phy::kAdjParams];
        need_fluid(w)->world().tape_backward(xbar_T, vbar_T, rhobar_T, xbar0, vbar0, g);
        if (dparams) std::copy(g, g + phy::kAdjParams, dparams);
        return 0;
    });
}

int32_t phy_fluid_tape_end(PhyWorldHandle* w) {
    return guard<int32_t>(-1, [&] {
        need_fluid(w)->world().tape_end();
        return 0;
    });
}

int32_t phy_fluid_sync(PhyWorldHandle* w) {
    return guard<int32_t>(-1, [&] {
        if (!w) throw std::invalid_argument("null world handle");
        World* W = as_world(w);
        for (size_t i = 0; i < W->subsystem_count(); ++i)
            if (auto* f = dynamic_cast<FluidSubsystem*>(W->get(i))) f->world().device().sync();
        return 0;
    });
}

int32_t phy_fluid_profile_begin(PhyWorldHandle* w) {
    return guard<int32_t>(-1, [&] {
        need_fluid(w)->world().device().profile_begin();
        return 0;
    });
}

int32_t phy_fluid_profile_end(PhyWorldHandle* w, double* ms4, uint32_t* steps) {
    return guard<int32_t>(-1, [&] {
        if (!ms4) throw std::invalid_argument("null out");
        uint32_t s = need_fluid(w)->world().device().profile_end(ms4);
        if (steps) *steps = s;
        return 0;
    });
}

int32_t phy_fluid_set_graph(PhyWorldHandle* w, int32_t on) {
    return guard<int32_t>(-1, [&] {
        need_fluid(w)->world().device().set_graph(on != 0);
        return 0;
    });
}

static void to_c(const phy::Tuning& t, PhyTuning* out) {
    out->density_list = t.density_list;
    out->force_list = t.force_list;
    out->split = t.split;
    out->cell_div = t.cell_div;
    out->block = t.block;
}

static int32_t tune(phy::DeviceFluid& d, double dt, uint32_t steps, const char* cache_path,
                    PhyTuning* out, double* ms12) {
    const bool cached = d.autotune((float)dt, steps, cache_path, ms12);
    if (out) to_c(d.tuning(), out);
    return cached ? 1 : 0;
}

static phy::Tuning from_c(const PhyTuning* t) {
    if (!t) throw std::invalid_argument("null tuning");
    phy::Tuning r;
    r.density_list = t->density_list;
    r.force_list = t->force_list;
    r.split = t->split;
    r.cell_div = t->cell_div;
    r.block = t->block;
    return r;
}

int32_t phy_fluid_get_tuning(PhyWorldHandle* w, PhyTuning* out) {
    return guard<int32_t>(-1, [&] {
        if (!out) throw std::invalid_argument("null out");
        to_c(need_fluid(w)->world().device().tuning(), out);
        return 0;
    });
}

int32_t phy_fluid_set_tuning(PhyWorldHandle* w, const PhyTuning* t) {
    return guard<int32_t>(-1, [&] {
        need_fluid(w)->world().device().set_tuning(from_c(t));
        return 0;
    });
}

int32_t phy_fluid_autotune(PhyWorldHandle* w, double dt, uint32_t steps, const char* cache_path,
                           PhyTuning* out, double* ms12) {
    return guard<int32_t>(-1, [&] {
        FluidWorld& f = need_fluid(w)->world();
        f.flush();
        int32_t r = tune(f.device(), dt, steps, cache_path, out, ms12);
        f.mark_state_changed();
        return r;
    });
}

int32_t phy_fluid_batch_get_tuning(PhyFluidBatch* b, PhyTuning* out) {
    return guard<int32_t>(-1, [&] {
        if (!out) throw std::invalid_argument("null out");
        to_c(need_batch(b)->device().tuning(), out);
        return 0;
    });
}

int32_t phy_fluid_batch_set_tuning(PhyFluidBatch* b, const PhyTuning* t) {
    return guard<int32_t>(-1, [&] {
        need_batch(b)->device().set_tuning(from_c(t));
        return 0;
    });
}

int32_t phy_fluid_batch_autotune(PhyFluidBatch* b, double dt, uint32_t steps, const char* cache_path,
                                 PhyTuning* out, double* ms12) {
    return guard<int32_t>(-1, [&] { return tune(need_batch(b)->device(), dt, steps, cache_path, out, ms12); });
}

int32_t phy_fluid_flat_export(PhyWorldHandle* w, PhyFlatInfo* info, float* pos4, float* vel4,
                              float* scalar4, uintptr_t n_cap, int32_t* cell_start, uintptr_t cell_cap,
                              int32_t* sorted) {
    return guard<int32_t>(-1, [&] {
        if (!info) throw std::invalid_argument("null info");
        FluidWorld& f = need_fluid(w)->world();
        const phy::SphParams& p = f.params();
        std::memset(info, 0, sizeof(*info));
        info->h = (float)p.h;
        info->rest_density = (float)p.rest_density;
        info->stiffness = (float)p.stiffness;
        info->shear_min = (float)p.shear_min;
        info->gravity[0] = (float)p.gravity.x;
        info->gravity[1] = (float)p.gravity.y;
        info->gravity[2] = (float)p.gravity.z;
        info->n_visc = (uint32_t)std::min<size_t>(p.visc_k.size(), PHY_MAX_MATERIALS);
        for (uint32_t m = 0; m < info->n_visc; ++m) info->visc_k[m] = (float)p.visc_k[m];
        for (size_t m = 0; m < p.visc_n.size() && m < PHY_MAX_MATERIALS; ++m) info->visc_n[m] = (float)p.visc_n[m];
        const bool room = n_cap > 0 || cell_cap > 0;
        FluidWorld::FlatShape sh;
        bool filled = false;
        if (room && pos4 && vel4 && scalar4 && cell_start && sorted)
            filled = f.flat_export(sh, pos4, vel4, scalar4, n_cap, cell_start, cell_cap, sorted);
        else
            f.flat_export(sh, nullptr, nullptr, nullptr, 0, nullptr, 0, nullptr);
        info->n = sh.n;
        info->ncells = sh.ncells;
        for (int a = 0; a < 3; ++a) {
            info->grid_min[a] = sh.grid_min[a];
            info->nc[a] = sh.nc[a];
        }
        return filled ? 0 : 1;
    });
}

int32_t phy_device_mem_info(uint64_t* free_bytes, uint64_t* total_bytes) {
    return guard<int32_t>(-1, [&] {
        size_t f = 0, t = 0;
        phy::hip_check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
        return 0;
    });
}

int32_t phy_fluid_diagnostics_f64_noalloc(const float* posm4, const float* vel4, uint64_t n, double* out5,
                                          void* stream) {
    return guard<int32_t>(-1, [&] {
        if (n > phy::kDiagNoAllocMaxParticles)
            throw std::invalid_argument("phy_fluid_diagnostics_f64_noalloc: n exceeds 2^24 particles");
        if (!out5 || (n > 0 && (!posm4 || !vel4)))
            throw std::invalid_argument("phy_fluid_diagnostics_f64_noalloc: null pointer");
        phy::launch_diag_noalloc(static_cast<hipStream_t>(stream), n, reinterpret_cast<const float4*>(posm4),
                                 reinterpret_cast<const float4*>(vel4), out5);
        return 0;
    });
}

int32_t phy_device_mem_stats(uint64_t* in_use, uint64_t* peak) {
    return guard<int32_t>(-1, [&] {
        if (in_use) *in_use = phy::mem_in_use();
        if (peak) *peak = phy::mem_peak();
        return 0;
    });
}

int32_t phy_device_mem_reset_peak(void) {
    return guard<int32_t>(-1, [&] {
        phy::mem_reset_peak();
        return 0;
    });
}

// ---------------------------------------------------------------------------
// batches

PhyFluidBatch* phy_fluid_batch_create(const PhySphParams* params, uint32_t n_scenes) {
    return guard<PhyFluidBatch*>(nullptr, [&] {
        if (!params || n_scenes == 0) throw std::invalid_argument("batch: need n_scenes >= 1 parameter sets");
        std::vector<phy::SphParams> ps;
        for (uint32_t s = 0; s < n_scenes; ++s) ps.push_back(from_c(params[s]));
        return reinterpret_cast<PhyFluidBatch*>(new FluidBatch(ps));
    });
}

void phy_fluid_batch_destroy(PhyFluidBatch* b) {
    guard<int>(0, [&] {
        delete as_batch(b);
        return 0;
    });
}

uint32_t phy_fluid_batch_scenes(PhyFluidBatch* b) {
    return guard<uint32_t>(0, [&] { return (uint32_t)need_batch(b)->scenes(); });
}

int32_t phy_fluid_batch_set_params(PhyFluidBatch* b, uint32_t scene, const PhySphParams* params) {
    return guard<int32_t>(-1, [&] {
        if (!params) throw std::invalid_argument("null params");
        need_batch(b)->set_params(scene, from_c(*params));
        return 0;
    });
}

int32_t phy_fluid_batch_get_params(PhyFluidBatch* b, uint32_t scene, PhySphParams* out) {
    return guard<int32_t>(-1, [&] {
        if (!out) throw std::invalid_argument("null out");
        to_c(need_batch(b)->params(scene), out);
        return 0;
    });
}

int32_t phy_fluid_batch_set_particles_f32(PhyFluidBatch* b, const uint32_t* counts, const float* pos3,
                                          const float* vel3, const float* mass,
                                          const uint32_t* material) {
    return guard<int32_t>(-1, [&] {
        FluidBatch* B = need_batch(b);
        if (!counts) throw std::invalid_argument("null counts");
        uint64_t n = 0;
        for (size_t s = 0; s < B->scenes(); ++s) n += counts[s];
        if (n && !pos3) throw std::invalid_argument("null positions");
        B->set_particles_f32(counts, pos3, vel3, mass, material);
        return 0;
    });
}

int32_t phy_fluid_batch_offsets(PhyFluidBatch* b, uint32_t* offsets) {
    return guard<int32_t>(-1, [&] {
        FluidBatch* B = need_batch(b);
        if (!offsets) throw std::invalid_argument("null out");
        for (size_t s = 0; s < B->scenes(); ++s) offsets[s] = B->device().offset(s);
        offsets[B->scenes()] = (uint32_t)B->device().size();
        return 0;
    });
}

int32_t phy_fluid_batch_step(PhyFluidBatch* b, double dt, uint32_t n) {
    return guard<int32_t>(-1, [&] {
        FluidBatch* B = need_batch(b);
        for (uint32_t i = 0; i < n; ++i) B->step(dt);
        return 0;
    });
}

int32_t phy_fluid_batch_step_checked(PhyFluidBatch* b, double dt, int32_t* status) {
    return guard<int32_t>(-1, [&] {
        FluidBatch* B = need_batch(b);
        if (!status) throw std::invalid_argument("null status");
        B->step_checked(dt, status);
        int32_t bad = 0;
        for (size_t s = 0; s < B->scenes(); ++s) bad += status[s] != 0;
        return bad;
    });
}

int32_t phy_fluid_batch_set_graph(PhyFluidBatch* b, int32_t on) {
    return guard<int32_t>(-1, [&] {
        need_batch(b)->device().set_graph(on != 0);
        return 0;
    });
}

int32_t phy_fluid_batch_set_keep_aux(PhyFluidBatch* b, int32_t on) {
    return guard<int32_t>(-1, [&] {
        need_batch(b)->device().set_keep_aux(on != 0);
        return 0;
    });
}

int32_t phy_fluid_batch_sync(PhyFluidBatch* b) {
    return guard<int32_t>(-1, [&] {
        need_batch(b)->device().sync();
        return 0;
    });
}

uintptr_t phy_fluid_batch_get_field_f32(PhyFluidBatch* b, int32_t field, float* buf, uintptr_t len) {
    return guard<uintptr_t>(0, [&] {
        int c = field_comps(field);
        if (!c) throw std::invalid_argument("unknown field");
        phy::DeviceFluid& d = need_batch(b)->device();
        size_t n = d.size();
        if (!buf || !n) return (uintptr_t)0;
        std::vector<float> tmp(n * c);
        d.download((phy::Field)field, tmp.data());
        uintptr_t np = std::min<uintptr_t>(n, len / c);
        std::memcpy(buf, tmp.data(), np * c * sizeof(float));
        return np;
    });
}

int32_t phy_fluid_batch_diagnostics_f64(PhyFluidBatch* b, double* out) {
    return guard<int32_t>(-1, [&] {
        if (!out) throw std::invalid_argument("null out");
        need_batch(b)->device().diagnostics(out);
        return 0;
    });
}

int32_t phy_fluid_batch_neighbor_stats(PhyFluidBatch* b, uint32_t* counts, uint64_t* sums,
                                       uint64_t* xors, uintptr_t len) {
    return guard<int32_t>(-1, [&] {
        phy::DeviceFluid& d = need_batch(b)->device();
        if (!counts || !sums || !xors || len < d.size()) throw std::invalid_argument("buffers too short");
        d.neighbor_stats(counts, sums, xors);
        return 0;
    });
}

int32_t phy_fluid_batch_gather_stats(PhyFluidBatch* b, int32_t pass, uint32_t* counts, uint64_t* sums,
                                     uint64_t* xors, uintptr_t len) {
    return guard<int32_t>(-1, [&] {
        phy::DeviceFluid& d = need_batch(b)->device();
        if (!counts || !sums || !xors || len < d.size()) throw std::invalid_argument("buffers too short");
        d.gather_stats(pass, counts, sums, xors);
        return 0;
    });
}

int32_t phy_fluid_batch_tape_begin(PhyFluidBatch* b, uint32_t checkpoint_every) {
    return guard<int32_t>(-1, [&] {
        need_batch(b)->device().tape_begin(checkpoint_every);
        return 0;
    });
}

uint32_t phy_fluid_batch_tape_steps(PhyFluidBatch* b) {
    return guard<uint32_t>(0, [&] { return (uint32_t)need_batch(b)->device().tape_steps(); });
}

int32_t phy_fluid_batch_tape_backward(PhyFluidBatch* b, const float* xbar_final, const float* vbar_final,
                                      const float* rhobar_last, float* xbar0, float* vbar0,
                                      double* dparams) {
    return guard<int32_t>(-1, [&] {
        phy::DeviceFluid& d = need_batch(b)->device();
        const size_t S = d.scenes();
        std::vector<double> g(S * phy::kAdjParams);
        d.tape_backward(xbar_final, vbar_final, rhobar_last, xbar0, vbar0, g.data());
        if (dparams)
            for (size_t s = 0; s < S; ++s) {
                dparams[2 * s] = g[s * phy::kAdjParams];
                dparams[2 * s + 1] = g[s * phy::kAdjParams + 1];
            }
        return 0;
    });
}

int32_t phy_fluid_batch_tape_backward_full(PhyFluidBatch* b, const float* xbar_final,
                                           const float* vbar_final, const float* rhobar_last,
                                           float* xbar0, float* vbar0, double* dparams) {
    return guard<int32_t>(-1, [&] {
        phy::DeviceFluid& d = need_batch(b)->device();
        std::vector<double> g(d.scenes() * phy::kAdjParams);
        d.tape_backward(xbar_final, vbar_final, rhobar_last, xbar0, vbar0, g.data());
        if (dparams) std::copy(g.begin(), g.end(), dparams);
        return 0;
    });
}

int32_t phy_fluid_batch_tape_end(PhyFluidBatch* b) {
    return guard<int32_t>(-1, [&] {
        need_batch(b)->device().tape_end();
        return 0;
    });
}

int32_t phy_fluid_batch_profile_begin(PhyFluidBatch* b) {
    return guard<int32_t>(-1, [&] {
        need_batch(b)->device().profile_begin();
        return 0;
    });
}

int32_t phy_fluid_batch_profile_end(PhyFluidBatch* b, double* ms4, uint32_t* steps) {
    return guard<int32_t>(-1, [&] {
        if (!ms4) throw std::invalid_argument("null out");
        uint32_t n = need_batch(b)->device().profile_end(ms4);
        if (steps) *steps = n;
        return 0;
    });
}

}  // extern "C"
