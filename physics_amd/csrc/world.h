// world.h — C++ host mirror of the reference's fluid-path interfaces:
//   SphParams<T>       crates/phy-fluid/src/sph/params.rs:12-85
//   Particle<T>        crates/phy-fluid/src/particle.rs:17-70
//   FluidWorld<T>      crates/phy-fluid/src/sph/world.rs:21-192
//   Subsystem<T>       crates/phy-core/src/world.rs:22-75
//   FluidSubsystem<T>  crates/phy-fluid/src/subsystem.rs:26-164
//   World<T>           crates/phy-core/src/world.rs:78-318
//   WorldError         crates/phy-core/src/lib.rs:38-55
// The host keeps f64 parameters and an f64 time; particle state lives on the
// GPU in FP32 (DeviceFluid) and is read back on request.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "fluid_device.h"

namespace phy {

struct Vec3 {
    double x = 0, y = 0, z = 0;
};

struct SphParams {
    double rest_density, stiffness, viscosity;
    std::vector<double> visc_k, visc_n;
    double shear_min, mass, h;
    Vec3 gravity, bounds_min, bounds_max;
    double boundary_damp, xsph_eps;
    static SphParams defaults();
    HostParams to_host() const;
};

struct Particle {
    Vec3 pos, vel, acc, body_acc;
    double rho = 0, p = 0, mass = 0;
    size_t material = 0;
    double mu_eff = 0;
    Vec3 xsph;
    static Particle make(Vec3 pos, double mass);
    static Particle with_material(Vec3 pos, double mass, size_t material);
};

// WorldError::NonFinite { subsystem, field } / Stalled { subsystem }
struct WorldError {
    enum Kind { NonFinite, Stalled } kind;
    const char* subsystem;
    const char* field;
};

class World;

class Subsystem {
   public:
    virtual ~Subsystem() = default;
    virtual void step(double dt) = 0;
    virtual void couple(World& world, double dt) {
        (void)world;
        (void)dt;
    }
    virtual const char* name() const { return "subsystem"; }
    virtual std::optional<WorldError> validate() { return std::nullopt; }
    virtual Vec3 total_momentum() { return Vec3{}; }
    virtual double kinetic_energy() { return 0.0; }
};

class FluidWorld {
   public:
    explicit FluidWorld(const SphParams& params);

    const SphParams& params() const { return params_; }
    void set_params(const SphParams& params);
    size_t len() const;
    bool is_empty() const { return len() == 0; }

    // world.rs:124-169 (host f64 lattice, then FP32 upload on the next step)
    void fill_box(Vec3 min, Vec3 max, double spacing, double margin);
    double lattice_mass(double s) const;
    void add_particle(const Particle& p);
    void set_particles(const std::vector<Particle>& ps);
    // Full particle records in original index order (downloads when the GPU
    // holds the newest state). acc/mu_eff/xsph are filled when keep_aux is on.
    std::vector<Particle> particles();
    // world.rs:178-180 (needs keep_aux before the step)
    double effective_viscosity(size_t i);

    void step(double dt);

    // Field readback in original order: f64 channel = exact (double) of FP32.
    void read_vec(Field f, double* out3);
    void read_vec_f32(Field f, float* out3);
    void read_scalar(Field f, double* out);
    void read_scalar_f32(Field f, float* out);
    // f64 / f32 particle upload (positions and velocities rounded to FP32)
    void upload(size_t n, const double* pos3, const double* vel3, const double* mass,
                const uint32_t* material);
    void upload_f32(size_t n, const float* pos3, const float* vel3, const float* mass,
                    const uint32_t* material);
    // body_acc accumulator (world.rs:356, 361): added once by the next step
    void set_body_acc(const double* b3);
    void get_body_acc(double* out3);

    // Coupling (world.rs:400-773) on the current state, as the reference's
    // FluidWorld::couple_bodies / couple_points / couple_heat (and the
    // velocity sampler couple_heat installs); see DeviceFluid::couple_*.
    // friction is accepted and unused, as in the reference (push_out ignores it).
    void couple_bodies(CoupleBody* bodies, size_t n, double dt, double friction);
    void couple_points(CouplePoint* pts, size_t n, double dt, double drag, double soft_density);
    void couple_heat(const HeatGrid& grid, double* src, double dt, double t_ref, double beta,
                     double heat_gain);
    void sample_velocity(const double* q3, size_t n, double* out3);

    void set_keep_aux(bool on);
    // adjoint tape (see DeviceFluid::tape_*); cotangents in original index
    // order, dparams: kAdjParams parameter gradients
    void tape_begin(uint32_t K);
    void tape_backward(const double* xbar_T, const double* vbar_T, const double* rhobar_T,
                       double* xbar0, double* vbar0, double* dparams);
    void tape_end();
    size_t tape_steps() const { return dev_.tape_steps(); }
    // true while the device state is exactly what the last step wrote, so the
    // watchdog fused into that step describes it
    bool watchdog_current() const { return last_was_step_; }
    int last_validate();
    void diagnostics(double out5[5]);
    void neighbor_stats(uint32_t* cnt, uint64_t* sum, uint64_t* xr);
    void gather_stats(int pass, uint32_t* cnt, uint64_t* sum, uint64_t* xr);
    DeviceFluid& device() { return dev_; }

    // FluidWorld::to_gpu_flat over Grid::to_flat (world.rs:782-859,
    // grid.rs:62-175), with the grid built from the current positions (as
    // after build_grid). Fills the shape always; fills the arrays (pos4, vel4,
    // scalar4 = (rho, p, mass, material): 4 floats per particle; sorted: n;
    // cell_start: ncells + 1) and returns true when the capacities suffice.
    struct FlatShape {
        uint64_t n, ncells;
        int64_t grid_min[3];
        uint64_t nc[3];
    };
    bool flat_export(FlatShape& shape, float* pos4, float* vel4, float* scalar4, uint64_t n_cap,
                     int32_t* cell_start, uint64_t cell_cap, int32_t* sorted);
    // push pending host-side particle edits to the device
    void flush() { flush_host(); }
    // the device state was changed outside step() (autotuner trials): the last
    // step's watchdog no longer describes it
    void mark_state_changed() { last_was_step_ = false; }

   private:
    void flush_host();  // push pending host particles to the device

    SphParams params_;
    DeviceFluid dev_;
    bool host_dirty_ = false;
    bool last_was_step_ = false;
    std::vector<Particle> host_;  // pending host-side edits (valid when host_dirty_)
};

class FluidSubsystem : public Subsystem {
   public:
    explicit FluidSubsystem(std::unique_ptr<FluidWorld> w);
    FluidWorld& world() { return *world_; }
    void step(double dt) override;
    // subsystem.rs:92-143: couples to HeatField / RigidSubsystem when present;
    // a fluid-only World has neither, so this is a no-op here.
    void couple(World& world, double dt) override;
    const char* name() const override { return "fluid"; }
    std::optional<WorldError> validate() override;
    Vec3 total_momentum() override;
    double kinetic_energy() override;

    double thermal_expansion = 0.0, heat_gain = 0.0, t_ref = 0.0, drag = 1.0, friction = 0.1;

   private:
    std::unique_ptr<FluidWorld> world_;
};

// Many independent FluidWorlds of uneven size stepped together on one GPU (no
// counterpart in the reference). Scene s behaves as FluidWorld(params[s])
// holding particles [offset(s), offset(s) + count(s)) of the concatenated
// arrays; all scenes share dt and advance with one launch per stage, the
// step replayed from a captured hipGraph.
class FluidBatch {
   public:
    explicit FluidBatch(const std::vector<SphParams>& params);
    size_t scenes() const { return params_.size(); }
    const SphParams& params(size_t s) const { return params_.at(s); }
    void set_params(size_t s, const SphParams& p);
    // counts[scenes()]; FP32 arrays concatenated in scene order (vel, mass,
    // material may be null: zero, SphParams::mass, 0)
    void set_particles_f32(const uint32_t* counts, const float* pos3, const float* vel3,
                           const float* mass, const uint32_t* material);
    void step(double dt) { dev_.step((float)dt); }
    // step, then FluidSubsystem::validate of every scene (subsystem.rs:74-90):
    // status[s] = 0 ok, 1 non-finite position, 2 non-finite velocity
    void step_checked(double dt, int32_t* status);
    DeviceFluid& device() { return dev_; }

   private:
    std::vector<SphParams> params_;
    DeviceFluid dev_;
};

class World {
   public:
    World() = default;
    void add_subsystem(std::unique_ptr<Subsystem> s) { subs_.push_back(std::move(s)); }
    size_t subsystem_count() const { return subs_.size(); }
    Subsystem* get(size_t i) { return i < subs_.size() ? subs_[i].get() : nullptr; }
    double time() const { return t_; }
    void set_time(double t) { t_ = t; }
    uint64_t seed() const { return seed_; }

    // world.rs:233-254: every step, then every couple (with the subsystem taken
    // out of the list), then t += dt.
    void step_skipping_seeded(double dt, const std::vector<bool>& skip, uint64_t seed);
    void step(double dt) { step_skipping_seeded(dt, {}, 0); }
    void step_seeded(double dt, uint64_t seed) { step_skipping_seeded(dt, {}, seed); }
    void step_skipping(double dt, const std::vector<bool>& skip) { step_skipping_seeded(dt, skip, 0); }
    // world.rs:282-293
    std::optional<WorldError> step_skipping_checked(double dt, const std::vector<bool>& skip);
    std::optional<WorldError> step_checked(double dt) { return step_skipping_checked(dt, {}); }

    Vec3 total_momentum();
    double kinetic_energy();

    // couple-phase helpers (world.rs:201-209)
    std::unique_ptr<Subsystem> remove(size_t i);
    void insert(size_t i, std::unique_ptr<Subsystem> s);

   private:
    std::vector<std::unique_ptr<Subsystem>> subs_;
    double t_ = 0.0;
    bool started_ = false;
    uint64_t seed_ = 0;
};

}  // namespace phy
