// fluid_device.h — device state and step sequencing of one FluidWorld on one GPU.
//
// Owns every HBM buffer of the world, one HIP stream (default priority), and
// runs FluidWorld::step (crates/phy-fluid/src/sph/world.rs:183-192) as
// domain -> counting sort -> density -> fused force/integrate/bounds.
// Particle data stays resident in cell order between steps; host transfers
// happen only on explicit upload / download calls.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <stdexcept>
#include <utility>
#include <string>
#include <vector>

#include "sph_couple.h"
#include "sph_kernels.h"
#include "sph_types.h"

namespace phy {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what);

// Bytes of device memory the library's buffers hold (every world and batch of
// the process) and the largest value since the last reset_peak.
void mem_note(long long delta_bytes);
unsigned long long mem_in_use();
unsigned long long mem_peak();
void mem_reset_peak();

// Device buffer of T elements (hipMalloc / hipFree).
template <class T>
class DevBuf {
   public:
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void ensure(size_t n) {
        if (n <= cap_) return;
        release();
        void* p = nullptr;
        hip_check(hipMalloc(&p, (n ? n : 1) * sizeof(T)), "hipMalloc");
        p_ = (T*)p;
        cap_ = n;
        mem_note((long long)((n ? n : 1) * sizeof(T)));
    }
    void release() {
        if (p_) {
            (void)hipFree(p_);
            mem_note(-(long long)((cap_ ? cap_ : 1) * sizeof(T)));
        }
        p_ = nullptr;
        cap_ = 0;
    }
    T* get() const { return p_; }
    size_t cap() const { return cap_; }
    void swap(DevBuf& o) {
        std::swap(p_, o.p_);
        std::swap(cap_, o.cap_);
    }

   private:
    T* p_ = nullptr;
    size_t cap_ = 0;
};

// Host-side f64 parameters, the SphParams<f64> of the reference
// (crates/phy-fluid/src/sph/params.rs:12-50), flattened.
struct HostParams {
    double rest_density, stiffness, viscosity;
    std::vector<double> visc_k, visc_n;
    double shear_min, mass, h;
    double gravity[3], bounds_min[3], bounds_max[3];
    double boundary_damp, xsph_eps;
};

// Derives the device constants: f64 neighbour threshold and FP32 coefficients
// (kernels.rs:22-35 computed in f64, rounded once).
DevParams make_dev_params(const HostParams& hp, int cdiv = 2);

// Particle fields the device can hand back in original index order.
enum Field : int {
    F_POS = 0,
    F_VEL = 1,
    F_ACC = 2,
    F_XSPH = 3,
    F_RHO = 4,
    F_P = 5,
    F_MASS = 6,
    F_MU_EFF = 7,
};

// One or more independent scenes on one GPU. A FluidWorld is a batch of one
// scene; a batch concatenates its scenes' particles (scene s owns the index
// range [offset(s), offset(s) + count(s))) and steps all of them with one
// launch per stage. Each scene keeps its own parameters, bounds, dense cell
// table and watchdog.
class DeviceFluid {
   public:
    explicit DeviceFluid(const HostParams& hp);
    explicit DeviceFluid(const std::vector<HostParams>& scenes);
    ~DeviceFluid();
    DeviceFluid(const DeviceFluid&) = delete;
    DeviceFluid& operator=(const DeviceFluid&) = delete;

    size_t scenes() const { return hps_.size(); }
    // set_params replaces the parameters of every scene
    void set_params(const HostParams& hp);
    void set_scene_params(size_t s, const HostParams& hp);
    const HostParams& params(size_t s = 0) const { return hps_.at(s); }
    size_t size() const { return n_; }
    uint32_t offset(size_t s) const { return scn_.at(s).off; }
    uint32_t count(size_t s) const { return scn_.at(s).n; }

    // Replace all particles of a one-scene world (original index order). Null
    // vel / material / body_acc mean zero / 0 / zero. Positions and velocities
    // are rounded to FP32.
    void upload(size_t n, const float* pos3, const float* vel3, const float* mass,
                const uint32_t* material, const float* body3);
    // Replace the particles of every scene: counts[scenes()], then the
    // scenes' particles concatenated in scene order (null vel / material as above).
    void upload_scenes(const uint32_t* counts, const float* pos3, const float* vel3,
                       const float* mass, const uint32_t* material);

    // Advance every scene one step of dt (FP32 on the device).
    void step(float dt);
    // While on, step() replays one captured hipGraph of the whole step
    // (recaptured when dt, the particle set, the parameters or the buffers
    // change). Profiling and steps that consume body_acc run eagerly.
    void set_graph(bool on);
    bool graph() const { return graph_on_; }

    // Copy a field to host memory in (concatenated) original index order (3
    // floats per particle for vector fields, 1 otherwise). Auxiliary fields
    // (acc, xsph, mu_eff) hold the last step's values only while keep_aux is on.
    void download(Field f, float* out);
    // material tags in (concatenated) original index order
    void download_material(uint32_t* out);

    void set_keep_aux(bool on);
    bool keep_aux() const { return keep_aux_; }
    bool aux_valid() const { return aux_valid_; }
    bool density_valid() const { return dens_valid_; }

    // Watchdog of the last step (subsystem.rs:74-90): 0 ok, 1 pos, 2 vel, for
    // scene 0 / for every scene.
    int last_validate();
    void last_validate_all(int32_t* status);
    // FP64 mass, momentum (3) and kinetic energy of the current state, 5 per scene.
    void diagnostics(double* out5);
    // Neighbour statistics of the density predicate on the current state
    // (original order): count, sum of neighbour indices, xor of hashed
    // indices; indices are scene-local.
    void neighbor_stats(uint32_t* cnt, uint64_t* sum, uint64_t* xr);
    // The same statistics from the step's own gathers (pass 0: density, j
    // with r < h, self included; pass 1: forces, j != i with 0 < r < h or a
    // NaN distance), run on the current state with the current tuning: the
    // neighbour sets the density and force sums actually run over.
    void gather_stats(int pass, uint32_t* cnt, uint64_t* sum, uint64_t* xr);

    hipStream_t stream() const { return st_; }
    void sync();

    // Adjoint tape. tape_begin(K) marks the current state as the start of a
    // recorded rollout; every later step is recorded and the state is
    // checkpointed every K steps. tape_backward() takes the cotangents of the
    // final x, v (3 per particle) and of the last step's rho (1 per particle),
    // any of them null for zero, all in original index order, and returns the
    // cotangents of the initial x, v and, per scene, kAdjParams parameter
    // gradients: [0] dL/d stiffness, [1 + m] dL/d visc_k[m], [1 +
    // kMaxMaterials + m] dL/d visc_n[m] (zero for m >= visc_k.len()). Each
    // segment is recomputed from its checkpoint into per-step slots, then the
    // adjoint runs backwards through it. The world state is left unchanged.
    // While recording, forces are computed by the shear-rate kernel variant
    // (the visc_n gradient needs the shear sum; for visc_n = 1 the values are
    // the same), so that recomputed steps match recorded ones bit for bit.
    void tape_begin(uint32_t K);
    void tape_backward(const float* xbar_T, const float* vbar_T, const float* rhobar_T,
                       float* xbar0, float* vbar0, double* dparams);
    void tape_end();
    size_t tape_steps() const { return tape_dt_.size(); }
    bool taping() const { return taping_; }

    // Launch choices of the gather kernels (sph_kernels.h Tuning).
    const Tuning& tuning() const { return tuning_; }
    void set_tuning(const Tuning& t);
    // Autotuner: times `steps` eager steps of the current state per
    // candidate with HIP events: first the cell size (h / 2 or h, whole
    // steps, default lists), then at that cell size the list lengths 8, 16,
    // 32 in one and in two launches (density and force timed separately; the
    // launch form with the smaller best sum and the fastest list length of
    // each pass in it), then the gathers' workgroup size (256 or 128 lanes).
    // Restores the state; rho / p / aux are undefined until the next step.
    // With a cache path, a stored choice for this device and workload class is
    // applied without timing (returns true), and a new choice is stored, so
    // repeated runs keep one summation order. ms[12] (optional): density
    // milliseconds per step of the six list trials (lists 8, 16, 32 combined,
    // then split), then force.
    bool autotune(float dt, uint32_t steps, const char* cache_path, double* ms);
    // "gfx950 n20 s0 shear0": device architecture, log2 particles, log2 scenes,
    // non-Newtonian flag (the cache key)
    std::string tuning_key() const;

    // Per-stage HIP-event timing on this world's stream: while enabled, every
    // step records events around sort (domain + counting sort), density and
    // force; profile_end() synchronises and returns the summed milliseconds
    // {sort, density, force, whole step} and the number of steps timed.
    void profile_begin();
    uint32_t profile_end(double ms4[4]);
    int device() const { return dev_; }

    // Pending body_acc (world.rs:356, 361), consumed by the next step. In
    // original index order, 3 per particle; all zero clears it.
    void set_body_acc(const float* b3);
    void get_body_acc(float* out3);
    bool has_body_acc() const { return has_body_; }

    // Coupling of a one-scene world (world.rs:400-773, sph_couple.h). They act
    // on the current state like the reference's calls after a step: particles
    // change in place (position, velocity, or body_acc), rho / p / aux keep
    // the last step's values, and a recording tape is invalidated.
    // bodies: velocities updated in place.
    void couple_bodies(CoupleBody* bodies, uint32_t nb, double dt);
    // points: forces written (untouched outside the bounds).
    void couple_points(CouplePoint* pts, uint32_t np, double dt, double drag, double soft_density);
    // body_acc += -g beta (T - t_ref); src[nx*ny*nz] += heat_gain |v| dt per
    // cell (host arrays; src may be null when heat_gain <= 0).
    void couple_heat(const HeatGrid& host_grid, double* src, double dt, double t_ref, double beta,
                     double heat_gain);
    // velocity of the particle nearest to each query (the sampler couple_heat
    // installs); zero without particles
    void sample_velocity(const double* q3, uint32_t nq, double* out3);

   private:
    struct StepBufs {
        float4 *posm_s, *velm_s, *body_s;
        float4* posl;  // consumed within the step: one buffer serves every slot
        uint32_t *id_s, *mat_s, *cell_s;
        float *rho, *prs, *mp;
        uint32_t* start;
        Domain* dom;
        float4 *acc, *xsph, *fvisc;
        float* mu;
    };
    struct Snapshot {
        DevBuf<float4> posm, vel;
        DevBuf<uint32_t> id, mat;
    };
    struct TapeSlot {
        DevBuf<float4> posm_s, velm_s, body_s, acc, xsph, fvisc;
        DevBuf<uint32_t> id_s, mat_s, cell_s, start;
        DevBuf<float> rho, prs, mp, mu;
        DevBuf<Domain> dom;
    };
    void init_stream();
    // a Domain array of S scenes, reset when (re)allocated: the domain pass
    // expects reset AABB words and leaves them reset
    void ensure_dom(DevBuf<Domain>& b, size_t S);
    void ensure_capacity(size_t n);
    void set_layout(const uint32_t* counts);
    void size_tables(const float* pos3);
    void push_params();
    void apply_cell_div(int old_div);
    SceneTable table(Domain* dom) const;
    StepBufs own_bufs();
    StepBufs slot_bufs(TapeSlot& t);
    // Tape slots keep only the used part of each step's cell table when the
    // table's capacity is large (huge bounds boxes), at one host sync per
    // recomputed step; small tables are kept whole, with no sync.
    bool compact_slots() const { return table_cap_ > (1ull << 22); }
    void run_domain_and_sort(const StepBufs& s);
    void step_core(float dt, const StepBufs& s);
    void capture(float dt);
    void drop_graph();
    void save_state(Snapshot& c);
    void load_state(const Snapshot& c);
    void tape_record(float dt);
    // A recording becomes unusable when the state, the parameters, the launch
    // tuning or the forces change outside a recorded step; tape_backward then
    // fails with the reason.
    void tape_invalidate(const char* why);
    CoupleParams couple_params() const;
    void one_scene(const char* what) const;
    // bins the current positions into the step's sorted buffers with a domain
    // of its own (the last step's watchdog stays intact) and builds inv
    void sort_current();

    std::vector<HostParams> hps_;
    std::vector<DevParams> Ps_;
    std::vector<float> m_uni_;  // per scene: the common particle mass, NaN when masses differ
    bool all_um_ = false;       // every scene has a common mass: the force pass reads frec_
    std::vector<SceneInfo> scn_;
    std::vector<uint32_t> tile_scene_;
    int ntiles_ = 0;
    bool any_shear_ = false;
    int dev_ = 0;
    hipStream_t st_ = nullptr;
    size_t n_ = 0;
    size_t cap_n_ = 0;
    unsigned long long table_cap_ = 0;  // cells of the concatenated table (incl. overflow cells)
    bool keep_aux_ = false, aux_valid_ = false, dens_valid_ = false;
    bool has_body_ = false;

    DevBuf<DevParams> d_params_;
    DevBuf<SceneInfo> d_info_;
    DevBuf<uint32_t> d_tiles_, table_total_;
    // current state (cell order after the first step)
    DevBuf<float4> posm_, vel_;
    DevBuf<uint32_t> id_, mat_;
    DevBuf<float4> body_;
    // sorted copies for the step
    DevBuf<float4> posm_s_, velm_s_, body_s_;
    DevBuf<float4> posl_;
    DevBuf<float4> frec_;  // (x, y, z, m p / rho^2) per sorted slot, consumed within a step
    DevBuf<uint32_t> id_s_, mat_s_, cell_s_;
    DevBuf<float> rho_, prs_, mp_;
    // sort scratch
    DevBuf<uint32_t> cell_, rank_, counts_, partial_, start_, tmp_, ctmp_;
    DevBuf<unsigned long long> keytmp_;
    // aux outputs
    DevBuf<float4> acc_, xsph_;
    DevBuf<float> mu_;
    // staging for readback
    DevBuf<float4> stage4_;
    DevBuf<float> stage1_;
    DevBuf<Domain> dom_;
    DevBuf<double> diag_part_, diag_out_;

    Tuning tuning_;
    bool graph_on_ = false, graph_dirty_ = true;
    float graph_dt_ = 0.f;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;

    bool prof_ = false;
    std::vector<hipEvent_t> prof_ev_;  // 4 per profiled step
    hipEvent_t ev();

    bool taping_ = false, tape_broken_ = false;
    std::string tape_broken_why_;
    uint32_t tape_K_ = 0;
    size_t tape_n_ = 0;
    std::vector<float> tape_dt_;
    std::vector<std::unique_ptr<Snapshot>> ckpt_;  // state before steps 0, K, 2K, ...
    std::vector<std::unique_ptr<TapeSlot>> slots_;
    DevBuf<float4> xbar_[2], vbar_[2], Fb_, Sb_, xb_loc_, vb_loc_, xb_acc_;
    DevBuf<float> rhob_loc_, rhob_, seed_;
    DevBuf<float2> mush_;
    DevBuf<double> adj_part_, adj_tot_;

    // coupling scratch
    DevBuf<Domain> cdom_;
    DevBuf<uint32_t> inv_, ccnt_, coff_, cn_, cpartial_;
    DevBuf<Contact> contacts_;
    DevBuf<CoupleBody> cbodies_;
    DevBuf<CouplePoint> cpoints_;
    DevBuf<double> cdisp_, vel64_, heat_u_, heat_src_, heat_q_, cq_;
    DevBuf<uint32_t> heat_qcell_, heat_flags_;
    DevBuf<unsigned long long> heat_lo_, heat_hi_, heat_max_;
    DevBuf<int> heat_scale_;
};

}  // namespace phy
