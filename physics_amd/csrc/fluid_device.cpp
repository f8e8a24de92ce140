// fluid_device.cpp — DeviceFluid: buffers, uploads and the step sequence.
#include "fluid_device.h"

#include <atomic>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <cfloat>
#include <cstddef>
#include <climits>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace phy {

namespace {
std::atomic<unsigned long long> g_mem_in_use{0}, g_mem_peak{0};
}

void mem_note(long long delta) {
    const unsigned long long now = g_mem_in_use.fetch_add((unsigned long long)delta) + (unsigned long long)delta;
    unsigned long long pk = g_mem_peak.load();
    while (delta > 0 && now > pk && !g_mem_peak.compare_exchange_weak(pk, now)) {
    }
}
unsigned long long mem_in_use() { return g_mem_in_use.load(); }
unsigned long long mem_peak() { return g_mem_peak.load(); }
void mem_reset_peak() { g_mem_peak.store(g_mem_in_use.load()); }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

// Cell key of the dense table (sph_device.h hkey: floor(cdiv fl(x (1/h))),
// the reference key for cdiv = 1, grid.rs:26-33) on the host; 0 where the key
// fails (such positions are binned to the overflow segment).
static long long host_key(float x, double inv, int cdiv) {
    const double v = (double)x * inv;
    const double f = std::floor(v * (double)cdiv);
    if (f >= -9223372036854775808.0 && f < 9223372036854775808.0 && v >= -9223372036854775808.0 &&
        v < 9223372036854775808.0)
        return (long long)f;
    return 0;
}

// The gathers address neighbour records with 32-bit byte offsets (16 bytes
// per particle) from a wave-uniform base: at most 2^27 particles per world or
// batch.
static constexpr uint64_t kMaxParticles = 1ull << 27;
static const char* const kTooMany = "phy-fluid: at most 2^27 particles per world or batch";

DevParams make_dev_params(const HostParams& hp, int cdiv) {
    if (cdiv != 1 && cdiv != 2) throw std::invalid_argument("phy-fluid: cells of h or h / 2 only");
    DevParams P;
    memset(&P, 0, sizeof(P));
    const double h = hp.h;
    P.h64 = h;
    P.inv_cell = 1.0 / h;
    P.cdiv = cdiv;
    P.cell64 = h / (double)cdiv;
    // smallest f64 s with sqrt(s) >= h: the reference's r < h is s < s_thresh
    double s = h * h;
    if (std::isfinite(s) && h > 0.0) {
        for (int it = 0; it < 64 && std::sqrt(std::nextafter(s, 0.0)) >= h; ++it)
            s = std::nextafter(s, 0.0);
        for (int it = 0; it < 64 && std::sqrt(s) < h; ++it) s = std::nextafter(s, INFINITY);
    }
    P.s_thresh = s;
    const double marg = std::ldexp(1.0, -19);
    P.s_lo = (float)(s * (1.0 - marg));
    P.s_hi = (float)(s * (1.0 + marg));
    if ((double)P.s_lo >= s * (1.0 - marg)) P.s_lo = std::nextafter(P.s_lo, 0.0f);
    if ((double)P.s_hi <= s * (1.0 + marg)) P.s_hi = std::nextafter(P.s_hi, INFINITY);
    // Column-local filter of the uniform scan: its FP32 terms stay finite and
    // its rounding below 1e-5 h^2 for h in [1e-15, 1e15]; elsewhere NaN, which
    // lets every candidate through.
    P.s_hq = (float)(s * (1.0 + std::ldexp(1.0, -16)));
    if ((double)P.s_hq <= s * (1.0 + std::ldexp(1.0, -16))) P.s_hq = std::nextafter(P.s_hq, INFINITY);
    if (!(h >= 1e-15 && h <= 1e15)) P.s_hq = NAN;
    // Band of the FP32 q = h^2 - s of the column density (an FMA chain: error
    // below 10u h^2 for s <= 2 h^2, u = 2^-24, and |q| > h^2 / 2 beyond):
    // outside +-q_band its sign is the reference's r < h. Same range as the
    // column-local filter; elsewhere infinite, so the exact predicate decides
    // every pair.
    P.q_band = (float)(s * std::ldexp(1.0, -19));
    if ((double)P.q_band <= s * std::ldexp(1.0, -19)) P.q_band = std::nextafter(P.q_band, INFINITY);
    if (!(h >= 1e-15 && h <= 1e15)) P.q_band = INFINITY;
    if (!(h > 0.0)) {  // degenerate support radius: only the exact path decides
        P.s_lo = -INFINITY;
        P.s_hi = INFINITY;
    }
    // z half-widths of the trimmed scan rows. A neighbour has real
    // dx^2 + dy^2 + dz^2 < h^2 (1 + 2^-50) (s64 < s_thresh plus the rounding of
    // s64). Two half-cell columns 2 apart hold points at least h/2 apart, less
    // the key rounding (|x| 2^-52 per point; the scan trims only domains with
    // keys below 2^40 in magnitude, so less than h 2^-11): (h/2)(1 - 2^-8) is a
    // safe gap. Cells of h: the outer rows are adjacent columns, with no gap,
    // so every row takes the full width. Degenerate h: no trimming (infinite
    // widths).
    {
        const double gap2 = 0.25 * (1.0 - std::ldexp(1.0, -8)) * (1.0 - std::ldexp(1.0, -8));
        const double top = 1.0 + std::ldexp(1.0, -40), up = 1.0 + std::ldexp(1.0, -30);
        const bool ok = std::isnormal(h) && h > 0.0 && std::isfinite(h * h * 4.0);
        P.zw[0] = ok ? h * std::sqrt(top) * up : INFINITY;
        P.zw[1] = ok ? h * std::sqrt(top - gap2) * up : INFINITY;
        P.zw[2] = ok ? h * std::sqrt(top - 2.0 * gap2) * up : INFINITY;
        if (cdiv == 1) P.zw[1] = P.zw[2] = P.zw[0];
    }
    // kernels.rs:22-35 in f64
    const double pi = 3.141592653589793;
    double h2 = h * h, h3 = h2 * h, h6 = h3 * h3, h9 = h6 * h3;
    P.h2_64 = h2;
    P.c6_64 = 315.0 / (64.0 * pi * h9);
    P.rest64 = hp.rest_density;
    P.stiff64 = hp.stiffness;
    P.h = (float)h;
    P.h2 = (float)h2;
    P.c6 = (float)(315.0 / (64.0 * pi * h9));
    P.cs = (float)(-45.0 / (pi * h6));
    P.cv = (float)(45.0 / (pi * h6));
    P.rest = (float)hp.rest_density;
    P.stiff = (float)hp.stiffness;
    P.shear_min = (float)hp.shear_min;
    P.r_eps = (float)1e-4;
    for (int a = 0; a < 3; ++a) {
        P.g[a] = (float)hp.gravity[a];
        P.lo[a] = (float)hp.bounds_min[a];
        P.hi[a] = (float)hp.bounds_max[a];
    }
    P.damp = (float)hp.boundary_damp;
    P.xsph_eps = (float)hp.xsph_eps;
    if (hp.visc_k.empty() || hp.visc_n.empty())
        throw std::invalid_argument("SphParams: visc_k and visc_n must be non-empty");
    if (hp.visc_k.size() > (size_t)kMaxMaterials || hp.visc_n.size() > (size_t)kMaxMaterials)
        throw std::invalid_argument("SphParams: at most 16 materials");
    // world.rs:319-325: m < visc_k.len() ? m : 0, then visc_n[m] (panics in the
    // reference when visc_n is shorter; rejected here up front)
    if (hp.visc_n.size() < hp.visc_k.size())
        throw std::invalid_argument("SphParams: visc_n shorter than visc_k");
    P.n_mat = (int32_t)hp.visc_k.size();
    P.newtonian = 1;
    for (size_t m = 0; m < hp.visc_k.size(); ++m) {
        P.visc_k[m] = (float)hp.visc_k[m];
        P.visc_n[m] = (float)hp.visc_n[m];
        if (hp.visc_n[m] != 1.0) P.newtonian = 0;
    }
    return P;
}

// Largest concatenated cell table of one world or batch (cells). When the
// scenes' capacities add up to more, they are scaled down: particles beyond a
// scene's dense domain go to its overflow segment (exact, but slower).
constexpr unsigned long long kMaxTableCells = 1ull << 28;

static unsigned long long box_cells(const long long* lo, const long long* hi) {
    double v = 1.0;
    for (int a = 0; a < 3; ++a) v *= (double)((unsigned long long)hi[a] - (unsigned long long)lo[a]) + 1.0;
    return v >= (double)kMaxDenseCells ? (unsigned long long)kMaxDenseCells : (unsigned long long)v;
}

DeviceFluid::DeviceFluid(const HostParams& hp) : DeviceFluid(std::vector<HostParams>{hp}) {}

DeviceFluid::DeviceFluid(const std::vector<HostParams>& scenes) {
    if (scenes.empty()) throw std::invalid_argument("phy-fluid: a batch needs at least one scene");
    if (scenes.size() > (1u << 24)) throw std::invalid_argument("phy-fluid: at most 2^24 scenes");
    std::vector<DevParams> ps;
    for (const HostParams& hp : scenes) ps.push_back(make_dev_params(hp, tuning_.cell_div));
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        throw HipError("phy-fluid: no HIP device available (this library runs on gfx950 only)");
    hip_check(hipGetDevice(&dev_), "hipGetDevice");
    hip_check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "hipStreamCreate");
    hps_ = scenes;
    Ps_ = ps;
    m_uni_.assign(scenes.size(), NAN);
    const size_t S = scenes.size();
    scn_.assign(S, SceneInfo{});
    for (size_t s = 0; s < S; ++s) {
        for (int a = 0; a < 3; ++a) {
            scn_[s].bkey[a] = host_key(Ps_[s].lo[a], Ps_[s].inv_cell, Ps_[s].cdiv);
            scn_[s].bkey[3 + a] = host_key(Ps_[s].hi[a], Ps_[s].inv_cell, Ps_[s].cdiv);
        }
    }
    ensure_dom(dom_, S);
    table_total_.ensure(1);
    partial_.ensure(kScanBlocks);
    diag_out_.ensure(5 * S);
    std::vector<uint32_t> zero(S, 0u);
    set_layout(zero.data());
    size_tables(nullptr);
    push_params();
}

DeviceFluid::~DeviceFluid() {
    drop_graph();
    for (hipEvent_t e : prof_ev_) (void)hipEventDestroy(e);
    if (st_) {
        (void)hipStreamSynchronize(st_);
        (void)hipStreamDestroy(st_);
    }
}

void DeviceFluid::sync() { hip_check(hipStreamSynchronize(st_), "hipStreamSynchronize"); }

// Scene ranges and the tile -> scene map for the given particle counts.
void DeviceFluid::set_layout(const uint32_t* counts) {
    uint64_t off = 0;
    uint32_t t = 0;
    tile_scene_.clear();
    for (size_t s = 0; s < scn_.size(); ++s) {
        const uint32_t n = counts[s];
        const uint32_t nt = (uint32_t)((n + (uint64_t)kTile - 1) / kTile);
        scn_[s].off = (uint32_t)off;
        scn_[s].n = n;
        scn_[s].tile0 = t;
        scn_[s].ntiles = nt;
        tile_scene_.insert(tile_scene_.end(), nt, (uint32_t)s);
        off += n;
        t += nt;
        if (off > kMaxParticles) throw std::invalid_argument(kTooMany);
    }
    n_ = (size_t)off;
    ntiles_ = (int)t;
    d_tiles_.ensure(tile_scene_.size() ? tile_scene_.size() : 1);
    if (!tile_scene_.empty())
        hip_check(hipMemcpyAsync(d_tiles_.get(), tile_scene_.data(), tile_scene_.size() * sizeof(uint32_t),
                                 hipMemcpyHostToDevice, st_), "upload tiles");
    sync();
    graph_dirty_ = true;
}

// Dense table capacity of every scene: the cells of its bounds key box (every
// finite particle is inside it after one step) and of its particles' key box
// (pos3: concatenated positions, or null to keep the current capacities as a
// floor). Grows the concatenated table when needed.
void DeviceFluid::size_tables(const float* pos3) {
    const size_t S = scn_.size();
    std::vector<unsigned long long> caps(S);
    double total = 0.0;
    for (size_t s = 0; s < S; ++s) {
        SceneInfo& si = scn_[s];
        unsigned long long need = box_cells(si.bkey, si.bkey + 3);
        if (pos3) {
            long long plo[3] = {LLONG_MAX, LLONG_MAX, LLONG_MAX}, phi[3] = {LLONG_MIN, LLONG_MIN, LLONG_MIN};
            bool any = false;
            for (uint32_t k = 0; k < si.n; ++k) {
                const float* p = pos3 + 3 * ((size_t)si.off + k);
                if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) continue;
                any = true;
                for (int a = 0; a < 3; ++a) {
                    long long key = host_key(p[a], Ps_[s].inv_cell, Ps_[s].cdiv);
                    plo[a] = std::min(plo[a], key);
                    phi[a] = std::max(phi[a], key);
                }
            }
            if (any) need = std::max(need, box_cells(plo, phi));
        } else {
            need = std::max<unsigned long long>(need, si.cap);
        }
        caps[s] = std::min<unsigned long long>(std::max<unsigned long long>(need, 1), kMaxDenseCells);
        total += (double)caps[s] + 1.0;
    }
    if (total > (double)kMaxTableCells) {
        const double f = ((double)kMaxTableCells - (double)S) / (total - (double)S);
        for (size_t s = 0; s < S; ++s)
            caps[s] = std::max<unsigned long long>(1, (unsigned long long)((double)caps[s] * f));
    }
    unsigned long long need_table = 0;
    for (size_t s = 0; s < S; ++s) {
        scn_[s].cap = caps[s];
        need_table += caps[s] + 1;
    }
    if (need_table > table_cap_ || !counts_.get()) {
        table_cap_ = std::max(need_table, table_cap_);
        counts_.release();
        start_.release();
        counts_.ensure(table_cap_);
        start_.ensure(table_cap_ + 1);
        hip_check(hipMemsetAsync(counts_.get(), 0, table_cap_ * sizeof(uint32_t), st_), "memset counts");
        sync();
        graph_dirty_ = true;
    }
}

void DeviceFluid::ensure_dom(DevBuf<Domain>& b, size_t S) {
    if (S <= b.cap() && b.get()) return;
    b.ensure(S);
    launch_domain_reset(st_, b.get(), (int)S);
}

void DeviceFluid::push_params() {
    const size_t S = scn_.size();
    any_shear_ = false;
    for (const DevParams& P : Ps_) any_shear_ = any_shear_ || !P.newtonian;
    all_um_ = S > 0;
    for (size_t s = 0; s < S; ++s) {
        Ps_[s].m_uni = m_uni_[s];
        all_um_ = all_um_ && m_uni_[s] == m_uni_[s];
    }
    d_params_.ensure(S);
    d_info_.ensure(S);
    hip_check(hipMemcpyAsync(d_params_.get(), Ps_.data(), S * sizeof(DevParams), hipMemcpyHostToDevice, st_), "params");
    hip_check(hipMemcpyAsync(d_info_.get(), scn_.data(), S * sizeof(SceneInfo), hipMemcpyHostToDevice, st_), "scenes");
    sync();
    graph_dirty_ = true;
}

void DeviceFluid::set_scene_params(size_t s, const HostParams& hp) {
    if (s >= scn_.size()) throw std::out_of_range("phy-fluid: scene index");
    DevParams P = make_dev_params(hp, tuning_.cell_div);
    Ps_[s] = P;
    hps_[s] = hp;
    for (int a = 0; a < 3; ++a) {
        scn_[s].bkey[a] = host_key(P.lo[a], P.inv_cell, P.cdiv);
        scn_[s].bkey[3 + a] = host_key(P.hi[a], P.inv_cell, P.cdiv);
    }
    size_tables(nullptr);
    push_params();
    dens_valid_ = aux_valid_ = false;
    tape_invalidate("the parameters changed");
}

void DeviceFluid::set_params(const HostParams& hp) {
    make_dev_params(hp, tuning_.cell_div);  // validate before changing anything
    for (size_t s = 0; s < scn_.size(); ++s) set_scene_params(s, hp);
}

void DeviceFluid::ensure_capacity(size_t n) {
    if (n > cap_n_) {
        size_t c = std::max<size_t>(n, 1);
        posm_.ensure(c); vel_.ensure(c); id_.ensure(c); mat_.ensure(c); body_.ensure(c);
        // the sorted and column-local positions keep kScanPad spare entries:
        // the uniform scan loads candidates eight at a time
        posm_s_.ensure(c + kScanPad); velm_s_.ensure(c); body_s_.ensure(c);
        posl_.ensure(c + kScanPad);
        frec_.ensure(c + kScanPad);  // read by the force pass's scans too
        id_s_.ensure(c); mat_s_.ensure(c); cell_s_.ensure(c);
        rho_.ensure(c); prs_.ensure(c); mp_.ensure(c);
        cell_.ensure(c); rank_.ensure(c); tmp_.ensure(c); ctmp_.ensure(c); keytmp_.ensure(c);
        stage4_.ensure(c); stage1_.ensure(c);
        if (keep_aux_) { acc_.ensure(c); xsph_.ensure(c); mu_.ensure(c); }
        cap_n_ = c;
        graph_dirty_ = true;
    }
    diag_part_.ensure((size_t)std::max(ntiles_, 1) * 5);
}

void DeviceFluid::upload(size_t n, const float* pos3, const float* vel3, const float* mass,
                         const uint32_t* material, const float* body3) {
    if (scn_.size() != 1) throw std::logic_error("phy-fluid: upload() needs a one-scene world");
    if (n > kMaxParticles) throw std::invalid_argument(kTooMany);
    const uint32_t c = (uint32_t)n;
    upload_scenes(&c, pos3, vel3, mass, material);
    has_body_ = false;
    if (body3 && n) {
        std::vector<float4> b4(n);
        for (size_t i = 0; i < n; ++i) {
            b4[i] = make_float4(body3[3 * i], body3[3 * i + 1], body3[3 * i + 2], 0.f);
            if (b4[i].x != 0.f || b4[i].y != 0.f || b4[i].z != 0.f) has_body_ = true;
        }
        if (has_body_) {
            hip_check(hipMemcpyAsync(body_.get(), b4.data(), n * sizeof(float4), hipMemcpyHostToDevice, st_), "upload body");
            sync();
        }
    }
}

void DeviceFluid::upload_scenes(const uint32_t* counts, const float* pos3, const float* vel3,
                                const float* mass, const uint32_t* material) {
    uint64_t total = 0;
    for (size_t s = 0; s < scn_.size(); ++s) total += counts[s];
    if (total > kMaxParticles) throw std::invalid_argument(kTooMany);
    set_layout(counts);
    ensure_capacity(n_);
    size_tables(pos3);
    // per scene the common mass (bit-equal masses), which the column density
    // factors out of its terms; NaN otherwise
    for (size_t s = 0; s < scn_.size(); ++s) {
        const float m0 = (float)hps_[s].mass;
        float mu = NAN;
        for (uint32_t k = 0; k < scn_[s].n; ++k) {
            const float m = mass ? mass[(size_t)scn_[s].off + k] : m0;
            if (k == 0) mu = m;
            else if (memcmp(&m, &mu, sizeof(float)) != 0) {
                mu = NAN;
                break;
            }
        }
        m_uni_[s] = mu;
    }
    push_params();
    const size_t n = n_;
    std::vector<float4> p4(n), v4(n);
    std::vector<uint32_t> ids(n), mats(n);
    for (size_t s = 0; s < scn_.size(); ++s) {
        const float m0 = (float)hps_[s].mass;
        for (uint32_t k = 0; k < scn_[s].n; ++k) {
            const size_t i = (size_t)scn_[s].off + k;
            p4[i] = make_float4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], mass ? mass[i] : m0);
            v4[i] = vel3 ? make_float4(vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2], 0.f)
                         : make_float4(0.f, 0.f, 0.f, 0.f);
            ids[i] = (uint32_t)i;
            mats[i] = material ? material[i] : 0u;
        }
    }
    if (n) {
        hip_check(hipMemcpyAsync(posm_.get(), p4.data(), n * sizeof(float4), hipMemcpyHostToDevice, st_), "upload pos");
        hip_check(hipMemcpyAsync(vel_.get(), v4.data(), n * sizeof(float4), hipMemcpyHostToDevice, st_), "upload vel");
        hip_check(hipMemcpyAsync(id_.get(), ids.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st_), "upload id");
        hip_check(hipMemcpyAsync(mat_.get(), mats.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st_), "upload mat");
    }
    has_body_ = false;
    sync();  // host vectors go out of scope
    dens_valid_ = aux_valid_ = false;
    tape_invalidate("the particles were replaced");
    graph_dirty_ = true;
}

void DeviceFluid::set_keep_aux(bool on) {
    keep_aux_ = on;
    if (on && cap_n_) {
        acc_.ensure(cap_n_);
        xsph_.ensure(cap_n_);
        mu_.ensure(cap_n_);
    }
    if (!on) aux_valid_ = false;
    graph_dirty_ = true;
}

SceneTable DeviceFluid::table(Domain* dom) const {
    SceneTable T;
    T.info = d_info_.get();
    T.tile_scene = d_tiles_.get();
    T.params = d_params_.get();
    T.dom = dom;
    T.table_total = table_total_.get();
    T.nscenes = (int)scn_.size();
    T.ntiles = ntiles_;
    T.n = (int)n_;
    return T;
}

DeviceFluid::StepBufs DeviceFluid::own_bufs() {
    StepBufs s;
    s.posm_s = posm_s_.get();
    s.velm_s = velm_s_.get();
    s.body_s = body_s_.get();
    s.posl = posl_.get();
    s.id_s = id_s_.get();
    s.mat_s = mat_s_.get();
    s.cell_s = cell_s_.get();
    s.rho = rho_.get();
    s.prs = prs_.get();
    s.mp = mp_.get();
    s.start = start_.get();
    s.dom = dom_.get();
    s.acc = keep_aux_ ? acc_.get() : nullptr;
    s.mu = keep_aux_ ? mu_.get() : nullptr;
    s.xsph = keep_aux_ ? xsph_.get() : nullptr;
    s.fvisc = nullptr;
    return s;
}

void DeviceFluid::run_domain_and_sort(const StepBufs& s) {
    const SceneTable T = table(s.dom);
    launch_domain(st_, T, posm_.get());
    SortBuffers b;
    b.posm_in = posm_.get();
    b.vel_in = vel_.get();
    b.id_in = id_.get();
    b.mat_in = mat_.get();
    b.body_in = has_body_ ? body_.get() : nullptr;
    b.cell = cell_.get();
    b.rank = rank_.get();
    b.counts = counts_.get();
    b.partial = partial_.get();
    b.start = s.start;
    b.tmp = tmp_.get();
    b.ctmp = ctmp_.get();
    b.keytmp = keytmp_.get();
    b.posm_out = s.posm_s;
    b.velm_out = s.velm_s;
    b.id_out = s.id_s;
    b.mat_out = s.mat_s;
    b.body_out = s.body_s;
    b.cell_out = s.cell_s;
    b.posl_out = s.posl;
    launch_sort(st_, T, b);
}

hipEvent_t DeviceFluid::ev() {
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    hip_check(hipEventRecord(e, st_), "hipEventRecord");
    prof_ev_.push_back(e);
    return e;
}

void DeviceFluid::profile_begin() {
    sync();
    for (hipEvent_t e : prof_ev_) (void)hipEventDestroy(e);
    prof_ev_.clear();
    prof_ = true;
}

uint32_t DeviceFluid::profile_end(double ms4[4]) {
    sync();
    for (int k = 0; k < 4; ++k) ms4[k] = 0.0;
    uint32_t steps = (uint32_t)(prof_ev_.size() / 4);
    for (uint32_t s = 0; s < steps; ++s) {
        hipEvent_t* e = &prof_ev_[4 * s];
        float a = 0, b = 0, c = 0;
        hip_check(hipEventElapsedTime(&a, e[0], e[1]), "elapsed");
        hip_check(hipEventElapsedTime(&b, e[1], e[2]), "elapsed");
        hip_check(hipEventElapsedTime(&c, e[2], e[3]), "elapsed");
        ms4[0] += a;
        ms4[1] += b;
        ms4[2] += c;
        ms4[3] += (double)a + b + c;
    }
    for (hipEvent_t e : prof_ev_) (void)hipEventDestroy(e);
    prof_ev_.clear();
    prof_ = false;
    return steps;
}

// One step into the given buffer set; the new state (x, v, id, material)
// lands in posm_ / vel_ / id_ / mat_ in this step's cell order.
void DeviceFluid::step_core(float dt, const StepBufs& s) {
    if (prof_) ev();
    run_domain_and_sort(s);
    if (prof_) ev();
    const SceneTable T = table(s.dom);
    // with every scene's masses equal the density pass writes each slot's (x,
    // y, z, m p / rho^2) and the force pass reads its neighbours from that
    // and velm (same values, two loads instead of three); one buffer serves
    // every buffer set, like posl
    float4* const frec = all_um_ ? frec_.get() : nullptr;
    launch_density(st_, T, s.posm_s, s.posl, s.cell_s, s.start, s.rho, s.prs, s.velm_s, s.mp, frec,
                   tuning_);
    ForceArgs a;
    a.posm = s.posm_s;
    a.posl = s.posl;
    a.velm = s.velm_s;
    a.rho = s.rho;
    a.prs = s.prs;
    a.mp = s.mp;
    a.frec = frec;
    a.cell = s.cell_s;
    a.start = s.start;
    a.id = s.id_s;
    a.mat = s.mat_s;
    a.body = has_body_ ? s.body_s : nullptr;
    a.posm_out = posm_.get();
    a.vel_out = vel_.get();
    a.id_out = id_.get();
    a.mat_out = mat_.get();
    a.acc_out = s.acc;
    a.mu_out = s.mu;
    a.xsph_out = s.xsph;
    a.fvisc_out = s.fvisc;
    if (prof_) ev();
    launch_force(st_, T, dt, a, any_shear_ || taping_, tuning_);
    if (prof_) ev();
    hip_check(hipGetLastError(), "step launch");
}

void DeviceFluid::drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
}

// Records the whole step (domain, sort, density, force: no memsets, no host
// synchronisation) into one graph.
void DeviceFluid::capture(float dt) {
    drop_graph();
    hip_check(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        step_core(dt, own_bufs());
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    hip_check(hipStreamEndCapture(st_, &graph_), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
    graph_dt_ = dt;
    graph_dirty_ = false;
}

void DeviceFluid::set_graph(bool on) {
    graph_on_ = on;
    if (!on) drop_graph();
}

void DeviceFluid::set_tuning(const Tuning& t) {
    if (!tuning_valid(t))
        throw std::invalid_argument(
            "tuning: list lengths must be 8, 16 or 32, split 0 or 1, cell_div 1 or 2 and block 128 or 256");
    // the force pass's list length and the cell size are part of the
    // summation order: recomputed segments would not reproduce the recorded
    // steps
    if (!(t.density_list == tuning_.density_list && t.force_list == tuning_.force_list &&
          t.split == tuning_.split && t.cell_div == tuning_.cell_div && t.block == tuning_.block))
        tape_invalidate("the launch tuning changed");
    const int old_div = tuning_.cell_div;
    tuning_ = t;
    graph_dirty_ = true;
    if (t.cell_div != old_div) apply_cell_div(old_div);
}

// New cell size: every scene's device constants and bounds keys, and its
// table capacity scaled by the change in cells per volume.
void DeviceFluid::apply_cell_div(int old_div) {
    const int cd = tuning_.cell_div;
    for (size_t s = 0; s < scn_.size(); ++s) {
        Ps_[s] = make_dev_params(hps_[s], cd);
        for (int a = 0; a < 3; ++a) {
            scn_[s].bkey[a] = host_key(Ps_[s].lo[a], Ps_[s].inv_cell, cd);
            scn_[s].bkey[3 + a] = host_key(Ps_[s].hi[a], Ps_[s].inv_cell, cd);
        }
        const double f = std::pow((double)cd / (double)old_div, 3.0);
        const double c = std::ceil((double)scn_[s].cap * f);
        scn_[s].cap = (unsigned long long)std::min(std::max(c, 1.0), (double)kMaxDenseCells);
    }
    size_tables(nullptr);
    push_params();
    dens_valid_ = aux_valid_ = false;
}

std::string DeviceFluid::tuning_key() const {
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, dev_), "hipGetDeviceProperties");
    std::string arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    auto lg = [](size_t v) {
        int k = 0;
        while (v > 1) {
            v >>= 1;
            ++k;
        }
        return k;
    };
    std::ostringstream o;
    o << arch << " n" << lg(n_) << " s" << lg(scn_.size()) << " shear" << (any_shear_ ? 1 : 0);
    return o.str();
}

// Cache lines: "<key> density_list=<L> force_list=<L> split=<0|1> cell_div=<1|2>
// block=<128|256>", one per key (lines without the last two fields take the
// defaults: cells of h / 2 and 128-lane workgroups).
static bool read_tuning(const std::string& path, const std::string& key, Tuning& t) {
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        if (line.compare(0, key.size() + 1, key + " ") != 0) continue;
        Tuning c;
        const int got = std::sscanf(line.c_str() + key.size(),
                                    " density_list=%d force_list=%d split=%d cell_div=%d block=%d",
                                    &c.density_list, &c.force_list, &c.split, &c.cell_div, &c.block);
        if ((got == 3 || got == 5) && tuning_valid(c)) {
            t = c;
            return true;
        }
    }
    return false;
}

static void write_tuning(const std::string& path, const std::string& key, const Tuning& t) {
    std::vector<std::string> keep;
    {
        std::ifstream f(path);
        std::string line;
        while (std::getline(f, line))
            if (!line.empty() && line.compare(0, key.size() + 1, key + " ") != 0) keep.push_back(line);
    }
    std::ostringstream o;
    o << key << " density_list=" << t.density_list << " force_list=" << t.force_list
      << " split=" << t.split << " cell_div=" << t.cell_div << " block=" << t.block;
    keep.push_back(o.str());
    const std::string tmp = path + ".tmp";
    {
        std::ofstream f(tmp, std::ios::trunc);
        if (!f) throw std::runtime_error("tuning cache: cannot write " + tmp);
        for (const std::string& l : keep) f << l << "\n";
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0)
        throw std::runtime_error("tuning cache: cannot replace " + path);
}

bool DeviceFluid::autotune(float dt, uint32_t steps, const char* cache_path, double* ms) {
    if (taping_) throw std::logic_error("autotune: not while recording a tape");
    if (prof_) throw std::logic_error("autotune: not while profiling");
    if (has_body_) throw std::logic_error("autotune: a pending body_acc would be consumed");
    const std::string key = tuning_key();
    if (cache_path && *cache_path) {
        Tuning t;
        if (read_tuning(cache_path, key, t)) {
            set_tuning(t);
            return true;
        }
    }
    if (n_ == 0 || steps == 0) return false;
    Snapshot snap;
    save_state(snap);
    // `steps` eager steps of the saved state under tuning t: milliseconds per
    // step {sort, density, force, whole step}
    auto trial = [&](const Tuning& t, double out[4]) {
        set_tuning(t);
        load_state(snap);
        profile_begin();
        for (uint32_t s = 0; s < steps; ++s) step_core(dt, own_bufs());
        double m4[4];
        const uint32_t nst = profile_end(m4);
        for (int k = 0; k < 4; ++k) out[k] = m4[k] / nst;
    };
    double m4[4];
    // 1. the cell size, with the default lists and launch form: whole steps
    // (the sort's table and both passes' candidate counts change with it)
    Tuning t;
    double best = 0.0;
    int best_div = 2;
    for (int cd : {2, 1}) {
        t.cell_div = cd;
        trial(t, m4);
        if (cd == 2 || m4[3] < best) {
            best = m4[3];
            best_div = cd;
        }
    }
    // 2. list lengths and launch form at that cell size: trial k runs list
    // length cand[k % 3] in both passes, launch form k / 3; the density and
    // force passes are timed separately and chosen apart
    const int cand[3] = {8, 16, 32};
    double dms[6], fms[6];
    for (int k = 0; k < 6; ++k) {
        Tuning c;
        c.cell_div = best_div;
        c.density_list = c.force_list = cand[k % 3];
        c.split = k / 3;
        trial(c, m4);
        dms[k] = m4[1];
        fms[k] = m4[2];
    }
    // one launch form for both passes: the one with the smaller best sum
    int best_split = 0;
    double best_sum = 0.0;
    int bd = 0, bf = 0;
    for (int sp = 0; sp < 2; ++sp) {
        int d0 = 3 * sp, f0 = 3 * sp;
        for (int k = 1; k < 3; ++k) {
            if (dms[3 * sp + k] < dms[d0]) d0 = 3 * sp + k;
            if (fms[3 * sp + k] < fms[f0]) f0 = 3 * sp + k;
        }
        const double sum = dms[d0] + fms[f0];
        if (sp == 0 || sum < best_sum) {
            best_sum = sum;
            best_split = sp;
            bd = d0;
            bf = f0;
        }
    }
    // 3. the gathers' workgroup size for that choice
    t = Tuning();
    t.cell_div = best_div;
    t.density_list = cand[bd % 3];
    t.force_list = cand[bf % 3];
    t.split = best_split;
    double b256 = 0.0;
    for (int bs : {256, 128}) {
        t.block = bs;
        trial(t, m4);
        const double g = m4[1] + m4[2];
        if (bs == 256) b256 = g;
        else if (!(g < b256)) t.block = 256;
    }
    set_tuning(t);
    load_state(snap);
    sync();
    dens_valid_ = aux_valid_ = false;
    if (ms)
        for (int k = 0; k < 6; ++k) {
            ms[k] = dms[k];
            ms[6 + k] = fms[k];
        }
    if (cache_path && *cache_path) write_tuning(cache_path, key, t);
    return false;
}

void DeviceFluid::step(float dt) {
    if (n_ == 0) return;  // world.rs:184-186
    if (graph_on_ && !prof_ && !has_body_) {
        if (graph_dirty_ || !graph_exec_ || !(graph_dt_ == dt)) capture(dt);
        hip_check(hipGraphLaunch(graph_exec_, st_), "hipGraphLaunch");
    } else {
        step_core(dt, own_bufs());
    }
    has_body_ = false;  // integrate zeroes body_acc (world.rs:361)
    dens_valid_ = true;
    aux_valid_ = keep_aux_;
    if (taping_) tape_record(dt);
}

// ---------------------------------------------------------------------------
// adjoint tape

void DeviceFluid::save_state(Snapshot& c) {
    const size_t n = n_ ? n_ : 1;
    c.posm.ensure(n);
    c.vel.ensure(n);
    c.id.ensure(n);
    c.mat.ensure(n);
    if (!n_) return;
    hip_check(hipMemcpyAsync(c.posm.get(), posm_.get(), n_ * sizeof(float4), hipMemcpyDeviceToDevice, st_), "ckpt");
    hip_check(hipMemcpyAsync(c.vel.get(), vel_.get(), n_ * sizeof(float4), hipMemcpyDeviceToDevice, st_), "ckpt");
    hip_check(hipMemcpyAsync(c.id.get(), id_.get(), n_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, st_), "ckpt");
    hip_check(hipMemcpyAsync(c.mat.get(), mat_.get(), n_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, st_), "ckpt");
}

void DeviceFluid::load_state(const Snapshot& c) {
    if (!n_) return;
    hip_check(hipMemcpyAsync(posm_.get(), c.posm.get(), n_ * sizeof(float4), hipMemcpyDeviceToDevice, st_), "restore");
    hip_check(hipMemcpyAsync(vel_.get(), c.vel.get(), n_ * sizeof(float4), hipMemcpyDeviceToDevice, st_), "restore");
    hip_check(hipMemcpyAsync(id_.get(), c.id.get(), n_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, st_), "restore");
    hip_check(hipMemcpyAsync(mat_.get(), c.mat.get(), n_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, st_), "restore");
}

void DeviceFluid::tape_begin(uint32_t K) {
    if (K == 0) throw std::invalid_argument("tape: checkpoint interval must be >= 1");
    if (has_body_) throw std::invalid_argument("tape: body_acc is not differentiated");
    tape_end();
    taping_ = true;
    graph_dirty_ = true;  // recorded steps use the shear-rate force variant
    tape_K_ = K;
    tape_n_ = n_;
    tape_dt_.clear();
    ckpt_.emplace_back(new Snapshot());
    save_state(*ckpt_.back());
}

void DeviceFluid::tape_record(float dt) {
    if (n_ != tape_n_) throw std::logic_error("tape: the particle set changed while recording");
    tape_dt_.push_back(dt);
    if (tape_dt_.size() % tape_K_ == 0) {
        ckpt_.emplace_back(new Snapshot());
        save_state(*ckpt_.back());
    }
}

void DeviceFluid::tape_invalidate(const char* why) {
    if (taping_ && !tape_broken_) {
        tape_broken_ = true;
        tape_broken_why_ = why;
    }
}

void DeviceFluid::tape_end() {
    ckpt_.clear();
    slots_.clear();
    tape_dt_.clear();
    if (taping_) graph_dirty_ = true;
    taping_ = false;
    tape_broken_ = false;
    tape_broken_why_.clear();
}

DeviceFluid::StepBufs DeviceFluid::slot_bufs(TapeSlot& t) {
    const size_t n = n_ ? n_ : 1;
    t.posm_s.ensure(n + kScanPad); t.velm_s.ensure(n); t.body_s.ensure(1);
    t.id_s.ensure(n); t.mat_s.ensure(n); t.cell_s.ensure(n);
    t.rho.ensure(n); t.prs.ensure(n); t.mp.ensure(n);
    t.acc.ensure(n); t.xsph.ensure(n); t.fvisc.ensure(n); t.mu.ensure(n);
    if (!compact_slots()) t.start.ensure(table_cap_ + 1);
    ensure_dom(t.dom, scn_.size());
    StepBufs s;
    s.posm_s = t.posm_s.get(); s.velm_s = t.velm_s.get(); s.body_s = t.body_s.get();
    s.posl = posl_.get();
    s.id_s = t.id_s.get(); s.mat_s = t.mat_s.get(); s.cell_s = t.cell_s.get();
    s.rho = t.rho.get(); s.prs = t.prs.get(); s.mp = t.mp.get();
    // with compact slots the sort uses the world's own full-size table and
    // the used part is copied into the slot after the step
    s.start = compact_slots() ? start_.get() : t.start.get();
    s.dom = t.dom.get();
    s.acc = t.acc.get(); s.mu = t.mu.get(); s.xsph = t.xsph.get(); s.fvisc = t.fvisc.get();
    return s;
}

void DeviceFluid::tape_backward(const float* xbar_T, const float* vbar_T, const float* rhobar_T,
                                float* xbar0, float* vbar0, double* dparams) {
    if (!taping_) throw std::logic_error("tape: no recording (call tape_begin first)");
    if (tape_broken_)
        throw std::logic_error("tape: the recording is invalid (" + tape_broken_why_ +
                               " while recording); call tape_end and record again");
    const size_t T = tape_dt_.size();
    const int n = (int)n_;
    const size_t S = scn_.size();
    const size_t NP = (size_t)kAdjParams * S;
    for (size_t k = 0; k < NP; ++k) dparams[k] = 0.0;
    if (T == 0 || n == 0) {
        for (size_t k = 0; k < 3 * n_; ++k) {
            if (xbar0) xbar0[k] = xbar_T ? xbar_T[k] : 0.f;
            if (vbar0) vbar0[k] = vbar_T ? vbar_T[k] : 0.f;
        }
        return;
    }
    const size_t nt = (size_t)std::max(ntiles_, 1);
    for (int b = 0; b < 2; ++b) { xbar_[b].ensure(n); vbar_[b].ensure(n); }
    Fb_.ensure(n); Sb_.ensure(n); xb_loc_.ensure(n); vb_loc_.ensure(n); xb_acc_.ensure(n);
    rhob_loc_.ensure(n); rhob_.ensure(n); seed_.ensure(n);
    mush_.ensure(n); adj_part_.ensure(nt * kAdjParams); adj_tot_.ensure(NP);
    {
        std::vector<float4> xs(n), vs(n);
        for (int i = 0; i < n; ++i) {
            xs[i] = xbar_T ? make_float4(xbar_T[3 * i], xbar_T[3 * i + 1], xbar_T[3 * i + 2], 0.f)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
            vs[i] = vbar_T ? make_float4(vbar_T[3 * i], vbar_T[3 * i + 1], vbar_T[3 * i + 2], 0.f)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        std::vector<double> z(NP, 0.0);
        hip_check(hipMemcpyAsync(xbar_[0].get(), xs.data(), n * sizeof(float4), hipMemcpyHostToDevice, st_), "adj");
        hip_check(hipMemcpyAsync(vbar_[0].get(), vs.data(), n * sizeof(float4), hipMemcpyHostToDevice, st_), "adj");
        if (rhobar_T)
            hip_check(hipMemcpyAsync(seed_.get(), rhobar_T, n * sizeof(float), hipMemcpyHostToDevice, st_), "adj");
        hip_check(hipMemcpyAsync(adj_tot_.get(), z.data(), NP * sizeof(double), hipMemcpyHostToDevice, st_), "adj");
        sync();
    }
    Snapshot final_state;
    save_state(final_state);
    const bool prof = prof_;
    prof_ = false;
    int cur = 0;
    const size_t K = tape_K_;
    const size_t nseg = (T + K - 1) / K;
    if (slots_.size() < std::min(K, T)) {
        slots_.clear();
        for (size_t k = 0; k < std::min(K, T); ++k) slots_.emplace_back(new TapeSlot());
    }
    for (size_t seg = nseg; seg-- > 0;) {
        load_state(*ckpt_[seg]);
        const size_t s0 = seg * K, s1 = std::min(T, s0 + K);
        for (size_t t = s0; t < s1; ++t) {
            TapeSlot& sl = *slots_[t - s0];
            step_core(tape_dt_[t], slot_bufs(sl));
            if (compact_slots()) {
                uint32_t total = 0;
                hip_check(hipMemcpyAsync(&total, table_total_.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, st_), "table size");
                sync();
                sl.start.ensure((size_t)total + 1);
                hip_check(hipMemcpyAsync(sl.start.get(), start_.get(), ((size_t)total + 1) * sizeof(uint32_t),
                                         hipMemcpyDeviceToDevice, st_), "slot table");
            }
        }
        for (size_t t = s1; t-- > s0;) {
            TapeSlot& sl = *slots_[t - s0];
            AdjSlot A;
            A.posm = sl.posm_s.get(); A.velm = sl.velm_s.get(); A.rho = sl.rho.get(); A.prs = sl.prs.get();
            A.cell = sl.cell_s.get(); A.start = sl.start.get(); A.id = sl.id_s.get();
            A.acc = sl.acc.get(); A.xsph = sl.xsph.get(); A.fvisc = sl.fvisc.get();
            A.mu = sl.mu.get(); A.mat = sl.mat_s.get(); A.dt = tape_dt_[t];
            AdjBuffers B;
            B.xbar_in = xbar_[cur].get(); B.vbar_in = vbar_[cur].get();
            B.Fb = Fb_.get(); B.Sb = Sb_.get(); B.xb_loc = xb_loc_.get(); B.vb_loc = vb_loc_.get();
            B.rhob_loc = rhob_loc_.get(); B.rhob = rhob_.get(); B.xb_acc = xb_acc_.get();
            B.xbar_out = xbar_[1 - cur].get(); B.vbar_out = vbar_[1 - cur].get();
            B.mush = mush_.get(); B.part = adj_part_.get(); B.totals = adj_tot_.get();
            launch_adjoint_step(st_, table(sl.dom.get()), A, B,
                                (t == T - 1 && rhobar_T) ? seed_.get() : nullptr, any_shear_);
            hip_check(hipGetLastError(), "adjoint launch");
            cur = 1 - cur;
        }
    }
    std::vector<float4> xs(n), vs(n);
    std::vector<double> tot(NP);
    hip_check(hipMemcpyAsync(xs.data(), xbar_[cur].get(), n * sizeof(float4), hipMemcpyDeviceToHost, st_), "adj out");
    hip_check(hipMemcpyAsync(vs.data(), vbar_[cur].get(), n * sizeof(float4), hipMemcpyDeviceToHost, st_), "adj out");
    hip_check(hipMemcpyAsync(tot.data(), adj_tot_.get(), NP * sizeof(double), hipMemcpyDeviceToHost, st_), "adj out");
    load_state(final_state);
    sync();
    for (int i = 0; i < n; ++i) {
        if (xbar0) { xbar0[3 * i] = xs[i].x; xbar0[3 * i + 1] = xs[i].y; xbar0[3 * i + 2] = xs[i].z; }
        if (vbar0) { vbar0[3 * i] = vs[i].x; vbar0[3 * i + 1] = vs[i].y; vbar0[3 * i + 2] = vs[i].z; }
    }
    for (size_t k = 0; k < NP; ++k) dparams[k] = tot[k];
    prof_ = prof;
    // the slot steps wrote only slot buffers and the state, which is restored;
    // rho / p / aux of the last forward step are untouched
}

void DeviceFluid::download(Field f, float* out) {
    const int n = (int)n_;
    if (n == 0) return;
    const uint32_t* id = id_.get();
    // rho/p/aux belong to the sorted order of the last step, which equals the
    // current order (the force pass writes the new state in that order)
    switch (f) {
        case F_POS:
        case F_VEL:
        case F_ACC:
        case F_XSPH:
        case F_MASS: {
            const float4* src = f == F_POS || f == F_MASS ? posm_.get()
                                : f == F_VEL              ? vel_.get()
                                : f == F_ACC              ? acc_.get()
                                                          : xsph_.get();
            if ((f == F_ACC || f == F_XSPH) && !aux_valid_)
                throw std::logic_error("phy-fluid: acc/xsph need keep_aux before the step");
            launch_unsort4(st_, n, id, src, stage4_.get());
            std::vector<float4> h(n);
            hip_check(hipMemcpyAsync(h.data(), stage4_.get(), n * sizeof(float4), hipMemcpyDeviceToHost, st_), "download");
            sync();
            for (int i = 0; i < n; ++i) {
                if (f == F_MASS) {
                    out[i] = h[i].w;
                } else {
                    out[3 * i] = h[i].x;
                    out[3 * i + 1] = h[i].y;
                    out[3 * i + 2] = h[i].z;
                }
            }
            return;
        }
        case F_RHO:
        case F_P:
        case F_MU_EFF: {
            if (f == F_MU_EFF && !aux_valid_)
                throw std::logic_error("phy-fluid: mu_eff needs keep_aux before the step");
            if ((f == F_RHO || f == F_P) && !dens_valid_)
                throw std::logic_error("phy-fluid: rho/p are defined after a step");
            const float* src = f == F_RHO ? rho_.get() : f == F_P ? prs_.get() : mu_.get();
            launch_unsort1(st_, n, id, src, stage1_.get());
            hip_check(hipMemcpyAsync(out, stage1_.get(), n * sizeof(float), hipMemcpyDeviceToHost, st_), "download");
            sync();
            return;
        }
    }
    throw std::invalid_argument("phy-fluid: unknown field");
}

void DeviceFluid::last_validate_all(int32_t* status) {
    const size_t S = scn_.size();
    if (n_ == 0) {
        for (size_t s = 0; s < S; ++s) status[s] = 0;
        return;
    }
    // only the two watchdog words of each scene's Domain
    static_assert(offsetof(Domain, bad_vel) == offsetof(Domain, bad_pos) + sizeof(uint32_t), "adjacent words");
    std::vector<uint32_t> w(2 * S);
    hip_check(hipMemcpy2DAsync(w.data(), 2 * sizeof(uint32_t),
                               reinterpret_cast<const char*>(dom_.get()) + offsetof(Domain, bad_pos),
                               sizeof(Domain), 2 * sizeof(uint32_t), S, hipMemcpyDeviceToHost, st_),
              "watchdog");
    sync();
    for (size_t s = 0; s < S; ++s) {
        // first particle in index order with a bad field; pos is checked before vel
        const uint32_t bp = w[2 * s], bv = w[2 * s + 1];
        if (bp == 0xffffffffu && bv == 0xffffffffu)
            status[s] = 0;
        else
            status[s] = bp <= bv ? 1 : 2;
    }
}

void DeviceFluid::download_material(uint32_t* out) {
    const int n = (int)n_;
    if (n == 0) return;
    // a 4-byte gather: the float unsort moves the tags' bits unchanged
    launch_unsort1(st_, n, id_.get(), reinterpret_cast<const float*>(mat_.get()), stage1_.get());
    hip_check(hipMemcpyAsync(out, stage1_.get(), n * sizeof(uint32_t), hipMemcpyDeviceToHost, st_), "download");
    sync();
}

int DeviceFluid::last_validate() {
    std::vector<int32_t> st(scn_.size());
    last_validate_all(st.data());
    return st[0];
}

void DeviceFluid::diagnostics(double* out5) {
    const size_t S = scn_.size();
    for (size_t k = 0; k < 5 * S; ++k) out5[k] = 0.0;
    if (n_ == 0) return;
    launch_diag(st_, table(dom_.get()), posm_.get(), vel_.get(), diag_part_.get(), diag_out_.get());
    hip_check(hipMemcpyAsync(out5, diag_out_.get(), 5 * S * sizeof(double), hipMemcpyDeviceToHost, st_), "diag");
    sync();
}

void DeviceFluid::neighbor_stats(uint32_t* cnt, uint64_t* sum, uint64_t* xr) {
    const int n = (int)n_;
    if (n == 0) return;
    // sort the current state (the same binning as a step) with a domain array
    // of its own, so the last step's watchdog survives, then count
    DevBuf<Domain> dom;
    ensure_dom(dom, scn_.size());
    StepBufs sb = own_bufs();
    sb.dom = dom.get();
    run_domain_and_sort(sb);
    DevBuf<uint32_t> c;
    DevBuf<unsigned long long> s, x;
    c.ensure(n);
    s.ensure(n);
    x.ensure(n);
    launch_neighbor_stats(st_, table(sb.dom), posm_s_.get(), cell_s_.get(), start_.get(), id_s_.get(),
                          c.get(), s.get(), x.get());
    hip_check(hipMemcpyAsync(cnt, c.get(), n * sizeof(uint32_t), hipMemcpyDeviceToHost, st_), "nstats");
    hip_check(hipMemcpyAsync(sum, s.get(), n * sizeof(uint64_t), hipMemcpyDeviceToHost, st_), "nstats");
    hip_check(hipMemcpyAsync(xr, x.get(), n * sizeof(uint64_t), hipMemcpyDeviceToHost, st_), "nstats");
    sync();
}

void DeviceFluid::gather_stats(int pass, uint32_t* cnt, uint64_t* sum, uint64_t* xr) {
    if (pass != 0 && pass != 1) throw std::invalid_argument("gather_stats: pass must be 0 (density) or 1 (force)");
    const int n = (int)n_;
    if (n == 0) return;
    // the step's binning into the step's sorted buffers, with a domain array
    // of its own (the last step's watchdog survives); rho / p / aux of the
    // last step are not touched, since the statistics kernels write only the
    // statistics
    DevBuf<Domain> dom;
    ensure_dom(dom, scn_.size());
    StepBufs sb = own_bufs();
    sb.dom = dom.get();
    run_domain_and_sort(sb);
    DevBuf<uint32_t> c;
    DevBuf<unsigned long long> s, x;
    c.ensure(n);
    s.ensure(n);
    x.ensure(n);
    ForceArgs a{};
    a.posm = sb.posm_s;
    a.posl = sb.posl;
    a.velm = sb.velm_s;
    a.rho = sb.rho;
    a.prs = sb.prs;
    a.mp = sb.mp;
    a.cell = sb.cell_s;
    a.start = sb.start;
    a.id = sb.id_s;
    a.mat = sb.mat_s;
    GatherStats g{sb.id_s, c.get(), s.get(), x.get()};
    // the force variant the step would launch (recorded steps sum the shear rate)
    launch_gather_stats(st_, table(sb.dom), pass, a, any_shear_ || taping_, tuning_, g);
    hip_check(hipGetLastError(), "gather_stats");
    hip_check(hipMemcpyAsync(cnt, c.get(), n * sizeof(uint32_t), hipMemcpyDeviceToHost, st_), "gstats");
    hip_check(hipMemcpyAsync(sum, s.get(), n * sizeof(uint64_t), hipMemcpyDeviceToHost, st_), "gstats");
    hip_check(hipMemcpyAsync(xr, x.get(), n * sizeof(uint64_t), hipMemcpyDeviceToHost, st_), "gstats");
    sync();
}

// ---------------------------------------------------------------------------
// pending body_acc and coupling

void DeviceFluid::set_body_acc(const float* b3) {
    if (n_ == 0) return;
    std::vector<float4> b4(n_);
    bool any = false;
    for (size_t i = 0; i < n_; ++i) {
        b4[i] = make_float4(b3[3 * i], b3[3 * i + 1], b3[3 * i + 2], 0.f);
        any = any || b4[i].x != 0.f || b4[i].y != 0.f || b4[i].z != 0.f;
    }
    has_body_ = any;
    if (!any) return;
    // body_acc is not differentiated: the recomputed segments would step
    // without it
    tape_invalidate("a body acceleration was set");
    stage4_.ensure(n_);
    hip_check(hipMemcpyAsync(stage4_.get(), b4.data(), n_ * sizeof(float4), hipMemcpyHostToDevice, st_), "body_acc");
    launch_gather_by_id4(st_, (int)n_, id_.get(), stage4_.get(), body_.get());
    hip_check(hipGetLastError(), "body_acc");
    sync();
}

void DeviceFluid::get_body_acc(float* out3) {
    if (!has_body_) {
        for (size_t k = 0; k < 3 * n_; ++k) out3[k] = 0.f;
        return;
    }
    launch_unsort4(st_, (int)n_, id_.get(), body_.get(), stage4_.get());
    std::vector<float4> h(n_);
    hip_check(hipMemcpyAsync(h.data(), stage4_.get(), n_ * sizeof(float4), hipMemcpyDeviceToHost, st_), "body_acc");
    sync();
    for (size_t i = 0; i < n_; ++i) {
        out3[3 * i] = h[i].x;
        out3[3 * i + 1] = h[i].y;
        out3[3 * i + 2] = h[i].z;
    }
}

CoupleParams DeviceFluid::couple_params() const {
    const HostParams& h = hps_[0];
    CoupleParams c;
    for (int a = 0; a < 3; ++a) {
        c.lo[a] = h.bounds_min[a];
        c.hi[a] = h.bounds_max[a];
        c.g[a] = h.gravity[a];
    }
    c.rest = h.rest_density;
    c.mass = h.mass;
    c.inv_cell = Ps_[0].inv_cell;
    c.h = h.h;
    return c;
}

void DeviceFluid::one_scene(const char* what) const {
    if (scn_.size() != 1) throw std::logic_error(std::string("phy-fluid: ") + what + " needs a one-scene world");
}

void DeviceFluid::sort_current() {
    ensure_dom(cdom_, scn_.size());
    StepBufs sb = own_bufs();
    sb.dom = cdom_.get();
    run_domain_and_sort(sb);
    inv_.ensure(n_);
    launch_inverse_perm(st_, (int)n_, id_.get(), inv_.get());
}

void DeviceFluid::couple_bodies(CoupleBody* bodies, uint32_t nb, double dt) {
    one_scene("couple_bodies");
    if (n_ == 0 || nb == 0) return;  // world.rs:415-417
    tape_invalidate("a coupling pass changed the particles");
    const int n = (int)n_;
    cbodies_.ensure(nb);
    cdisp_.ensure(nb);
    ccnt_.ensure(n_);
    coff_.ensure(n_ + 1);
    cn_.ensure(1);
    cpartial_.ensure(kScanBlocks);
    const uint32_t n32 = (uint32_t)n_;
    std::vector<double> zero(nb, 0.0);
    hip_check(hipMemcpyAsync(cbodies_.get(), bodies, nb * sizeof(CoupleBody), hipMemcpyHostToDevice, st_), "bodies");
    hip_check(hipMemcpyAsync(cdisp_.get(), zero.data(), nb * sizeof(double), hipMemcpyHostToDevice, st_), "bodies");
    hip_check(hipMemcpyAsync(cn_.get(), &n32, sizeof(uint32_t), hipMemcpyHostToDevice, st_), "bodies");
    launch_contact_count(st_, n, posm_.get(), id_.get(), cbodies_.get(), (int)nb, ccnt_.get());
    launch_exclusive_scan(st_, ccnt_.get(), cn_.get(), cpartial_.get(), coff_.get());
    uint32_t K = 0;
    hip_check(hipMemcpyAsync(&K, coff_.get() + n_, sizeof(uint32_t), hipMemcpyDeviceToHost, st_), "contacts");
    sync();
    contacts_.ensure(K ? K : 1);
    if (K > 0)
        launch_contact_fill(st_, n, posm_.get(), id_.get(), cbodies_.get(), (int)nb, coff_.get(), contacts_.get());
    // with no contact only the buoyancy pass runs (every displaced volume is zero)
    launch_contact_resolve(st_, contacts_.get(), coff_.get() + n_, posm_.get(), vel_.get(), cbodies_.get(), (int)nb,
                           cdisp_.get(), couple_params(), dt);
    hip_check(hipGetLastError(), "couple_bodies");
    hip_check(hipMemcpyAsync(bodies, cbodies_.get(), nb * sizeof(CoupleBody), hipMemcpyDeviceToHost, st_), "bodies");
    sync();
}

void DeviceFluid::couple_points(CouplePoint* pts, uint32_t np, double dt, double drag, double soft_density) {
    one_scene("couple_points");
    if (np == 0 || n_ == 0) return;  // world.rs:692-694
    tape_invalidate("a coupling pass changed the particles");
    const int n = (int)n_;
    sort_current();
    vel64_.ensure(3 * n_);
    cpoints_.ensure(np);
    launch_widen_vel(st_, n, vel_.get(), vel64_.get());
    hip_check(hipMemcpyAsync(cpoints_.get(), pts, np * sizeof(CouplePoint), hipMemcpyHostToDevice, st_), "points");
    launch_couple_points(st_, cpoints_.get(), (int)np, couple_params(), dt, drag, soft_density, posm_s_.get(),
                         id_s_.get(), start_.get(), cdom_.get(), inv_.get(), vel64_.get());
    launch_narrow_vel(st_, n, vel64_.get(), vel_.get());
    hip_check(hipGetLastError(), "couple_points");
    hip_check(hipMemcpyAsync(pts, cpoints_.get(), np * sizeof(CouplePoint), hipMemcpyDeviceToHost, st_), "points");
    sync();
}

void DeviceFluid::couple_heat(const HeatGrid& hgh, double* src, double dt, double t_ref, double beta,
                              double heat_gain) {
    one_scene("couple_heat");
    if (n_ == 0) return;  // world.rs:565-567
    if (hgh.nx == 0 || hgh.ny == 0 || hgh.nz == 0) throw std::invalid_argument("couple_heat: empty heat grid");
    const uint64_t ncell = (uint64_t)hgh.nx * hgh.ny * hgh.nz;
    if (ncell >= 0xffffffffull) throw std::invalid_argument("couple_heat: heat grid too large");
    if (heat_gain > 0.0 && !src) throw std::invalid_argument("couple_heat: heat_gain > 0 needs a source grid");
    tape_invalidate("a coupling pass changed the particles");
    const int n = (int)n_;
    if (!has_body_) {
        hip_check(hipMemsetAsync(body_.get(), 0, n_ * sizeof(float4), st_), "body_acc");
        has_body_ = true;
    }
    heat_u_.ensure(ncell);
    hip_check(hipMemcpyAsync(heat_u_.get(), hgh.u, ncell * sizeof(double), hipMemcpyHostToDevice, st_), "heat");
    HeatGrid hg = hgh;
    hg.u = heat_u_.get();
    HeatScratch hs;
    heat_q_.ensure(n_);
    heat_qcell_.ensure(n_);
    heat_max_.ensure(1);
    heat_scale_.ensure(1);
    heat_lo_.ensure(ncell);
    heat_hi_.ensure(ncell);
    heat_flags_.ensure(ncell);
    hs.q = heat_q_.get();
    hs.qcell = heat_qcell_.get();
    hs.qmax = heat_max_.get();
    hs.scale = heat_scale_.get();
    hs.lo = heat_lo_.get();
    hs.hi = heat_hi_.get();
    hs.flags = heat_flags_.get();
    hip_check(hipMemsetAsync(hs.qmax, 0, sizeof(unsigned long long), st_), "heat");
    launch_heat_body(st_, n, posm_.get(), vel_.get(), body_.get(), couple_params(), hg, dt, t_ref, beta,
                     heat_gain, hs);
    if (heat_gain > 0.0) {
        heat_src_.ensure(ncell);
        hip_check(hipMemsetAsync(hs.lo, 0, ncell * sizeof(unsigned long long), st_), "heat");
        hip_check(hipMemsetAsync(hs.hi, 0, ncell * sizeof(unsigned long long), st_), "heat");
        hip_check(hipMemsetAsync(hs.flags, 0, ncell * sizeof(uint32_t), st_), "heat");
        hip_check(hipMemcpyAsync(heat_src_.get(), src, ncell * sizeof(double), hipMemcpyHostToDevice, st_), "heat");
        launch_heat_source(st_, n, dt, hs, (uint32_t)ncell, heat_src_.get());
        hip_check(hipMemcpyAsync(src, heat_src_.get(), ncell * sizeof(double), hipMemcpyDeviceToHost, st_), "heat");
    }
    hip_check(hipGetLastError(), "couple_heat");
    sync();
}

void DeviceFluid::sample_velocity(const double* q3, uint32_t nq, double* out3) {
    one_scene("sample_velocity");
    if (nq == 0) return;
    if (n_ == 0) {
        for (size_t k = 0; k < 3 * (size_t)nq; ++k) out3[k] = 0.0;
        return;
    }
    sort_current();
    cq_.ensure(6 * (size_t)nq);
    hip_check(hipMemcpyAsync(cq_.get(), q3, 3 * (size_t)nq * sizeof(double), hipMemcpyHostToDevice, st_), "sampler");
    launch_nearest_velocity(st_, (int)nq, cq_.get(), cq_.get() + 3 * (size_t)nq, posm_s_.get(), id_s_.get(),
                            start_.get(), cdom_.get(), inv_.get(), vel_.get(), couple_params());
    hip_check(hipGetLastError(), "sampler");
    hip_check(hipMemcpyAsync(out3, cq_.get() + 3 * (size_t)nq, 3 * (size_t)nq * sizeof(double),
                             hipMemcpyDeviceToHost, st_), "sampler");
    sync();
}

}  // namespace phy
