// archive.cpp — JSON world archive (see archive.h for the layout).
#include "archive.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <utility>
#include <vector>

namespace phy {
namespace {

// ---------------------------------------------------------------------------
// minimal JSON value: null, bool, integer (u64 / i64), double, string, array,
// object (insertion order kept for writing)

struct Json {
    enum Kind { Null, Bool, UInt, Int, Float, Str, Arr, Obj } kind = Null;
    bool b = false;
    uint64_t u = 0;
    int64_t i = 0;
    double f = 0.0;
    std::string s;
    std::vector<Json> a;
    std::vector<std::pair<std::string, Json>> o;

    static Json uint(uint64_t v) {
        Json j;
        j.kind = UInt;
        j.u = v;
        return j;
    }
    static Json array() {
        Json j;
        j.kind = Arr;
        return j;
    }
    static Json object() {
        Json j;
        j.kind = Obj;
        return j;
    }
    Json& add(const std::string& k, Json v) {
        o.emplace_back(k, std::move(v));
        return o.back().second;
    }
    const Json* get(const std::string& k) const {
        if (kind != Obj) return nullptr;
        for (const auto& kv : o)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    const Json& at(const std::string& k) const {
        const Json* p = get(k);
        if (!p) throw std::runtime_error("archive: missing field \"" + k + "\"");
        return *p;
    }
};

[[noreturn]] void bad(const std::string& what) { throw std::runtime_error("archive: " + what); }

class Parser {
   public:
    explicit Parser(const std::string& t) : t_(t) {}
    Json parse() {
        Json v = value();
        ws();
        if (p_ != t_.size()) bad("trailing characters");
        return v;
    }

   private:
    const std::string& t_;
    size_t p_ = 0;
    // Nesting depth of the open objects / arrays. serde_json stops at 128
    // levels and reports an error; so does this reader, instead of recursing
    // until the stack overflows.
    int depth_ = 0;
    static constexpr int kMaxDepth = 128;
    struct Nest {
        int& d;
        explicit Nest(int& depth) : d(depth) {
            if (++d > kMaxDepth) bad("nesting too deep (more than 128 levels)");
        }
        ~Nest() { --d; }
    };

    void ws() {
        while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\n' || t_[p_] == '\r' || t_[p_] == '\t')) ++p_;
    }
    char peek() {
        ws();
        if (p_ >= t_.size()) bad("unexpected end of input");
        return t_[p_];
    }
    void expect(char c) {
        if (peek() != c) bad(std::string("expected '") + c + "'");
        ++p_;
    }
    bool lit(const char* w) {
        size_t n = std::strlen(w);
        if (t_.compare(p_, n, w) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }
    Json value() {
        char c = peek();
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') {
            Json j;
            j.kind = Json::Str;
            j.s = string();
            return j;
        }
        Json j;
        if (lit("null")) return j;
        if (lit("true")) {
            j.kind = Json::Bool;
            j.b = true;
            return j;
        }
        if (lit("false")) {
            j.kind = Json::Bool;
            return j;
        }
        return number();
    }
    Json object() {
        Nest nest(depth_);
        expect('{');
        Json j = Json::object();
        if (peek() == '}') {
            ++p_;
            return j;
        }
        for (;;) {
            if (peek() != '"') bad("expected a key");
            std::string k = string();
            expect(':');
            j.add(k, value());
            char c = peek();
            ++p_;
            if (c == '}') return j;
            if (c != ',') bad("expected ',' or '}'");
        }
    }
    Json array() {
        Nest nest(depth_);
        expect('[');
        Json j = Json::array();
        if (peek() == ']') {
            ++p_;
            return j;
        }
        for (;;) {
            j.a.push_back(value());
            char c = peek();
            ++p_;
            if (c == ']') return j;
            if (c != ',') bad("expected ',' or ']'");
        }
    }
    std::string string() {
        expect('"');
        std::string out;
        while (p_ < t_.size() && t_[p_] != '"') {
            char c = t_[p_++];
            if (c == '\\') {
                if (p_ >= t_.size()) bad("bad escape");
                char e = t_[p_++];
                switch (e) {
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'n': out += '\n'; break;
                    case 'r': out += '\r'; break;
                    case 't': out += '\t'; break;
                    case 'u': {
                        if (p_ + 4 > t_.size()) bad("bad \\u escape");
                        unsigned cp = (unsigned)std::stoul(t_.substr(p_, 4), nullptr, 16);
                        p_ += 4;
                        if (cp < 0x80) {
                            out += (char)cp;
                        } else if (cp < 0x800) {
                            out += (char)(0xc0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3f));
                        } else {
                            out += (char)(0xe0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3f));
                            out += (char)(0x80 | (cp & 0x3f));
                        }
                        break;
                    }
                    default: bad("bad escape");
                }
            } else {
                out += c;
            }
        }
        if (p_ >= t_.size()) bad("unterminated string");
        ++p_;
        return out;
    }
    Json number() {
        size_t b = p_;
        if (p_ < t_.size() && (t_[p_] == '-' || t_[p_] == '+')) ++p_;
        bool isf = false;
        while (p_ < t_.size()) {
            char c = t_[p_];
            if (c >= '0' && c <= '9') {
                ++p_;
            } else if (c == '.' || c == 'e' || c == 'E' || c == '-' || c == '+') {
                isf = true;
                ++p_;
            } else {
                break;
            }
        }
        if (p_ == b) bad("unexpected character");
        const std::string txt = t_.substr(b, p_ - b);
        Json j;
        try {
            if (!isf && txt[0] != '-') {
                j.kind = Json::UInt;
                j.u = std::stoull(txt);
            } else if (!isf) {
                j.kind = Json::Int;
                j.i = std::stoll(txt);
            } else {
                j.kind = Json::Float;
                j.f = std::stod(txt);
            }
        } catch (const std::exception&) {
            bad("bad number " + txt);
        }
        return j;
    }
};

void write(const Json& j, std::string& out, int indent) {
    auto nl = [&](int k) {
        out += '\n';
        out.append((size_t)(2 * k), ' ');
    };
    switch (j.kind) {
        case Json::Null: out += "null"; break;
        case Json::Bool: out += j.b ? "true" : "false"; break;
        case Json::UInt: out += std::to_string(j.u); break;
        case Json::Int: out += std::to_string(j.i); break;
        case Json::Float: {
            char buf[64];
            std::snprintf(buf, sizeof buf, "%.17g", j.f);
            out += buf;
            break;
        }
        case Json::Str: {
            out += '"';
            for (char c : j.s) {
                if (c == '"' || c == '\\') out += '\\';
                out += c;
            }
            out += '"';
            break;
        }
        case Json::Arr: {
            if (j.a.empty()) {
                out += "[]";
                break;
            }
            out += '[';
            for (size_t k = 0; k < j.a.size(); ++k) {
                nl(indent + 1);
                write(j.a[k], out, indent + 1);
                if (k + 1 < j.a.size()) out += ',';
            }
            nl(indent);
            out += ']';
            break;
        }
        case Json::Obj: {
            if (j.o.empty()) {
                out += "{}";
                break;
            }
            out += '{';
            for (size_t k = 0; k < j.o.size(); ++k) {
                nl(indent + 1);
                Json key;
                key.kind = Json::Str;
                key.s = j.o[k].first;
                write(key, out, indent + 1);
                out += ": ";
                write(j.o[k].second, out, indent + 1);
                if (k + 1 < j.o.size()) out += ',';
            }
            nl(indent);
            out += '}';
            break;
        }
    }
}

// ---------------------------------------------------------------------------
// field codecs

uint64_t bits_of(double v) {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    return b;
}
double from_bits(uint64_t b) {
    double v;
    std::memcpy(&v, &b, 8);
    return v;
}

Json f64(double v) {  // {"__f64": bits} (archive.rs:31-60)
    Json j = Json::object();
    j.add("__f64", Json::uint(bits_of(v)));
    return j;
}

Json vec3(const Vec3& v) {  // [bits, bits, bits] (serde_geom)
    Json j = Json::array();
    j.a.push_back(Json::uint(bits_of(v.x)));
    j.a.push_back(Json::uint(bits_of(v.y)));
    j.a.push_back(Json::uint(bits_of(v.z)));
    return j;
}

double read_f64(const Json& j, const char* what) {
    if (j.kind == Json::Obj) {
        const Json* b = j.get("__f64");
        if (!b || b->kind != Json::UInt) bad(std::string("bad f64 field ") + what);
        double v = from_bits(b->u);
        // archive.rs:77-81 (Number::from_f64(..).expect("finite f64"))
        if (!std::isfinite(v)) bad(std::string("non-finite f64 in field ") + what);
        return v;
    }
    if (j.kind == Json::Float) return j.f;
    if (j.kind == Json::UInt) return (double)j.u;
    if (j.kind == Json::Int) return (double)j.i;
    bad(std::string("expected a number in field ") + what);
}

uint64_t read_uint(const Json& j, const char* what) {
    if (j.kind != Json::UInt) bad(std::string("expected an unsigned integer in field ") + what);
    return j.u;
}

Vec3 read_vec3(const Json& j, const char* what) {
    if (j.kind != Json::Arr || j.a.size() != 3) bad(std::string("expected three bit patterns in field ") + what);
    return Vec3{from_bits(read_uint(j.a[0], what)), from_bits(read_uint(j.a[1], what)),
                from_bits(read_uint(j.a[2], what))};
}

std::vector<double> read_f64s(const Json& j, const char* what) {
    if (j.kind != Json::Arr) bad(std::string("expected an array in field ") + what);
    std::vector<double> v;
    for (const Json& e : j.a) v.push_back(read_f64(e, what));
    return v;
}

Json params_json(const SphParams& p) {
    Json j = Json::object();
    j.add("rest_density", f64(p.rest_density));
    j.add("stiffness", f64(p.stiffness));
    j.add("viscosity", f64(p.viscosity));
    Json k = Json::array(), n = Json::array();
    for (double v : p.visc_k) k.a.push_back(f64(v));
    for (double v : p.visc_n) n.a.push_back(f64(v));
    j.add("visc_k", k);
    j.add("visc_n", n);
    j.add("shear_min", f64(p.shear_min));
    j.add("mass", f64(p.mass));
    j.add("h", f64(p.h));
    j.add("gravity", vec3(p.gravity));
    j.add("bounds_min", vec3(p.bounds_min));
    j.add("bounds_max", vec3(p.bounds_max));
    j.add("boundary_damp", f64(p.boundary_damp));
    j.add("xsph_eps", f64(p.xsph_eps));
    return j;
}

SphParams params_from(const Json& j) {
    SphParams p;
    p.rest_density = read_f64(j.at("rest_density"), "rest_density");
    p.stiffness = read_f64(j.at("stiffness"), "stiffness");
    p.viscosity = read_f64(j.at("viscosity"), "viscosity");
    p.visc_k = read_f64s(j.at("visc_k"), "visc_k");
    p.visc_n = read_f64s(j.at("visc_n"), "visc_n");
    p.shear_min = read_f64(j.at("shear_min"), "shear_min");
    p.mass = read_f64(j.at("mass"), "mass");
    p.h = read_f64(j.at("h"), "h");
    p.gravity = read_vec3(j.at("gravity"), "gravity");
    p.bounds_min = read_vec3(j.at("bounds_min"), "bounds_min");
    p.bounds_max = read_vec3(j.at("bounds_max"), "bounds_max");
    p.boundary_damp = read_f64(j.at("boundary_damp"), "boundary_damp");
    p.xsph_eps = read_f64(j.at("xsph_eps"), "xsph_eps");
    return p;
}

// Particle fields in serde order; mu_eff and xsph are not archived
// (particle.rs:40-45, serde(skip)).
Json particle_json(const Particle& q) {
    Json j = Json::object();
    j.add("pos", vec3(q.pos));
    j.add("vel", vec3(q.vel));
    j.add("acc", vec3(q.acc));
    j.add("body_acc", vec3(q.body_acc));
    j.add("rho", f64(q.rho));
    j.add("p", f64(q.p));
    j.add("mass", f64(q.mass));
    j.add("material", Json::uint(q.material));
    return j;
}

Particle particle_from(const Json& j) {
    Particle q;
    q.pos = read_vec3(j.at("pos"), "pos");
    q.vel = read_vec3(j.at("vel"), "vel");
    q.acc = read_vec3(j.at("acc"), "acc");
    q.body_acc = read_vec3(j.at("body_acc"), "body_acc");
    q.rho = read_f64(j.at("rho"), "rho");
    q.p = read_f64(j.at("p"), "p");
    q.mass = read_f64(j.at("mass"), "mass");
    const uint64_t m = read_uint(j.at("material"), "material");
    // tags beyond the table select material 0 in the reference (world.rs:319-325)
    q.material = m > 0xffffffffull ? 0xffffffffull : m;
    return q;
}

}  // namespace

std::string save_world_json(World& w) {
    Json root = Json::object();
    root.add("version", Json::uint(1));
    root.add("time", f64(w.time()));
    Json subs = Json::array();
    for (size_t i = 0; i < w.subsystem_count(); ++i) {
        auto* f = dynamic_cast<FluidSubsystem*>(w.get(i));
        if (!f) continue;  // archive.rs:189-211: unknown kinds are skipped
        FluidWorld& fw = f->world();
        Json world = Json::object();
        world.add("params", params_json(fw.params()));
        Json ps = Json::array();
        for (const Particle& q : fw.particles()) ps.a.push_back(particle_json(q));
        world.add("particles", ps);
        Json fs = Json::object();
        fs.add("world", world);
        fs.add("thermal_expansion", f64(f->thermal_expansion));
        fs.add("heat_gain", f64(f->heat_gain));
        fs.add("t_ref", f64(f->t_ref));
        fs.add("drag", f64(f->drag));
        fs.add("friction", f64(f->friction));
        Json tagged = Json::object();
        tagged.add("Fluid", fs);
        subs.a.push_back(tagged);
    }
    root.add("subsystems", subs);
    std::string out;
    write(root, out, 0);
    return out;
}

std::unique_ptr<World> load_world_json(const std::string& json) {
    const Json root = Parser(json).parse();
    if (root.kind != Json::Obj) bad("top level is not an object");
    read_uint(root.at("version"), "version");
    const double t = read_f64(root.at("time"), "time");
    const Json& subs = root.at("subsystems");
    if (subs.kind != Json::Arr) bad("subsystems is not an array");
    auto w = std::make_unique<World>();
    for (const Json& s : subs.a) {
        if (s.kind != Json::Obj || s.o.size() != 1) bad("a subsystem entry is not a tagged object");
        const std::string& tag = s.o[0].first;
        if (tag != "Fluid") bad("subsystem kind \"" + tag + "\" is not hosted by this library (fluid only)");
        const Json& fs = s.o[0].second;
        const Json& world = fs.at("world");
        auto fw = std::make_unique<FluidWorld>(params_from(world.at("params")));
        const Json& ps = world.at("particles");
        if (ps.kind != Json::Arr) bad("particles is not an array");
        std::vector<Particle> parts;
        parts.reserve(ps.a.size());
        for (const Json& q : ps.a) parts.push_back(particle_from(q));
        fw->set_particles(parts);
        auto sub = std::make_unique<FluidSubsystem>(std::move(fw));
        sub->thermal_expansion = read_f64(fs.at("thermal_expansion"), "thermal_expansion");
        sub->heat_gain = read_f64(fs.at("heat_gain"), "heat_gain");
        sub->t_ref = read_f64(fs.at("t_ref"), "t_ref");
        sub->drag = read_f64(fs.at("drag"), "drag");
        sub->friction = read_f64(fs.at("friction"), "friction");
        w->add_subsystem(std::move(sub));
    }
    w->set_time(t);
    return w;
}

}  // namespace phy
