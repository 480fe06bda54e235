// Byte-compatible Structure checkpoint codec (SWST around SWBC and SWSC).
//
// Formats (little-endian throughout, FNV-1a 64 checksums over payloads):
//   SWST envelope (structure_checkpoint_persistence.cpp:31-37, 301-310):
//     "SWST", u16 protocol 1, u16 0, u64 payload size, u64 checksum, payload
//   SWST payload (writePayload, :229-257): u32 1, u32 0, u32 version 2, u32 0,
//     u64 definition fingerprint, u64 accepted steps, f64 time, u64 N,
//     N x (position, previous, velocity), u64 N, N x pending force,
//     last applied force, u64 body size, SWBC bytes, u8 suspension flag
//     [u64 size, SWSC bytes]
//   SWBC envelope (soft_body_checkpoint_persistence.cpp:31-44, 895-909), with
//     payload writeMutableState (:645-703): u32 1, u32 0, u64 topology
//     fingerprint, six counts, N x (position, previous, velocity, force),
//     T x pressure, C x multiplier, M x (multiplier, resultant, residual),
//     D x multiplier, the (empty) contact multipliers and records, the
//     contact diagnostics block (:430-453) and four (empty) audit key lists.
//   SWSC envelope and payload (suspension_checkpoint_persistence.cpp:575-644,
//     752-800): see the suspension section below.
// A body built by Structure carries no contact state: its contact
// diagnostics are the default block (not registered, solve succeeded, every
// count and vector zero, minimum gap +infinity).
#include "checkpoint_codec.hpp"

#include "schedule.hpp"

#include <algorithm>
#include <bit>
#include <cstdio>
#include <map>
#include <set>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace swx {

void membraneCompliance(const double* mat, double stiffnessScale, double out[6]);

namespace {

constexpr std::uint64_t kFnvBasis = 14695981039346656037ULL;
constexpr std::uint64_t kFnvPrime = 1099511628211ULL;
constexpr std::uint16_t kStructureProtocol = 1;
constexpr std::uint16_t kBodyProtocol = 1;
constexpr std::uint32_t kStateVersion = 1;
constexpr std::uint32_t kCheckpointVersion = 2;
constexpr std::size_t kEnvelopeBytes = 24;
constexpr std::size_t kContactDiagnosticsBytes = 307;

std::uint64_t fnv(const std::uint8_t* data, std::size_t size) {
    std::uint64_t h = kFnvBasis;
    for (std::size_t i = 0; i < size; ++i) {
        h ^= data[i];
        h *= kFnvPrime;
    }
    return h;
}

class Out {
public:
    explicit Out(std::vector<std::uint8_t>& bytes) : bytes_(bytes) {}
    void u8(std::uint8_t v) { bytes_.push_back(v); }
    void u16(std::uint16_t v) { integer(v, 2); }
    void u32(std::uint32_t v) { integer(v, 4); }
    void u64(std::uint64_t v) { integer(v, 8); }
    void raw(double v) { u64(std::bit_cast<std::uint64_t>(v)); }
    void f64(double v, const char* what) {
        if (!std::isfinite(v)) throw std::invalid_argument(std::string(what) + " is non-finite");
        raw(v);
    }
    void v3(const double* v, const char* what) {
        for (int k = 0; k < 3; ++k) f64(v[k], what);
    }
    void bytes(const std::vector<std::uint8_t>& b) { bytes_.insert(bytes_.end(), b.begin(), b.end()); }

private:
    void integer(std::uint64_t v, int n) {
        for (int i = 0; i < n; ++i) {
            bytes_.push_back(static_cast<std::uint8_t>(v & 0xffu));
            v >>= 8;
        }
    }
    std::vector<std::uint8_t>& bytes_;
};

class In {
public:
    In(const std::uint8_t* data, std::size_t size, const char* section) : data_(data), size_(size), section_(section) {}
    std::uint64_t integer(int n) {
        need(static_cast<std::size_t>(n));
        std::uint64_t v = 0;
        for (int i = 0; i < n; ++i) v |= static_cast<std::uint64_t>(data_[at_ + i]) << (8 * i);
        at_ += static_cast<std::size_t>(n);
        return v;
    }
    std::uint8_t u8() { return static_cast<std::uint8_t>(integer(1)); }
    std::uint16_t u16() { return static_cast<std::uint16_t>(integer(2)); }
    std::uint32_t u32() { return static_cast<std::uint32_t>(integer(4)); }
    std::uint64_t u64() { return integer(8); }
    double raw() { return std::bit_cast<double>(u64()); }
    double f64() {
        const double v = raw();
        if (!std::isfinite(v)) fail("holds a non-finite value");
        return v;
    }
    void v3(double* out) {
        for (int k = 0; k < 3; ++k) out[k] = f64();
    }
    // A count that must equal the topology's.
    void exactCount(std::size_t expected, const char* what) {
        if (u64() != expected) throw std::invalid_argument(std::string(section_) + " " + what + " does not match topology");
    }
    const std::uint8_t* take(std::size_t n) {
        need(n);
        const std::uint8_t* p = data_ + at_;
        at_ += n;
        return p;
    }
    bool atEnd() const { return at_ == size_; }
    [[noreturn]] void fail(const std::string& why) const {
        throw std::invalid_argument(std::string(section_) + " " + why);
    }

private:
    void need(std::size_t n) {
        if (n > size_ - at_) fail("is truncated");
    }
    const std::uint8_t* data_;
    std::size_t size_;
    std::size_t at_ = 0;
    const char* section_;
};

// CheckpointTopologyHasher (soft_body.cpp:40-63): FNV-1a over the
// little-endian bytes of each 64-bit value.
class TopologyHash {
public:
    void add(std::uint64_t v) {
        for (int i = 0; i < 8; ++i) {
            h_ ^= static_cast<std::uint8_t>(v & 0xffu);
            h_ *= kFnvPrime;
            v >>= 8;
        }
    }
    void add(double v) { add(std::bit_cast<std::uint64_t>(v)); }
    std::uint64_t value() const { return h_; }

private:
    std::uint64_t h_ = kFnvBasis;
};

// softwing::ConstraintKind of a Structure constraint kind
// (soft_body.h:90-101; structure.cpp:568-595).
std::uint64_t bodyConstraintKind(std::uint8_t kind) {
    switch (kind) {
    case SW_CONSTRAINT_DISTANCE: return 0;
    case SW_CONSTRAINT_CABLE: return 1;
    case SW_CONSTRAINT_SUSPENSION_TIE: return 2;
    case SW_CONSTRAINT_SEAM_STITCH: return 3;
    default: throw std::invalid_argument("Unknown structure constraint kind");
    }
}

void envelope(std::vector<std::uint8_t>& out, const char magic[4], std::uint16_t protocol,
              const std::vector<std::uint8_t>& payload) {
    Out w(out);
    for (int i = 0; i < 4; ++i) w.u8(static_cast<std::uint8_t>(magic[i]));
    w.u16(protocol);
    w.u16(0);
    w.u64(payload.size());
    w.u64(fnv(payload.data(), payload.size()));
    w.bytes(payload);
}

// The payload of an envelope, after checking magic, protocol and checksum
// (deserializeStructureCheckpoint :378-429 and its SoftBody counterpart).
std::pair<const std::uint8_t*, std::size_t> openEnvelope(const std::uint8_t* bytes, std::size_t size,
                                                         const char magic[4], std::uint16_t protocol,
                                                         const char* section) {
    In r(bytes, size, section);
    if (size < kEnvelopeBytes) r.fail("envelope is truncated");
    for (int i = 0; i < 4; ++i)
        if (r.u8() != static_cast<std::uint8_t>(magic[i])) r.fail("magic is invalid");
    if (r.u16() != protocol) r.fail("protocol version is unsupported");
    if (r.u16() != 0) r.fail("reserved bits are nonzero");
    const std::uint64_t length = r.u64();
    const std::uint64_t checksum = r.u64();
    if (length > size - kEnvelopeBytes) r.fail("payload is truncated");
    if (length < size - kEnvelopeBytes) r.fail("has trailing data");
    const std::uint8_t* payload = r.take(static_cast<std::size_t>(length));
    if (fnv(payload, static_cast<std::size_t>(length)) != checksum) r.fail("checksum does not match");
    return {payload, static_cast<std::size_t>(length)};
}

std::vector<std::uint8_t> encodeBody(const Definition& d, const CheckpointImage& image) {
    const DeviceWing::State& s = image.state;
    std::vector<std::uint8_t> payload;
    Out w(payload);
    w.u32(kStateVersion);
    w.u32(0);
    w.u64(softBodyTopologyFingerprint(d));
    w.u64(d.nodeCount());
    w.u64(d.triangleCount());
    w.u64(d.constraintCount());
    w.u64(d.membraneCount());
    w.u64(d.dihedralCount());
    w.u64(0);  // contact pairs
    const char* bad = "SoftBody checkpoint contains invalid mutable state:";
    const double zero[3] = {0.0, 0.0, 0.0};
    for (std::size_t i = 0; i < d.nodeCount(); ++i) {
        w.v3(&s.position[3 * i], bad);
        w.v3(&s.previous[3 * i], bad);
        w.v3(&s.velocity[3 * i], bad);
        // Structure::step applies the pending forces inside the step and the
        // body holds none between steps.
        w.v3(zero, bad);
    }
    for (std::size_t t = 0; t < d.triangleCount(); ++t) w.f64(s.pressure[t], bad);
    for (std::size_t c = 0; c < d.constraintCount(); ++c) w.f64(s.constraintLambda[c], bad);
    for (std::size_t m = 0; m < d.membraneCount(); ++m) {
        w.v3(&s.membraneLambda[3 * m], bad);
        w.v3(&s.membraneResultant[3 * m], bad);
        w.f64(s.membraneResidual[m], bad);
    }
    for (std::size_t h = 0; h < d.dihedralCount(); ++h) w.f64(s.dihedralLambda[h], bad);
    w.u64(0);  // contact multipliers
    w.u64(0);  // contact records
    // Default ContactDiagnostics (writeDiagnostics, :430-453).
    w.u8(0);  // registered
    w.u8(1);  // solveSucceeded
    for (int k = 0; k < 6; ++k) w.u64(0);  // possible .. indeterminate counts
    w.raw(std::numeric_limits<double>::infinity());  // minimumGap
    for (int k = 0; k < 3; ++k) w.raw(0.0);  // penetration, normal and friction residuals
    for (int k = 0; k < 18; ++k) w.raw(0.0);  // six impulse and moment vectors
    w.raw(0.0);  // frictionWork
    w.u32(0);    // maximumCcdIterations
    w.u32(0);    // maximumIntervalSubdivisions
    w.u8(0);     // hasFailure
    w.u64(0);    // failure key: pair, kind, 7 padding bytes, six primitives
    for (int k = 0; k < 8; ++k) w.u8(0);
    for (int k = 0; k < 6; ++k) w.u64(0);
    for (int k = 0; k < 4; ++k) w.u64(0);  // audit key lists
    std::vector<std::uint8_t> out;
    envelope(out, "SWBC", kBodyProtocol, payload);
    return out;
}

void decodeBody(const Definition& d, const std::uint8_t* bytes, std::size_t size, DeviceWing::State& s) {
    const auto [payload, length] = openEnvelope(bytes, size, "SWBC", kBodyProtocol, "SoftBody checkpoint");
    In r(payload, length, "SoftBody checkpoint");
    if (r.u32() != kStateVersion || r.u32() != 0) r.fail("contains invalid mutable state");
    if (r.u64() != softBodyTopologyFingerprint(d)) r.fail("topology does not match");
    r.exactCount(d.nodeCount(), "node count");
    r.exactCount(d.triangleCount(), "triangle count");
    r.exactCount(d.constraintCount(), "constraint count");
    r.exactCount(d.membraneCount(), "membrane count");
    r.exactCount(d.dihedralCount(), "dihedral count");
    r.exactCount(0, "contact pair count");
    for (std::size_t i = 0; i < d.nodeCount(); ++i) {
        r.v3(&s.position[3 * i]);
        r.v3(&s.previous[3 * i]);
        r.v3(&s.velocity[3 * i]);
        double force[3];
        r.v3(force);  // replaced by the pending forces at the next step
    }
    for (std::size_t t = 0; t < d.triangleCount(); ++t) s.pressure[t] = r.f64();
    for (std::size_t c = 0; c < d.constraintCount(); ++c) s.constraintLambda[c] = r.f64();
    for (std::size_t m = 0; m < d.membraneCount(); ++m) {
        r.v3(&s.membraneLambda[3 * m]);
        r.v3(&s.membraneResultant[3 * m]);
        s.membraneResidual[m] = r.f64();
    }
    for (std::size_t h = 0; h < d.dihedralCount(); ++h) s.dihedralLambda[h] = r.f64();
    // With no contact pair, every contact key is invalid (validKey), so the
    // multiplier and record lists are empty and the diagnostics carry no
    // failure.
    if (r.u64() != 0 || r.u64() != 0) r.fail("contains invalid mutable state");
    const std::uint8_t* diag = r.take(kContactDiagnosticsBytes);
    if (diag[0] > 1 || diag[1] > 1 || diag[kContactDiagnosticsBytes - 65] != 0)
        r.fail("contains invalid mutable state");
    for (int k = 0; k < 4; ++k)
        if (r.u64() != 0) r.fail("contains invalid mutable state");
    if (!r.atEnd()) r.fail("has trailing data");
}

// ---- The suspension section (SWSC) ------------------------------------
//
// SuspensionSystem::checkpoint (suspension.cpp:1725-1752) of the suspension
// that Structure's buildSuspension (structure.cpp:459-548) makes: identifiers
// "attachment:<id>", "junction:<id>", "harness:<id>", "segment:<id>", one
// group "role:<role>" per segment, every collection in string-id order,
// free ground mode, no controls. Its diagnostics are those finishSubstep
// (suspension.cpp:2481-2610) leaves after the last substep, restated here
// from the committed state with the reference's operations in its order.

constexpr std::uint16_t kSuspensionProtocol = 1;
constexpr std::uint64_t kSuspensionSchema = 1;
const char* const kProvenance = "simwing-structure-resolved";
constexpr std::uint64_t kPhaseDiagnosticsEnum = 9;  // SuspensionPhase::Diagnostics

// softwing::Vec3 arithmetic (vec3.h): component-wise, left to right.
struct V3 {
    double x = 0.0, y = 0.0, z = 0.0;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
V3 operator*(double s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
V3& operator+=(V3& a, V3 b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    return a;
}
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double length(V3 a) { return std::sqrt(dot(a, a)); }
V3 normalized(V3 a) {
    const double m = length(a);
    return m > 0.0 ? a / m : V3{};
}

// Mat3 and quaternion helpers (suspension.cpp:600-660).
struct M3 {
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double operator()(int r, int c) const { return a[3 * r + c]; }
};
V3 operator*(const M3& m, V3 v) {
    return {m(0, 0) * v.x + m(0, 1) * v.y + m(0, 2) * v.z, m(1, 0) * v.x + m(1, 1) * v.y + m(1, 2) * v.z,
            m(2, 0) * v.x + m(2, 1) * v.y + m(2, 2) * v.z};
}
M3 operator*(const M3& x, const M3& y) {
    M3 r;
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col)
            for (int k = 0; k < 3; ++k) r.a[3 * row + col] += x(row, k) * y(k, col);
    return r;
}
M3 transpose(const M3& m) {
    M3 r;
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) r.a[3 * row + col] = m(col, row);
    return r;
}
struct Q4 {
    double w = 1.0, x = 0.0, y = 0.0, z = 0.0;
};
Q4 normalizedCanonical(const Q4& q) {
    const double n2 = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
    if (!std::isfinite(q.w) || !std::isfinite(q.x) || !std::isfinite(q.y) || !std::isfinite(q.z) || !(n2 > 1.0e-28))
        throw std::invalid_argument("payload orientation is non-finite or near zero");
    const double inv = 1.0 / std::sqrt(n2);
    Q4 r{q.w * inv, q.x * inv, q.y * inv, q.z * inv};
    const bool flip = r.w < 0.0 || (r.w == 0.0 && (r.x < 0.0 || (r.x == 0.0 && (r.y < 0.0 || (r.y == 0.0 && r.z < 0.0)))));
    if (flip) r = {-r.w, -r.x, -r.y, -r.z};
    return r;
}
M3 rotationMatrix(const Q4& orientation) {
    const Q4 q = normalizedCanonical(orientation);
    const double xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const double xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    const double wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    return M3{{1.0 - 2.0 * (yy + zz), 2.0 * (xy - wz), 2.0 * (xz + wy), 2.0 * (xy + wz), 1.0 - 2.0 * (xx + zz),
               2.0 * (yz - wx), 2.0 * (xz - wy), 2.0 * (yz + wx), 1.0 - 2.0 * (xx + yy)}};
}

struct PayloadState {
    V3 com;
    Q4 q;
    V3 v, w;
};

// The runtime suspension of a definition: string-id ordered collections,
// runtime endpoints and attachment paths (suspension.cpp:1371-1512).
struct RuntimeSuspension {
    struct Endpoint {
        std::uint64_t kind = 0;  // softwing::SuspensionEndpointKind
        std::string id;
        long node = -1;  // body node, or -1
        long hang = -1;  // runtime hang-point index, or -1
    };
    struct Segment {
        std::string id;
        std::uint32_t def = 0;  // definition index
        Endpoint from, to;
        std::vector<std::string> paths;
        std::string group;
    };
    std::vector<std::pair<std::string, std::uint64_t>> attachments;  // id, node
    std::vector<std::pair<std::string, std::uint64_t>> junctions;    // id, node
    std::vector<std::pair<std::string, std::uint32_t>> hangs;        // id, definition harness index
    std::vector<Segment> segments;
};

std::string endpointName(std::uint8_t kind) {
    switch (kind) {
    case SW_ENDPOINT_ATTACHMENT: return "attachment:";
    case SW_ENDPOINT_JUNCTION: return "junction:";
    case SW_ENDPOINT_HARNESS: return "harness:";
    default: throw std::invalid_argument("Unknown structure suspension endpoint kind");
    }
}

RuntimeSuspension runtimeSuspension(const Definition& d) {
    const SuspensionDef& s = *d.suspension;
    RuntimeSuspension r;
    for (std::size_t k = 0; k < s.attachmentIds.size(); ++k)
        r.attachments.emplace_back("attachment:" + std::to_string(s.attachmentIds[k]), s.attachmentNodes[k]);
    for (std::size_t k = 0; k < s.junctionIds.size(); ++k)
        r.junctions.emplace_back("junction:" + std::to_string(s.junctionIds[k]), s.junctionNodes[k]);
    for (std::size_t k = 0; k < s.harnessIds.size(); ++k)
        r.hangs.emplace_back("harness:" + std::to_string(s.harnessIds[k]), static_cast<std::uint32_t>(k));
    std::sort(r.attachments.begin(), r.attachments.end());
    std::sort(r.junctions.begin(), r.junctions.end());
    std::sort(r.hangs.begin(), r.hangs.end());
    std::map<std::string, RuntimeSuspension::Endpoint> endpoints;
    for (const auto& [id, node] : r.attachments) endpoints[id] = {0, id, static_cast<long>(node), -1};
    for (const auto& [id, node] : r.junctions) endpoints[id] = {1, id, static_cast<long>(node), -1};
    for (std::size_t h = 0; h < r.hangs.size(); ++h) endpoints[r.hangs[h].first] = {2, r.hangs[h].first, -1, static_cast<long>(h)};
    std::vector<std::pair<std::string, std::uint32_t>> order;
    for (std::size_t k = 0; k < s.segmentIds.size(); ++k)
        order.emplace_back("segment:" + std::to_string(s.segmentIds[k]), static_cast<std::uint32_t>(k));
    std::sort(order.begin(), order.end());
    for (const auto& [id, k] : order) {
        RuntimeSuspension::Segment g;
        g.id = id;
        g.def = k;
        g.from = endpoints.at(endpointName(s.segmentFromKind[k]) + std::to_string(s.segmentFromId[k]));
        g.to = endpoints.at(endpointName(s.segmentToKind[k]) + std::to_string(s.segmentToId[k]));
        g.group = "role:" + std::to_string(s.segmentRole[k]);
        r.segments.push_back(std::move(g));
    }
    // Attachment paths (suspension.cpp:1462-1496): the attachments whose
    // walk along first-matching outgoing segments reaches the segment.
    for (auto& g : r.segments) {
        for (const auto& [attachment, node] : r.attachments) {
            (void)node;
            std::pair<std::uint64_t, std::string> at{0, attachment};
            std::set<std::pair<std::uint64_t, std::string>> visited;
            while (visited.insert(at).second) {
                const auto found = std::find_if(r.segments.begin(), r.segments.end(), [&](const auto& c) {
                    return c.from.kind == at.first && c.from.id == at.second;
                });
                if (found == r.segments.end()) break;
                if (found->id == g.id) {
                    g.paths.push_back(attachment);
                    break;
                }
                at = {found->to.kind, found->to.id};
            }
        }
    }
    return r;
}

// std::quoted in the classic locale.
std::string quoted(const std::string& text) {
    std::string out = "\"";
    for (char c : text) {
        if (c == '"' || c == '\\') out += '\\';
        out += c;
    }
    return out + "\"";
}
// A double as `ostream << std::scientific << std::setprecision(17)` prints it.
std::string sci(double v) {
    char buffer[64];
    std::snprintf(buffer, sizeof(buffer), "%.17e", v);
    return buffer;
}

// serializeSuspensionDefinition (suspension.cpp:2642-2760) of the
// definition buildSuspension makes.
std::string suspensionDefinitionText(const Definition& d, const RuntimeSuspension& r) {
    const SuspensionDef& s = *d.suspension;
    std::string o = "SOFTWING_SUSPENSION 1\n";
    o += "IDENTIFIER \"simwing-structure\"\n";
    o += "DESCRIPTION \"Resolved scene-v2 suspension and rigid pilot\"\n";
    o += "FRAME \"SI right-handed +X chord +Y span +Z up\"\n";
    o += "PROVENANCE 1\n";
    o += "PROV " + quoted(kProvenance) + " \"authoritative scene-v2 stable topology\"\n";
    o += "ATTACHMENTS " + std::to_string(r.attachments.size()) + "\n";
    for (const auto& [id, node] : r.attachments)
        o += "ATTACH " + quoted(id) + " " + quoted("resolved-node:" + std::to_string(node)) + " " + sci(0.0) + " " +
             sci(0.0) + " centre \"\" " + quoted(kProvenance) + "\n";
    o += "JUNCTIONS " + std::to_string(r.junctions.size()) + "\n";
    for (const auto& [id, node] : r.junctions) {
        const double inverseMass = 1.0 / d.nodeMass[node];
        o += "JUNCTION " + quoted(id) + " " + sci(d.nodePosition[3 * node]) + " " + sci(d.nodePosition[3 * node + 1]) +
             " " + sci(d.nodePosition[3 * node + 2]) + " " + sci(1.0 / inverseMass) + " centre " + quoted(kProvenance) +
             "\n";
    }
    o += "PAYLOAD " + sci(s.pilotMass) + " " + sci(s.comLocal[0]) + " " + sci(s.comLocal[1]) + " " + sci(s.comLocal[2]);
    const double inertia[9] = {s.inertia[0], 0.0, 0.0, 0.0, s.inertia[1], 0.0, 0.0, 0.0, s.inertia[2]};
    for (double v : inertia) o += " " + sci(v);
    const Q4 q0 = normalizedCanonical({s.orientation[0], s.orientation[1], s.orientation[2], s.orientation[3]});
    for (double v : {s.comWorld[0], s.comWorld[1], s.comWorld[2], q0.w, q0.x, q0.y, q0.z, s.velocity[0], s.velocity[1],
                     s.velocity[2], s.angularVelocity[0], s.angularVelocity[1], s.angularVelocity[2]})
        o += " " + sci(v);
    o += " " + quoted(kProvenance) + "\n";
    o += "HANG_POINTS " + std::to_string(r.hangs.size()) + "\n";
    for (const auto& [id, k] : r.hangs)
        o += "HANG " + quoted(id) + " " + sci(s.harnessLocal[3 * k]) + " " + sci(s.harnessLocal[3 * k + 1]) + " " +
             sci(s.harnessLocal[3 * k + 2]) + " centre " + quoted(kProvenance) + "\n";
    o += "SUPPORT_POINTS 0\n";
    o += "SEGMENTS " + std::to_string(r.segments.size()) + "\n";
    static const char* const kinds[3] = {"attachment", "junction", "hang-point"};
    for (const auto& g : r.segments)
        o += "SEGMENT " + quoted(g.id) + " " + kinds[g.from.kind] + " " + quoted(g.from.id) + " " + kinds[g.to.kind] +
             " " + quoted(g.to.id) + " " + sci(s.segmentRest[g.def]) + " " + sci(s.segmentStiffness[g.def]) + " " +
             sci(s.segmentDamping[g.def]) + " centre 1 " + quoted(g.group) + " " + quoted(kProvenance) + "\n";
    o += "CONTROLS 0\n";
    o += "SOLVER " + std::to_string(s.solverIterations) + " " + sci(s.attachmentTolerance) + " " +
         sci(s.minimumLineLength) + " " + sci(s.maximumLineResidual) + " " + sci(s.maximumControlWork) + "\n";
    o += "GROUND free " + sci(0.0) + " " + sci(0.0) + " " + sci(1.0) + " " + sci(0.0) + " " + sci(0.0) + " " +
         sci(0.02) + "\n";
    o += "END\n";
    return o;
}

// CheckpointHasher (suspension.cpp:49-89).
class Hasher {
public:
    void byte(std::uint8_t v) {
        h_ ^= v;
        h_ *= kFnvPrime;
    }
    void boolean(bool v) { byte(v ? 1 : 0); }
    void u64(std::uint64_t v) {
        for (int i = 0; i < 8; ++i) {
            byte(static_cast<std::uint8_t>(v & 0xffu));
            v >>= 8;
        }
    }
    void count(std::uint64_t v) { u64(v); }
    void real(double v) { u64(std::bit_cast<std::uint64_t>(v)); }
    void text(const std::string& t) {
        count(t.size());
        for (unsigned char c : t) byte(c);
    }
    void vec(V3 v) {
        real(v.x);
        real(v.y);
        real(v.z);
    }
    void enumeration(std::uint64_t v) { u64(v); }  // 64-bit in the hash
    std::uint64_t value() const { return h_; }

private:
    std::uint64_t h_ = kFnvBasis;
};

// The persistence writer (suspension_checkpoint_persistence.cpp): the same
// stream as the hasher except that enumerations are one byte and values
// must be finite.
class Writer {
public:
    explicit Writer(std::vector<std::uint8_t>& bytes) : out_(bytes) {}
    void boolean(bool v) { out_.u8(v ? 1 : 0); }
    void count(std::uint64_t v) { out_.u64(v); }
    void u64(std::uint64_t v) { out_.u64(v); }
    void real(double v) { out_.f64(v, "suspension checkpoint contains invalid state:"); }
    void text(const std::string& t) {
        out_.u64(t.size());
        for (unsigned char c : t) out_.u8(c);
    }
    void vec(V3 v) {
        real(v.x);
        real(v.y);
        real(v.z);
    }
    void enumeration(std::uint64_t v) { out_.u8(static_cast<std::uint8_t>(v)); }
    Out& raw() { return out_; }

private:
    Out out_;
};

struct SegmentDiag {
    std::string id;
    RuntimeSuspension::Endpoint from, to;  // kind and id only
    std::vector<std::string> paths, groups;
    double length = 0, commanded = 0, stretch = 0, strain = 0;
    bool taut = false;
    double multiplier = 0, tension = 0, residual = 0, energy = 0;
    V3 dampingImpulse;
    double dampingWork = 0, controlWork = 0;
    V3 fromImpulse, toImpulse, fromMoment, toMoment;
};
struct Wrench {
    V3 force, moment;
};
struct PayloadDiag {
    PayloadState state;  // default: zero com, identity orientation
    V3 linear, angular;
    double translational = 0, rotational = 0, gravitational = 0;
    Wrench applied, line, ground, anchor;
};
struct SuspensionDiag {
    bool registered = true, converged = false, allSlack = true, anchored = false, grounded = false, failed = false;
    std::uint64_t attachments = 0, junctions = 0, segments = 0, taut = 0, slack = 0, iterations = 0;
    double maxResidual = 0, maxPenetration = 0, energy = 0, dampingWork = 0, controlWork = 0;
    V3 netImpulse, netMoment;
    Wrench fixed, canopy, ground;
    std::vector<std::pair<std::string, V3>> attachmentLoads, groupLoads;
    std::string provenance = kProvenance;
    std::uint64_t phase = kPhaseDiagnosticsEnum;
    std::string failureEntity;
};

struct SuspensionContent {
    std::uint64_t topology = 0;
    PayloadState payload, previous;
    std::vector<V3> hangs;  // commanded = base hang points, runtime order
    std::vector<std::string> segmentIds;
    std::vector<double> commanded, lambda;
    Wrench applied;
    V3 gravity;
    std::vector<double> controlWork;  // pending, per segment
    std::vector<SegmentDiag> segments;
    PayloadDiag payloadDiag;
    SuspensionDiag diag, committed;
};

template <class Sink>
void emitPayload(Sink& k, const PayloadState& p) {
    k.vec(p.com);
    k.real(p.q.w);
    k.real(p.q.x);
    k.real(p.q.y);
    k.real(p.q.z);
    k.vec(p.v);
    k.vec(p.w);
}
template <class Sink>
void emitWrench(Sink& k, const Wrench& w) {
    k.vec(w.force);
    k.vec(w.moment);
}
template <class Sink>
void emitDiagnostics(Sink& k, const SuspensionDiag& d) {
    k.boolean(d.registered);
    k.boolean(d.converged);
    k.boolean(d.allSlack);
    k.boolean(d.anchored);
    k.boolean(d.grounded);
    k.boolean(d.failed);
    k.count(d.attachments);
    k.count(d.junctions);
    k.count(d.segments);
    k.count(d.taut);
    k.count(d.slack);
    k.count(d.iterations);
    k.real(d.maxResidual);
    k.real(d.maxPenetration);
    k.real(d.energy);
    k.real(d.dampingWork);
    k.real(d.controlWork);
    k.vec(d.netImpulse);
    k.vec(d.netMoment);
    emitWrench(k, d.fixed);
    emitWrench(k, d.canopy);
    emitWrench(k, d.ground);
    for (const auto* loads : {&d.attachmentLoads, &d.groupLoads}) {
        k.count(loads->size());
        for (const auto& [id, load] : *loads) {
            k.text(id);
            k.vec(load);
        }
    }
    k.text(d.provenance);
    k.enumeration(d.phase);
    k.text(d.failureEntity);
}

// Everything after the state fingerprint (writeState,
// suspension_checkpoint_persistence.cpp:585-644; checkpointStateFingerprintOf,
// suspension.cpp:198-235, hashes the same sequence after the topology).
template <class Sink>
void emitState(Sink& k, const SuspensionContent& c) {
    emitPayload(k, c.payload);
    emitPayload(k, c.previous);
    k.count(c.hangs.size());
    for (const V3& h : c.hangs) k.vec(h);
    k.count(c.segmentIds.size());
    for (std::size_t i = 0; i < c.segmentIds.size(); ++i) {
        k.text(c.segmentIds[i]);
        k.real(c.commanded[i]);
        k.real(c.lambda[i]);
    }
    k.count(0);  // controls
    k.count(0);  // ground multipliers (free mode: no support points)
    emitWrench(k, c.applied);
    k.vec(c.gravity);
    k.count(c.controlWork.size());
    for (double w : c.controlWork) k.real(w);
    k.count(c.segments.size());
    for (const SegmentDiag& g : c.segments) {
        k.text(g.id);
        k.enumeration(g.from.kind);
        k.text(g.from.id);
        k.enumeration(g.to.kind);
        k.text(g.to.id);
        for (const auto* list : {&g.paths, &g.groups}) {
            k.count(list->size());
            for (const std::string& t : *list) k.text(t);
        }
        k.real(g.length);
        k.real(g.commanded);
        k.real(g.stretch);
        k.real(g.strain);
        k.boolean(g.taut);
        k.real(g.multiplier);
        k.real(g.tension);
        k.real(g.residual);
        k.real(g.energy);
        k.vec(g.dampingImpulse);
        k.real(g.dampingWork);
        k.real(g.controlWork);
        k.vec(g.fromImpulse);
        k.vec(g.toImpulse);
        k.vec(g.fromMoment);
        k.vec(g.toMoment);
    }
    const PayloadDiag& p = c.payloadDiag;
    emitPayload(k, p.state);
    k.vec(p.linear);
    k.vec(p.angular);
    k.real(p.translational);
    k.real(p.rotational);
    k.real(p.gravitational);
    emitWrench(k, p.applied);
    emitWrench(k, p.line);
    emitWrench(k, p.ground);
    emitWrench(k, p.anchor);
    emitDiagnostics(k, c.diag);
    emitDiagnostics(k, c.committed);
}

// SuspensionSystem::checkpointTopologyFingerprint (suspension.cpp:1597-1723).
std::uint64_t suspensionTopologyFingerprint(const Definition& d, const RuntimeSuspension& r) {
    Hasher h;
    h.text("SOFTWING_SUSPENSION_CHECKPOINT_TOPOLOGY");
    h.count(kSuspensionSchema);
    h.text(suspensionDefinitionText(d, r));
    h.count(d.nodeCount());
    for (std::size_t i = 0; i < d.nodeCount(); ++i) h.real(d.nodeFixed[i] ? 0.0 : 1.0 / d.nodeMass[i]);
    h.count(d.triangleCount());
    for (std::size_t k = 0; k < 3 * d.triangleCount(); ++k) h.count(d.triangleNodes[k]);
    h.count(d.constraintCount());
    for (std::size_t c = 0; c < d.constraintCount(); ++c) {
        const bool stitch = d.constraintKind[c] == SW_CONSTRAINT_SEAM_STITCH;
        h.count(d.constraintNodes[2 * c]);
        h.count(d.constraintNodes[2 * c + 1]);
        h.real(stitch ? 0.0 : d.constraintRest[c]);
        h.real(stitch ? 0.0 : d.constraintCompliance[c]);
        h.enumeration(bodyConstraintKind(d.constraintKind[c]));
    }
    h.count(d.membraneCount());
    for (std::size_t m = 0; m < d.membraneCount(); ++m) {
        h.count(d.membraneTriangle[m]);
        const double* chart = &d.membraneChart[6 * m];
        for (int k = 0; k < 6; ++k) h.real(chart[k]);
        const double* mat = &d.membraneMaterial[8 * m];
        for (int k = 0; k < 8; ++k) h.real(mat[k]);
        h.enumeration(static_cast<std::uint64_t>(d.membraneRole[m] - 1));
        const double r00 = chart[2] - chart[0], r01 = chart[4] - chart[0];
        const double r10 = chart[3] - chart[1], r11 = chart[5] - chart[1];
        const double det = r00 * r11 - r01 * r10;
        const double inv = 1.0 / det;
        h.real(0.5 * det);
        h.real(r11 * inv);
        h.real(-r01 * inv);
        h.real(-r10 * inv);
        h.real(r00 * inv);
        double compliance[6];
        membraneCompliance(mat, 1.0, compliance);
        for (int k = 0; k < 6; ++k) h.real(compliance[k]);
    }
    h.count(d.dihedralCount());
    for (std::size_t k = 0; k < d.dihedralCount(); ++k) {
        for (int c = 0; c < 4; ++c) h.count(d.dihedralNodes[4 * k + c]);
        h.real(d.dihedralRestAngle[k]);
        h.real(d.dihedralCompliance[k]);
    }
    h.boolean(false);  // interior pressure partitions
    h.count(0);        // contact surfaces
    h.count(0);        // contact lines
    h.count(0);        // contact pairs
    h.count(r.attachments.size());
    for (const auto& [id, node] : r.attachments) {
        h.text(id);
        h.text("resolved-node:" + std::to_string(node));
        h.real(0.0);
        h.real(0.0);
        h.count(node);
        h.vec({d.nodePosition[3 * node], d.nodePosition[3 * node + 1], d.nodePosition[3 * node + 2]});
        h.text(kProvenance);
    }
    h.count(r.junctions.size());
    for (const auto& [id, node] : r.junctions) {
        (void)id;
        h.count(node);
    }
    const SuspensionDef& s = *d.suspension;
    h.count(r.hangs.size());
    for (const auto& [id, k] : r.hangs) {
        (void)id;
        h.vec({s.harnessLocal[3 * k], s.harnessLocal[3 * k + 1], s.harnessLocal[3 * k + 2]});
    }
    const auto endpoint = [&](const RuntimeSuspension::Endpoint& e) {
        h.enumeration(e.kind);
        h.text(e.id);
        h.boolean(e.node >= 0);
        if (e.node >= 0) h.count(static_cast<std::uint64_t>(e.node));
        h.boolean(e.hang >= 0);
        if (e.hang >= 0) h.count(static_cast<std::uint64_t>(e.hang));
    };
    h.count(r.segments.size());
    for (const auto& g : r.segments) {
        h.text(g.id);
        endpoint(g.from);
        endpoint(g.to);
        h.count(g.paths.size());
        for (const std::string& p : g.paths) h.text(p);
    }
    return h.value();
}

// The suspension's checkpoint content for the committed state of `image`.
SuspensionContent suspensionContent(const Definition& d, const CheckpointImage& image) {
    const SuspensionDef& s = *d.suspension;
    const RuntimeSuspension r = runtimeSuspension(d);
    const std::vector<double>& p = image.state.payload;
    SuspensionContent c;
    c.topology = suspensionTopologyFingerprint(d, r);
    c.payload = {{p[0], p[1], p[2]}, {p[3], p[4], p[5], p[6]}, {p[7], p[8], p[9]}, {p[10], p[11], p[12]}};
    c.previous = {{p[13], p[14], p[15]}, {p[16], p[17], p[18], p[19]}, {p[20], p[21], p[22]}, {p[23], p[24], p[25]}};
    for (const auto& [id, k] : r.hangs) {
        (void)id;
        c.hangs.push_back({s.harnessLocal[3 * k], s.harnessLocal[3 * k + 1], s.harnessLocal[3 * k + 2]});
    }
    for (std::size_t i = 0; i < r.segments.size(); ++i) {
        c.segmentIds.push_back(r.segments[i].id);
        c.commanded.push_back(s.segmentRest[r.segments[i].def]);
        c.lambda.push_back(image.state.segmentLambda[i]);
    }
    c.controlWork.assign(r.segments.size(), 0.0);
    SuspensionDiag& diag = c.diag;
    diag.attachments = r.attachments.size();
    diag.junctions = r.junctions.size();
    diag.segments = r.segments.size();
    if (!image.suspensionStepped) {
        c.segments.assign(r.segments.size(), SegmentDiag{});
        c.committed = c.diag;
        return c;
    }
    for (std::size_t i = 0; i < r.segments.size(); ++i)
        if (s.segmentDamping[r.segments[i].def] > 0.0)
            throw std::invalid_argument(
                "the damping impulses of damped suspension lines are not recorded for the checkpoint format");
    const double dt = image.substepDt;
    c.gravity = {image.gravity[0], image.gravity[1], image.gravity[2]};
    diag.iterations = static_cast<std::uint64_t>(image.lineIterations) * r.segments.size();
    diag.maxResidual = image.maxResidual;
    // finishSubstep (suspension.cpp:2481-2610).
    const PayloadState& ps = c.payload;
    const M3 rotation = rotationMatrix(ps.q);
    const M3 inertia{{s.inertia[0], 0.0, 0.0, 0.0, s.inertia[1], 0.0, 0.0, 0.0, s.inertia[2]}};
    const M3 world = rotation * inertia * transpose(rotation);
    PayloadDiag& pd = c.payloadDiag;
    pd.state = ps;
    pd.linear = s.pilotMass * ps.v;
    pd.angular = world * ps.w;
    pd.translational = 0.5 * s.pilotMass * dot(ps.v, ps.v);
    pd.rotational = 0.5 * dot(ps.w, pd.angular);
    pd.gravitational = -s.pilotMass * dot(c.gravity, ps.com);
    const V3 comLocal{s.comLocal[0], s.comLocal[1], s.comLocal[2]};
    const auto position = [&](const RuntimeSuspension::Endpoint& e) -> V3 {
        if (e.node >= 0) {
            const double* x = &image.state.position[3 * static_cast<std::size_t>(e.node)];
            return {x[0], x[1], x[2]};
        }
        return ps.com + rotationMatrix(ps.q) * (c.hangs[static_cast<std::size_t>(e.hang)] - comLocal);
    };
    std::map<std::string, V3> attachmentLoads, groupLoads;
    for (std::size_t i = 0; i < r.segments.size(); ++i) {
        const RuntimeSuspension::Segment& g = r.segments[i];
        SegmentDiag rec;
        rec.id = g.id;
        rec.from = {g.from.kind, g.from.id, -1, -1};
        rec.to = {g.to.kind, g.to.id, -1, -1};
        rec.paths = g.paths;
        rec.groups = {g.group};
        const V3 from = position(g.from), to = position(g.to);
        const V3 direction = normalized(to - from);
        rec.length = length(to - from);
        rec.commanded = c.commanded[i];
        rec.stretch = std::max(0.0, rec.length - rec.commanded);
        rec.strain = rec.stretch / rec.commanded;
        rec.multiplier = c.lambda[i];
        rec.tension = rec.multiplier / (dt * dt);
        rec.taut = rec.multiplier > 0.0;
        const double stiffness = s.segmentStiffness[g.def];
        rec.residual = std::abs(rec.stretch - rec.multiplier / (stiffness * dt * dt));
        rec.energy = 0.5 * stiffness * rec.stretch * rec.stretch;
        rec.controlWork = 0.0;
        rec.fromImpulse = direction * (rec.multiplier / dt);
        rec.toImpulse = -rec.fromImpulse;
        rec.fromMoment = cross(from, rec.fromImpulse);
        rec.toMoment = cross(to, rec.toImpulse);
        diag.netImpulse += rec.fromImpulse + rec.toImpulse;
        diag.netMoment += rec.fromMoment + rec.toMoment;
        diag.energy += rec.energy;
        if (rec.taut) ++diag.taut;
        else ++diag.slack;
        if (g.from.kind == 0) {
            attachmentLoads[g.from.id] += direction * rec.tension;
            if (g.from.node >= 0 && d.nodeFixed[static_cast<std::size_t>(g.from.node)]) {
                const V3 reaction = -direction * rec.tension;
                diag.canopy.force += reaction;
                diag.canopy.moment += cross(from, reaction);
            }
        }
        groupLoads[g.group] += direction * rec.tension;
        if (g.from.hang >= 0 || g.to.hang >= 0) {
            const V3 impulse = g.to.hang >= 0 ? rec.toImpulse : rec.fromImpulse;
            const V3 point = g.to.hang >= 0 ? to : from;
            pd.line.force += impulse / dt;
            pd.line.moment += cross(point - ps.com, impulse / dt);
        }
        c.segments.push_back(std::move(rec));
    }
    diag.allSlack = diag.taut == 0;
    for (const auto& [id, load] : attachmentLoads) diag.attachmentLoads.emplace_back(id, load);
    for (const auto& [id, load] : groupLoads) diag.groupLoads.emplace_back(id, load);
    diag.fixed = diag.canopy;
    diag.fixed.force += pd.anchor.force;
    diag.fixed.moment += pd.anchor.moment + cross(ps.com, pd.anchor.force);
    diag.converged = diag.maxResidual <= s.maximumLineResidual;
    c.committed = diag;
    return c;
}

std::vector<std::uint8_t> encodeSuspension(const Definition& d, const CheckpointImage& image) {
    const SuspensionContent c = suspensionContent(d, image);
    Hasher h;
    h.count(kSuspensionSchema);
    h.u64(c.topology);
    emitState(h, c);
    std::vector<std::uint8_t> payload;
    Writer w(payload);
    w.raw().u32(kStateVersion);
    w.raw().u32(0);
    w.count(kSuspensionSchema);
    w.u64(c.topology);
    w.u64(h.value());
    emitState(w, c);
    std::vector<std::uint8_t> out;
    envelope(out, "SWSC", kSuspensionProtocol, payload);
    return out;
}

// Reads the fields of an SWSC payload after its topology and state
// fingerprints (readState, suspension_checkpoint_persistence.cpp:646-750).
SuspensionContent parseSuspensionState(In& in) {
    SuspensionContent c;
    const auto text = [&]() {
        const std::uint64_t n = in.u64();
        const std::uint8_t* t = in.take(static_cast<std::size_t>(n));
        return std::string(reinterpret_cast<const char*>(t), static_cast<std::size_t>(n));
    };
    const auto vec = [&]() {
        V3 v;
        v.x = in.f64();
        v.y = in.f64();
        v.z = in.f64();
        return v;
    };
    const auto boolean = [&]() {
        const std::uint8_t b = in.u8();
        if (b > 1) in.fail("contains invalid state");
        return b != 0;
    };
    const auto payload = [&]() {
        PayloadState p;
        p.com = vec();
        p.q.w = in.f64();
        p.q.x = in.f64();
        p.q.y = in.f64();
        p.q.z = in.f64();
        p.v = vec();
        p.w = vec();
        return p;
    };
    const auto wrench = [&]() {
        Wrench w;
        w.force = vec();
        w.moment = vec();
        return w;
    };
    const auto diagnostics = [&]() {
        SuspensionDiag d;
        d.registered = boolean();
        d.converged = boolean();
        d.allSlack = boolean();
        d.anchored = boolean();
        d.grounded = boolean();
        d.failed = boolean();
        d.attachments = in.u64();
        d.junctions = in.u64();
        d.segments = in.u64();
        d.taut = in.u64();
        d.slack = in.u64();
        d.iterations = in.u64();
        d.maxResidual = in.f64();
        d.maxPenetration = in.f64();
        d.energy = in.f64();
        d.dampingWork = in.f64();
        d.controlWork = in.f64();
        d.netImpulse = vec();
        d.netMoment = vec();
        d.fixed = wrench();
        d.canopy = wrench();
        d.ground = wrench();
        for (auto* loads : {&d.attachmentLoads, &d.groupLoads}) {
            const std::uint64_t n = in.u64();
            for (std::uint64_t k = 0; k < n; ++k) {
                std::string id = text();
                loads->emplace_back(std::move(id), vec());
            }
        }
        d.provenance = text();
        d.phase = in.u8();
        if (d.phase > kPhaseDiagnosticsEnum) in.fail("contains invalid state");
        d.failureEntity = text();
        return d;
    };
    c.payload = payload();
    c.previous = payload();
    const std::uint64_t hangs = in.u64();
    for (std::uint64_t k = 0; k < hangs; ++k) c.hangs.push_back(vec());
    const std::uint64_t segments = in.u64();
    for (std::uint64_t k = 0; k < segments; ++k) {
        c.segmentIds.push_back(text());
        c.commanded.push_back(in.f64());
        c.lambda.push_back(in.f64());
    }
    if (in.u64() != 0) in.fail("holds controls this build does not run");
    if (in.u64() != 0) in.fail("holds ground multipliers this build does not run");
    c.applied = wrench();
    c.gravity = vec();
    const std::uint64_t work = in.u64();
    for (std::uint64_t k = 0; k < work; ++k) c.controlWork.push_back(in.f64());
    const std::uint64_t records = in.u64();
    for (std::uint64_t k = 0; k < records; ++k) {
        SegmentDiag g;
        g.id = text();
        g.from.kind = in.u8();
        g.from.id = text();
        g.to.kind = in.u8();
        g.to.id = text();
        for (auto* list : {&g.paths, &g.groups}) {
            const std::uint64_t n = in.u64();
            for (std::uint64_t j = 0; j < n; ++j) list->push_back(text());
        }
        g.length = in.f64();
        g.commanded = in.f64();
        g.stretch = in.f64();
        g.strain = in.f64();
        g.taut = boolean();
        g.multiplier = in.f64();
        g.tension = in.f64();
        g.residual = in.f64();
        g.energy = in.f64();
        g.dampingImpulse = vec();
        g.dampingWork = in.f64();
        g.controlWork = in.f64();
        g.fromImpulse = vec();
        g.toImpulse = vec();
        g.fromMoment = vec();
        g.toMoment = vec();
        c.segments.push_back(std::move(g));
    }
    PayloadDiag& p = c.payloadDiag;
    p.state = payload();
    p.linear = vec();
    p.angular = vec();
    p.translational = in.f64();
    p.rotational = in.f64();
    p.gravitational = in.f64();
    p.applied = wrench();
    p.line = wrench();
    p.ground = wrench();
    p.anchor = wrench();
    c.diag = diagnostics();
    c.committed = diagnostics();
    return c;
}

// Reads an SWSC section (deserializeSuspensionCheckpoint,
// suspension_checkpoint_persistence.cpp:802-921) into the image: payload and
// previous payload state, multipliers, and for suspensionState() the
// diagnostics' per-segment length, commanded rest length and tension. The
// section must belong to this definition (topology fingerprint), carry its
// own state fingerprint, and command the base hang points and rest lengths
// (the structure runs no controls).
void decodeSuspension(const Definition& d, const std::uint8_t* bytes, std::size_t size, CheckpointImage& image) {
    const auto [payload, length] =
        openEnvelope(bytes, size, "SWSC", kSuspensionProtocol, "suspension checkpoint");
    const RuntimeSuspension r = runtimeSuspension(d);
    const SuspensionDef& s = *d.suspension;
    In in(payload, length, "suspension checkpoint");
    if (in.u32() != kStateVersion || in.u32() != 0 || in.u64() != kSuspensionSchema)
        in.fail("contains invalid state");
    const std::uint64_t topology = in.u64();
    if (topology != suspensionTopologyFingerprint(d, r)) in.fail("topology does not match");
    const std::uint64_t stateFingerprint = in.u64();
    SuspensionContent c = parseSuspensionState(in);
    if (!in.atEnd()) in.fail("has trailing data");
    c.topology = topology;
    Hasher h;
    h.count(kSuspensionSchema);
    h.u64(c.topology);
    emitState(h, c);
    if (h.value() != stateFingerprint) in.fail("state fingerprint is invalid");
    if (c.hangs.size() != r.hangs.size() || c.segmentIds.size() != r.segments.size() ||
        c.controlWork.size() != r.segments.size() || c.segments.size() != r.segments.size())
        in.fail("counts do not match topology");
    for (std::size_t k = 0; k < r.hangs.size(); ++k) {
        const std::uint32_t def = r.hangs[k].second;
        if (c.hangs[k].x != s.harnessLocal[3 * def] || c.hangs[k].y != s.harnessLocal[3 * def + 1] ||
            c.hangs[k].z != s.harnessLocal[3 * def + 2])
            in.fail("commands hang points this structure does not hold");
    }
    bool stepped = true;
    for (std::size_t i = 0; i < r.segments.size(); ++i) {
        if (c.segmentIds[i] != r.segments[i].id) in.fail("segment topology does not match");
        if (c.commanded[i] != s.segmentRest[r.segments[i].def])
            in.fail("commands rest lengths this structure does not hold");
        stepped = stepped && c.segments[i].id == r.segments[i].id;
    }
    std::vector<double>& p = image.state.payload;
    const PayloadState* states[2] = {&c.payload, &c.previous};
    for (int which = 0; which < 2; ++which) {
        const PayloadState& ps = *states[which];
        double* out = which == 0 ? p.data() : p.data() + 13;
        const double pose[7] = {ps.com.x, ps.com.y, ps.com.z, ps.q.w, ps.q.x, ps.q.y, ps.q.z};
        for (int k = 0; k < 7; ++k) out[k] = pose[k];
        double* velocity = which == 0 ? p.data() + 7 : p.data() + 20;
        const double vw[6] = {ps.v.x, ps.v.y, ps.v.z, ps.w.x, ps.w.y, ps.w.z};
        for (int k = 0; k < 6; ++k) velocity[k] = vw[k];
    }
    image.state.segmentLambda = c.lambda;
    image.suspensionStepped = stepped;
    image.gravity[0] = c.gravity.x;
    image.gravity[1] = c.gravity.y;
    image.gravity[2] = c.gravity.z;
    image.maxResidual = c.diag.maxResidual;
    image.lineIterations = r.segments.empty() ? 0 : static_cast<int>(c.diag.iterations / r.segments.size());
    image.suspensionSection.assign(bytes, bytes + size);
    image.sectionLength.clear();
    image.sectionCommanded.clear();
    image.sectionTension.clear();
    if (stepped)
        for (const SegmentDiag& g : c.segments) {
            image.sectionLength.push_back(g.length);
            image.sectionCommanded.push_back(g.commanded);
            image.sectionTension.push_back(g.tension);
        }
}

}  // namespace

std::uint64_t softBodyTopologyFingerprint(const Definition& d) {
    TopologyHash h;
    h.add(std::uint64_t{1});
    h.add(static_cast<std::uint64_t>(d.nodeCount()));
    for (std::size_t i = 0; i < d.nodeCount(); ++i) h.add(d.nodeFixed[i] ? 0.0 : 1.0 / d.nodeMass[i]);
    h.add(static_cast<std::uint64_t>(d.triangleCount()));
    for (std::size_t k = 0; k < 3 * d.triangleCount(); ++k) h.add(d.triangleNodes[k]);
    h.add(static_cast<std::uint64_t>(d.constraintCount()));
    for (std::size_t c = 0; c < d.constraintCount(); ++c) {
        const bool stitch = d.constraintKind[c] == SW_CONSTRAINT_SEAM_STITCH;
        h.add(d.constraintNodes[2 * c]);
        h.add(d.constraintNodes[2 * c + 1]);
        h.add(stitch ? 0.0 : d.constraintRest[c]);  // addSeamStitchConstraint: zero rest and compliance
        h.add(stitch ? 0.0 : d.constraintCompliance[c]);
        h.add(bodyConstraintKind(d.constraintKind[c]));
    }
    h.add(static_cast<std::uint64_t>(d.membraneCount()));
    for (std::size_t m = 0; m < d.membraneCount(); ++m) {
        h.add(d.membraneTriangle[m]);
        const double* chart = &d.membraneChart[6 * m];
        for (int k = 0; k < 6; ++k) h.add(chart[k]);
        const double* mat = &d.membraneMaterial[8 * m];
        for (int k = 0; k < 8; ++k) h.add(mat[k]);
        h.add(static_cast<std::uint64_t>(d.membraneRole[m] - 1));  // softwing::MaterialRole
        // addMembraneElements (soft_body.cpp:664-689): reference area and
        // inverse reference matrix of the chart edges, material compliance.
        const double r00 = chart[2] - chart[0], r01 = chart[4] - chart[0];
        const double r10 = chart[3] - chart[1], r11 = chart[5] - chart[1];
        const double det = r00 * r11 - r01 * r10;
        const double inv = 1.0 / det;
        h.add(0.5 * det);
        h.add(r11 * inv);
        h.add(-r01 * inv);
        h.add(-r10 * inv);
        h.add(r00 * inv);
        double compliance[6];
        membraneCompliance(mat, 1.0, compliance);
        for (int k = 0; k < 6; ++k) h.add(compliance[k]);
    }
    h.add(static_cast<std::uint64_t>(d.dihedralCount()));
    for (std::size_t k = 0; k < d.dihedralCount(); ++k) {
        for (int c = 0; c < 4; ++c) h.add(d.dihedralNodes[4 * k + c]);
        h.add(d.dihedralRestAngle[k]);
        h.add(d.dihedralCompliance[k]);
    }
    h.add(std::uint64_t{0});  // interior pressure partitions
    h.add(std::uint64_t{0});  // contact surfaces
    h.add(std::uint64_t{0});  // contact lines
    h.add(std::uint64_t{0});  // contact pairs
    h.add(std::uint64_t{0});  // aerodynamic registration
    return h.value();
}

std::vector<std::uint8_t> encodeStructureCheckpoint(const Definition& d, const CheckpointImage& image) {
    if (d.sphereCount() > 0)
        throw std::invalid_argument("contact spheres have no counterpart in the reference checkpoint format");
    const std::vector<std::uint8_t> body = encodeBody(d, image);
    std::vector<std::uint8_t> suspension;
    if (d.suspension)
        suspension = image.suspensionSection.empty() ? encodeSuspension(d, image) : image.suspensionSection;
    std::vector<std::uint8_t> payload;
    Out w(payload);
    const char* bad = "Structure checkpoint payload is invalid:";
    w.u32(kStateVersion);
    w.u32(0);
    w.u32(kCheckpointVersion);
    w.u32(0);
    w.u64(image.fingerprint);
    w.u64(image.acceptedSteps);
    w.f64(image.time, bad);
    w.u64(d.nodeCount());
    for (std::size_t i = 0; i < d.nodeCount(); ++i) {
        w.v3(&image.state.position[3 * i], bad);
        w.v3(&image.state.previous[3 * i], bad);
        w.v3(&image.state.velocity[3 * i], bad);
    }
    w.u64(d.nodeCount());
    for (std::size_t i = 0; i < d.nodeCount(); ++i) w.v3(&image.pending[3 * i], bad);
    w.v3(image.lastApplied, bad);
    w.u64(body.size());
    w.bytes(body);
    w.u8(d.suspension ? 1 : 0);
    if (d.suspension) {
        w.u64(suspension.size());
        w.bytes(suspension);
    }
    std::vector<std::uint8_t> out;
    envelope(out, "SWST", kStructureProtocol, payload);
    return out;
}

void decodeStructureCheckpoint(const Definition& d, std::uint64_t fingerprint, const std::uint8_t* bytes,
                               std::size_t size, CheckpointImage& image) {
    if (d.sphereCount() > 0)
        throw std::invalid_argument("contact spheres have no counterpart in the reference checkpoint format");
    const auto [payload, length] = openEnvelope(bytes, size, "SWST", kStructureProtocol, "Structure checkpoint");
    In r(payload, length, "Structure checkpoint");
    const std::uint32_t state = r.u32(), reserved = r.u32(), version = r.u32(), reserved2 = r.u32();
    if (state != kStateVersion || version != kCheckpointVersion) r.fail("state version is unsupported");
    if (reserved != 0 || reserved2 != 0) r.fail("payload reserved bits are nonzero");
    image.fingerprint = r.u64();
    if (image.fingerprint != fingerprint) r.fail("belongs to a different definition");
    image.acceptedSteps = r.u64();
    image.time = r.f64();
    r.exactCount(d.nodeCount(), "node count");
    std::vector<double> x(3 * d.nodeCount()), p(3 * d.nodeCount()), v(3 * d.nodeCount());
    for (std::size_t i = 0; i < d.nodeCount(); ++i) {
        r.v3(&x[3 * i]);
        r.v3(&p[3 * i]);
        r.v3(&v[3 * i]);
    }
    r.exactCount(d.nodeCount(), "force count");
    image.pending.assign(3 * d.nodeCount(), 0.0);
    for (std::size_t i = 0; i < d.nodeCount(); ++i) r.v3(&image.pending[3 * i]);
    r.v3(image.lastApplied);
    const std::uint64_t bodySize = r.u64();
    const std::uint8_t* body = r.take(static_cast<std::size_t>(bodySize));
    decodeBody(d, body, static_cast<std::size_t>(bodySize), image.state);
    const std::uint8_t flag = r.u8();
    if (flag > 1) r.fail("suspension marker is invalid");
    if ((flag == 1) != d.suspension.has_value()) r.fail("suspension topology does not match");
    if (flag == 1) {
        const std::uint64_t sectionSize = r.u64();
        const std::uint8_t* section = r.take(static_cast<std::size_t>(sectionSize));
        decodeSuspension(d, section, static_cast<std::size_t>(sectionSize), image);
    }
    if (!r.atEnd()) r.fail("payload has trailing data");
    // The public nodes restate the body's (deserializeStructureCheckpoint
    // :501-511).
    if (x != image.state.position || p != image.state.previous || v != image.state.velocity)
        throw std::invalid_argument("Structure checkpoint public and solver state disagree");
}

}  // namespace swx
