// Byte-compatible Structure checkpoint codec (SWST around SWBC).
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
// A body built by Structure carries no contact state: its contact
// diagnostics are the default block (not registered, solve succeeded, every
// count and vector zero, minimum gap +infinity).
#include "checkpoint_codec.hpp"

#include <bit>
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
    if (d.suspension)
        throw std::invalid_argument("the suspension section of the checkpoint format is not written by this codec");
    const std::vector<std::uint8_t> body = encodeBody(d, image);
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
    w.u8(0);  // no suspension
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
    if (flag == 1) r.fail("suspension section is not read by this codec");
    if (!r.atEnd()) r.fail("payload has trailing data");
    // The public nodes restate the body's (deserializeStructureCheckpoint
    // :501-511).
    if (x != image.state.position || p != image.state.previous || v != image.state.velocity)
        throw std::invalid_argument("Structure checkpoint public and solver state disagree");
}

}  // namespace swx
