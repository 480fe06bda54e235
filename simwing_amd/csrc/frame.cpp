// Structure diagnostic frames (src/viewer/structure_frame.cpp:480-666) and
// their SWVF encoding (src/viewer/viewer_protocol.cpp:286-376, 690-882).
#include "frame.hpp"

#include <bit>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <unordered_set>

namespace swx {
namespace {

// ProtocolLimits defaults (viewer_protocol.h:17-28).
constexpr std::uint64_t kMaxFrameBytes = 256ULL * 1024ULL * 1024ULL;
constexpr std::uint32_t kMaxVertices = 5'000'000;
constexpr std::uint32_t kMaxTriangles = 10'000'000;
constexpr std::uint32_t kMaxLines = 10'000'000;
constexpr std::uint32_t kMaxFieldValues = 10'000'000;
constexpr std::uint32_t kMaxStringBytes = 1024;
constexpr std::uint16_t kFrameVersion = 1;  // kFrameProtocolVersion (viewer_protocol.h:12)
constexpr std::size_t kEnvelopeBytes = 16;  // magic, version, reserved, payload size

enum Association : std::uint8_t { kGlobal = 0, kVertex = 1, kLine = 3 };

// The field contract of buildStructureFrame (structure_frame.cpp:614-660),
// in emission order.
struct FieldName {
    const char* name;
    const char* unit;
};
constexpr FieldName kLineScalars[2] = {{"structure.constraint_length", "m"},
                                       {"structure.constraint_violation", "m"}};
constexpr FieldName kGlobalScalars[8] = {{"structure.dynamic_mass", "kg"},
                                         {"structure.kinetic_energy", "J"},
                                         {"structure.maximum_distance_error", "m"},
                                         {"structure.maximum_cable_extension", "m"},
                                         {"structure.maximum_absolute_membrane_strain", "1"},
                                         {"structure.maximum_membrane_residual", "1"},
                                         {"structure.maximum_contact_penetration", "m"},
                                         {"structure.maximum_suspension_residual", "m"}};
constexpr FieldName kVertexVectors[2] = {{"structure.velocity", "m/s"}, {"structure.pending_external_force", "N"}};
constexpr FieldName kGlobalVectors[4] = {{"structure.total_pending_external_force", "N"},
                                         {"structure.last_applied_external_force", "N"},
                                         {"structure.center_of_mass", "m"},
                                         {"structure.linear_momentum", "kg*m/s"}};

std::size_t harnessCount(const Definition& d) { return d.suspension ? d.suspension->harnessIds.size() : 0; }
std::size_t segmentCount(const Definition& d) { return d.suspension ? d.suspension->segmentIds.size() : 0; }

[[noreturn]] void invalidFrame(const std::string& message) {
    throw std::invalid_argument("Invalid structure diagnostic frame: " + message);
}

void checkString(const char* value, const char* label, bool nonempty) {
    if (!value) throw std::invalid_argument(std::string(label) + " is NULL");
    const std::size_t size = std::strlen(value);
    if (nonempty && size == 0) invalidFrame(std::string(label) + " must not be empty");
    if (size > kMaxStringBytes) invalidFrame(std::string(label) + " exceeds the string-size limit");
}

void requireUnique(const std::uint64_t* ids, std::size_t count, const char* label) {
    std::unordered_set<std::uint64_t> seen;
    seen.reserve(count);
    for (std::size_t i = 0; i < count; ++i)
        if (ids[i] == 0 || !seen.insert(ids[i]).second)
            throw std::invalid_argument(std::string("Structure frame ") + label + " IDs must be nonzero and unique");
}

bool finite3(const double* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

// Little-endian writer of BufferWriter's encodings (viewer_protocol.cpp:
// 150-199).
struct Writer {
    std::vector<std::uint8_t>& out;
    template <class Integer>
    void integer(Integer v) {
        for (std::size_t i = 0; i < sizeof(Integer); ++i) {
            out.push_back(static_cast<std::uint8_t>(v & 0xffU));
            v = static_cast<Integer>(v >> 8U);
        }
    }
    void u8(std::uint8_t v) { out.push_back(v); }
    void u32(std::uint32_t v) { integer(v); }
    void u64(std::uint64_t v) { integer(v); }
    void real(double v) { u64(std::bit_cast<std::uint64_t>(v)); }
    void vec(const double* v) {
        real(v[0]);
        real(v[1]);
        real(v[2]);
    }
    void string(const char* s) {
        const std::size_t n = std::strlen(s);
        u32(static_cast<std::uint32_t>(n));
        out.insert(out.end(), s, s + n);
    }
};

std::size_t fieldHeader(const FieldName& f) { return 4 + std::strlen(f.name) + 4 + std::strlen(f.unit) + 1 + 4; }

}  // namespace

std::vector<std::uint32_t> frameLineVertices(const Definition& d, const sw_frame_mapping& m) {
    const std::size_t n = d.nodeCount(), h = harnessCount(d), c = d.constraintCount(), s = segmentCount(d);
    constexpr std::size_t maxVertices = std::numeric_limits<std::uint32_t>::max();
    if (h > maxVertices || n > maxVertices - h)
        throw std::invalid_argument("Structure frame has too many vertices for the viewer protocol");
    if (m.vertex_count != n + h || m.triangle_count != d.triangleCount() || m.line_count != c + s)
        throw std::invalid_argument("Structure frame mapping sizes do not match structure topology");
    if ((m.vertex_count && !m.vertex_ids) ||
        (m.triangle_count && (!m.triangle_ids || !m.triangle_negative_region || !m.triangle_positive_region)) ||
        (m.line_count && (!m.line_ids || !m.line_roles)))
        throw std::invalid_argument("frame mapping array is NULL");
    requireUnique(m.vertex_ids, m.vertex_count, "vertex");
    requireUnique(m.triangle_ids, m.triangle_count, "triangle");
    requireUnique(m.line_ids, m.line_count, "line");
    for (std::size_t t = 0; t < m.triangle_count; ++t)
        if (m.triangle_negative_region[t] == 0 || m.triangle_positive_region[t] == 0)
            throw std::invalid_argument("Structure frame triangle needs nonzero side regions");
    std::vector<std::uint32_t> vertex(2 * (c + s));
    for (std::size_t k = 0; k < c; ++k) {
        vertex[2 * k] = static_cast<std::uint32_t>(d.constraintNodes[2 * k]);
        vertex[2 * k + 1] = static_cast<std::uint32_t>(d.constraintNodes[2 * k + 1]);
    }
    if (s) {
        const SuspensionDef& sd = *d.suspension;
        const auto endpoint = [&](std::uint8_t kind, std::uint64_t id) -> std::size_t {
            if (kind == SW_ENDPOINT_ATTACHMENT) {
                for (std::size_t a = 0; a < sd.attachmentIds.size(); ++a)
                    if (sd.attachmentIds[a] == id) return sd.attachmentNodes[a];
            } else if (kind == SW_ENDPOINT_JUNCTION) {
                for (std::size_t j = 0; j < sd.junctionIds.size(); ++j)
                    if (sd.junctionIds[j] == id) return sd.junctionNodes[j];
            } else if (kind == SW_ENDPOINT_HARNESS) {
                for (std::size_t p = 0; p < sd.harnessIds.size(); ++p)
                    if (sd.harnessIds[p] == id) return n + p;
            }
            throw std::invalid_argument("Scene suspension endpoint mapping is incomplete");
        };
        for (std::size_t k = 0; k < s; ++k) {
            // buildStructureFrame matches each segment line to the segment of
            // the same stable id (structure_frame.cpp:566-573).
            if (m.line_ids[c + k] != sd.segmentIds[k])
                throw std::logic_error("Structure suspension state and frame mapping disagree");
            const std::size_t a = endpoint(sd.segmentFromKind[k], sd.segmentFromId[k]);
            const std::size_t b = endpoint(sd.segmentToKind[k], sd.segmentToId[k]);
            if (a >= n + h || b >= n + h || a == b)
                throw std::invalid_argument("Structure frame suspension line has invalid endpoints");
            vertex[2 * (c + k)] = static_cast<std::uint32_t>(a);
            vertex[2 * (c + k) + 1] = static_cast<std::uint32_t>(b);
        }
    }
    return vertex;
}

std::size_t frameBytes(const Definition& d, const sw_frame_mapping& m, const sw_frame_context& c) {
    checkString(c.scene_checksum, "scene checksum", true);
    checkString(c.solver_commit, "solver commit", true);
    std::size_t bytes = kEnvelopeBytes;
    bytes += 4 + std::strlen(c.scene_checksum) + 4 + std::strlen(c.solver_commit);
    bytes += 8 + 8 + 8 + 4;            // step, time, time step, coupling iteration
    bytes += 5 * 8;                    // coupling residuals
    bytes += 8 + 24 + 8 + 24 + 24 + 8; // conservation values
    bytes += 4 + m.vertex_count * (8 + 24);
    bytes += 4 + d.triangleCount() * (8 + 12 + 16);
    bytes += 4 + m.line_count * (8 + 12);
    bytes += 4 + 4;  // no contact or sealing markers
    bytes += 4;
    for (const FieldName& f : kLineScalars) bytes += fieldHeader(f) + 8 * m.line_count;
    for (const FieldName& f : kGlobalScalars) bytes += fieldHeader(f) + 8;
    bytes += 4;
    for (const FieldName& f : kVertexVectors) bytes += fieldHeader(f) + 24 * m.vertex_count;
    for (const FieldName& f : kGlobalVectors) bytes += fieldHeader(f) + 24;
    return bytes;
}

std::vector<std::uint8_t> encodeFrame(const Definition& d, const sw_frame_mapping& m, const sw_frame_context& c,
                                      const FrameFields& f) {
    const std::vector<std::uint32_t> lineVertex = frameLineVertices(d, m);
    const std::size_t size = frameBytes(d, m, c);
    const std::size_t v = m.vertex_count, lines = m.line_count, t = d.triangleCount();
    if (!f.diagnostics || (v && (!f.vertexPosition || !f.vertexVelocity || !f.vertexPendingForce)) ||
        (lines && (!f.lineLength || !f.lineViolation)))
        throw std::invalid_argument("frame field array is NULL");
    const sw_diagnostics& g = *f.diagnostics;

    // validateFrame (viewer_protocol.cpp:690-836).
    if (v > kMaxVertices) invalidFrame("vertex count exceeds the protocol limit");
    if (t > kMaxTriangles) invalidFrame("triangle count exceeds the protocol limit");
    if (lines > kMaxLines) invalidFrame("line count exceeds the protocol limit");
    if (v > kMaxFieldValues || lines > kMaxFieldValues) invalidFrame("field value count exceeds the protocol limit");
    if (!std::isfinite(f.simulationTime) || f.simulationTime < 0.0 || !std::isfinite(c.time_step) || c.time_step < 0.0)
        invalidFrame("simulation time and time step must be finite and nonnegative");
    bool residualsFinite = std::isfinite(c.fluid_mass) && c.fluid_mass >= 0.0 && std::isfinite(c.total_energy) &&
                           std::isfinite(c.interface_power_residual) && finite3(c.total_momentum) &&
                           finite3(c.interface_force_residual) && finite3(c.interface_moment_residual);
    for (double r : c.coupling_residuals) residualsFinite = residualsFinite && std::isfinite(r);
    if (!residualsFinite) invalidFrame("residual and conservation values must be finite");
    for (std::size_t i = 0; i < v; ++i)
        if (!finite3(f.vertexPosition + 3 * i)) invalidFrame("vertex position is not finite");
    for (std::size_t k = 0; k < t; ++k) {
        const std::uint64_t* tri = &d.triangleNodes[3 * k];
        if (tri[0] >= v || tri[1] >= v || tri[2] >= v || tri[0] == tri[1] || tri[1] == tri[2] || tri[2] == tri[0])
            invalidFrame("triangle has invalid vertices or side-region IDs");
    }
    for (std::size_t k = 0; k < lines; ++k)
        if (lineVertex[2 * k] >= v || lineVertex[2 * k + 1] >= v || lineVertex[2 * k] == lineVertex[2 * k + 1])
            invalidFrame("line has invalid vertex indices");
    const double globals[8] = {g.total_dynamic_mass,           g.kinetic_energy,
                               g.maximum_distance_error,       g.maximum_cable_extension,
                               g.maximum_absolute_membrane_strain, g.maximum_membrane_residual,
                               g.maximum_contact_penetration,  g.maximum_suspension_residual};
    const double* globalVectors[4] = {g.pending_external_force, g.last_applied_external_force, g.center_of_mass,
                                      g.linear_momentum};
    for (std::size_t k = 0; k < lines; ++k)
        if (!std::isfinite(f.lineLength[k]) || !std::isfinite(f.lineViolation[k]))
            invalidFrame("scalar field contains a non-finite value");
    for (double x : globals)
        if (!std::isfinite(x)) invalidFrame("scalar field contains a non-finite value");
    for (std::size_t i = 0; i < v; ++i)
        if (!finite3(f.vertexVelocity + 3 * i) || !finite3(f.vertexPendingForce + 3 * i))
            invalidFrame("vector field contains a non-finite value");
    for (const double* x : globalVectors)
        if (!finite3(x)) invalidFrame("vector field contains a non-finite value");
    if (size > kMaxFrameBytes) invalidFrame("serialized diagnostic frame exceeds the byte limit");

    std::vector<std::uint8_t> bytes;
    bytes.reserve(size);
    Writer w{bytes};
    // serializeFrame's envelope (viewer_protocol.cpp:854-873).
    for (char ch : {'S', 'W', 'V', 'F'}) w.u8(static_cast<std::uint8_t>(ch));
    w.integer<std::uint16_t>(kFrameVersion);
    w.integer<std::uint16_t>(0);
    w.u64(static_cast<std::uint64_t>(size - kEnvelopeBytes));
    // writeFramePayload (viewer_protocol.cpp:286-376).
    w.string(c.scene_checksum);
    w.string(c.solver_commit);
    w.u64(f.step);
    w.real(f.simulationTime);
    w.real(c.time_step);
    w.u32(c.coupling_iteration);
    for (double r : c.coupling_residuals) w.real(r);
    w.real(c.fluid_mass);
    w.vec(c.total_momentum);
    w.real(c.total_energy);
    w.vec(c.interface_force_residual);
    w.vec(c.interface_moment_residual);
    w.real(c.interface_power_residual);
    w.u32(static_cast<std::uint32_t>(v));
    for (std::size_t i = 0; i < v; ++i) {
        w.u64(m.vertex_ids[i]);
        w.vec(f.vertexPosition + 3 * i);
    }
    w.u32(static_cast<std::uint32_t>(t));
    for (std::size_t k = 0; k < t; ++k) {
        w.u64(m.triangle_ids[k]);
        for (int corner = 0; corner < 3; ++corner) w.u32(static_cast<std::uint32_t>(d.triangleNodes[3 * k + corner]));
        w.u64(m.triangle_negative_region[k]);
        w.u64(m.triangle_positive_region[k]);
    }
    w.u32(static_cast<std::uint32_t>(lines));
    for (std::size_t k = 0; k < lines; ++k) {
        w.u64(m.line_ids[k]);
        w.u32(lineVertex[2 * k]);
        w.u32(lineVertex[2 * k + 1]);
        w.u32(m.line_roles[k]);
    }
    w.u32(0);  // contact markers
    w.u32(0);  // sealing markers
    w.u32(2 + 8);
    const double* lineValues[2] = {f.lineLength, f.lineViolation};
    for (int k = 0; k < 2; ++k) {
        w.string(kLineScalars[k].name);
        w.string(kLineScalars[k].unit);
        w.u8(kLine);
        w.u32(static_cast<std::uint32_t>(lines));
        for (std::size_t i = 0; i < lines; ++i) w.real(lineValues[k][i]);
    }
    for (int k = 0; k < 8; ++k) {
        w.string(kGlobalScalars[k].name);
        w.string(kGlobalScalars[k].unit);
        w.u8(kGlobal);
        w.u32(1);
        w.real(globals[k]);
    }
    w.u32(2 + 4);
    const double* vertexValues[2] = {f.vertexVelocity, f.vertexPendingForce};
    for (int k = 0; k < 2; ++k) {
        w.string(kVertexVectors[k].name);
        w.string(kVertexVectors[k].unit);
        w.u8(kVertex);
        w.u32(static_cast<std::uint32_t>(v));
        for (std::size_t i = 0; i < v; ++i) w.vec(vertexValues[k] + 3 * i);
    }
    for (int k = 0; k < 4; ++k) {
        w.string(kGlobalVectors[k].name);
        w.string(kGlobalVectors[k].unit);
        w.u8(kGlobal);
        w.u32(1);
        w.vec(globalVectors[k]);
    }
    if (bytes.size() != size) throw std::logic_error("frame encoder size mismatch");
    return bytes;
}

}  // namespace swx
