// Scene-v2 compact binary stream (src/fsi/scene.cpp:1051-1418).
#include "scene_io.hpp"

#include "assembly.hpp"

#include <algorithm>
#include <bit>
#include <cstring>
#include <stdexcept>
#include <string>

namespace swx {
namespace {

// Stream constants (scene.cpp:23-28, scene.h:20).
constexpr char kMagic[8] = {'S', 'I', 'M', 'W', 'S', 'C', 'N', '2'};
constexpr std::uint32_t kBinaryVersion = 3;
constexpr std::uint32_t kMaxEntities = 10'000'000;
constexpr std::uint32_t kMaxStringBytes = 1'048'576;
constexpr std::size_t kMaxDecodedBytes = 256ULL * 1024ULL * 1024ULL;

[[noreturn]] void fail(const std::string& message) { throw std::invalid_argument(message); }

struct Writer {
    std::vector<std::uint8_t> out;
    void u8(std::uint8_t v) { out.push_back(v); }
    void u32(std::uint32_t v) {
        for (int i = 0; i < 4; ++i) out.push_back(static_cast<std::uint8_t>((v >> (8 * i)) & 0xffU));
    }
    void u64(std::uint64_t v) {
        for (int i = 0; i < 8; ++i) out.push_back(static_cast<std::uint8_t>((v >> (8 * i)) & 0xffU));
    }
    void real(double v) { u64(std::bit_cast<std::uint64_t>(v)); }
    void v2(const V2& v) {
        real(v.x);
        real(v.y);
    }
    void v3(const V3& v) {
        real(v.x);
        real(v.y);
        real(v.z);
    }
    void string(const std::string& s) {
        if (s.size() > kMaxStringBytes) fail("scene string exceeds the binary safety limit");
        u32(static_cast<std::uint32_t>(s.size()));
        out.insert(out.end(), s.begin(), s.end());
    }
    void count(std::size_t n) {
        if (n > kMaxEntities) fail("scene entity collection exceeds the binary safety limit");
        u32(static_cast<std::uint32_t>(n));
    }
};

struct Reader {
    const std::uint8_t* data;
    std::size_t size;
    std::size_t at = 0;
    std::size_t remaining = kMaxDecodedBytes;

    void need(std::size_t n) {
        if (n > size - at) fail("truncated scene stream");
    }
    std::uint8_t u8() {
        need(1);
        return data[at++];
    }
    std::uint32_t u32() {
        need(4);
        std::uint32_t v = 0;
        for (int i = 0; i < 4; ++i) v |= static_cast<std::uint32_t>(data[at++]) << (8 * i);
        return v;
    }
    std::uint64_t u64() {
        need(8);
        std::uint64_t v = 0;
        for (int i = 0; i < 8; ++i) v |= static_cast<std::uint64_t>(data[at++]) << (8 * i);
        return v;
    }
    double real() { return std::bit_cast<double>(u64()); }
    V2 v2() {
        const double x = real();
        const double y = real();
        return {x, y};
    }
    V3 v3() {
        const double x = real();
        const double y = real();
        const double z = real();
        return {x, y, z};
    }
    void claim(std::size_t n, std::size_t element, const std::string& label) {
        if (element != 0 && n > remaining / element) fail(label + " exceeds the scene decoded-allocation limit");
        remaining -= n * element;
    }
    std::string string() {
        const std::uint32_t n = u32();
        if (n > kMaxStringBytes) fail("scene string exceeds the binary safety limit");
        claim(n, 1, "scene string");
        need(n);
        std::string s(reinterpret_cast<const char*>(data + at), n);
        at += n;
        return s;
    }
    template <class T>
    void resize(std::vector<T>& values, const std::string& label) {
        const std::uint32_t n = u32();
        if (n > kMaxEntities) fail(label + " count exceeds the binary safety limit");
        claim(n, sizeof(T), label);
        values.resize(n);
    }
};

template <class T>
std::vector<T> sortedById(const std::vector<T>& values) {
    std::vector<T> out = values;
    std::stable_sort(out.begin(), out.end(), [](const T& a, const T& b) { return a.id < b.id; });
    return out;
}

// withinBinarySafetyLimits (scene.cpp:218-287): counts and strings.
void checkLimits(const Scene& s) {
    const auto count = [](std::size_t n, const char* label) {
        if (n > kMaxEntities) fail(std::string(label) + " exceeds the scene binary allocation limit");
    };
    const auto text = [](const std::string& v, const char* label) {
        if (v.size() > kMaxStringBytes)
            fail(std::string(label) + " exceeds the scene binary string/allocation limit");
    };
    count(s.regions.size(), "region count");
    count(s.vertices.size(), "vertex count");
    count(s.fabricMaterials.size(), "fabric-material count");
    count(s.seamMaterials.size(), "seam-material count");
    count(s.triangles.size(), "triangle count");
    count(s.openings.size(), "opening count");
    count(s.seams.size(), "seam count");
    count(s.lineMaterials.size(), "line-material count");
    count(s.pilots.size(), "pilot count");
    count(s.junctions.size(), "suspension-junction count");
    count(s.attachments.size(), "attachment count");
    count(s.lines.size(), "suspension-line count");
    text(s.metadata.designChecksum, "design checksum");
    text(s.metadata.exporterVersion, "exporter version");
    for (const auto& r : s.regions) text(r.name, "region name");
    for (const auto& m : s.fabricMaterials) text(m.name, "fabric-material name");
    for (const auto& m : s.seamMaterials) text(m.name, "seam-material name");
    for (const auto& m : s.lineMaterials) text(m.name, "line-material name");
    for (const auto& p : s.pilots) text(p.name, "pilot name");
}

}  // namespace

std::vector<std::uint8_t> writeSceneBinary(const Scene& scene) {
    validateSceneStructure(scene);
    checkLimits(scene);
    Writer w;
    w.out.insert(w.out.end(), kMagic, kMagic + 8);
    w.u32(kBinaryVersion);
    const SceneMetadata& m = scene.metadata;
    w.u32(m.schemaMajor);
    w.u32(m.schemaMinor);
    w.u8(m.lengthUnit);
    w.u8(m.handedness);
    w.u8(m.upAxis);
    w.real(m.sourceLengthToMeters);
    w.string(m.designChecksum);
    w.string(m.exporterVersion);
    const auto regions = sortedById(scene.regions);
    w.count(regions.size());
    for (const Region& r : regions) {
        w.u64(r.id);
        w.u8(static_cast<std::uint8_t>(r.kind));
        w.string(r.name);
    }
    const auto vertices = sortedById(scene.vertices);
    w.count(vertices.size());
    for (const Vertex& v : vertices) {
        w.u64(v.id);
        w.v3(v.position);
    }
    const auto fabric = sortedById(scene.fabricMaterials);
    w.count(fabric.size());
    for (const FabricMaterial& f : fabric) {
        w.u64(f.id);
        w.string(f.name);
        w.real(f.warpStiffness);
        w.real(f.weftStiffness);
        w.real(f.shearStiffness);
        w.real(f.bendingStiffness);
        w.real(f.arealDensity);
        w.real(f.dampingSeconds);
        w.real(f.porosity);
        w.real(f.permeability);
    }
    const auto seamMaterials = sortedById(scene.seamMaterials);
    w.count(seamMaterials.size());
    for (const SeamMaterial& f : seamMaterials) {
        w.u64(f.id);
        w.string(f.name);
        w.real(f.linearDensity);
        w.real(f.axialStiffness);
    }
    const auto triangles = sortedById(scene.triangles);
    w.count(triangles.size());
    for (const Triangle& t : triangles) {
        w.u64(t.id);
        for (StableId v : t.vertexIds) w.u64(v);
        for (const V2& c : t.chart) w.v2(c);
        w.u64(t.negativeRegion);
        w.u64(t.positiveRegion);
        w.u64(t.materialId);
        w.u64(t.sheetId);
        w.u8(static_cast<std::uint8_t>(t.role));
    }
    const auto openings = sortedById(scene.openings);
    w.count(openings.size());
    for (const Opening& o : openings) {
        w.u64(o.id);
        w.count(o.orderedVertexIds.size());
        for (StableId v : o.orderedVertexIds) w.u64(v);
        w.u64(o.negativeRegion);
        w.u64(o.positiveRegion);
        w.u8(o.role);
        w.count(o.capTriangles.size());
        for (const auto& t : o.capTriangles)
            for (StableId v : t) w.u64(v);
    }
    const auto seams = sortedById(scene.seams);
    w.count(seams.size());
    for (const Seam& s : seams) {
        w.u64(s.id);
        w.u64(s.materialId);
        w.count(s.first.size());
        for (StableId v : s.first) w.u64(v);
        w.count(s.second.size());
        for (StableId v : s.second) w.u64(v);
    }
    const auto lineMaterials = sortedById(scene.lineMaterials);
    w.count(lineMaterials.size());
    for (const LineMaterial& l : lineMaterials) {
        w.u64(l.id);
        w.string(l.name);
        w.real(l.diameter);
        w.real(l.linearDensity);
        w.real(l.axialStiffness);
        w.real(l.drag);
    }
    const auto pilots = sortedById(scene.pilots);
    w.count(pilots.size());
    for (const Pilot& p : pilots) {
        w.u64(p.id);
        w.string(p.name);
        w.real(p.mass);
        w.v3(p.centerOfMass);
        w.v3(p.linearVelocity);
        w.real(p.bodyToWorld.w);
        w.real(p.bodyToWorld.x);
        w.real(p.bodyToWorld.y);
        w.real(p.bodyToWorld.z);
        w.v3(p.principalInertia);
    }
    const auto junctions = sortedById(scene.junctions);
    w.count(junctions.size());
    for (const Junction& j : junctions) {
        w.u64(j.id);
        w.v3(j.position);
        w.real(j.mass);
        w.u8(j.fixed ? 1 : 0);
    }
    const auto attachments = sortedById(scene.attachments);
    w.count(attachments.size());
    for (const Attachment& a : attachments) {
        w.u64(a.id);
        w.u8(static_cast<std::uint8_t>(a.kind));
        w.u64(a.vertexId);
        w.u64(a.pilotId);
        w.v3(a.pilotLocal);
        w.u64(a.junctionId);
    }
    const auto lines = sortedById(scene.lines);
    w.count(lines.size());
    for (const Line& l : lines) {
        w.u64(l.id);
        w.u64(l.startAttachment);
        w.u64(l.endAttachment);
        w.u64(l.materialId);
        w.real(l.restLength);
        w.u8(static_cast<std::uint8_t>(l.role));
    }
    return std::move(w.out);
}

Scene readSceneBinary(const std::uint8_t* bytes, std::size_t size) {
    if (!bytes && size) fail("scene stream is NULL");
    Reader r{bytes, size};
    r.need(8);
    if (std::memcmp(bytes, kMagic, 8) != 0) fail("scene stream has an invalid magic header");
    r.at = 8;
    const std::uint32_t version = r.u32();
    if (version != 2 && version != kBinaryVersion) fail("scene stream uses an unsupported binary version");
    Scene s;
    s.metadata.schemaMajor = r.u32();
    s.metadata.schemaMinor = r.u32();
    s.metadata.lengthUnit = r.u8();
    s.metadata.handedness = r.u8();
    s.metadata.upAxis = r.u8();
    s.metadata.sourceLengthToMeters = r.real();
    s.metadata.designChecksum = r.string();
    s.metadata.exporterVersion = r.string();
    r.resize(s.regions, "region");
    for (Region& x : s.regions) {
        x.id = r.u64();
        x.kind = static_cast<RegionKind>(r.u8());
        x.name = r.string();
    }
    r.resize(s.vertices, "vertex");
    for (Vertex& x : s.vertices) {
        x.id = r.u64();
        x.position = r.v3();
    }
    r.resize(s.fabricMaterials, "fabric material");
    for (FabricMaterial& x : s.fabricMaterials) {
        x.id = r.u64();
        x.name = r.string();
        x.warpStiffness = r.real();
        x.weftStiffness = r.real();
        x.shearStiffness = r.real();
        x.bendingStiffness = r.real();
        x.arealDensity = r.real();
        x.dampingSeconds = r.real();
        x.porosity = r.real();
        x.permeability = r.real();
    }
    r.resize(s.seamMaterials, "seam material");
    for (SeamMaterial& x : s.seamMaterials) {
        x.id = r.u64();
        x.name = r.string();
        x.linearDensity = r.real();
        x.axialStiffness = r.real();
    }
    r.resize(s.triangles, "triangle");
    for (Triangle& x : s.triangles) {
        x.id = r.u64();
        for (StableId& v : x.vertexIds) v = r.u64();
        for (V2& c : x.chart) c = r.v2();
        x.negativeRegion = r.u64();
        x.positiveRegion = r.u64();
        x.materialId = r.u64();
        x.sheetId = r.u64();
        x.role = static_cast<SurfaceRole>(r.u8());
    }
    r.resize(s.openings, "opening");
    for (Opening& x : s.openings) {
        x.id = r.u64();
        r.resize(x.orderedVertexIds, "opening boundary vertex");
        for (StableId& v : x.orderedVertexIds) v = r.u64();
        x.negativeRegion = r.u64();
        x.positiveRegion = r.u64();
        x.role = r.u8();
        if (version >= 3) {
            r.resize(x.capTriangles, "opening cap triangle");
            for (auto& t : x.capTriangles)
                for (StableId& v : t) v = r.u64();
        }
    }
    r.resize(s.seams, "seam");
    for (Seam& x : s.seams) {
        x.id = r.u64();
        x.materialId = r.u64();
        r.resize(x.first, "first seam-chain vertex");
        for (StableId& v : x.first) v = r.u64();
        r.resize(x.second, "second seam-chain vertex");
        for (StableId& v : x.second) v = r.u64();
    }
    r.resize(s.lineMaterials, "line material");
    for (LineMaterial& x : s.lineMaterials) {
        x.id = r.u64();
        x.name = r.string();
        x.diameter = r.real();
        x.linearDensity = r.real();
        x.axialStiffness = r.real();
        x.drag = r.real();
    }
    r.resize(s.pilots, "pilot");
    for (Pilot& x : s.pilots) {
        x.id = r.u64();
        x.name = r.string();
        x.mass = r.real();
        x.centerOfMass = r.v3();
        x.linearVelocity = r.v3();
        x.bodyToWorld.w = r.real();
        x.bodyToWorld.x = r.real();
        x.bodyToWorld.y = r.real();
        x.bodyToWorld.z = r.real();
        x.principalInertia = r.v3();
    }
    r.resize(s.junctions, "suspension junction");
    for (Junction& x : s.junctions) {
        x.id = r.u64();
        x.position = r.v3();
        x.mass = r.real();
        const std::uint8_t fixed = r.u8();
        if (fixed > 1) fail("suspension junction fixed flag is invalid");
        x.fixed = fixed != 0;
    }
    r.resize(s.attachments, "attachment");
    for (Attachment& x : s.attachments) {
        x.id = r.u64();
        x.kind = static_cast<AttachmentKind>(r.u8());
        x.vertexId = r.u64();
        x.pilotId = r.u64();
        x.pilotLocal = r.v3();
        x.junctionId = r.u64();
    }
    r.resize(s.lines, "suspension line");
    for (Line& x : s.lines) {
        x.id = r.u64();
        x.startAttachment = r.u64();
        x.endAttachment = r.u64();
        x.materialId = r.u64();
        x.restLength = r.real();
        x.role = static_cast<LineRole>(r.u8());
    }
    if (r.at != size) fail("scene stream contains trailing data");
    validateSceneStructure(s);
    return s;
}

}  // namespace swx
