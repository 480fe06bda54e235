// Scene -> structural definition. Every ordering rule and every floating
// point expression that reaches the solver follows
// src/fsi/scene_structure.cpp (cited per step) so the tables are bit-equal
// to the reference's; the data structures are this build's own.
#include "assembly.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <queue>
#include <set>
#include <tuple>
#include <unordered_map>
#include <unordered_set>

namespace swx {
namespace {

[[noreturn]] void reject(const std::string& message) {
    throw std::invalid_argument("scene assembly rejected: " + message);
}

template <typename T>
std::vector<const T*> byId(const std::vector<T>& items) {
    std::vector<const T*> out;
    out.reserve(items.size());
    for (const T& item : items) out.push_back(&item);
    std::sort(out.begin(), out.end(), [](const T* a, const T* b) { return a->id < b->id; });
    return out;
}

template <typename T>
void requireUniqueIds(const std::vector<T>& items, const char* label) {
    std::unordered_set<StableId> seen;
    for (const T& item : items) {
        if (item.id == 0) reject(std::string(label) + " stable ID must be non-zero");
        if (!seen.insert(item.id).second) reject(std::string(label) + " stable ID is duplicated");
    }
}

bool finite(const V3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
V3 minus(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 crossOf(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
double dotOf(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double normOf(const V3& a) { return std::sqrt(dotOf(a, a)); }
V3 times(const V3& v, double s) { return {s * v.x, s * v.y, s * v.z}; }

// Scene validation: the subset of validateScene (scene.cpp:552-1040) that
// the structural conversion depends on.
void validate(const Scene& s) {
    requireUniqueIds(s.regions, "region");
    requireUniqueIds(s.vertices, "vertex");
    requireUniqueIds(s.fabricMaterials, "fabric material");
    requireUniqueIds(s.seamMaterials, "seam material");
    requireUniqueIds(s.triangles, "triangle");
    requireUniqueIds(s.seams, "seam");
    requireUniqueIds(s.lineMaterials, "line material");
    requireUniqueIds(s.pilots, "pilot");
    requireUniqueIds(s.junctions, "suspension junction");
    requireUniqueIds(s.attachments, "attachment");
    requireUniqueIds(s.lines, "suspension line");
    std::unordered_map<StableId, const Vertex*> vertices;
    for (const Vertex& v : s.vertices) {
        if (!finite(v.position)) reject("vertex position is not finite");
        vertices.emplace(v.id, &v);
    }
    std::unordered_set<StableId> regions, fabric, seamMaterials, lineMaterials, pilots, junctions,
        attachments;
    std::size_t outside = 0;
    for (const Region& r : s.regions) {
        regions.insert(r.id);
        if (r.kind == RegionKind::Outside) ++outside;
    }
    if (outside != 1) reject("scene must contain exactly one outside fluid region");
    for (const FabricMaterial& m : s.fabricMaterials) {
        fabric.insert(m.id);
        const double values[] = {m.warpStiffness, m.weftStiffness, m.shearStiffness, m.bendingStiffness,
                                 m.arealDensity, m.dampingSeconds, m.porosity, m.permeability};
        for (double v : values)
            if (!std::isfinite(v)) reject("fabric material contains a non-finite value");
        if (!(m.warpStiffness > 0.0) || !(m.weftStiffness > 0.0) || !(m.shearStiffness > 0.0) ||
            m.bendingStiffness < 0.0 || !(m.arealDensity > 0.0) || m.dampingSeconds < 0.0 ||
            m.porosity < 0.0 || m.porosity > 1.0 || m.permeability < 0.0)
            reject("fabric material has an invalid physical value");
    }
    for (const SeamMaterial& m : s.seamMaterials) {
        seamMaterials.insert(m.id);
        if (!(m.linearDensity > 0.0) || !(m.axialStiffness > 0.0) || !std::isfinite(m.linearDensity) ||
            !std::isfinite(m.axialStiffness))
            reject("seam material density and axial stiffness must be positive");
    }
    std::set<std::array<StableId, 3>> faces;
    for (const Triangle& t : s.triangles) {
        for (const V2& c : t.chart)
            if (!std::isfinite(c.x) || !std::isfinite(c.y)) reject("triangle material coordinate is not finite");
        std::set<StableId> local(t.vertexIds.begin(), t.vertexIds.end());
        for (StableId v : t.vertexIds)
            if (!vertices.count(v)) reject("triangle references a missing vertex");
        if (local.size() != 3) reject("triangle has repeated vertices");
        const V3 a = vertices[t.vertexIds[0]]->position, b = vertices[t.vertexIds[1]]->position,
                 c = vertices[t.vertexIds[2]]->position;
        const V3 ab = minus(b, a), ac = minus(c, a), bc = minus(c, b);
        const double longest = std::max({dotOf(ab, ab), dotOf(ac, ac), dotOf(bc, bc)});
        const V3 n = crossOf(ab, ac);
        const double crossSquared = dotOf(n, n);
        if (!(longest > 0.0) || !(crossSquared > 1.0e-24 * longest * longest))
            reject("triangle has effectively collinear vertices");
        if (!fabric.count(t.materialId)) reject("triangle references a missing fabric material");
        if (t.sheetId == 0) reject("triangle fabric sheet ID must be non-zero");
        if (!regions.count(t.negativeRegion) || !regions.count(t.positiveRegion))
            reject("triangle references a missing side region");
        const bool internal = t.role == SurfaceRole::Diagonal || t.role == SurfaceRole::MiniRib;
        if (t.negativeRegion == t.positiveRegion && !internal)
            reject("skin and rib triangle side regions must be distinct");
        std::array<StableId, 3> key = t.vertexIds;
        std::sort(key.begin(), key.end());
        if (!faces.insert(key).second) reject("triangle duplicates an existing geometric face");
    }
    std::set<std::array<StableId, 2>> seamPairs;
    for (const Seam& seam : s.seams) {
        bool invalid = seam.first.size() < 2 || seam.first.size() != seam.second.size();
        std::unordered_set<StableId> c0(seam.first.begin(), seam.first.end());
        std::unordered_set<StableId> c1(seam.second.begin(), seam.second.end());
        invalid = invalid || c0.size() != seam.first.size() || c1.size() != seam.second.size();
        for (StableId v : seam.first)
            if (!vertices.count(v)) reject("seam chain references a missing vertex");
        for (StableId v : seam.second)
            if (!vertices.count(v)) reject("seam chain references a missing vertex");
        std::size_t distinct = 0;
        if (!invalid) {
            for (StableId v : c0) {
                if (!c1.count(v)) continue;
                const bool sharedFirst = seam.first.front() == v && seam.second.front() == v;
                const bool sharedLast = seam.first.back() == v && seam.second.back() == v;
                if (!sharedFirst && !sharedLast) invalid = true;
            }
            for (std::size_t k = 0; k < seam.first.size(); ++k) {
                std::array<StableId, 2> pair{seam.first[k], seam.second[k]};
                if (pair[0] != pair[1]) ++distinct;
                std::sort(pair.begin(), pair.end());
                if (!seamPairs.insert(pair).second) invalid = true;
            }
        }
        if (invalid || distinct == 0) reject("seam needs equal-length unique ordered chains");
        if (!seamMaterials.count(seam.materialId)) reject("seam references a missing seam material");
    }
    for (const LineMaterial& m : s.lineMaterials) {
        lineMaterials.insert(m.id);
        if (!(m.diameter > 0.0) || !(m.linearDensity > 0.0) || !(m.axialStiffness > 0.0) || m.drag < 0.0 ||
            !std::isfinite(m.diameter) || !std::isfinite(m.linearDensity) ||
            !std::isfinite(m.axialStiffness) || !std::isfinite(m.drag))
            reject("line material has an invalid physical value");
    }
    for (const Pilot& p : s.pilots) {
        pilots.insert(p.id);
        const double q2 = p.bodyToWorld.w * p.bodyToWorld.w + p.bodyToWorld.x * p.bodyToWorld.x +
                          p.bodyToWorld.y * p.bodyToWorld.y + p.bodyToWorld.z * p.bodyToWorld.z;
        if (!std::isfinite(p.mass) || !finite(p.centerOfMass) || !finite(p.linearVelocity) ||
            !finite(p.principalInertia) || !(p.mass > 0.0) || !(p.principalInertia.x > 0.0) ||
            !(p.principalInertia.y > 0.0) || !(p.principalInertia.z > 0.0) || std::abs(q2 - 1.0) > 1.0e-6)
            reject("pilot mass/inertia must be positive and orientation normalized");
    }
    for (const Junction& j : s.junctions) {
        junctions.insert(j.id);
        if (!finite(j.position) || !std::isfinite(j.mass) || j.mass < 0.0 || (!j.fixed && !(j.mass > 0.0)))
            reject("dynamic suspension junction requires positive explicit mass");
        if (vertices.count(j.id)) reject("suspension junction ID must not collide with a surface vertex ID");
    }
    for (const Attachment& a : s.attachments) {
        attachments.insert(a.id);
        const bool zeroLocal = a.pilotLocal.x == 0.0 && a.pilotLocal.y == 0.0 && a.pilotLocal.z == 0.0;
        switch (a.kind) {
        case AttachmentKind::SurfaceVertex:
            if (!vertices.count(a.vertexId) || a.pilotId != 0 || a.junctionId != 0 || !zeroLocal)
                reject("surface attachment must reference one existing vertex");
            break;
        case AttachmentKind::PilotHarness:
            if (!pilots.count(a.pilotId)) reject("pilot attachment references a missing pilot");
            if (a.vertexId != 0 || a.junctionId != 0 || !finite(a.pilotLocal))
                reject("pilot attachment target or local position is invalid");
            break;
        case AttachmentKind::SuspensionJunction:
            if (!junctions.count(a.junctionId) || a.vertexId != 0 || a.pilotId != 0 || !zeroLocal)
                reject("junction attachment must reference one existing suspension junction");
            break;
        default:
            reject("attachment kind is invalid");
        }
    }
    for (const Line& l : s.lines) {
        if (!attachments.count(l.startAttachment) || !attachments.count(l.endAttachment))
            reject("suspension line references a missing attachment");
        if (l.startAttachment == l.endAttachment || !std::isfinite(l.restLength) || !(l.restLength > 0.0))
            reject("suspension line endpoints and rest length are invalid");
        if (!lineMaterials.count(l.materialId)) reject("suspension line references a missing line material");
    }
}

// Diagonal specialisation of the membrane stiffness inverse's conditioning
// test (scene_structure.cpp:206-223).
bool invertibleMaterial(const FabricMaterial& m) {
    const double det = m.warpStiffness * m.weftStiffness * m.shearStiffness;
    const double scale = std::max({m.warpStiffness, m.weftStiffness, m.shearStiffness});
    const double tol = 128.0 * std::numeric_limits<double>::epsilon() * scale * scale * scale;
    return scale > 0.0 && std::isfinite(det) && std::isfinite(tol) && det > tol &&
           std::isfinite(1.0 / m.warpStiffness) && std::isfinite(1.0 / m.weftStiffness) &&
           std::isfinite(1.0 / m.shearStiffness);
}

// Positive, well-conditioned chart (scene_structure.cpp:225-250).
bool usableChart(const Triangle& t) {
    const double e1x = t.chart[1].x - t.chart[0].x, e1y = t.chart[1].y - t.chart[0].y;
    const double e2x = t.chart[2].x - t.chart[0].x, e2y = t.chart[2].y - t.chart[0].y;
    const double det = e1x * e2y - e2x * e1y;
    const double e3x = t.chart[2].x - t.chart[1].x, e3y = t.chart[2].y - t.chart[1].y;
    const double longest = std::max({e1x * e1x + e1y * e1y, e2x * e2x + e2y * e2y, e3x * e3x + e3y * e3y});
    const double tol = 64.0 * std::numeric_limits<double>::epsilon() * longest;
    return longest > 0.0 && std::isfinite(det) && std::isfinite(tol) && det > tol;
}

double chartArea(const Triangle& t) {
    const double e1x = t.chart[1].x - t.chart[0].x, e1y = t.chart[1].y - t.chart[0].y;
    const double e2x = t.chart[2].x - t.chart[0].x, e2y = t.chart[2].y - t.chart[0].y;
    return 0.5 * (e1x * e2y - e2x * e1y);
}

struct HalfEdge {
    const Triangle* triangle;
    StableId from, to, opposite;
};

// Sheet-keyed hinges (scene_structure.cpp:278-422): one hinge per two-face,
// consistently oriented edge of one sheet, visited in (sheet, min, max)
// order; series strip compliance from the material charts and the signed
// spatial rest angle.
void assembleHinges(const std::vector<const Triangle*>& triangles,
                    const std::unordered_map<StableId, V3>& positions,
                    const std::unordered_map<StableId, const FabricMaterial*>& materials,
                    const std::unordered_map<StableId, std::size_t>& nodeOf, Assembly& out) {
    std::map<std::array<StableId, 3>, std::vector<HalfEdge>> edges;
    for (const Triangle* t : triangles) {
        for (int corner = 0; corner < 3; ++corner) {
            const StableId from = t->vertexIds[corner];
            const StableId to = t->vertexIds[(corner + 1) % 3];
            const StableId opposite = t->vertexIds[(corner + 2) % 3];
            const std::array<StableId, 3> key{t->sheetId, std::min(from, to), std::max(from, to)};
            edges[key].push_back({t, from, to, opposite});
        }
    }
    Definition& d = out.definition;
    for (const auto& [key, incidences] : edges) {
        if (incidences.size() == 1) continue;  // a boundary: no bending constraint
        // UnsupportedBendingTopology diagnostics (scene_structure.cpp:311-335),
        // with the reference's message; the reference goes on collecting
        // diagnostics and then rejects the scene with the first of them.
        const std::string edgeName = std::to_string(key[1]) + "-" + std::to_string(key[2]);
        if (incidences.size() != 2)
            reject("non-manifold edge " + edgeName + " in sheet " + std::to_string(key[0]) +
                   " cannot define one fabric bending hinge");
        const HalfEdge& e0 = incidences[0];
        const HalfEdge& e1 = incidences[1];
        if (e0.from != e1.to || e0.to != e1.from)
            reject("inconsistently oriented edge " + edgeName + " in sheet " + std::to_string(key[0]) +
                   " cannot define a signed fabric bending hinge");
        const double rigidity0 = materials.at(e0.triangle->materialId)->bendingStiffness;
        const double rigidity1 = materials.at(e1.triangle->materialId)->bendingStiffness;
        if (!(rigidity0 > 0.0) || !(rigidity1 > 0.0)) continue;

        const V3 a = positions.at(e0.from), b = positions.at(e0.to);
        const V3 c = positions.at(e0.opposite), dd = positions.at(e1.opposite);
        const V3 edge = minus(b, a);
        const double edgeLength = normOf(edge);
        const V3 area0 = crossOf(edge, minus(c, a));
        const V3 area1 = crossOf(minus(dd, a), edge);
        const double area0Length = normOf(area0);
        const double area1Length = normOf(area1);
        const auto strip = [](const HalfEdge& h) {
            V2 pa{}, pb{}, pc{};
            for (int k = 0; k < 3; ++k) {
                if (h.triangle->vertexIds[k] == h.from) pa = h.triangle->chart[k];
                if (h.triangle->vertexIds[k] == h.to) pb = h.triangle->chart[k];
                if (h.triangle->vertexIds[k] == h.opposite) pc = h.triangle->chart[k];
            }
            const double ex = pb.x - pa.x, ey = pb.y - pa.y;
            const double ox = pc.x - pa.x, oy = pc.y - pa.y;
            const double length = std::hypot(ex, ey);
            const double altitude = std::abs(ex * oy - ey * ox) / length;
            return std::array<double, 2>{length, altitude};
        };
        const std::array<double, 2> s0 = strip(e0), s1 = strip(e1);
        const double compliance = s0[1] / (2.0 * rigidity0 * s0[0]) + s1[1] / (2.0 * rigidity1 * s1[0]);
        const V3 edgeUnit = times(edge, 1.0 / edgeLength);
        const V3 n0 = times(area0, 1.0 / area0Length);
        const V3 n1 = times(area1, 1.0 / area1Length);
        const double cosine = std::clamp(dotOf(n0, n1), -1.0, 1.0);
        const double sine = dotOf(edgeUnit, crossOf(n0, n1));
        const double rest = std::atan2(sine, cosine);
        if (!std::isfinite(compliance) || compliance < 0.0 || !std::isfinite(rest))
            reject("fabric bending hinge is outside the representable XPBD range");
        for (StableId v : {e0.from, e0.to, e0.opposite, e1.opposite}) d.dihedralNodes.push_back(nodeOf.at(v));
        d.dihedralRestAngle.push_back(rest);
        d.dihedralCompliance.push_back(compliance);
        out.dihedralTriangleIds.push_back({e0.triangle->id, e1.triangle->id});
    }
}

// Endpoint key of the suspension graph (scene_structure.cpp:564-590):
// kind (1 surface attachment, 2 junction, 3 harness) then stable id.
using EndpointKey = std::pair<std::uint8_t, StableId>;

EndpointKey keyOf(const Attachment& a) {
    switch (a.kind) {
    case AttachmentKind::SurfaceVertex:
        return {SW_ENDPOINT_ATTACHMENT, a.id};
    case AttachmentKind::SuspensionJunction:
        return {SW_ENDPOINT_JUNCTION, a.junctionId};
    case AttachmentKind::PilotHarness:
        return {SW_ENDPOINT_HARNESS, a.id};
    }
    throw std::logic_error("validated attachment has an unknown kind");
}

// Rigid-pilot suspension tree (scene_structure.cpp:592-822): validate the
// forest, orient every line from its canopy leaf toward its harness root,
// and list attachments/junctions/harness points in endpoint-key order and
// segments in line-id order.
void assemblePilot(const Pilot& pilot, const std::vector<const Attachment*>& attachments,
                   const std::vector<const Line*>& lines,
                   const std::unordered_map<StableId, const Junction*>& junctions,
                   const std::unordered_map<StableId, std::size_t>& nodeOf,
                   const std::unordered_map<StableId, const LineMaterial*>& lineMaterials,
                   const AssemblySettings& settings, Assembly& out) {
    if (lines.empty()) reject("rigid pilot requires a connected suspension tree");
    std::unordered_map<StableId, const Attachment*> attachmentById;
    for (const Attachment* a : attachments) attachmentById.emplace(a->id, a);
    using Neighbour = std::pair<EndpointKey, const Line*>;
    std::map<EndpointKey, std::vector<Neighbour>> adjacency;
    std::set<std::pair<EndpointKey, EndpointKey>> pairs;
    for (const Line* line : lines) {
        const EndpointKey first = keyOf(*attachmentById.at(line->startAttachment));
        const EndpointKey second = keyOf(*attachmentById.at(line->endAttachment));
        const auto canonical = std::minmax(first, second);
        if (first == second || !pairs.insert({canonical.first, canonical.second}).second)
            reject("suspension tree contains a self-edge or duplicate endpoint pair");
        adjacency[first].push_back({second, line});
        adjacency[second].push_back({first, line});
    }
    for (auto& [endpoint, neighbours] : adjacency) {
        std::sort(neighbours.begin(), neighbours.end(), [](const Neighbour& l, const Neighbour& r) {
            return std::tuple(l.first.first, l.first.second, l.second->id) <
                   std::tuple(r.first.first, r.first.second, r.second->id);
        });
        if (endpoint.first == SW_ENDPOINT_ATTACHMENT && neighbours.size() != 1)
            reject("surface suspension attachment must be a tree leaf");
        if (endpoint.first == SW_ENDPOINT_JUNCTION) {
            const Junction* j = junctions.at(endpoint.second);
            if (j->fixed || neighbours.size() < 2) reject("pilot suspension junction must be a dynamic inner node");
        }
    }
    std::set<EndpointKey> visited;
    std::map<StableId, std::pair<EndpointKey, EndpointKey>> oriented;
    for (const auto& [start, ignored] : adjacency) {
        (void)ignored;
        if (visited.count(start)) continue;
        std::vector<EndpointKey> component;
        std::queue<EndpointKey> pending;
        pending.push(start);
        visited.insert(start);
        std::size_t degreeSum = 0;
        while (!pending.empty()) {
            const EndpointKey current = pending.front();
            pending.pop();
            component.push_back(current);
            degreeSum += adjacency.at(current).size();
            for (const Neighbour& n : adjacency.at(current))
                if (visited.insert(n.first).second) pending.push(n.first);
        }
        std::vector<EndpointKey> roots;
        std::size_t leaves = 0;
        for (const EndpointKey& e : component) {
            if (e.first == SW_ENDPOINT_HARNESS) roots.push_back(e);
            else if (e.first == SW_ENDPOINT_ATTACHMENT) ++leaves;
        }
        if (roots.size() != 1 || leaves == 0 || degreeSum / 2 != component.size() - 1)
            reject("each suspension component must be a tree with one harness root and a surface leaf");
        std::set<EndpointKey> rooted{roots.front()};
        std::queue<EndpointKey> outward;
        outward.push(roots.front());
        while (!outward.empty()) {
            const EndpointKey parent = outward.front();
            outward.pop();
            for (const Neighbour& n : adjacency.at(parent)) {
                if (!rooted.insert(n.first).second) continue;
                oriented.emplace(n.second->id, std::pair{n.first, parent});
                outward.push(n.first);
            }
        }
    }
    if (oriented.size() != lines.size()) reject("suspension tree could not be oriented toward the pilot");

    SuspensionDef s;
    s.pilotStableId = pilot.id;
    s.pilotMass = pilot.mass;
    s.comWorld = {pilot.centerOfMass.x, pilot.centerOfMass.y, pilot.centerOfMass.z};
    s.velocity = {pilot.linearVelocity.x, pilot.linearVelocity.y, pilot.linearVelocity.z};
    s.orientation = {pilot.bodyToWorld.w, pilot.bodyToWorld.x, pilot.bodyToWorld.y, pilot.bodyToWorld.z};
    s.inertia = {pilot.principalInertia.x, pilot.principalInertia.y, pilot.principalInertia.z};
    s.solverIterations = settings.suspensionIterations;
    s.attachmentTolerance = settings.suspensionAttachmentTolerance;
    s.minimumLineLength = settings.suspensionMinimumLineLength;
    s.maximumLineResidual = settings.suspensionMaximumLineResidual;
    s.maximumControlWork = settings.suspensionMaximumControlWork;
    for (const auto& [endpoint, neighbours] : adjacency) {
        (void)neighbours;
        if (endpoint.first == SW_ENDPOINT_ATTACHMENT) {
            s.attachmentIds.push_back(endpoint.second);
            s.attachmentNodes.push_back(nodeOf.at(attachmentById.at(endpoint.second)->vertexId));
        } else if (endpoint.first == SW_ENDPOINT_JUNCTION) {
            s.junctionIds.push_back(endpoint.second);
            s.junctionNodes.push_back(nodeOf.at(endpoint.second));
        } else {
            const Attachment* a = attachmentById.at(endpoint.second);
            if (a->pilotId != pilot.id) reject("harness root belongs to a different pilot");
            s.harnessIds.push_back(endpoint.second);
            s.harnessLocal.insert(s.harnessLocal.end(), {a->pilotLocal.x, a->pilotLocal.y, a->pilotLocal.z});
        }
    }
    for (const Line* line : lines) {
        const auto& o = oriented.at(line->id);
        s.segmentIds.push_back(line->id);
        s.segmentFromKind.push_back(o.first.first);
        s.segmentFromId.push_back(o.first.second);
        s.segmentToKind.push_back(o.second.first);
        s.segmentToId.push_back(o.second.second);
        s.segmentRest.push_back(line->restLength);
        s.segmentStiffness.push_back(lineMaterials.at(line->materialId)->axialStiffness);
        s.segmentDamping.push_back(0.0);
        s.segmentRole.push_back(static_cast<std::uint32_t>(line->role));
    }
    out.definition.suspension = std::move(s);
}

// Seams (scene_structure.cpp:440-562): zero-rest stitches for distinct
// pairs, then one half-rigidity rail per chain segment (chain 0, chain 1),
// and each centreline segment's mass split over its four nodes.
void assembleSeams(const std::vector<const Seam*>& seams,
                   const std::unordered_map<StableId, const SeamMaterial*>& materials,
                   const std::unordered_map<StableId, V3>& positions,
                   const std::unordered_map<StableId, std::size_t>& nodeOf, const AssemblySettings& settings,
                   std::vector<long double>& lumped, long double& totalSeamMass, Assembly& out) {
    Definition& d = out.definition;
    for (const Seam* seam : seams) {
        const SeamMaterial* material = materials.at(seam->materialId);
        const std::size_t count = seam->first.size();
        for (std::size_t k = 0; k < count; ++k) {
            const double separation = normOf(minus(positions.at(seam->second[k]), positions.at(seam->first[k])));
            if (!std::isfinite(separation) || separation > settings.seamCoincidenceTolerance)
                reject("paired seam vertices are not coincident within the structural tolerance");
        }
        std::array<std::vector<double>, 2> lengths;
        for (int chain = 0; chain < 2; ++chain) {
            const auto& ids = chain == 0 ? seam->first : seam->second;
            for (std::size_t k = 0; k + 1 < count; ++k) {
                const double length = normOf(minus(positions.at(ids[k + 1]), positions.at(ids[k])));
                const double compliance = 2.0 * length / material->axialStiffness;
                if (!(length > 0.0) || !std::isfinite(length) || !std::isfinite(compliance))
                    reject("seam segment stiffness cannot be represented as finite XPBD compliance");
                lengths[chain].push_back(length);
            }
        }
        for (std::size_t k = 0; k < count; ++k) {
            if (seam->first[k] == seam->second[k]) continue;
            d.constraintKind.push_back(SW_CONSTRAINT_SEAM_STITCH);
            d.constraintNodes.push_back(nodeOf.at(seam->first[k]));
            d.constraintNodes.push_back(nodeOf.at(seam->second[k]));
            d.constraintRest.push_back(0.0);
            d.constraintCompliance.push_back(0.0);
        }
        for (int chain = 0; chain < 2; ++chain) {
            const auto& ids = chain == 0 ? seam->first : seam->second;
            for (std::size_t k = 0; k + 1 < count; ++k) {
                const double length = lengths[chain][k];
                d.constraintKind.push_back(SW_CONSTRAINT_DISTANCE);
                d.constraintNodes.push_back(nodeOf.at(ids[k]));
                d.constraintNodes.push_back(nodeOf.at(ids[k + 1]));
                d.constraintRest.push_back(length);
                d.constraintCompliance.push_back(2.0 * length / material->axialStiffness);
            }
        }
        for (std::size_t k = 0; k + 1 < count; ++k) {
            const double centreline = 0.5 * (lengths[0][k] + lengths[1][k]);
            const double segmentMass = centreline * material->linearDensity;
            const long double share = static_cast<long double>(segmentMass) / 4.0L;
            for (StableId id : {seam->first[k], seam->first[k + 1], seam->second[k], seam->second[k + 1]})
                lumped[nodeOf.at(id)] += share;
            totalSeamMass += static_cast<long double>(segmentMass);
        }
    }
}

}  // namespace

void validateSceneStructure(const Scene& scene) { validate(scene); }

Assembly assembleScene(const Scene& scene, const AssemblySettings& settings,
                       const std::vector<ContactSphere>& spheres) {
    validate(scene);
    if (!(settings.seamCoincidenceTolerance >= 0.0) || !std::isfinite(settings.seamCoincidenceTolerance))
        reject("invalid seam coincidence tolerance");
    if (settings.suspensionIterations <= 0 || !(settings.suspensionAttachmentTolerance > 0.0) ||
        !(settings.suspensionMinimumLineLength > 0.0) || settings.suspensionMaximumLineResidual < 0.0 ||
        settings.suspensionMaximumControlWork < 0.0)
        reject("invalid suspension solver settings");
    if (scene.pilots.size() > 1) reject("one Structure instance supports exactly one rigid pilot");

    std::unordered_set<StableId> fabricVertices;
    for (const Triangle& t : scene.triangles) fabricVertices.insert(t.vertexIds.begin(), t.vertexIds.end());
    for (const Seam& seam : scene.seams) {
        for (StableId v : seam.first)
            if (!fabricVertices.count(v)) reject("seam vertices must belong to fabric triangles");
        for (StableId v : seam.second)
            if (!fabricVertices.count(v)) reject("seam vertices must belong to fabric triangles");
    }
    const auto attachments = byId(scene.attachments);
    std::unordered_map<StableId, const Attachment*> attachmentById;
    for (const Attachment* a : attachments) attachmentById.emplace(a->id, a);
    const auto lines = byId(scene.lines);
    for (const Line* line : lines) {
        const Attachment* start = attachmentById.at(line->startAttachment);
        const Attachment* end = attachmentById.at(line->endAttachment);
        const bool harness = start->kind == AttachmentKind::PilotHarness || end->kind == AttachmentKind::PilotHarness;
        if (scene.pilots.empty() && harness) reject("a pilot harness line requires exactly one rigid pilot");
        if (!harness) {
            const StableId s0 = start->kind == AttachmentKind::SurfaceVertex ? start->vertexId : start->junctionId;
            const StableId s1 = end->kind == AttachmentKind::SurfaceVertex ? end->vertexId : end->junctionId;
            if (start->kind == end->kind && s0 == s1)
                reject("suspension line endpoints must be distinct structural nodes");
        }
    }

    // Nodes: structural vertices (used by a triangle or a surface
    // attachment) in id order, then every junction in id order
    // (scene_structure.cpp:1058-1094).
    std::set<StableId> structural(fabricVertices.begin(), fabricVertices.end());
    for (const Attachment* a : attachments)
        if (a->kind == AttachmentKind::SurfaceVertex) structural.insert(a->vertexId);
    Assembly out;
    Definition& d = out.definition;
    std::unordered_map<StableId, std::size_t> nodeOf;
    std::unordered_map<StableId, V3> positions;
    for (const Vertex* v : byId(scene.vertices)) {
        if (!structural.count(v->id)) continue;
        nodeOf.emplace(v->id, out.nodeVertexIds.size());
        positions.emplace(v->id, v->position);
        out.nodeVertexIds.push_back(v->id);
        d.nodePosition.insert(d.nodePosition.end(), {v->position.x, v->position.y, v->position.z});
        d.nodeMass.push_back(0.0);
        d.nodeFixed.push_back(0);
    }
    std::unordered_map<StableId, const Junction*> junctionById;
    for (const Junction* j : byId(scene.junctions)) {
        nodeOf.emplace(j->id, d.nodeMass.size());
        junctionById.emplace(j->id, j);
        out.nodeJunctionIds.push_back(j->id);
        d.nodePosition.insert(d.nodePosition.end(), {j->position.x, j->position.y, j->position.z});
        d.nodeMass.push_back(j->mass);
        d.nodeFixed.push_back(j->fixed ? 1 : 0);
    }

    std::unordered_map<StableId, const FabricMaterial*> materials;
    for (const FabricMaterial* m : byId(scene.fabricMaterials)) {
        materials.emplace(m->id, m);
        if (!invertibleMaterial(*m)) reject("fabric stiffness is too ill-conditioned for the XPBD membrane inverse");
    }

    // Triangles, one Bulk membrane each, and long-double lumped fabric mass
    // (scene_structure.cpp:1114-1172).
    std::vector<long double> lumped(out.nodeVertexIds.size(), 0.0L);
    long double totalFabric = 0.0L;
    long double totalSeam = 0.0L;
    const auto triangles = byId(scene.triangles);
    for (const Triangle* t : triangles) {
        if (!usableChart(*t)) reject("triangle material coordinates are singular or ill-conditioned");
        const std::size_t index = d.triangleCount();
        std::array<std::size_t, 3> nodes{};
        for (int k = 0; k < 3; ++k) {
            nodes[k] = nodeOf.at(t->vertexIds[k]);
            d.triangleNodes.push_back(nodes[k]);
        }
        out.triangleIds.push_back(t->id);
        const FabricMaterial* m = materials.at(t->materialId);
        d.membraneTriangle.push_back(index);
        for (int k = 0; k < 3; ++k) d.membraneChart.insert(d.membraneChart.end(), {t->chart[k].x, t->chart[k].y});
        d.membraneMaterial.insert(d.membraneMaterial.end(),
                                  {m->warpStiffness, m->weftStiffness, 0.0, m->shearStiffness, 0.0, 0.0,
                                   m->dampingSeconds, 1.0});
        d.membraneRole.push_back(SW_ROLE_BULK);
        const double mass = chartArea(*t) * m->arealDensity;
        if (!(mass > 0.0) || !std::isfinite(mass)) reject("triangle fabric mass is non-finite or underflows to zero");
        const long double share = static_cast<long double>(mass) / 3.0L;
        for (std::size_t node : nodes) lumped[node] += share;
        totalFabric += static_cast<long double>(mass);
    }
    assembleHinges(triangles, positions, materials, nodeOf, out);

    std::unordered_map<StableId, const LineMaterial*> lineMaterials;
    for (const LineMaterial* m : byId(scene.lineMaterials)) lineMaterials.emplace(m->id, m);
    if (scene.pilots.empty()) {
        for (const Line* line : lines) {
            const Attachment* start = attachmentById.at(line->startAttachment);
            const Attachment* end = attachmentById.at(line->endAttachment);
            const LineMaterial* m = lineMaterials.at(line->materialId);
            const double compliance = line->restLength / m->axialStiffness;
            if (!std::isfinite(compliance) || compliance < 0.0)
                reject("line stiffness cannot be represented as finite XPBD compliance");
            d.constraintKind.push_back(SW_CONSTRAINT_CABLE);
            d.constraintNodes.push_back(
                nodeOf.at(start->kind == AttachmentKind::SurfaceVertex ? start->vertexId : start->junctionId));
            d.constraintNodes.push_back(
                nodeOf.at(end->kind == AttachmentKind::SurfaceVertex ? end->vertexId : end->junctionId));
            d.constraintRest.push_back(line->restLength);
            d.constraintCompliance.push_back(compliance);
        }
    } else {
        assemblePilot(scene.pilots.front(), attachments, lines, junctionById, nodeOf, lineMaterials, settings, out);
    }

    std::unordered_map<StableId, const SeamMaterial*> seamMaterials;
    for (const SeamMaterial* m : byId(scene.seamMaterials)) seamMaterials.emplace(m->id, m);
    assembleSeams(byId(scene.seams), seamMaterials, positions, nodeOf, settings, lumped, totalSeam, out);

    for (std::size_t i = 0; i < lumped.size(); ++i) {
        const double mass = static_cast<double>(lumped[i]);
        if (!(mass > 0.0) || !std::isfinite(mass))
            reject("scene vertex has no finite positive lumped fabric or seam mass");
        d.nodeMass[i] = mass;
    }
    out.totalFabricMass = static_cast<double>(totalFabric);
    out.totalSeamMass = static_cast<double>(totalSeam);
    if (!std::isfinite(out.totalFabricMass) || !std::isfinite(out.totalSeamMass))
        reject("total mass exceeds the representable structure range");

    for (const ContactSphere& sphere : spheres) {
        d.sphereCentre.push_back(nodeOf.at(sphere.centreVertex));
        d.sphereRadius.push_back(sphere.radius);
        d.sphereCompliance.push_back(settings.contactSphereCompliance);
        for (StableId v : sphere.vertices) d.sphereNodes.push_back(nodeOf.at(v));
        d.sphereOffsets.push_back(d.sphereNodes.size());
    }
    return out;
}

}  // namespace swx
