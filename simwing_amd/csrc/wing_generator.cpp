// Seeded synthetic ram-air wing (the benchmark workload generator).
//
// Geometry: an elliptic-arc planform of span `span` with an elliptic chord
// distribution (floored at `tipChordFloor`), a straight quarter-chord line
// and a closed-trailing-edge NACA 4-digit symmetric section. The upper and
// lower skins are structured grids sharing their leading- and trailing-edge
// rows (welded across sheets, so no hinge forms there). Every `ribEvery`
// span stations a rib sheet with its own vertices (upper and lower rows, one
// triangle strip) is sewn to both skins by two seams. Suspension: `lineCount` lines of
// `segmentsPerLine` segments from lower-skin attachment points on the line
// ribs to one of two harness points of a rigid pilot.
//
// Stable-ID layout (it fixes node, triangle and segment order):
//  * vertices are numbered station by station: the station's skin vertices
//    (upper row then lower interior row), then its rib vertices if any;
//  * triangles are numbered station strip by station strip (upper skin,
//    lower skin, then the rib at the strip's first station);
//  * suspension line IDs are (1 + hop) * 100000 + line, so every segment
//    string "segment:<id>" has six digits and the reference's string order
//    is hop-major: all first hops, then all second hops, ... then the hops
//    that reach the pilot.
#include "scene.hpp"

#include <cmath>
#include <cstddef>
#include <stdexcept>

namespace swx {

double seriesExp(double x) {
    if (!(x == x)) return x;
    int halvings = 0;
    double r = x;
    while (r > 0x1.0p-8 || r < -0x1.0p-8) {
        r *= 0.5;
        ++halvings;
        if (halvings > 2000) break;
    }
    double term = 1.0;
    double sum = 1.0;
    for (int k = 1; k < 30; ++k) {
        term = term * r / static_cast<double>(k);
        sum += term;
        if (term < 1.0e-18 * sum && term > -1.0e-18 * sum) break;
    }
    for (int i = 0; i < halvings; ++i) sum *= sum;
    return sum;
}

namespace {

// Natural log by Newton iteration on seriesExp (x > 0).
double seriesLog(double x) {
    double y = 0.0;
    double scaled = x;
    while (scaled > 2.0) {
        scaled *= 0.5;
        y += 0.69314718055994530942;
    }
    while (scaled < 0.5) {
        scaled *= 2.0;
        y -= 0.69314718055994530942;
    }
    double z = 0.0;
    for (int it = 0; it < 60; ++it) {
        const double e = seriesExp(z);
        const double next = z + 2.0 * (scaled - e) / (scaled + e);
        if (next == z) break;
        z = next;
    }
    return y + z;
}

V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 crossv(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
double dotv(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double lengthv(const V3& a) { return std::sqrt(dotv(a, a)); }

// Approximate standard normal draw: Irwin-Hall sum of 12 uniforms minus 6.
double normalDraw(SplitMix64& rng) {
    double sum = 0.0;
    for (int k = 0; k < 12; ++k) sum += rng.uniform();
    return sum - 6.0;
}

// Isometric material chart of a triangle: a at the origin, b on +x.
std::array<V2, 3> flatChart(const V3& a, const V3& b, const V3& c) {
    const V3 ab = sub(b, a);
    const V3 ac = sub(c, a);
    const double lab = lengthv(ab);
    const V3 n = crossv(ab, ac);
    return {V2{0.0, 0.0}, V2{lab, 0.0},
            V2{dotv(ac, ab) / lab, lengthv(n) / lab}};
}

struct Builder {
    GeneratedWing out;
    std::vector<V3> positionById{V3{}};  // index = vertex id (id 0 unused)

    StableId addVertex(const V3& p) {
        const StableId id = positionById.size();
        positionById.push_back(p);
        out.scene.vertices.push_back({id, p});
        return id;
    }
};

}  // namespace

GeneratedWing generateWing(const WingParams& P) {
    if (P.spanNodes < 3 || P.chordNodes < 4 || P.ribEvery < 1 ||
        P.segmentsPerLine < 2 || P.segmentsPerLine > 8 || P.lineCount < 1 ||
        P.lineCount >= 100000 || P.lineRows < 1 || P.lineRows > 4 ||
        !(P.span > 0.0) || !(P.rootChord > 0.0) || !(P.thickness > 0.0) ||
        !(P.complianceMin > 0.0) || !(P.complianceMax >= P.complianceMin) ||
        !(P.lineRestScale > 0.0) || P.contactSpheres < 0) {
        throw std::invalid_argument("invalid synthetic wing parameters");
    }
    SplitMix64 rng(P.seed);
    Builder B;
    Scene& S = B.out.scene;
    const int nS = P.spanNodes;
    const int nC = P.chordNodes;

    // Stations.
    std::vector<double> ys(nS), zs(nS), chords(nS), les(nS);
    for (int i = 0; i < nS; ++i) {
        const double s = -1.0 + 2.0 * static_cast<double>(i) / static_cast<double>(nS - 1);
        double root = 1.0 - s * s;
        if (root < 0.0) root = 0.0;
        const double elliptic = P.rootChord * std::sqrt(root);
        chords[i] = elliptic > P.tipChordFloor ? elliptic : P.tipChordFloor;
        ys[i] = 0.5 * P.span * s;
        // Circular arc about the pilot: radius = line length, so lines meet
        // the canopy at right angles.
        const double radius = P.lineLength;
        zs[i] = std::sqrt(radius * radius - ys[i] * ys[i]) - radius;
        les[i] = 0.25 * (P.rootChord - chords[i]);
    }
    // Uniform chordwise stations: no clustering, so no sub-millimetre
    // leading-edge elements with vanishing lumped mass.
    std::vector<double> us(nC);
    for (int j = 0; j < nC; ++j) us[j] = static_cast<double>(j) / static_cast<double>(nC - 1);
    const auto halfThickness = [&](double chord, double u) {
        const double su = std::sqrt(u);
        return 5.0 * P.thickness * chord *
               (0.2969 * su - 0.1260 * u - 0.3516 * u * u + 0.2843 * u * u * u -
                0.1036 * u * u * u * u);
    };
    const auto surfacePoint = [&](int i, int j, int side) {
        const double x = les[i] + us[j] * chords[i];
        double dz = 0.0;
        if (j != 0 && j != nC - 1 && side != 0) {
            dz = static_cast<double>(side) * halfThickness(chords[i], us[j]);
        }
        return V3{x, ys[i], zs[i] + dz};
    };

    // Rib stations.
    std::vector<int> ribOfStation(nS, -1);
    std::vector<int> ribStations;
    for (int i = 0; i < nS; i += P.ribEvery) {
        ribOfStation[i] = static_cast<int>(ribStations.size());
        ribStations.push_back(i);
    }
    // Both wing tips are closed by a rib (no free, unsupported skin edge).
    if (ribOfStation[nS - 1] < 0) {
        ribOfStation[nS - 1] = static_cast<int>(ribStations.size());
        ribStations.push_back(nS - 1);
    }
    const int nRibs = static_cast<int>(ribStations.size());

    // Regions: one outside region, one cell per inter-rib bay.
    S.regions.push_back({1, RegionKind::Outside, "outside"});
    for (int k = 0; k <= nRibs; ++k) {
        S.regions.push_back({static_cast<StableId>(2 + k), RegionKind::Cell,
                             "cell-" + std::to_string(k)});
    }
    const auto cellOfStation = [&](int i) {
        int k = 0;
        for (int r = 0; r < nRibs; ++r) {
            if (ribStations[r] <= i) k = r + 1;
        }
        return static_cast<StableId>(2 + k);
    };

    // Vertices, station by station.
    std::vector<std::vector<StableId>> upper(nS, std::vector<StableId>(nC, 0));
    std::vector<std::vector<StableId>> lower(nS, std::vector<StableId>(nC, 0));
    struct RibVertices {
        std::vector<StableId> up, low;
    };
    std::vector<RibVertices> ribs(nRibs);
    for (int i = 0; i < nS; ++i) {
        for (int j = 0; j < nC; ++j) upper[i][j] = B.addVertex(surfacePoint(i, j, +1));
        for (int j = 1; j < nC - 1; ++j) lower[i][j] = B.addVertex(surfacePoint(i, j, -1));
        lower[i][0] = upper[i][0];
        lower[i][nC - 1] = upper[i][nC - 1];
        const int r = ribOfStation[i];
        if (r >= 0) {
            RibVertices& rv = ribs[r];
            rv.up.assign(nC, 0);
            rv.low.assign(nC, 0);
            for (int j = 0; j < nC; ++j) rv.up[j] = B.addVertex(surfacePoint(i, j, +1));
            for (int j = 1; j < nC - 1; ++j) rv.low[j] = B.addVertex(surfacePoint(i, j, -1));
            rv.low[0] = rv.up[0];
            rv.low[nC - 1] = rv.up[nC - 1];
        }
    }

    // Triangles.
    const double logRatio = seriesLog(P.complianceMax / P.complianceMin);
    const auto addTriangle = [&](StableId a, StableId b, StableId c, StableId sheet,
                                 SurfaceRole role, StableId negative, StableId positive,
                                 double pressure, const V3& outward) {
        const V3& pa = B.positionById[a];
        V3 pb = B.positionById[b];
        V3 pc = B.positionById[c];
        if (dotv(crossv(sub(pb, pa), sub(pc, pa)), outward) < 0.0) {
            std::swap(b, c);
            std::swap(pb, pc);
        }
        Triangle t;
        t.id = S.triangles.size() + 1;
        t.vertexIds = {a, b, c};
        t.chart = flatChart(pa, pb, pc);
        t.negativeRegion = negative;
        t.positiveRegion = positive;
        t.materialId = t.id;
        t.sheetId = sheet;
        t.role = role;
        S.triangles.push_back(t);

        const double compliance = P.complianceMin * seriesExp(rng.uniform() * logRatio);
        FabricMaterial m;
        m.id = t.id;
        m.name = "fabric-" + std::to_string(t.id);
        m.warpStiffness = 1.0 / compliance;
        m.weftStiffness = 1.0 / compliance;
        m.shearStiffness = 0.5 / compliance;
        m.bendingStiffness = P.bendingRigidity;
        m.arealDensity = P.fabricDensity;
        S.fabricMaterials.push_back(m);

        double p = pressure;
        if (pressure != 0.0 && P.pressureStd > 0.0) {
            p = pressure + P.pressureStd * normalDraw(rng);
            if (p < 0.0) p = 0.0;
        }
        B.out.facePressure.push_back(p);
    };
    const V3 up{0.0, 0.0, 1.0};
    const V3 down{0.0, 0.0, -1.0};
    const V3 right{0.0, 1.0, 0.0};
    // Rib: one strip between its upper and lower rows (every rib vertex is
    // sewn to a skin, so no rib node is free to flap out of plane).
    const auto addRib = [&](int r) {
        const RibVertices& rv = ribs[r];
        const StableId sheet = 3 + static_cast<StableId>(r);
        const StableId left = r == 0 ? 1 : static_cast<StableId>(2 + r);
        const StableId rightCell = static_cast<StableId>(3 + r);
        for (int j = 0; j < nC - 1; ++j) {
            const StableId u0 = rv.up[j], u1 = rv.up[j + 1];
            const StableId l0 = rv.low[j], l1 = rv.low[j + 1];
            if (j == 0) {
                addTriangle(u0, u1, l1, sheet, SurfaceRole::Rib, left, rightCell, 0.0, right);
            } else if (j == nC - 2) {
                addTriangle(u0, u1, l0, sheet, SurfaceRole::Rib, left, rightCell, 0.0, right);
            } else {
                addTriangle(u0, u1, l1, sheet, SurfaceRole::Rib, left, rightCell, 0.0, right);
                addTriangle(u0, l1, l0, sheet, SurfaceRole::Rib, left, rightCell, 0.0, right);
            }
        }
    };
    for (int i = 0; i + 1 < nS; ++i) {
        const StableId cell = cellOfStation(i);
        for (int j = 0; j + 1 < nC; ++j) {
            addTriangle(upper[i][j], upper[i][j + 1], upper[i + 1][j + 1], 1,
                        SurfaceRole::Skin, cell, 1, P.pressureMean, up);
            addTriangle(upper[i][j], upper[i + 1][j + 1], upper[i + 1][j], 1,
                        SurfaceRole::Skin, cell, 1, P.pressureMean, up);
        }
        for (int j = 0; j + 1 < nC; ++j) {
            addTriangle(lower[i][j], lower[i][j + 1], lower[i + 1][j + 1], 2,
                        SurfaceRole::Skin, cell, 1, P.pressureMean, down);
            addTriangle(lower[i][j], lower[i + 1][j + 1], lower[i + 1][j], 2,
                        SurfaceRole::Skin, cell, 1, P.pressureMean, down);
        }
        if (ribOfStation[i] >= 0) addRib(ribOfStation[i]);
    }
    if (ribOfStation[nS - 1] >= 0) addRib(ribOfStation[nS - 1]);

    // Seams: rib upper row to the upper skin row (leading and trailing edge
    // pairs included), rib lower interior row to the lower skin row.
    S.seamMaterials.push_back({1, "seam-tape", P.seamLinearDensity, P.seamAxialStiffness});
    for (int r = 0; r < nRibs; ++r) {
        const int i = ribStations[r];
        Seam upperSeam;
        upperSeam.id = 2 * static_cast<StableId>(r) + 1;
        upperSeam.materialId = 1;
        for (int j = 0; j < nC; ++j) {
            upperSeam.first.push_back(upper[i][j]);
            upperSeam.second.push_back(ribs[r].up[j]);
        }
        Seam lowerSeam;
        lowerSeam.id = 2 * static_cast<StableId>(r) + 2;
        lowerSeam.materialId = 1;
        for (int j = 1; j < nC - 1; ++j) {
            lowerSeam.first.push_back(lower[i][j]);
            lowerSeam.second.push_back(ribs[r].low[j]);
        }
        S.seams.push_back(std::move(upperSeam));
        S.seams.push_back(std::move(lowerSeam));
    }

    // Suspension.
    const double rowFractions[4] = {0.12, 0.38, 0.64, 0.85};
    const int perRow = (P.lineCount + P.lineRows - 1) / P.lineRows;
    Pilot pilot;
    pilot.id = 1;
    pilot.name = "pilot";
    pilot.mass = P.pilotMass;
    // Lines `lineRestScale` times their nominal length: the pilot hangs that
    // much lower (or higher), so every line starts at its rest length and no
    // wing is released with its lines pre-stretched.
    pilot.centerOfMass = V3{0.35 * P.rootChord, 0.0, -P.lineLength - 0.35 - (P.lineRestScale - 1.0) * P.lineLength};
    pilot.principalInertia = P.pilotInertia;
    // Tethered wing: no rigid pilot; the lines end at two fixed riser
    // junctions where the harness points would be (the scene then carries
    // its lines as load-path Cable constraints).
    if (!P.tether) S.pilots.push_back(pilot);
    S.lineMaterials.push_back({1, "line", 0.0015, P.lineLinearDensity, P.lineAxialStiffness, 1.0});

    const StableId junctionBase = B.positionById.size() + 1000;
    const StableId harnessLeft = static_cast<StableId>(P.lineCount) + 1;
    const StableId harnessRight = static_cast<StableId>(P.lineCount) + 2;
    const V3 localLeft{0.0, -0.22, 0.35};
    const V3 localRight{0.0, 0.22, 0.35};
    Attachment hl;
    hl.id = harnessLeft;
    hl.kind = AttachmentKind::PilotHarness;
    hl.pilotId = 1;
    hl.pilotLocal = localLeft;
    Attachment hr = hl;
    hr.id = harnessRight;
    hr.pilotLocal = localRight;
    const StableId riserBase =
        junctionBase + static_cast<StableId>(P.lineCount) * static_cast<StableId>(P.segmentsPerLine) + 1;
    if (P.tether) {
        hl.kind = AttachmentKind::SuspensionJunction;
        hl.pilotId = 0;
        hl.pilotLocal = V3{};
        hl.junctionId = riserBase;
        hr.kind = AttachmentKind::SuspensionJunction;
        hr.pilotId = 0;
        hr.pilotLocal = V3{};
        hr.junctionId = riserBase + 1;
    }
    std::vector<Attachment> surfaceAttachments, junctionAttachments;
    StableId nextJunctionAttachment = static_cast<StableId>(P.lineCount) + 3;
    const int segs = P.segmentsPerLine;
    std::vector<std::vector<StableId>> lineNodes(P.lineCount);  // attachment ids per hop
    std::vector<std::vector<V3>> linePoints(P.lineCount);
    for (int l = 0; l < P.lineCount; ++l) {
        const int row = l % P.lineRows;
        const int pick = l / P.lineRows;
        const int r = perRow > 1 ? static_cast<int>(
                                        (static_cast<long long>(pick) * (nRibs - 1) +
                                         (perRow - 1) / 2) /
                                        (perRow - 1))
                                  : nRibs / 2;
        const int i = ribStations[r];
        int bestJ = 1;
        double bestD = 1.0e300;
        for (int j = 1; j < nC - 1; ++j) {
            const double d = us[j] > rowFractions[row] ? us[j] - rowFractions[row]
                                                       : rowFractions[row] - us[j];
            if (d < bestD) {
                bestD = d;
                bestJ = j;
            }
        }
        const StableId vertex = lower[i][bestJ];
        const V3 a = B.positionById[vertex];
        const bool isLeft = ys[i] < 0.0;
        const V3 local = isLeft ? localLeft : localRight;
        const V3 h{pilot.centerOfMass.x + local.x, pilot.centerOfMass.y + local.y,
                   pilot.centerOfMass.z + local.z};
        Attachment sa;
        sa.id = static_cast<StableId>(l) + 1;
        sa.kind = AttachmentKind::SurfaceVertex;
        sa.vertexId = vertex;
        surfaceAttachments.push_back(sa);
        lineNodes[l].push_back(sa.id);
        linePoints[l].push_back(a);
        for (int k = 1; k < segs; ++k) {
            const double f = static_cast<double>(k) / static_cast<double>(segs);
            const V3 p{a.x + f * (h.x - a.x), a.y + f * (h.y - a.y), a.z + f * (h.z - a.z)};
            Junction jn;
            jn.id = junctionBase + static_cast<StableId>(l) * static_cast<StableId>(segs) +
                    static_cast<StableId>(k);
            jn.position = p;
            S.junctions.push_back(jn);
            Attachment ja;
            ja.id = nextJunctionAttachment++;
            ja.kind = AttachmentKind::SuspensionJunction;
            ja.junctionId = jn.id;
            junctionAttachments.push_back(ja);
            lineNodes[l].push_back(ja.id);
            linePoints[l].push_back(p);
        }
        lineNodes[l].push_back(isLeft ? harnessLeft : harnessRight);
        linePoints[l].push_back(h);
    }
    // Reinforcement patches: the fabric of every triangle within two rings
    // of a line attachment is `attachmentReinforcement` times denser (the
    // load tape of a real attachment point), so the attachment nodes are not
    // the lightest nodes of the mesh while carrying the line loads.
    if (P.attachmentReinforcement != 1.0) {
        std::vector<std::vector<std::size_t>> trianglesOf(B.positionById.size());
        for (std::size_t t = 0; t < S.triangles.size(); ++t)
            for (StableId v : S.triangles[t].vertexIds) trianglesOf[v].push_back(t);
        std::vector<char> reinforced(S.triangles.size(), 0);
        for (const Attachment& a : surfaceAttachments) {
            std::vector<StableId> ring{a.vertexId};
            for (std::size_t t : trianglesOf[a.vertexId])
                for (StableId v : S.triangles[t].vertexIds) ring.push_back(v);
            for (StableId v : ring)
                for (std::size_t t : trianglesOf[v]) reinforced[t] = 1;
        }
        for (std::size_t t = 0; t < S.triangles.size(); ++t)
            if (reinforced[t]) S.fabricMaterials[t].arealDensity *= P.attachmentReinforcement;
    }
    // Junction masses: one segment of line each, nudged upward until the
    // reciprocal of its reciprocal inverse mass is exact (the reference
    // re-derives a junction's mass from the node's inverse mass).
    for (int l = 0; l < P.lineCount; ++l) {
        for (int k = 1; k < segs; ++k) {
            Junction& jn = S.junctions[static_cast<std::size_t>(l) * (segs - 1) + (k - 1)];
            const double segment = lengthv(sub(linePoints[l][k + 1], linePoints[l][k]));
            double mass = P.lineLinearDensity * segment;
            if (mass < 1.0e-4) mass = 1.0e-4;
            for (int guard = 0; guard < 64; ++guard) {
                const double w = 1.0 / mass;
                if (1.0 / (1.0 / w) == w) break;
                mass = std::nextafter(mass, 1.0e300);
            }
            jn.mass = mass;
        }
    }
    if (P.tether) {
        for (int side = 0; side < 2; ++side) {
            const V3& local = side == 0 ? localLeft : localRight;
            Junction riser;
            riser.id = riserBase + static_cast<StableId>(side);
            riser.position = V3{pilot.centerOfMass.x + local.x, pilot.centerOfMass.y + local.y,
                                pilot.centerOfMass.z + local.z};
            riser.mass = 0.5 * P.pilotMass;
            riser.fixed = P.tether == 1;
            S.junctions.push_back(riser);
        }
    }
    for (const Attachment& a : surfaceAttachments) S.attachments.push_back(a);
    S.attachments.push_back(hl);
    S.attachments.push_back(hr);
    for (const Attachment& a : junctionAttachments) S.attachments.push_back(a);
    for (int hop = 0; hop < segs; ++hop) {
        for (int l = 0; l < P.lineCount; ++l) {
            Line line;
            line.id = static_cast<StableId>(1 + hop) * 100000 + static_cast<StableId>(l);
            line.startAttachment = lineNodes[l][hop];
            line.endAttachment = lineNodes[l][hop + 1];
            line.materialId = 1;
            line.restLength = lengthv(sub(linePoints[l][hop + 1], linePoints[l][hop]));
            line.role = LineRole::Suspension;
            S.lines.push_back(line);
        }
    }

    // Contact spheres: centred on evenly spaced lower-skin attachment
    // vertices, holding off the upper-skin vertices within four radii (the
    // skins must not collapse through the line attachment). The radius is
    // capped below the closest candidate's rest distance, so every sphere
    // starts inactive.
    for (int c = 0; c < P.contactSpheres && c < P.lineCount; ++c) {
        ContactSphere sphere;
        sphere.centreVertex =
            surfaceAttachments[static_cast<std::size_t>(c) * (P.lineCount / P.contactSpheres)].vertexId;
        const V3 centre = B.positionById[sphere.centreVertex];
        double closest = 1.0e300;
        for (int i = 0; i < nS; ++i) {
            for (int j = 1; j + 1 < nC; ++j) {
                const StableId v = upper[i][j];
                const double d = lengthv(sub(B.positionById[v], centre));
                if (d < 4.0 * P.contactRadius) {
                    sphere.vertices.push_back(v);
                    if (d < closest) closest = d;
                }
            }
        }
        sphere.radius = P.contactRadius < 0.8 * closest ? P.contactRadius : 0.8 * closest;
        B.out.contactSpheres.push_back(std::move(sphere));
    }
    return std::move(B.out);
}

}  // namespace swx

namespace swx {

WingParams wingPreset(const std::string& name) {
    WingParams p;
    if (name == "small") return p;
    if (name == "tiny") {
        p.spanNodes = 6;
        p.chordNodes = 5;
        p.span = 2.0;
        p.rootChord = 0.5;
        p.lineRows = 2;
        p.lineCount = 4;
        p.segmentsPerLine = 3;
        p.lineLength = 2.0;
        p.pilotMass = 20.0;
        p.pilotInertia = V3{2.0, 2.0, 2.0};
        p.seed = 0x5EED0000ULL + 100;
        return p;
    }
    if (name == "export") {
        p.seed = 0x5EED0001ULL;
        p.spanNodes = 176;
        p.chordNodes = 82;
        p.lineRows = 2;
        p.lineCount = 50;
        p.pressureMean = 60.0;
        p.pressureStd = 15.0;
        p.contactSpheres = 8;
        return p;
    }
    if (name == "tiny-tethered" || name == "small-tethered") {
        // The same wings without a rigid pilot, their lines ending at two
        // fixed riser junctions (load-path cables).
        WingParams t = wingPreset(name == "tiny-tethered" ? "tiny" : "small");
        t.tether = 1;
        return t;
    }
    if (name == "tiny-free" || name == "small-free") {
        // ... ending at two dynamic riser junctions of half the pilot's
        // mass each: every node is dynamic and there is no suspension, so
        // the reference restores the free membrane's momentum.
        WingParams t = wingPreset(name == "tiny-free" ? "tiny" : "small");
        t.tether = 2;
        return t;
    }
    if (name == "small-stiff") {
        // The cfg0 wing with reinforced (stiff) hinges, 1 N m of bending
        // rigidity: under the flight loads its hinges stay within 0.75 rad
        // of rest and the motion settles, so last-bit differences stay at
        // roundoff over hundreds of steps (the long-horizon parity wing).
        WingParams t = wingPreset("small");
        t.bendingRigidity = 1.0;
        return t;
    }
    if (name == "variant10k") {
        p.seed = 0x5EED0003ULL;
        p.spanNodes = 64;
        p.chordNodes = 34;
        return p;
    }
    throw std::invalid_argument("unknown wing preset: " + name);
}

void applyOverrides(WingParams& p, const std::string& overrides) {
    std::size_t start = 0;
    while (start < overrides.size()) {
        std::size_t end = overrides.find('&', start);
        if (end == std::string::npos) end = overrides.size();
        const std::string item = overrides.substr(start, end - start);
        start = end + 1;
        const std::size_t eq = item.find('=');
        if (eq == std::string::npos) throw std::invalid_argument("override must be key=value: " + item);
        const std::string key = item.substr(0, eq);
        const double value = std::stod(item.substr(eq + 1));
        double* reals[] = {&p.span, &p.rootChord, &p.tipChordFloor, &p.thickness, &p.lineLength,
                           &p.lineAxialStiffness, &p.lineLinearDensity, &p.lineRestScale, &p.fabricDensity,
                           &p.attachmentReinforcement, &p.complianceMin, &p.complianceMax, &p.bendingRigidity,
                           &p.pressureMean, &p.pressureStd, &p.seamAxialStiffness, &p.seamLinearDensity,
                           &p.pilotMass, &p.contactRadius};
        const char* realNames[] = {"span", "rootChord", "tipChordFloor", "thickness", "lineLength",
                                   "lineAxialStiffness", "lineLinearDensity", "lineRestScale", "fabricDensity",
                                   "attachmentReinforcement", "complianceMin", "complianceMax", "bendingRigidity",
                                   "pressureMean", "pressureStd", "seamAxialStiffness", "seamLinearDensity",
                                   "pilotMass", "contactRadius"};
        int* ints[] = {&p.spanNodes, &p.chordNodes, &p.ribEvery, &p.lineRows, &p.lineCount, &p.segmentsPerLine,
                       &p.contactSpheres, &p.tether};
        const char* intNames[] = {"spanNodes", "chordNodes", "ribEvery", "lineRows", "lineCount", "segmentsPerLine",
                                  "contactSpheres", "tether"};
        bool found = false;
        for (std::size_t k = 0; k < sizeof(reals) / sizeof(reals[0]); ++k)
            if (key == realNames[k]) {
                *reals[k] = value;
                found = true;
            }
        for (std::size_t k = 0; k < sizeof(ints) / sizeof(ints[0]); ++k)
            if (key == intNames[k]) {
                *ints[k] = static_cast<int>(value);
                found = true;
            }
        if (key == "seed") {
            p.seed = static_cast<std::uint64_t>(std::stoull(item.substr(eq + 1)));
            found = true;
        }
        if (!found) throw std::invalid_argument("unknown wing parameter: " + key);
    }
}

std::vector<WingParams> batchPresets(std::uint64_t seed, int count, double minNodes,
                                     double maxNodes, bool quantiles) {
    if (count < 1 || !(minNodes > 0.0) || !(maxNodes >= minNodes)) {
        throw std::invalid_argument("invalid batch preset request");
    }
    SplitMix64 rng(seed);
    double logRatio = 0.0;
    {
        // log(maxNodes / minNodes) by Newton iteration on seriesExp.
        const double x = maxNodes / minNodes;
        double z = 0.0;
        for (int it = 0; it < 200; ++it) {
            const double e = seriesExp(z);
            const double next = z + 2.0 * (x - e) / (x + e);
            if (next == z) break;
            z = next;
        }
        logRatio = z;
    }
    std::vector<WingParams> out;
    for (int w = 0; w < count; ++w) {
        WingParams p;
        p.seed = seed + 0x1000ULL * static_cast<std::uint64_t>(w + 1);
        // Node counts log-uniform: random draws, or the distribution's
        // (w + 1/2)/count quantiles (the sum then sits at its expectation).
        const double u = rng.uniform();
        const double nodes = minNodes * seriesExp((quantiles ? (w + 0.5) / count : u) * logRatio);
        // nodes ~ 3 * span * chord for a 2:1 grid with ribs every 2 stations.
        int chord = static_cast<int>(std::sqrt(nodes / 6.0) + 0.5);
        if (chord < 16) chord = 16;
        p.chordNodes = chord;
        p.spanNodes = 2 * chord;
        p.span = 6.0 + 6.0 * rng.uniform();
        p.rootChord = p.span / 8.0;
        p.pressureMean = 20.0 + 80.0 * rng.uniform();
        p.lineRestScale = 0.97 + 0.06 * rng.uniform();
        out.push_back(p);
    }
    return out;
}

}  // namespace swx

namespace swx {

V3 flightLoadCentre(const double* position, const std::uint8_t* fixed, std::size_t nodeCount, const V3* pilot) {
    if (pilot) return *pilot;
    V3 sum{};
    std::size_t count = 0;
    for (std::size_t i = 0; i < nodeCount; ++i) {
        if (!fixed[i]) continue;
        sum.x += position[3 * i];
        sum.y += position[3 * i + 1];
        sum.z += position[3 * i + 2];
        ++count;
    }
    if (count == 0) return V3{};
    const double inv = 1.0 / static_cast<double>(count);
    return V3{sum.x * inv, sum.y * inv, sum.z * inv};
}

void flightLoads(const double* position, const double* mass, std::size_t nodeCount,
                 std::size_t surfaceNodes, double totalWeight, const V3& centre, double* out) {
    double surfaceMass = 0.0;
    for (std::size_t i = 0; i < surfaceNodes; ++i) surfaceMass += mass[i];
    double vertical = 0.0;
    for (std::size_t i = 0; i < surfaceNodes; ++i) {
        const double dx = position[3 * i] - centre.x;
        const double dy = position[3 * i + 1] - centre.y;
        const double dz = position[3 * i + 2] - centre.z;
        const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
        vertical += (mass[i] / surfaceMass) * (dz / r);
    }
    const double scale = totalWeight / vertical;
    for (std::size_t i = 0; i < nodeCount; ++i) {
        if (i >= surfaceNodes) {
            out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = 0.0;
            continue;
        }
        const double dx = position[3 * i] - centre.x;
        const double dy = position[3 * i + 1] - centre.y;
        const double dz = position[3 * i + 2] - centre.z;
        const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
        const double f = scale * (mass[i] / surfaceMass) / r;
        out[3 * i] = f * dx;
        out[3 * i + 1] = f * dy;
        out[3 * i + 2] = f * dz;
    }
}

}  // namespace swx
