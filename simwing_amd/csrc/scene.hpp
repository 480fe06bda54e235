// Scene-v2 input model for the XPBD structural substep, and the seeded
// synthetic ram-air wing generator that produces the benchmark workloads.
//
// The entity collections and their fields follow the reference's scene-v2
// model (src/fsi/scene.h:12-306) because the assembly contract (node order,
// dihedral keys, seam pairing, suspension orientation) is defined over them.
// Only what the structural path reads is kept; openings, region names and
// exporter metadata are carried so a scene stays valid for the reference's
// validateScene (src/fsi/scene.cpp:552).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace swx {

using StableId = std::uint64_t;

struct V2 {
    double x = 0.0;
    double y = 0.0;
};

struct V3 {
    double x = 0.0;
    double y = 0.0;
    double z = 0.0;
};

struct Quat {
    double w = 1.0;
    double x = 0.0;
    double y = 0.0;
    double z = 0.0;
};

enum class RegionKind : std::uint8_t { Outside = 1, Cell = 2 };
enum class SurfaceRole : std::uint8_t { Skin = 1, Rib = 2, Diagonal = 3, MiniRib = 4 };
enum class AttachmentKind : std::uint8_t {
    SurfaceVertex = 1,
    PilotHarness = 2,
    SuspensionJunction = 3,
};
enum class LineRole : std::uint8_t { Suspension = 1, Riser = 2, Brake = 3, Harness = 4 };

struct Region {
    StableId id = 0;
    RegionKind kind = RegionKind::Cell;
    std::string name;
};

struct Vertex {
    StableId id = 0;
    V3 position;
};

struct FabricMaterial {
    StableId id = 0;
    std::string name;
    double warpStiffness = 0.0;    // N/m
    double weftStiffness = 0.0;    // N/m
    double shearStiffness = 0.0;   // N/m
    double bendingStiffness = 0.0; // N*m
    double arealDensity = 0.0;     // kg/m^2
    double dampingSeconds = 0.0;
    double porosity = 0.0;
    double permeability = 0.0;
};

struct SeamMaterial {
    StableId id = 0;
    std::string name;
    double linearDensity = 0.0;  // kg/m
    double axialStiffness = 0.0; // N
};

struct Triangle {
    StableId id = 0;
    std::array<StableId, 3> vertexIds{};
    std::array<V2, 3> chart{};  // undeformed material coordinates, metres
    StableId negativeRegion = 0;
    StableId positiveRegion = 0;
    StableId materialId = 0;
    StableId sheetId = 0;
    SurfaceRole role = SurfaceRole::Skin;
};

struct Seam {
    StableId id = 0;
    StableId materialId = 0;
    std::vector<StableId> first;
    std::vector<StableId> second;
};

struct LineMaterial {
    StableId id = 0;
    std::string name;
    double diameter = 0.0;
    double linearDensity = 0.0;
    double axialStiffness = 0.0; // N
    double drag = 0.0;
};

struct Pilot {
    StableId id = 0;
    std::string name;
    double mass = 0.0;
    V3 centerOfMass;
    V3 linearVelocity;
    Quat bodyToWorld;
    V3 principalInertia;
};

struct Junction {
    StableId id = 0;
    V3 position;
    double mass = 0.0;
    bool fixed = false;
};

struct Attachment {
    StableId id = 0;
    AttachmentKind kind = AttachmentKind::SurfaceVertex;
    StableId vertexId = 0;
    StableId pilotId = 0;
    V3 pilotLocal;
    StableId junctionId = 0;
};

struct Line {
    StableId id = 0;
    StableId startAttachment = 0;
    StableId endAttachment = 0;
    StableId materialId = 0;
    double restLength = 0.0;
    LineRole role = LineRole::Suspension;
};

// Scene provenance (SceneMetadata, scene.h:52-64); the defaults are the
// synthetic generator's. Geometry is always SI.
struct SceneMetadata {
    std::uint32_t schemaMajor = 2;
    std::uint32_t schemaMinor = 2;
    std::uint8_t lengthUnit = 1;  // metre
    std::uint8_t handedness = 1;  // right-handed
    std::uint8_t upAxis = 1;      // +z
    double sourceLengthToMeters = 1.0;
    std::string designChecksum = "synthetic-wing";
    std::string exporterVersion = "simwing-amd-generator-1";
};

// A fluid opening (scene.h:129-145). Carried through the binary codec; the
// structural path does not read it.
struct Opening {
    StableId id = 0;
    std::vector<StableId> orderedVertexIds;
    StableId negativeRegion = 0;
    StableId positiveRegion = 0;
    std::uint8_t role = 1;  // intake
    std::vector<std::array<StableId, 3>> capTriangles;
};

struct Scene {
    SceneMetadata metadata;
    std::vector<Region> regions;
    std::vector<Vertex> vertices;
    std::vector<FabricMaterial> fabricMaterials;
    std::vector<SeamMaterial> seamMaterials;
    std::vector<Triangle> triangles;
    std::vector<Opening> openings;
    std::vector<Seam> seams;
    std::vector<LineMaterial> lineMaterials;
    std::vector<Pilot> pilots;
    std::vector<Junction> junctions;
    std::vector<Attachment> attachments;
    std::vector<Line> lines;
};

// Contact sphere of the build-defined sphere-contact family. The sphere is
// centred on a structural vertex and pushes the listed vertices out to the
// radius (a unilateral node-node minimum-distance constraint).
struct ContactSphere {
    StableId centreVertex = 0;
    double radius = 0.0;
    std::vector<StableId> vertices;
};

// Parameters of the seeded synthetic wing. Defaults are the small
// CPU-runnable wing: 32 x 16 nodes per skin, 16 ribs, 48 lines of 6
// segments, 90 kg rigid pilot.
struct WingParams {
    std::uint64_t seed = 0x5EED0000ULL;
    int spanNodes = 32;
    int chordNodes = 16;
    int ribEvery = 2;
    double span = 8.0;
    double rootChord = 1.0;
    double tipChordFloor = 0.35;
    double thickness = 0.12;
    int lineRows = 3;                // attachment rows per line rib (A, B, C)
    int lineCount = 48;
    int segmentsPerLine = 6;
    double lineLength = 7.0;         // canopy centre to the harness points
    double lineAxialStiffness = 2.0e4;
    double lineLinearDensity = 0.002;
    double lineRestScale = 1.0;
    double fabricDensity = 0.040;
    double attachmentReinforcement = 20.0;  // density factor near line attachments
    double complianceMin = 1.0e-7;   // per-triangle stretch compliance, m/N
    double complianceMax = 1.0e-6;
    double bendingRigidity = 1.0e-3; // N*m
    double pressureMean = 50.0;      // Pa
    double pressureStd = 0.0;        // 0 = uniform jump
    double seamAxialStiffness = 5.0e4;
    double seamLinearDensity = 0.01;
    double pilotMass = 90.0;
    V3 pilotInertia{8.0, 8.0, 8.0};
    int contactSpheres = 0;
    double contactRadius = 0.05;
    int tether = 0;  // no pilot; lines end at two riser junctions, 1: fixed, 2: dynamic
};

struct GeneratedWing {
    Scene scene;
    // Per-triangle pressure jump (Pa) in ascending triangle-id order, signed
    // by the stored winding: positive pushes along the right-hand normal.
    std::vector<double> facePressure;
    std::vector<ContactSphere> contactSpheres;
};

// Deterministic generator: SplitMix64 draws and only + - * / sqrt, plus a
// series exponential, so the same parameters give the same bits everywhere.
[[nodiscard]] GeneratedWing generateWing(const WingParams& params);

// SplitMix64 stream used by the generator (and by tests that need the same
// draws).
struct SplitMix64 {
    std::uint64_t state;
    explicit SplitMix64(std::uint64_t seed) : state(seed) {}
    std::uint64_t next() {
        std::uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    // Uniform in [0, 1) with 53 random bits.
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

// exp(x) from a reduced Taylor series: x is halved until |x| < 2^-8, the
// series is summed to 1e-17 relative, and the result is squared back up.
// Uses only + - * / so the value is identical on every host.
[[nodiscard]] double seriesExp(double x);

} // namespace swx

namespace swx {

// Named workloads: "tiny" (unit-test wing), "small" (the 1.5k-node,
// 3k-triangle CPU-runnable wing), "export" (the ~70k-triangle real-export
// scale wing with per-triangle pressure N(60,15) and 8 contact spheres) and
// "variant10k" (a ~10k-triangle wing). Throws on an unknown name.
[[nodiscard]] WingParams wingPreset(const std::string& name);

// Per-wing parameters of an uneven batch: node counts log-uniform in
// [minNodes, maxNodes] (drawn, or with `quantiles` the distribution's
// (w + 1/2)/count quantiles), span U(6, 12) m, pressure U(20, 100) Pa and
// line rest-length scale U(0.97, 1.03), all drawn from `seed`.
[[nodiscard]] std::vector<WingParams> batchPresets(std::uint64_t seed, int count, double minNodes,
                                                   double maxNodes, bool quantiles = false);

// Applies "key=value&key=value" overrides (WingParams field names) to a
// preset; throws std::invalid_argument on an unknown key or bad value.
void applyOverrides(WingParams& params, const std::string& overrides);

// Steady flight loads of the workload: the whole system's weight
// `totalWeight` (N) carried by the first `surfaceNodes` nodes in proportion
// to their mass, each pointing radially away from `centre` (the pilot), and
// scaled so the vertical components sum to `totalWeight`. Nodes past
// `surfaceNodes` (line junctions) get zero. `position` is 3*nodeCount,
// `out` is 3*nodeCount. This is the surface-flight mass-weighted
// preview-load pattern (surface_flight_case.cpp:187-214) with a radial
// rather than a uniform direction.
// Centre of the flight loads: the pilot's centre of mass when there is one
// (`pilot` non-null), else the mean position of the fixed nodes (the
// tethered wing's riser points), else the origin.
V3 flightLoadCentre(const double* position, const std::uint8_t* fixed, std::size_t nodeCount, const V3* pilot);

void flightLoads(const double* position, const double* mass, std::size_t nodeCount,
                 std::size_t surfaceNodes, double totalWeight, const V3& centre, double* out);

}  // namespace swx
