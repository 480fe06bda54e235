// Owning, SoA form of the structural definition (the reference's
// StructureDefinition, src/fsi/structure.h:30-176) and its validation and
// fingerprint.
#pragma once

#include "../../include/simwing_xpbd.h"

#include <array>
#include <cstddef>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace swx {

// softwing::SuspensionError (suspension.h:40-52) as a C++ exception of this
// library; phase uses the same numbering as softwing::SuspensionPhase.
class SuspensionFailure : public std::runtime_error {
public:
    SuspensionFailure(int phase, std::string entity, const std::string& message)
        : std::runtime_error(message), phase_(phase), entity_(std::move(entity)) {}
    int phase() const { return phase_; }
    const std::string& entity() const { return entity_; }

private:
    int phase_;
    std::string entity_;
};

enum SuspensionPhaseCode : int {
    kPhaseParse = 0,
    kPhaseValidation = 1,
    kPhaseAttachmentResolution = 2,
    kPhaseGraphConstruction = 3,
    kPhaseControl = 4,
    kPhasePrediction = 5,
    kPhaseLineSolve = 6,
    kPhaseGroundSolve = 7,
    kPhaseCertification = 8,
    kPhaseDiagnostics = 9,
};

struct SuspensionDef {
    std::uint64_t pilotStableId = 0;
    double pilotMass = 0.0;
    std::array<double, 3> comLocal{};
    std::array<double, 3> comWorld{};
    std::array<double, 3> velocity{};
    std::array<double, 3> angularVelocity{};
    std::array<double, 4> orientation{1.0, 0.0, 0.0, 0.0};
    std::array<double, 3> inertia{};
    std::vector<std::uint64_t> attachmentIds, attachmentNodes;
    std::vector<std::uint64_t> junctionIds, junctionNodes;
    std::vector<std::uint64_t> harnessIds;
    std::vector<double> harnessLocal;  // 3 per point
    std::vector<std::uint64_t> segmentIds, segmentFromId, segmentToId;
    std::vector<std::uint8_t> segmentFromKind, segmentToKind;
    std::vector<double> segmentRest, segmentStiffness, segmentDamping;
    std::vector<std::uint32_t> segmentRole;
    int solverIterations = 12;
    double attachmentTolerance = 1.0e-10;
    double minimumLineLength = 1.0e-10;
    double maximumLineResidual = 2.0e-4;
    double maximumControlWork = 1.0e6;
};

struct Definition {
    std::vector<double> nodePosition;  // 3N
    std::vector<double> nodeMass;
    std::vector<std::uint8_t> nodeFixed;
    std::vector<std::uint64_t> triangleNodes;  // 3T
    std::vector<std::uint8_t> constraintKind;
    std::vector<std::uint64_t> constraintNodes;  // 2C
    std::vector<double> constraintRest, constraintCompliance;
    std::vector<std::uint64_t> membraneTriangle;
    std::vector<double> membraneChart;     // 6M
    std::vector<double> membraneMaterial;  // 8M
    std::vector<std::uint8_t> membraneRole;
    std::vector<std::uint64_t> dihedralNodes;  // 4D
    std::vector<double> dihedralRestAngle, dihedralCompliance;
    std::optional<SuspensionDef> suspension;
    std::vector<std::uint64_t> sphereCentre;
    std::vector<double> sphereRadius, sphereCompliance;
    std::vector<std::uint64_t> sphereOffsets{0};
    std::vector<std::uint64_t> sphereNodes;

    std::size_t nodeCount() const { return nodeMass.size(); }
    std::size_t triangleCount() const { return triangleNodes.size() / 3; }
    std::size_t constraintCount() const { return constraintKind.size(); }
    std::size_t membraneCount() const { return membraneTriangle.size(); }
    std::size_t dihedralCount() const { return dihedralRestAngle.size(); }
    std::size_t sphereCount() const { return sphereCentre.size(); }

    // Borrowed C view; `suspensionView` must outlive the returned struct.
    sw_definition view(sw_suspension_definition* suspensionView) const;
    static Definition fromView(const sw_definition& view);
};

// Structure's validateDefinition (structure.cpp:102-291), including the
// membrane material validation of softwing (membrane.cpp:298-329).
// Throws std::invalid_argument.
void validateDefinition(const Definition& d);

// FNV-1a definition fingerprint (structure.cpp:326-431), extended with the
// contact spheres when present.
std::uint64_t definitionFingerprint(const Definition& d);

}  // namespace swx
