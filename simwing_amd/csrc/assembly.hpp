// Scene-v2 -> XPBD structural definition, integer-exact with the
// reference's assembleSceneStructure (src/fsi/scene_structure.cpp:894-1282).
#pragma once

#include "definition.hpp"
#include "scene.hpp"

namespace swx {

struct AssemblySettings {
    double seamCoincidenceTolerance = 1.0e-9;
    int suspensionIterations = 12;
    double suspensionAttachmentTolerance = 1.0e-10;
    double suspensionMinimumLineLength = 1.0e-10;
    double suspensionMaximumLineResidual = 2.0e-4;
    double suspensionMaximumControlWork = 1.0e6;
    double contactSphereCompliance = 0.0;
};

struct Assembly {
    Definition definition;
    std::vector<StableId> nodeVertexIds;
    std::vector<StableId> nodeJunctionIds;
    std::vector<StableId> triangleIds;
    std::vector<std::array<StableId, 2>> dihedralTriangleIds;
    double totalFabricMass = 0.0;
    double totalSeamMass = 0.0;
};

// Throws std::invalid_argument with the first assembly diagnostic when the
// scene is rejected (the reference returns the diagnostics instead; the
// C-ABI reports them as SW_INVALID_ARGUMENT). Contact spheres, when given,
// are resolved to node indices with `settings.contactSphereCompliance`.
Assembly assembleScene(const Scene& scene, const AssemblySettings& settings,
                       const std::vector<ContactSphere>& spheres = {});

// The scene checks assembleScene applies first (the subset of validateScene,
// scene.cpp:552-1040, that the structural conversion depends on). Throws
// std::invalid_argument.
void validateSceneStructure(const Scene& scene);

}  // namespace swx
