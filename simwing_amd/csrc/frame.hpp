// Structure diagnostic frames: the frame simwing::viewer::buildStructureFrame
// builds from a committed Structure state (src/viewer/structure_frame.cpp:
// 480-666) and its SWVF byte form (serializeFrame, viewer_protocol.cpp:
// 286-376, 690-882).
#pragma once

#include "definition.hpp"

#include <cstdint>
#include <vector>

namespace swx {

// The per-vertex and per-line arrays of one frame (vertices: the N nodes
// then the H harness points; lines: the C constraints then the S segments;
// all in definition order) and the global diagnostics.
struct FrameFields {
    std::uint64_t step = 0;
    double simulationTime = 0.0;
    const double* vertexPosition = nullptr;      // 3 (N + H)
    const double* vertexVelocity = nullptr;      // 3 (N + H)
    const double* vertexPendingForce = nullptr;  // 3 (N + H)
    const double* lineLength = nullptr;          // C + S
    const double* lineViolation = nullptr;       // C + S
    const sw_diagnostics* diagnostics = nullptr;
};

// Checks the mapping as StructureFrameMapping's constructor does
// (structure_frame.cpp:205-264) and returns the frame's line endpoints
// (2 (C + S)): constraints by their nodes, segments by the endpoint rule of
// makeStructureFrameMapping (structure_frame.cpp:427-474). Throws
// std::invalid_argument.
std::vector<std::uint32_t> frameLineVertices(const Definition& d, const sw_frame_mapping& m);

// Exact SWVF size of the frame of `d` under `m` and `c`.
std::size_t frameBytes(const Definition& d, const sw_frame_mapping& m, const sw_frame_context& c);

// Builds and encodes the frame; throws std::invalid_argument where
// buildStructureFrame or validateFrame reject it.
std::vector<std::uint8_t> encodeFrame(const Definition& d, const sw_frame_mapping& m, const sw_frame_context& c,
                                      const FrameFields& f);

}  // namespace swx
