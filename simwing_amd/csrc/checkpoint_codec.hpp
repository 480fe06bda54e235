// Byte-compatible Structure checkpoint codec: the SWST envelope of
// serializeStructureCheckpoint (structure_checkpoint_persistence.cpp:229-333)
// around the SWBC SoftBody section (soft_body_checkpoint_persistence.cpp:
// 645-703, 858-919) and the SWSC suspension section
// (suspension_checkpoint_persistence.cpp:575-644, 752-800), written from and
// read into the device wing's committed state. Host code only.
#pragma once

#include "definition.hpp"
#include "device_wing.hpp"

#include <cstdint>
#include <string>
#include <vector>

namespace swx {

// What a Structure checkpoint holds (structure.h:254-271) in this build's
// terms: the committed solver state in definition order, the pending
// per-node external forces and the last applied total force.
struct CheckpointImage {
    std::uint64_t fingerprint = 0;
    std::uint64_t acceptedSteps = 0;
    double time = 0.0;
    DeviceWing::State state;
    std::vector<double> pending;  // 3N
    double lastApplied[3] = {0.0, 0.0, 0.0};
    // The suspension's last finished substep, from which its diagnostics are
    // restated (SuspensionSystem::finishSubstep, suspension.cpp:2481-2610):
    // whether one has finished, its length, gravity, line iterations and
    // certified maximum line residual.
    bool suspensionStepped = false;
    double substepDt = 0.0;
    double gravity[3] = {0.0, 0.0, 0.0};
    int lineIterations = 0;
    double maxResidual = 0.0;
    // A suspension section read from bytes and not stepped since: written
    // back verbatim (its diagnostics are the reader's, as the reference's
    // restore keeps them), with the per-segment length, commanded rest
    // length and tension of its diagnostics in runtime order (empty before
    // the first substep) for suspensionState().
    std::vector<std::uint8_t> suspensionSection;
    std::vector<double> sectionLength, sectionCommanded, sectionTension;
};

// SoftBody::checkpointTopologyFingerprint (soft_body.cpp:820-917) of the body
// that Structure's buildBody (structure.cpp:554-634) makes of `d`.
std::uint64_t softBodyTopologyFingerprint(const Definition& d);

// The SWST bytes of `image` for a structure of definition `d`. Throws
// std::invalid_argument where the reference's serializer fails (non-finite
// values), where the definition has no reference counterpart (contact
// spheres) or where the suspension's diagnostics hold what the device does
// not record (damping impulses of damped lines).
std::vector<std::uint8_t> encodeStructureCheckpoint(const Definition& d, const CheckpointImage& image);

// Reads SWST bytes written for definition `d` (deserializeStructureCheckpoint,
// structure_checkpoint_persistence.cpp:335-520). `image` enters holding a
// state of the right shape (the structure's own) and leaves with every
// field the bytes carry. Throws std::invalid_argument with the reference's
// reason on malformed, mismatched or unsupported input.
void decodeStructureCheckpoint(const Definition& d, std::uint64_t fingerprint, const std::uint8_t* bytes,
                               std::size_t size, CheckpointImage& image);

}  // namespace swx
