// Byte-compatible Structure checkpoint codec: the SWST envelope of
// serializeStructureCheckpoint (structure_checkpoint_persistence.cpp:229-333)
// around the SWBC SoftBody section (soft_body_checkpoint_persistence.cpp:
// 645-703, 858-919), written from and read into the device wing's committed
// state. Host code only.
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
};

// SoftBody::checkpointTopologyFingerprint (soft_body.cpp:820-917) of the body
// that Structure's buildBody (structure.cpp:554-634) makes of `d`.
std::uint64_t softBodyTopologyFingerprint(const Definition& d);

// The SWST bytes of `image` for a structure of definition `d`. Throws
// std::invalid_argument where the reference's serializer fails (non-finite
// values) or where the definition has no reference counterpart (contact
// spheres) or a part this codec does not write (a suspension).
std::vector<std::uint8_t> encodeStructureCheckpoint(const Definition& d, const CheckpointImage& image);

// Reads SWST bytes written for definition `d` (deserializeStructureCheckpoint,
// structure_checkpoint_persistence.cpp:335-520). `image` enters holding a
// state of the right shape (the structure's own) and leaves with every
// field the bytes carry. Throws std::invalid_argument with the reference's
// reason on malformed, mismatched or unsupported input.
void decodeStructureCheckpoint(const Definition& d, std::uint64_t fingerprint, const std::uint8_t* bytes,
                               std::size_t size, CheckpointImage& image);

}  // namespace swx
