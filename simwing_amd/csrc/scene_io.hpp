// The scene-v2 compact binary stream (writeScene / readScene,
// src/fsi/scene.cpp:1051-1223, 1225-1418): little-endian, every top-level
// collection in stable-ID order.
#pragma once

#include "scene.hpp"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace swx {

// writeScene's bytes for `scene` (binary version 3). Throws
// std::invalid_argument where writeScene fails (safety limits, or a scene
// the structural validation rejects).
std::vector<std::uint8_t> writeSceneBinary(const Scene& scene);

// readScene: binary versions 2 and 3. Throws std::invalid_argument with
// readScene's message on a malformed stream (magic, version, counts, limits,
// truncation, trailing data) and where the structural validation rejects the
// decoded scene.
Scene readSceneBinary(const std::uint8_t* bytes, std::size_t size);

}  // namespace swx
