// Sweep schedules of the XPBD substep: colour orders for the parallel
// families and dependency levels for the serial ones.
//
// Colour orders (integer-exact with the reference):
//  * distance constraints: soft_body.cpp:1978-2075 (greedy first-fit over
//    constraint index, ceil((2*maxDegree+1)/64)-word node masks, colours
//    ranked by size desc then colour asc);
//  * membranes: soft_body.cpp:2163-2252 (one 64-bit node mask, at most 64
//    colours);
//  * dihedral hinges: this build's order (the reference sweeps hinges
//    serially in insertion order, soft_body.cpp:1353-1365): greedy
//    first-fit over hinge index with ceil((4*maxDegree-3)/64)-word masks,
//    ranked like the others.
// Inside a colour no two items share a node, so a colour is one parallel
// phase with exactly the serial sweep's result.
//
// Serial families (suspension segments in "segment:<id>" string order,
// suspension.cpp:996-1005 and 2197-2274; the seam-stitch closing pass in
// constraint order, soft_body.cpp:1389-1405; contact-sphere pairs in
// definition order) run as dependency levels: an item's level is one more
// than the highest level among earlier items it conflicts with (shares a
// written node, or both move the rigid payload). Items of one level commute,
// so running levels in order reproduces the serial sweep exactly.
#pragma once

#include "definition.hpp"

#include <cstdint>
#include <string>
#include <vector>

namespace swx {

struct ColourOrder {
    std::vector<std::uint32_t> order;    // item indices, colour-major
    std::vector<std::uint32_t> offsets;  // colourCount + 1
};

ColourOrder colourDistance(const Definition& d);
ColourOrder colourMembranes(const Definition& d);
ColourOrder colourDihedrals(const Definition& d);

// Dependency levels of a serial sequence. `writes[i]` / `reads[i]` are the
// node sets of item i (reads include writes); `payload[i]` marks items that
// move the rigid payload. Returns the level of every item.
std::vector<std::uint32_t> dependencyLevels(const std::vector<std::vector<std::uint32_t>>& writes,
                                            const std::vector<std::vector<std::uint32_t>>& reads,
                                            const std::vector<std::uint8_t>& payload, std::size_t nodeCount);

// Runtime suspension layout: segments in string-id order with resolved
// endpoints (node index, or -(hang point + 1) for a payload hang point in
// "harness:<id>" string order).
struct SuspensionLayout {
    std::vector<std::uint32_t> segmentOrder;  // runtime position -> definition index
    std::vector<std::int32_t> from, to;       // per runtime segment
    std::vector<std::uint32_t> hangOrder;     // runtime hang point -> definition harness index
};
SuspensionLayout suspensionLayout(const Definition& d);

// Load-path order of the Cable and SuspensionTie constraints
// (SoftBody::loadPathConstraintOrder, soft_body.cpp:528-629): constraint
// indices sorted stably by the larger breadth-first depth of their two
// endpoints over the load-path graph, counted from the structural nodes
// (fixed, or touched by a triangle, a distance or stitch constraint or a
// hinge); a component without one is rooted at its node of smallest
// inverse mass (first in node order). The cable sweep pairs project this
// order backwards, then forwards (soft_body.cpp:1265-1288).
std::vector<std::uint32_t> loadPathOrder(const Definition& d);

}  // namespace swx
