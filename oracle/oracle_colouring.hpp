// ORACLE (test infrastructure only; never linked into the product library).
//
// CPU restatement of the greedy colourings that fix the sweep order of the
// XPBD substep, plus the build-defined dihedral colouring.
//
//  * distanceColouring  follows soft_body.cpp:1978-2075 (SoftBody::
//    constraintColouring): greedy first-fit over constraint index with a
//    ceil((2*maxDegree+1)/64)-word mask per node, colours ranked by (size
//    desc, colour asc), constraints in ascending index inside a colour, and
//    the count of leading colours of size >= 1024 (the parallel colours).
//  * membraneColouring  follows soft_body.cpp:2163-2252 (SoftBody::
//    membraneColouring): same scheme with one 64-bit mask per node (throws
//    at 64 colours), parallel colours of size >= 128.
//  * dihedralColouring  is the build's own order for the four-node hinges
//    (the reference sweeps hinges serially in insertion order,
//    soft_body.cpp:1353-1365): greedy first-fit over hinge index with a
//    ceil((4*maxDegree-3)/64)-word mask per node, ranked like the others.
//    Inserting hinges into the reference in this order makes its serial
//    hinge sweep identical to a coloured one.
//
// Pinned by: tests compare distanceColouring with the reference's public
// constraintColouringView() fixture, and all three with brute-force
// checks of the greedy definition (first-fit minimality and node
// disjointness) on tiny graphs.
#pragma once

#include <algorithm>
#include <bit>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace oracle {

struct Colouring {
    std::vector<std::size_t> order;    // item indices, colour-major
    std::vector<std::size_t> offsets;  // colourCount + 1 entries
    std::size_t parallelColours = 0;
};

// items: itemCount x arity node indices, row-major.
inline Colouring greedyColouring(const std::vector<std::size_t>& items, std::size_t arity,
                                 std::size_t nodeCount, std::size_t wordsOverride,
                                 std::size_t barrierSize, bool singleWord) {
    const std::size_t count = arity ? items.size() / arity : 0;
    std::size_t words = 1;
    std::size_t colourBound = 64;
    if (!singleWord) {
        std::vector<std::size_t> degree(nodeCount, 0);
        for (std::size_t v : items) ++degree[v];
        std::size_t maxDegree = 0;
        for (std::size_t d : degree) maxDegree = std::max(maxDegree, d);
        colourBound = wordsOverride;  // caller passes the bound formula's value
        if (arity == 2) colourBound = 2 * maxDegree + 1;
        if (arity == 4) colourBound = maxDegree == 0 ? 1 : 4 * maxDegree - 3;
        words = std::max<std::size_t>(1, (colourBound + 63) / 64);
    }
    std::vector<std::uint64_t> nodeColours(nodeCount * words, 0);
    std::vector<std::size_t> colourOf(count, 0);
    std::size_t colourCount = 0;
    for (std::size_t index = 0; index < count; ++index) {
        const std::size_t* nodes = &items[index * arity];
        std::size_t colour = 0;
        for (std::size_t word = 0; word < words; ++word) {
            std::uint64_t used = 0;
            for (std::size_t k = 0; k < arity; ++k) used |= nodeColours[nodes[k] * words + word];
            if (used != ~std::uint64_t{0}) {
                colour = 64 * word + static_cast<std::size_t>(std::countr_one(used));
                break;
            }
            colour = 64 * (word + 1);
        }
        if (colour >= colourBound) {
            if (singleWord) throw std::invalid_argument("Membrane colouring exceeded 64 colours");
            throw std::logic_error("Constraint colouring exceeded its degree bound");
        }
        const std::uint64_t bit = std::uint64_t{1} << (colour % 64);
        for (std::size_t k = 0; k < arity; ++k) nodeColours[nodes[k] * words + colour / 64] |= bit;
        colourOf[index] = colour;
        colourCount = std::max(colourCount, colour + 1);
    }
    std::vector<std::size_t> sizes(colourCount, 0);
    for (std::size_t c : colourOf) ++sizes[c];
    std::vector<std::size_t> byDescendingSize(colourCount);
    for (std::size_t c = 0; c < colourCount; ++c) byDescendingSize[c] = c;
    std::sort(byDescendingSize.begin(), byDescendingSize.end(), [&](std::size_t l, std::size_t r) {
        if (sizes[l] != sizes[r]) return sizes[l] > sizes[r];
        return l < r;
    });
    std::vector<std::size_t> rankOfColour(colourCount, 0);
    for (std::size_t rank = 0; rank < colourCount; ++rank) rankOfColour[byDescendingSize[rank]] = rank;
    Colouring result;
    result.offsets.assign(colourCount + 1, 0);
    for (std::size_t rank = 0; rank < colourCount; ++rank)
        result.offsets[rank + 1] = result.offsets[rank] + sizes[byDescendingSize[rank]];
    result.order.assign(count, 0);
    std::vector<std::size_t> cursor(result.offsets.begin(), result.offsets.end() - 1);
    for (std::size_t index = 0; index < count; ++index)
        result.order[cursor[rankOfColour[colourOf[index]]]++] = index;
    std::size_t parallel = colourCount;
    while (parallel > 0 && result.offsets[parallel] - result.offsets[parallel - 1] < barrierSize)
        --parallel;
    result.parallelColours = parallel;
    return result;
}

// pairs: constraintCount x 2 node indices.
inline Colouring distanceColouring(const std::vector<std::size_t>& pairs, std::size_t nodeCount) {
    return greedyColouring(pairs, 2, nodeCount, 0, 1024, false);
}

// triangles: elementCount x 3 node indices (the element's triangle corners).
inline Colouring membraneColouring(const std::vector<std::size_t>& triangles, std::size_t nodeCount) {
    Colouring c = greedyColouring(triangles, 3, nodeCount, 0, 128, true);
    // Node-disjointness audit, as the reference performs after ordering.
    std::vector<std::size_t> last(nodeCount, static_cast<std::size_t>(-1));
    for (std::size_t colour = 0; colour + 1 < c.offsets.size(); ++colour) {
        for (std::size_t slot = c.offsets[colour]; slot < c.offsets[colour + 1]; ++slot) {
            for (std::size_t k = 0; k < 3; ++k) {
                const std::size_t node = triangles[c.order[slot] * 3 + k];
                if (last[node] == colour) throw std::logic_error("Membrane colouring is not node-disjoint");
                last[node] = colour;
            }
        }
    }
    return c;
}

// hinges: dihedralCount x 4 node indices.
inline Colouring dihedralColouring(const std::vector<std::size_t>& hinges, std::size_t nodeCount) {
    return greedyColouring(hinges, 4, nodeCount, 0, 1, false);
}

}  // namespace oracle
