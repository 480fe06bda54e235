#include "schedule.hpp"

#include <algorithm>
#include <limits>
#include <bit>
#include <map>
#include <numeric>
#include <stdexcept>

namespace swx {
namespace {

// Greedy first-fit colouring of items touching `arity` nodes each, with
// `words` 64-bit masks per node and a hard colour bound; colours are then
// ranked by (size desc, colour asc) and items listed colour-major in
// ascending index.
ColourOrder greedy(const std::uint64_t* nodes, std::size_t count, std::size_t arity, std::size_t nodeCount,
                   std::size_t words, std::size_t bound, const char* overflowMessage, bool invalidArgument) {
    std::vector<std::uint64_t> masks(nodeCount * words, 0);
    std::vector<std::uint32_t> colourOf(count);
    std::size_t colours = 0;
    for (std::size_t item = 0; item < count; ++item) {
        const std::uint64_t* v = nodes + item * arity;
        std::size_t colour = 64 * words;
        for (std::size_t w = 0; w < words; ++w) {
            std::uint64_t used = 0;
            for (std::size_t k = 0; k < arity; ++k) used |= masks[v[k] * words + w];
            if (used != ~std::uint64_t{0}) {
                colour = 64 * w + static_cast<std::size_t>(std::countr_one(used));
                break;
            }
        }
        if (colour >= bound) {
            if (invalidArgument) throw std::invalid_argument(overflowMessage);
            throw std::logic_error(overflowMessage);
        }
        for (std::size_t k = 0; k < arity; ++k) masks[v[k] * words + colour / 64] |= std::uint64_t{1} << (colour % 64);
        colourOf[item] = static_cast<std::uint32_t>(colour);
        colours = std::max(colours, colour + 1);
    }
    std::vector<std::uint32_t> size(colours, 0);
    for (auto c : colourOf) ++size[c];
    std::vector<std::uint32_t> rank(colours);
    std::iota(rank.begin(), rank.end(), 0u);
    std::stable_sort(rank.begin(), rank.end(), [&](std::uint32_t a, std::uint32_t b) { return size[a] > size[b]; });
    std::vector<std::uint32_t> slotOfColour(colours);
    ColourOrder out;
    out.offsets.assign(colours + 1, 0);
    for (std::size_t r = 0; r < colours; ++r) {
        slotOfColour[rank[r]] = static_cast<std::uint32_t>(r);
        out.offsets[r + 1] = out.offsets[r] + size[rank[r]];
    }
    std::vector<std::uint32_t> cursor(out.offsets.begin(), out.offsets.end() - 1);
    out.order.resize(count);
    for (std::size_t item = 0; item < count; ++item)
        out.order[cursor[slotOfColour[colourOf[item]]]++] = static_cast<std::uint32_t>(item);
    return out;
}

std::size_t maxDegree(const std::vector<std::uint64_t>& nodes, std::size_t nodeCount) {
    std::vector<std::uint32_t> degree(nodeCount, 0);
    std::uint32_t best = 0;
    for (auto v : nodes) best = std::max(best, ++degree[v]);
    return best;
}

}  // namespace

ColourOrder colourDistance(const Definition& d) {
    const std::size_t bound = 2 * maxDegree(d.constraintNodes, d.nodeCount()) + 1;
    const std::size_t words = std::max<std::size_t>(1, (bound + 63) / 64);
    return greedy(d.constraintNodes.data(), d.constraintCount(), 2, d.nodeCount(), words, bound,
                  "Constraint colouring exceeded its degree bound", false);
}

ColourOrder colourMembranes(const Definition& d) {
    std::vector<std::uint64_t> corners;
    corners.reserve(3 * d.membraneCount());
    for (auto t : d.membraneTriangle)
        corners.insert(corners.end(), {d.triangleNodes[3 * t], d.triangleNodes[3 * t + 1], d.triangleNodes[3 * t + 2]});
    return greedy(corners.data(), d.membraneCount(), 3, d.nodeCount(), 1, 64,
                  "Membrane colouring exceeded 64 colours", true);
}

ColourOrder colourDihedrals(const Definition& d) {
    const std::size_t degree = maxDegree(d.dihedralNodes, d.nodeCount());
    const std::size_t bound = degree == 0 ? 1 : 4 * degree - 3;
    const std::size_t words = std::max<std::size_t>(1, (bound + 63) / 64);
    return greedy(d.dihedralNodes.data(), d.dihedralCount(), 4, d.nodeCount(), words, bound,
                  "Dihedral colouring exceeded its degree bound", false);
}

std::vector<std::uint32_t> dependencyLevels(const std::vector<std::vector<std::uint32_t>>& writes,
                                            const std::vector<std::vector<std::uint32_t>>& reads,
                                            const std::vector<std::uint8_t>& payload, std::size_t nodeCount) {
    // lastWrite[n]: highest level that wrote node n; lastRead[n]: highest
    // level that read it. An item must follow every earlier writer of what
    // it reads or writes, and every earlier reader of what it writes.
    std::vector<std::uint32_t> lastWrite(nodeCount, 0), lastRead(nodeCount, 0);
    std::uint32_t lastPayload = 0;
    std::vector<std::uint32_t> level(writes.size());
    for (std::size_t i = 0; i < writes.size(); ++i) {
        std::uint32_t after = 0;
        for (auto n : reads[i]) after = std::max(after, lastWrite[n]);
        for (auto n : writes[i]) after = std::max({after, lastWrite[n], lastRead[n]});
        if (payload[i]) after = std::max(after, lastPayload);
        const std::uint32_t mine = after + 1;
        level[i] = mine;
        for (auto n : reads[i]) lastRead[n] = std::max(lastRead[n], mine);
        for (auto n : writes[i]) lastWrite[n] = std::max(lastWrite[n], mine);
        if (payload[i]) lastPayload = mine;
    }
    return level;
}

std::vector<std::uint32_t> loadPathOrder(const Definition& d) {
    const std::size_t n = d.nodeCount();
    std::vector<char> structural(n, 0);
    std::vector<std::vector<std::uint32_t>> neighbours(n);
    std::vector<std::uint32_t> path;
    const auto inverseMass = [&](std::size_t i) { return d.nodeFixed[i] != 0 ? 0.0 : 1.0 / d.nodeMass[i]; };
    for (std::size_t i = 0; i < n; ++i) structural[i] = inverseMass(i) == 0.0 ? 1 : 0;
    for (std::uint64_t v : d.triangleNodes) structural[v] = 1;
    for (std::size_t c = 0; c < d.constraintCount(); ++c) {
        const std::uint32_t a = static_cast<std::uint32_t>(d.constraintNodes[2 * c]);
        const std::uint32_t b = static_cast<std::uint32_t>(d.constraintNodes[2 * c + 1]);
        const std::uint8_t kind = d.constraintKind[c];
        if (kind == SW_CONSTRAINT_DISTANCE || kind == SW_CONSTRAINT_SEAM_STITCH) {
            structural[a] = structural[b] = 1;
            continue;
        }
        path.push_back(static_cast<std::uint32_t>(c));
        neighbours[a].push_back(b);
        neighbours[b].push_back(a);
    }
    for (std::uint64_t v : d.dihedralNodes) structural[v] = 1;
    constexpr std::uint32_t unknown = 0xffffffffu;
    std::vector<std::uint32_t> depth(n, unknown);
    std::vector<std::uint32_t> queue;
    const auto flood = [&](std::size_t head) {
        for (; head < queue.size(); ++head) {
            const std::uint32_t x = queue[head];
            for (std::uint32_t y : neighbours[x])
                if (depth[y] == unknown) {
                    depth[y] = depth[x] + 1;
                    queue.push_back(y);
                }
        }
    };
    for (std::size_t i = 0; i < n; ++i)
        if (structural[i] && !neighbours[i].empty()) {
            depth[i] = 0;
            queue.push_back(static_cast<std::uint32_t>(i));
        }
    flood(0);
    // Components with no structural node: rooted at the heaviest node (the
    // smallest inverse mass; the first such node on ties).
    for (;;) {
        std::size_t root = n;
        double lightest = std::numeric_limits<double>::infinity();
        for (std::size_t i = 0; i < n; ++i)
            if (depth[i] == unknown && !neighbours[i].empty() && inverseMass(i) < lightest) {
                root = i;
                lightest = inverseMass(i);
            }
        if (root == n) break;
        depth[root] = 0;
        const std::size_t head = queue.size();
        queue.push_back(static_cast<std::uint32_t>(root));
        flood(head);
    }
    const auto key = [&](std::uint32_t c) {
        return std::max(depth[d.constraintNodes[2 * c]], depth[d.constraintNodes[2 * c + 1]]);
    };
    std::stable_sort(path.begin(), path.end(), [&](std::uint32_t x, std::uint32_t y) { return key(x) < key(y); });
    return path;
}

SuspensionLayout suspensionLayout(const Definition& d) {
    SuspensionLayout out;
    if (!d.suspension) return out;
    const SuspensionDef& s = *d.suspension;
    const auto kindName = [](std::uint8_t kind) -> std::string {
        switch (kind) {
        case SW_ENDPOINT_ATTACHMENT:
            return "attachment:";
        case SW_ENDPOINT_JUNCTION:
            return "junction:";
        case SW_ENDPOINT_HARNESS:
            return "harness:";
        }
        throw std::invalid_argument("Unknown structure suspension endpoint kind");
    };
    // String-id order of segments and hang points (the reference normalises
    // every collection by its string id).
    std::vector<std::pair<std::string, std::uint32_t>> segments, hangs;
    for (std::size_t k = 0; k < s.segmentIds.size(); ++k)
        segments.emplace_back("segment:" + std::to_string(s.segmentIds[k]), static_cast<std::uint32_t>(k));
    for (std::size_t k = 0; k < s.harnessIds.size(); ++k)
        hangs.emplace_back("harness:" + std::to_string(s.harnessIds[k]), static_cast<std::uint32_t>(k));
    std::sort(segments.begin(), segments.end());
    std::sort(hangs.begin(), hangs.end());
    std::map<std::string, std::int32_t> endpoint;
    for (std::size_t k = 0; k < s.attachmentIds.size(); ++k)
        endpoint["attachment:" + std::to_string(s.attachmentIds[k])] = static_cast<std::int32_t>(s.attachmentNodes[k]);
    for (std::size_t k = 0; k < s.junctionIds.size(); ++k)
        endpoint["junction:" + std::to_string(s.junctionIds[k])] = static_cast<std::int32_t>(s.junctionNodes[k]);
    for (std::size_t r = 0; r < hangs.size(); ++r) {
        endpoint[hangs[r].first] = -static_cast<std::int32_t>(r) - 1;
        out.hangOrder.push_back(hangs[r].second);
    }
    for (const auto& [name, index] : segments) {
        (void)name;
        const std::string from = kindName(s.segmentFromKind[index]) + std::to_string(s.segmentFromId[index]);
        const std::string to = kindName(s.segmentToKind[index]) + std::to_string(s.segmentToId[index]);
        const auto f = endpoint.find(from), t = endpoint.find(to);
        if (f == endpoint.end() || t == endpoint.end())
            throw SuspensionFailure(kPhaseGraphConstruction, f == endpoint.end() ? from : to,
                                    "suspension segment endpoint is not resolved");
        if (s.segmentFromKind[index] == SW_ENDPOINT_HARNESS || s.segmentToKind[index] == SW_ENDPOINT_ATTACHMENT)
            throw SuspensionFailure(kPhaseValidation, name, "invalid directed endpoint kinds");
        out.segmentOrder.push_back(index);
        out.from.push_back(f->second);
        out.to.push_back(t->second);
    }
    return out;
}

}  // namespace swx
