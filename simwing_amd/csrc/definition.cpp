#include "definition.hpp"

#include <algorithm>
#include <bit>
#include <cmath>
#include <limits>

namespace swx {
namespace {

bool finite3(const double* v) {
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

template <typename T>
std::vector<T> copyOf(const T* data, std::size_t count) {
    if (count == 0) return {};
    if (!data) throw std::invalid_argument("definition array is null");
    return std::vector<T>(data, data + count);
}

// Symmetric 3x3 positive-definiteness by equilibrated Cholesky pivots
// (membrane.cpp:186-206).
bool positiveDefinite(double xx, double yy, double zz, double xy, double xz, double yz) {
    if (!std::isfinite(xx) || !std::isfinite(yy) || !std::isfinite(zz) || !std::isfinite(xy) ||
        !std::isfinite(xz) || !std::isfinite(yz) || !(xx > 0.0) || !(yy > 0.0) || !(zz > 0.0))
        return false;
    const double sx = std::sqrt(xx), sy = std::sqrt(yy), sz = std::sqrt(zz);
    const double l10 = xy / sx / sy;
    const double l20 = xz / sx / sz;
    const double p1 = 1.0 - l10 * l10;
    if (!std::isfinite(p1) || !(p1 > 0.0)) return false;
    const double l11 = std::sqrt(p1);
    const double l21 = (yz / sy / sz - l20 * l10) / l11;
    const double p2 = 1.0 - l20 * l20 - l21 * l21;
    return std::isfinite(p2) && p2 > 0.0;
}

// validateOrthotropicMembraneMaterial (membrane.cpp:298-329).
void validateMaterial(const double* m) {
    for (int k = 0; k < 8; ++k)
        if (!std::isfinite(m[k])) throw std::invalid_argument("Membrane material values must be finite");
    const double warp = m[0], weft = m[1], coupling = m[2], shear = m[3];
    const double preW = m[4], preT = m[5], damping = m[6], ratio = m[7];
    if (!(warp > 0.0) || !(weft > 0.0) || !(shear > 0.0) || coupling < 0.0 ||
        !(coupling * coupling < warp * weft) || preW < 0.0 || preW >= 0.25 || preT < 0.0 ||
        preT >= 0.25 || damping < 0.0 || !(ratio > 0.0) || ratio > 1.0 ||
        !positiveDefinite(warp, weft, shear, coupling, 0.0, 0.0)) {
        throw std::invalid_argument("Invalid orthotropic membrane material");
    }
}

class Fnv {
public:
    void add(std::uint64_t v) {
        for (int byte = 0; byte < 8; ++byte) {
            value_ ^= static_cast<std::uint8_t>(v & 0xffU);
            value_ *= 1099511628211ULL;
            v >>= 8U;
        }
    }
    void add(double v) { add(std::bit_cast<std::uint64_t>(v)); }
    std::uint64_t value() const { return value_; }

private:
    std::uint64_t value_ = 14695981039346656037ULL;
};

}  // namespace

sw_definition Definition::view(sw_suspension_definition* sv) const {
    sw_definition v{};
    v.node_count = nodeCount();
    v.node_position = nodePosition.data();
    v.node_mass = nodeMass.data();
    v.node_fixed = nodeFixed.data();
    v.triangle_count = triangleCount();
    v.triangle_nodes = triangleNodes.data();
    v.constraint_count = constraintCount();
    v.constraint_kind = constraintKind.data();
    v.constraint_nodes = constraintNodes.data();
    v.constraint_rest = constraintRest.data();
    v.constraint_compliance = constraintCompliance.data();
    v.membrane_count = membraneCount();
    v.membrane_triangle = membraneTriangle.data();
    v.membrane_chart = membraneChart.data();
    v.membrane_material = membraneMaterial.data();
    v.membrane_role = membraneRole.data();
    v.dihedral_count = dihedralCount();
    v.dihedral_nodes = dihedralNodes.data();
    v.dihedral_rest_angle = dihedralRestAngle.data();
    v.dihedral_compliance = dihedralCompliance.data();
    v.sphere_count = sphereCount();
    v.sphere_centre = sphereCentre.data();
    v.sphere_radius = sphereRadius.data();
    v.sphere_compliance = sphereCompliance.data();
    v.sphere_offsets = sphereOffsets.data();
    v.sphere_nodes = sphereNodes.data();
    v.suspension = nullptr;
    if (suspension && sv) {
        const SuspensionDef& s = *suspension;
        *sv = sw_suspension_definition{};
        sv->pilot_stable_id = s.pilotStableId;
        sv->pilot_mass = s.pilotMass;
        for (int k = 0; k < 3; ++k) {
            sv->pilot_com_local[k] = s.comLocal[k];
            sv->pilot_com_world[k] = s.comWorld[k];
            sv->pilot_velocity[k] = s.velocity[k];
            sv->pilot_angular_velocity[k] = s.angularVelocity[k];
            sv->pilot_inertia[k] = s.inertia[k];
        }
        for (int k = 0; k < 4; ++k) sv->pilot_orientation[k] = s.orientation[k];
        sv->attachment_count = s.attachmentIds.size();
        sv->attachment_ids = s.attachmentIds.data();
        sv->attachment_nodes = s.attachmentNodes.data();
        sv->junction_count = s.junctionIds.size();
        sv->junction_ids = s.junctionIds.data();
        sv->junction_nodes = s.junctionNodes.data();
        sv->harness_count = s.harnessIds.size();
        sv->harness_ids = s.harnessIds.data();
        sv->harness_local = s.harnessLocal.data();
        sv->segment_count = s.segmentIds.size();
        sv->segment_ids = s.segmentIds.data();
        sv->segment_from_kind = s.segmentFromKind.data();
        sv->segment_from_id = s.segmentFromId.data();
        sv->segment_to_kind = s.segmentToKind.data();
        sv->segment_to_id = s.segmentToId.data();
        sv->segment_rest = s.segmentRest.data();
        sv->segment_stiffness = s.segmentStiffness.data();
        sv->segment_damping = s.segmentDamping.data();
        sv->segment_role = s.segmentRole.data();
        sv->solver_iterations = s.solverIterations;
        sv->attachment_tolerance = s.attachmentTolerance;
        sv->minimum_line_length = s.minimumLineLength;
        sv->maximum_line_residual = s.maximumLineResidual;
        sv->maximum_control_work = s.maximumControlWork;
        v.suspension = sv;
    }
    return v;
}

Definition Definition::fromView(const sw_definition& v) {
    Definition d;
    d.nodePosition = copyOf(v.node_position, 3 * v.node_count);
    d.nodeMass = copyOf(v.node_mass, v.node_count);
    d.nodeFixed = copyOf(v.node_fixed, v.node_count);
    d.triangleNodes = copyOf(v.triangle_nodes, 3 * v.triangle_count);
    d.constraintKind = copyOf(v.constraint_kind, v.constraint_count);
    d.constraintNodes = copyOf(v.constraint_nodes, 2 * v.constraint_count);
    d.constraintRest = copyOf(v.constraint_rest, v.constraint_count);
    d.constraintCompliance = copyOf(v.constraint_compliance, v.constraint_count);
    d.membraneTriangle = copyOf(v.membrane_triangle, v.membrane_count);
    d.membraneChart = copyOf(v.membrane_chart, 6 * v.membrane_count);
    d.membraneMaterial = copyOf(v.membrane_material, 8 * v.membrane_count);
    if (v.membrane_role) d.membraneRole = copyOf(v.membrane_role, v.membrane_count);
    else d.membraneRole.assign(v.membrane_count, SW_ROLE_BULK);
    d.dihedralNodes = copyOf(v.dihedral_nodes, 4 * v.dihedral_count);
    d.dihedralRestAngle = copyOf(v.dihedral_rest_angle, v.dihedral_count);
    d.dihedralCompliance = copyOf(v.dihedral_compliance, v.dihedral_count);
    if (v.sphere_count) {
        d.sphereCentre = copyOf(v.sphere_centre, v.sphere_count);
        d.sphereRadius = copyOf(v.sphere_radius, v.sphere_count);
        d.sphereCompliance = copyOf(v.sphere_compliance, v.sphere_count);
        d.sphereOffsets = copyOf(v.sphere_offsets, v.sphere_count + 1);
        if (d.sphereOffsets.front() != 0) throw std::invalid_argument("Contact sphere offsets must start at zero");
        d.sphereNodes = copyOf(v.sphere_nodes, d.sphereOffsets.back());
    }
    if (v.suspension) {
        const sw_suspension_definition& s = *v.suspension;
        SuspensionDef o;
        o.pilotStableId = s.pilot_stable_id;
        o.pilotMass = s.pilot_mass;
        for (int k = 0; k < 3; ++k) {
            o.comLocal[k] = s.pilot_com_local[k];
            o.comWorld[k] = s.pilot_com_world[k];
            o.velocity[k] = s.pilot_velocity[k];
            o.angularVelocity[k] = s.pilot_angular_velocity[k];
            o.inertia[k] = s.pilot_inertia[k];
        }
        for (int k = 0; k < 4; ++k) o.orientation[k] = s.pilot_orientation[k];
        o.attachmentIds = copyOf(s.attachment_ids, s.attachment_count);
        o.attachmentNodes = copyOf(s.attachment_nodes, s.attachment_count);
        o.junctionIds = copyOf(s.junction_ids, s.junction_count);
        o.junctionNodes = copyOf(s.junction_nodes, s.junction_count);
        o.harnessIds = copyOf(s.harness_ids, s.harness_count);
        o.harnessLocal = copyOf(s.harness_local, 3 * s.harness_count);
        o.segmentIds = copyOf(s.segment_ids, s.segment_count);
        o.segmentFromKind = copyOf(s.segment_from_kind, s.segment_count);
        o.segmentFromId = copyOf(s.segment_from_id, s.segment_count);
        o.segmentToKind = copyOf(s.segment_to_kind, s.segment_count);
        o.segmentToId = copyOf(s.segment_to_id, s.segment_count);
        o.segmentRest = copyOf(s.segment_rest, s.segment_count);
        o.segmentStiffness = copyOf(s.segment_stiffness, s.segment_count);
        o.segmentDamping = copyOf(s.segment_damping, s.segment_count);
        if (s.segment_role) o.segmentRole = copyOf(s.segment_role, s.segment_count);
        else o.segmentRole.assign(s.segment_count, 0);
        o.solverIterations = s.solver_iterations;
        o.attachmentTolerance = s.attachment_tolerance;
        o.minimumLineLength = s.minimum_line_length;
        o.maximumLineResidual = s.maximum_line_residual;
        o.maximumControlWork = s.maximum_control_work;
        d.suspension = std::move(o);
    }
    return d;
}

void validateDefinition(const Definition& d) {
    const std::size_t n = d.nodeCount();
    if (d.nodePosition.size() != 3 * n || d.nodeFixed.size() != n)
        throw std::invalid_argument("Invalid structure node definition");
    for (std::size_t i = 0; i < n; ++i) {
        const double m = d.nodeMass[i];
        if (!finite3(&d.nodePosition[3 * i]) || !std::isfinite(m) || m < 0.0 ||
            (!d.nodeFixed[i] && !(m > 0.0)))
            throw std::invalid_argument("Invalid structure node definition");
    }
    const auto requireNode = [n](std::uint64_t node, const char* entity) {
        if (node >= n) throw std::invalid_argument(std::string(entity) + " references an unknown node");
    };
    for (std::size_t t = 0; t < d.triangleCount(); ++t) {
        const std::uint64_t* v = &d.triangleNodes[3 * t];
        for (int k = 0; k < 3; ++k) requireNode(v[k], "Triangle");
        if (v[0] == v[1] || v[1] == v[2] || v[2] == v[0]) throw std::invalid_argument("Triangle repeats a node");
    }
    for (std::size_t c = 0; c < d.constraintCount(); ++c) {
        const std::uint64_t a = d.constraintNodes[2 * c], b = d.constraintNodes[2 * c + 1];
        requireNode(a, "Constraint");
        requireNode(b, "Constraint");
        const double rest = d.constraintRest[c], comp = d.constraintCompliance[c];
        if (a == b || !std::isfinite(rest) || rest < 0.0 || !std::isfinite(comp) || comp < 0.0)
            throw std::invalid_argument("Invalid structure constraint");
        switch (d.constraintKind[c]) {
        case SW_CONSTRAINT_DISTANCE:
        case SW_CONSTRAINT_CABLE:
        case SW_CONSTRAINT_SUSPENSION_TIE:
            break;
        case SW_CONSTRAINT_SEAM_STITCH:
            if (rest != 0.0 || comp != 0.0) throw std::invalid_argument("Seam stitch must be rigid and zero-rest");
            break;
        default:
            throw std::invalid_argument("Unknown structure constraint kind");
        }
    }
    for (std::size_t m = 0; m < d.membraneCount(); ++m) {
        if (d.membraneTriangle[m] >= d.triangleCount())
            throw std::invalid_argument("Membrane references an unknown triangle");
        for (int k = 0; k < 6; ++k)
            if (!std::isfinite(d.membraneChart[6 * m + k]))
                throw std::invalid_argument("Membrane material coordinates must be finite");
        validateMaterial(&d.membraneMaterial[8 * m]);
        if (d.membraneRole[m] < SW_ROLE_BULK || d.membraneRole[m] > SW_ROLE_REINFORCEMENT)
            throw std::invalid_argument("Unknown structure material role");
    }
    for (std::size_t h = 0; h < d.dihedralCount(); ++h) {
        std::array<std::uint64_t, 4> u{d.dihedralNodes[4 * h], d.dihedralNodes[4 * h + 1],
                                       d.dihedralNodes[4 * h + 2], d.dihedralNodes[4 * h + 3]};
        for (auto v : u) requireNode(v, "Dihedral");
        std::sort(u.begin(), u.end());
        if (std::adjacent_find(u.begin(), u.end()) != u.end() || !std::isfinite(d.dihedralRestAngle[h]) ||
            !std::isfinite(d.dihedralCompliance[h]) || d.dihedralCompliance[h] < 0.0)
            throw std::invalid_argument("Invalid dihedral definition");
    }
    if (d.sphereOffsets.size() != d.sphereCount() + 1)
        throw std::invalid_argument("Invalid contact sphere definition");
    for (std::size_t k = 0; k < d.sphereCount(); ++k) {
        requireNode(d.sphereCentre[k], "Contact sphere");
        if (!(d.sphereRadius[k] > 0.0) || !std::isfinite(d.sphereRadius[k]) ||
            !(d.sphereCompliance[k] >= 0.0) || !std::isfinite(d.sphereCompliance[k]) ||
            d.sphereOffsets[k + 1] < d.sphereOffsets[k])
            throw std::invalid_argument("Invalid contact sphere definition");
        for (std::uint64_t s = d.sphereOffsets[k]; s < d.sphereOffsets[k + 1]; ++s) {
            requireNode(d.sphereNodes[s], "Contact sphere");
            if (d.sphereNodes[s] == d.sphereCentre[k])
                throw std::invalid_argument("Contact sphere holds off its own centre");
        }
    }
    if (d.suspension) {
        const SuspensionDef& s = *d.suspension;
        const auto finiteArray = [](const std::array<double, 3>& a) { return finite3(a.data()); };
        const bool orientationFinite = std::isfinite(s.orientation[0]) && std::isfinite(s.orientation[1]) &&
                                       std::isfinite(s.orientation[2]) && std::isfinite(s.orientation[3]);
        if (s.pilotStableId == 0 || !(s.pilotMass > 0.0) || !std::isfinite(s.pilotMass) ||
            !finiteArray(s.comLocal) || !finiteArray(s.comWorld) || !finiteArray(s.velocity) ||
            !finiteArray(s.angularVelocity) || !orientationFinite || !(s.inertia[0] > 0.0) ||
            !(s.inertia[1] > 0.0) || !(s.inertia[2] > 0.0) || !finiteArray(s.inertia) ||
            s.attachmentIds.empty() || s.harnessIds.empty() || s.segmentIds.empty() ||
            s.solverIterations <= 0 || !(s.attachmentTolerance > 0.0) ||
            !std::isfinite(s.attachmentTolerance) || !(s.minimumLineLength > 0.0) ||
            !std::isfinite(s.minimumLineLength) || s.maximumLineResidual < 0.0 ||
            !std::isfinite(s.maximumLineResidual) || s.maximumControlWork < 0.0 ||
            !std::isfinite(s.maximumControlWork)) {
            throw std::invalid_argument("Invalid structure suspension definition");
        }
        const auto requireStableIds = [](std::vector<std::uint64_t> ids, const char* label) {
            for (auto id : ids)
                if (id == 0) throw std::invalid_argument(std::string(label) + " stable ID must be nonzero");
            std::sort(ids.begin(), ids.end());
            if (std::adjacent_find(ids.begin(), ids.end()) != ids.end())
                throw std::invalid_argument(std::string(label) + " stable IDs must be unique");
        };
        requireStableIds(s.attachmentIds, "Suspension attachment");
        requireStableIds(s.junctionIds, "Suspension junction");
        requireStableIds(s.harnessIds, "Pilot harness");
        requireStableIds(s.segmentIds, "Suspension segment");
        for (auto node : s.attachmentNodes) requireNode(node, "Suspension attachment");
        for (auto node : s.junctionNodes) {
            requireNode(node, "Suspension junction");
            if (d.nodeFixed[node]) throw std::invalid_argument("Suspension junction must be a dynamic node");
        }
        for (std::size_t h = 0; h < s.harnessIds.size(); ++h)
            if (!finite3(&s.harnessLocal[3 * h])) throw std::invalid_argument("Pilot harness position must be finite");
        for (std::size_t g = 0; g < s.segmentIds.size(); ++g) {
            if (s.segmentFromId[g] == 0 || s.segmentToId[g] == 0 ||
                (s.segmentFromKind[g] == s.segmentToKind[g] && s.segmentFromId[g] == s.segmentToId[g]) ||
                !(s.segmentRest[g] > 0.0) || !std::isfinite(s.segmentRest[g]) ||
                !(s.segmentStiffness[g] > 0.0) || !std::isfinite(s.segmentStiffness[g]) ||
                s.segmentDamping[g] < 0.0 || !std::isfinite(s.segmentDamping[g]))
                throw std::invalid_argument("Invalid structure suspension segment");
        }
    }
}

std::uint64_t definitionFingerprint(const Definition& d) {
    Fnv h;
    h.add(std::uint64_t{2});
    h.add(static_cast<std::uint64_t>(d.nodeCount()));
    for (std::size_t i = 0; i < d.nodeCount(); ++i) {
        h.add(d.nodePosition[3 * i]);
        h.add(d.nodePosition[3 * i + 1]);
        h.add(d.nodePosition[3 * i + 2]);
        h.add(d.nodeMass[i]);
        h.add(d.nodeFixed[i] ? std::uint64_t{1} : std::uint64_t{0});
    }
    h.add(static_cast<std::uint64_t>(d.triangleCount()));
    for (auto v : d.triangleNodes) h.add(v);
    h.add(static_cast<std::uint64_t>(d.constraintCount()));
    for (std::size_t c = 0; c < d.constraintCount(); ++c) {
        h.add(static_cast<std::uint64_t>(d.constraintKind[c]));
        h.add(d.constraintNodes[2 * c]);
        h.add(d.constraintNodes[2 * c + 1]);
        h.add(d.constraintRest[c]);
        h.add(d.constraintCompliance[c]);
    }
    h.add(static_cast<std::uint64_t>(d.membraneCount()));
    for (std::size_t m = 0; m < d.membraneCount(); ++m) {
        h.add(d.membraneTriangle[m]);
        for (int k = 0; k < 6; ++k) h.add(d.membraneChart[6 * m + k]);
        for (int k = 0; k < 8; ++k) h.add(d.membraneMaterial[8 * m + k]);
        h.add(static_cast<std::uint64_t>(d.membraneRole[m]));
    }
    h.add(static_cast<std::uint64_t>(d.dihedralCount()));
    for (std::size_t k = 0; k < d.dihedralCount(); ++k) {
        for (int c = 0; c < 4; ++c) h.add(d.dihedralNodes[4 * k + c]);
        h.add(d.dihedralRestAngle[k]);
        h.add(d.dihedralCompliance[k]);
    }
    h.add(std::uint64_t{0});  // no exhaustive fabric self-contact
    h.add(d.suspension ? std::uint64_t{1} : std::uint64_t{0});
    if (d.suspension) {
        const SuspensionDef& s = *d.suspension;
        h.add(s.pilotStableId);
        h.add(s.pilotMass);
        for (double v : s.comLocal) h.add(v);
        for (double v : s.comWorld) h.add(v);
        for (double v : s.velocity) h.add(v);
        for (double v : s.angularVelocity) h.add(v);
        for (double v : s.orientation) h.add(v);
        for (double v : s.inertia) h.add(v);
        h.add(static_cast<std::uint64_t>(s.attachmentIds.size()));
        for (std::size_t k = 0; k < s.attachmentIds.size(); ++k) {
            h.add(s.attachmentIds[k]);
            h.add(s.attachmentNodes[k]);
        }
        h.add(static_cast<std::uint64_t>(s.junctionIds.size()));
        for (std::size_t k = 0; k < s.junctionIds.size(); ++k) {
            h.add(s.junctionIds[k]);
            h.add(s.junctionNodes[k]);
        }
        h.add(static_cast<std::uint64_t>(s.harnessIds.size()));
        for (std::size_t k = 0; k < s.harnessIds.size(); ++k) {
            h.add(s.harnessIds[k]);
            for (int c = 0; c < 3; ++c) h.add(s.harnessLocal[3 * k + c]);
        }
        h.add(static_cast<std::uint64_t>(s.segmentIds.size()));
        for (std::size_t k = 0; k < s.segmentIds.size(); ++k) {
            h.add(s.segmentIds[k]);
            h.add(static_cast<std::uint64_t>(s.segmentFromKind[k]));
            h.add(s.segmentFromId[k]);
            h.add(static_cast<std::uint64_t>(s.segmentToKind[k]));
            h.add(s.segmentToId[k]);
            h.add(s.segmentRest[k]);
            h.add(s.segmentStiffness[k]);
            h.add(s.segmentDamping[k]);
            h.add(static_cast<std::uint64_t>(s.segmentRole[k]));
        }
        h.add(static_cast<std::uint64_t>(s.solverIterations));
        h.add(s.attachmentTolerance);
        h.add(s.minimumLineLength);
        h.add(s.maximumLineResidual);
        h.add(s.maximumControlWork);
    }
    if (d.sphereCount()) {
        h.add(static_cast<std::uint64_t>(d.sphereCount()));
        for (std::size_t k = 0; k < d.sphereCount(); ++k) {
            h.add(d.sphereCentre[k]);
            h.add(d.sphereRadius[k]);
            h.add(d.sphereCompliance[k]);
            for (std::uint64_t s = d.sphereOffsets[k]; s < d.sphereOffsets[k + 1]; ++s) h.add(d.sphereNodes[s]);
        }
    }
    return h.value();
}

}  // namespace swx
