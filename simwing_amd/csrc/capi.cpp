// extern "C" boundary of the library (include/simwing_xpbd.h): the
// Structure adapter semantics (src/fsi/structure.cpp:682-1025) over the
// device wing, scene generation and assembly, and error translation.
#include "checkpoint_codec.hpp"
#include "../../include/simwing_xpbd.h"

#include "assembly.hpp"
#include "definition.hpp"
#include "frame.hpp"
#include "device_wing.hpp"
#include "scene.hpp"
#include "scene_io.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string lastError;
thread_local int lastPhase = -1;
thread_local std::string lastEntity;

template <typename F>
sw_status guarded(F&& body) {
    try {
        body();
        return SW_OK;
    } catch (const swx::SuspensionFailure& e) {
        lastError = e.what();
        lastPhase = e.phase();
        lastEntity = e.entity();
        return SW_SUSPENSION_ERROR;
    } catch (const std::out_of_range& e) {
        lastError = e.what();
        return SW_OUT_OF_RANGE;
    } catch (const std::invalid_argument& e) {
        lastError = e.what();
        return SW_INVALID_ARGUMENT;
    } catch (const std::logic_error& e) {
        lastError = e.what();
        return SW_LOGIC_ERROR;
    } catch (const std::bad_alloc& e) {
        lastError = e.what();
        return SW_BAD_ALLOC;
    } catch (const std::exception& e) {
        lastError = e.what();
        return SW_DEVICE_ERROR;
    } catch (...) {
        lastError = "unknown failure";
        return SW_LOGIC_ERROR;
    }
}

// Rethrows the failure a nested C entry point reported, as the exception
// that produced its status.
void check(sw_status status) {
    if (status == SW_OK) return;
    const std::string message = lastError;
    switch (status) {
    case SW_SUSPENSION_ERROR: throw swx::SuspensionFailure(lastPhase, lastEntity, message);
    case SW_OUT_OF_RANGE: throw std::out_of_range(message);
    case SW_INVALID_ARGUMENT: throw std::invalid_argument(message);
    case SW_BAD_ALLOC: throw std::bad_alloc();
    case SW_DEVICE_ERROR: throw std::runtime_error(message);
    default: throw std::logic_error(message);
    }
}

void requirePointer(const void* p, const char* what) {
    if (!p) throw std::invalid_argument(std::string(what) + " must not be null");
}

}  // namespace

// What the suspension's checkpoint section restates beyond the device state:
// the last substep's gravity and line iterations, and a section read from
// bytes (kept verbatim until the next step) with its diagnostics'
// per-segment length, commanded rest length and tension (runtime order).
struct SuspensionRecord {
    double gravity[3] = {0.0, 0.0, 0.0};
    int lineIterations = 0;
    std::vector<std::uint8_t> section;
    std::vector<double> length, commanded, tension;
};

struct sw_checkpoint {
    std::uint64_t fingerprint = 0;
    std::uint64_t acceptedSteps = 0;
    double time = 0.0;
    swx::DeviceWing::State state;
    std::vector<double> pending;
    double lastApplied[3] = {0.0, 0.0, 0.0};
    double maxResidual = 0.0;
    double lastSubstepDt = 0.0;
    bool suspensionStepped = false;
    SuspensionRecord record;
};

// Host side of the step profile: visit counts and wall-clock totals.
struct ProfileCounts {
    bool on = false;
    double totalNs = 0.0;
    std::uint64_t substeps = 0, iterations = 0, distance = 0, cable = 0, membrane = 0, bending = 0, suspension = 0,
                  contact = 0;
};

struct sw_structure {
    ProfileCounts profile;
    swx::Definition definition;
    std::uint64_t fingerprint = 0;
    std::unique_ptr<swx::DeviceWing> wing;
    int device = 0;
    std::vector<double> pending;
    bool pendingDirty = true;
    double lastApplied[3] = {0.0, 0.0, 0.0};
    std::uint64_t acceptedSteps = 0;
    double time = 0.0;
    double maxResidual = 0.0;
    // Substep length of the last accepted step, and whether the suspension
    // has finished a substep (its per-segment diagnostics exist,
    // suspension.cpp:2481-2540; before that suspensionState() reports rest
    // lengths and zero tension, structure.cpp:765-768).
    double lastSubstepDt = 0.0;
    bool suspensionStepped = false;
    // The gravity and line iterations of the last substep, and a
    // suspension checkpoint section restored from bytes and not stepped
    // since, for the checkpoint codec and suspensionState().
    SuspensionRecord record;
    std::vector<std::string> segmentNames;  // runtime order
    bool hasLoadPath = false;
};

namespace {

// What the adjoints do not differentiate: load-path cable pairs interleaved
// with the iterations, a free membrane's momentum restoration, axial line
// damping and compression-softened membranes.
void requireDifferentiable(const sw_structure& s) {
    if (s.wing->interleavedCablePairs() > 0)
        throw std::invalid_argument(
            "the adjoint reverses closing load-path cable pairs only (cableConstraintSweepPairs <= 3)");
    if (s.wing->preservesMomentum())
        throw std::invalid_argument("the adjoint does not differentiate free-membrane momentum restoration");
    if (s.wing->dampedSegments() > 0)
        throw std::invalid_argument("the adjoint does not differentiate axial line damping");
    if (s.wing->compressedMembranes())
        throw std::invalid_argument("the adjoint does not differentiate compressionStiffnessRatio other than 1");
}

// SoftBody::stepCoupled / step setting checks (soft_body.cpp:981-1050).
// `serialAllowed`: the caller can run the serial sweep (workerThreads = 0);
// the batch and the adjoints differentiate or batch the coloured sweep.
swx::StepArgs prepare(sw_structure& s, const sw_step_settings& st, bool serialAllowed = false) {
    const bool coupled = s.definition.suspension.has_value();
    const auto finite3 = [](const double* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); };
    if (!(st.time_step > 0.0) || !std::isfinite(st.time_step) || st.substeps <= 0 || st.constraint_iterations < 0 ||
        st.cable_sweep_pairs < 0 || (coupled && st.constraint_iterations == 0) || st.velocity_damping < 0.0 ||
        !std::isfinite(st.velocity_damping) || !finite3(st.gravity) || !finite3(st.damping_reference_velocity)) {
        throw std::invalid_argument(coupled ? "Invalid coupled SoftBody step settings" : "Invalid SoftBody step settings");
    }
    if (coupled && st.constraint_iterations != s.definition.suspension->solverIterations)
        throw std::invalid_argument(
            "Coupled structural iterations must equal the registered suspension lineIterations contract");
    if (st.worker_threads == 0 && !serialAllowed)
        throw std::invalid_argument(
            "workerThreads 0 selects the serial index-order sweep, which batched steps and adjoint runs do not "
            "take; use workerThreads >= 1 (the coloured sweep)");
    // workerThreads selects the sweep order (structure.h:216-218): 0 the
    // serial definition order, any other count the coloured order.
    s.wing->setSerialSweep(st.worker_threads == 0);
    swx::StepArgs a{};
    a.dt = st.time_step / st.substeps;
    a.invDt = 1.0 / a.dt;
    a.invDt2 = 1.0 / (a.dt * a.dt);
    a.damping = std::exp(-st.velocity_damping * a.dt);
    for (int k = 0; k < 3; ++k) {
        a.gravity[k] = st.gravity[k];
        a.reference[k] = st.damping_reference_velocity[k];
    }
    a.maxLineResidual = coupled ? s.definition.suspension->maximumLineResidual : 0.0;
    a.iterations = st.constraint_iterations;
    // Load-path cable sweep pairs (soft_body.cpp:1296-1315): a quarter of
    // them interleaved before the iterations' general sweeps, the rest
    // closing after them.
    s.wing->setCablePairs(st.cable_sweep_pairs, st.constraint_iterations);
    const int interleaved = s.wing->interleavedCablePairs();
    const int iterations = std::max(1, st.constraint_iterations);
    a.cableBase = interleaved / iterations;
    a.cableRem = interleaved % iterations;
    return a;
}

[[noreturn]] void throwDeviceError(sw_structure& s, const swx::DeviceError& e) {
    if (e.code == SW_SUSPENSION_ERROR) {
        std::string entity = "payload";
        if (e.kind == 1 && e.index >= 0 && static_cast<std::size_t>(e.index) < s.segmentNames.size())
            entity = s.segmentNames[e.index];
        if (e.kind == 3) entity = "line-residual";
        const char* what = e.phase == 8 ? "suspension certification failed"
                           : e.phase == 6 ? "line geometry is non-finite, scale-degenerate or invalid"
                                          : "payload state became non-finite or degenerate";
        throw swx::SuspensionFailure(e.phase, entity, what);
    }
    if (e.code == SW_INVALID_ARGUMENT && e.phase == 9)
        throw std::invalid_argument("SymmetricMatrix3 is singular or ill-conditioned (momentum restoration)");
    if (e.code == SW_INVALID_ARGUMENT)
        throw std::invalid_argument("SymmetricMatrix3 is singular or ill-conditioned; membrane element slot=" +
                                    std::to_string(e.index));
    if (e.code == SW_DEVICE_ERROR)
        throw std::runtime_error("device synchronisation timed out; the step was rolled back");
    throw std::logic_error("device step failed with code " + std::to_string(e.code));
}

void runSteps(sw_structure& s, const sw_step_settings& settings, int count, sw_diagnostics* out);

// Bookkeeping of accepted steps' last substep (solo and batched steps).
void noteSubsteps(sw_structure& s, double dt, const sw_step_settings& settings) {
    s.lastSubstepDt = dt;
    s.suspensionStepped = s.suspensionStepped || s.definition.suspension.has_value();
    for (int k = 0; k < 3; ++k) s.record.gravity[k] = settings.gravity[k];
    s.record.lineIterations = settings.constraint_iterations;
    s.record.section.clear();
    s.record.length.clear();
    s.record.commanded.clear();
    s.record.tension.clear();
}

void fillDiagnostics(sw_structure& s, sw_diagnostics* out) {
    if (!out) return;
    const swx::DeviceWing::Reductions r = s.wing->reduce();
    std::memset(out, 0, sizeof(*out));
    const swx::Definition& d = s.definition;
    out->node_count = d.nodeCount();
    out->triangle_count = d.triangleCount();
    out->constraint_count = d.constraintCount();
    out->membrane_count = d.membraneCount();
    out->dihedral_count = d.dihedralCount();
    out->dynamic_node_count = static_cast<std::size_t>(r.dynamicNodes);
    double mass = r.mass;
    double weighted[3] = {r.weighted[0], r.weighted[1], r.weighted[2]};
    double momentum[3] = {r.momentum[0], r.momentum[1], r.momentum[2]};
    double kinetic = r.kinetic;
    bool finite = r.finite != 0;
    if (d.suspension) {
        const swx::DeviceWing::State st = s.wing->download();
        const double m = d.suspension->pilotMass;
        const double* com = &st.payload[0];
        const double* vel = &st.payload[7];
        const double* w = &st.payload[10];
        for (int k = 0; k < 3; ++k) {
            weighted[k] += com[k] * m;
            momentum[k] += m * vel[k];
        }
        mass += m;
        // 0.5 m |v|^2 + 0.5 w . (R I R^T w)
        const double qw = st.payload[3], qx = st.payload[4], qy = st.payload[5], qz = st.payload[6];
        const double R[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy),
                             2 * (qx * qy + qw * qz),     1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx),
                             2 * (qx * qz - qw * qy),     2 * (qy * qz + qw * qx),     1 - 2 * (qx * qx + qy * qy)};
        double Rtw[3];
        for (int i = 0; i < 3; ++i) Rtw[i] = R[0 + i] * w[0] + R[3 + i] * w[1] + R[6 + i] * w[2];
        double rot = 0.0;
        for (int i = 0; i < 3; ++i) rot += d.suspension->inertia[i] * Rtw[i] * Rtw[i];
        kinetic += 0.5 * m * (vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]) + 0.5 * rot;
        out->suspension_segment_count = d.suspension->segmentIds.size();
        out->maximum_suspension_residual = s.maxResidual;
        for (int k = 0; k < 3; ++k) finite = finite && std::isfinite(com[k]) && std::isfinite(vel[k]) && std::isfinite(w[k]);
    }
    out->total_dynamic_mass = mass;
    for (int k = 0; k < 3; ++k) {
        out->center_of_mass[k] = mass > 0.0 ? weighted[k] * (1.0 / mass) : 0.0;
        out->linear_momentum[k] = momentum[k];
        out->last_applied_external_force[k] = s.lastApplied[k];
    }
    for (std::size_t i = 0; i < s.pending.size() / 3; ++i)
        for (int k = 0; k < 3; ++k) out->pending_external_force[k] += s.pending[3 * i + k];
    out->kinetic_energy = kinetic;
    out->maximum_distance_error = r.maxDistanceError;
    out->maximum_cable_extension = r.maxCableExtension;
    out->maximum_absolute_membrane_strain = r.maxStrain;
    out->maximum_membrane_residual = r.maxResidual;
    out->finite = (finite && std::isfinite(kinetic) && std::isfinite(r.maxDistanceError) &&
                   std::isfinite(r.maxStrain) && std::isfinite(r.maxResidual))
                      ? 1
                      : 0;
}

// Visits of `count` steps, by StepPerformanceProfile's rules
// (soft_body.cpp:1078, 1222-1245, 1389-1404), and their wall-clock time.
void countProfile(sw_structure& s, const sw_step_settings& settings, int count,
                  std::chrono::steady_clock::time_point started) {
    const swx::Definition& d = s.definition;
    std::uint64_t loadPath = 0, stitches = 0;
    for (std::uint8_t kind : d.constraintKind) {
        if (kind == SW_CONSTRAINT_CABLE || kind == SW_CONSTRAINT_SUSPENSION_TIE) ++loadPath;
        if (kind == SW_CONSTRAINT_SEAM_STITCH) ++stitches;
    }
    const std::uint64_t substeps = static_cast<std::uint64_t>(count) * static_cast<std::uint64_t>(settings.substeps);
    const std::uint64_t it = static_cast<std::uint64_t>(settings.constraint_iterations);
    ProfileCounts& p = s.profile;
    p.substeps += substeps;
    p.iterations += substeps * it;
    p.distance += substeps * (it * d.constraintCount() + (it > 0 ? stitches : 0));
    p.cable += substeps * 2ULL * static_cast<std::uint64_t>(settings.cable_sweep_pairs) * loadPath;
    p.membrane += substeps * it * d.membraneCount();
    p.bending += substeps * it * d.dihedralCount();
    if (d.suspension) p.suspension += substeps * it * d.suspension->segmentRest.size();
    if (!d.sphereNodes.empty()) p.contact += substeps * it;
    p.totalNs += std::chrono::duration<double, std::nano>(std::chrono::steady_clock::now() - started).count();
}

void runSteps(sw_structure& s, const sw_step_settings& settings, int count, sw_diagnostics* out) {
    if (count < 1) throw std::invalid_argument("step count must be positive");
    const swx::StepArgs args = prepare(s, settings, true);
    if (s.pendingDirty) {
        s.wing->setExternalForces(s.pending);
        s.pendingDirty = false;
    }
    s.wing->clearError();
    const auto started = std::chrono::steady_clock::now();
    for (int step = 0; step < count; ++step) {
        s.wing->snapshot();
        for (int sub = 0; sub < settings.substeps; ++sub) s.wing->launchSubstep(args);
    }
    double residual = 0.0;
    const swx::DeviceError e = s.wing->readError(&residual);
    if (s.profile.on) countProfile(s, settings, count, started);
    // Steps accepted: all of them, or (after a failure) those before the
    // failing one, whose pre-step snapshot is the last one taken.
    int accepted = count;
    if (e.code != 0) accepted = std::max(0, s.wing->snapshotsTaken() - 1);
    const auto commit = [&](int steps) {
        if (steps <= 0) return;
        double applied[3] = {0.0, 0.0, 0.0};
        for (std::size_t i = 0; i < s.pending.size() / 3; ++i)
            for (int k = 0; k < 3; ++k) applied[k] += s.pending[3 * i + k];
        for (int k = 0; k < 3; ++k) s.lastApplied[k] = applied[k];
        std::fill(s.pending.begin(), s.pending.end(), 0.0);
        s.pendingDirty = true;
        s.acceptedSteps += static_cast<std::uint64_t>(steps);
        for (int step = 0; step < steps; ++step) s.time += settings.time_step;
        noteSubsteps(s, args.dt, settings);
    };
    if (e.code != 0) {
        s.wing->restoreSnapshot();
        s.wing->clearError();
        s.wing->synchronize();
        commit(accepted);
        throwDeviceError(s, e);
    }
    commit(count);
    s.maxResidual = residual;
    fillDiagnostics(s, out);
}

}  // namespace

extern "C" {

const char* sw_last_error(void) { return lastError.c_str(); }
int sw_last_suspension_phase(void) { return lastPhase; }
const char* sw_last_suspension_entity(void) { return lastEntity.c_str(); }

sw_status sw_structure_create(const sw_definition* definition, int device, sw_structure** out) {
    return guarded([&] {
        requirePointer(definition, "definition");
        requirePointer(out, "out");
        *out = nullptr;
        auto s = std::make_unique<sw_structure>();
        s->definition = swx::Definition::fromView(*definition);
        swx::validateDefinition(s->definition);
        s->fingerprint = swx::definitionFingerprint(s->definition);
        for (std::size_t c = 0; c < s->definition.constraintCount(); ++c) {
            const auto kind = s->definition.constraintKind[c];
            if (kind == SW_CONSTRAINT_CABLE || kind == SW_CONSTRAINT_SUSPENSION_TIE) s->hasLoadPath = true;
        }
        s->wing = std::make_unique<swx::DeviceWing>(s->definition, device);
        s->device = device;
        s->pending.assign(3 * s->definition.nodeCount(), 0.0);
        if (s->definition.suspension) {
            const auto& layout = s->wing->suspension();
            for (std::uint32_t def : layout.segmentOrder)
                s->segmentNames.push_back("segment:" + std::to_string(s->definition.suspension->segmentIds[def]));
        }
        *out = s.release();
    });
}

void sw_structure_destroy(sw_structure* structure) { delete structure; }

sw_status sw_structure_fingerprint(const sw_structure* s, uint64_t* fingerprint) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(fingerprint, "fingerprint");
        *fingerprint = s->fingerprint;
    });
}

sw_status sw_structure_progress(const sw_structure* s, uint64_t* accepted, double* time) {
    return guarded([&] {
        requirePointer(s, "structure");
        if (accepted) *accepted = s->acceptedSteps;
        if (time) *time = s->time;
    });
}

sw_status sw_structure_clear_external_forces(sw_structure* s) {
    return guarded([&] {
        requirePointer(s, "structure");
        std::fill(s->pending.begin(), s->pending.end(), 0.0);
        s->pendingDirty = true;
    });
}

sw_status sw_structure_add_external_force(sw_structure* s, size_t node, const double force[3]) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(force, "force");
        if (node >= s->definition.nodeCount()) throw std::out_of_range("Structure force node is out of range");
        if (!std::isfinite(force[0]) || !std::isfinite(force[1]) || !std::isfinite(force[2]))
            throw std::invalid_argument("Structure force must be finite");
        for (int k = 0; k < 3; ++k) s->pending[3 * node + k] = s->pending[3 * node + k] + force[k];
        s->pendingDirty = true;
    });
}

sw_status sw_structure_set_external_forces(sw_structure* s, const double* forces, size_t nodeCount) {
    return guarded([&] {
        requirePointer(s, "structure");
        if (nodeCount != s->definition.nodeCount())
            throw std::invalid_argument("Structure force count must equal the node count");
        requirePointer(forces, "forces");
        for (std::size_t i = 0; i < 3 * nodeCount; ++i)
            if (!std::isfinite(forces[i])) throw std::invalid_argument("Structure forces must be finite");
        std::copy(forces, forces + 3 * nodeCount, s->pending.begin());
        s->pendingDirty = true;
    });
}

sw_status sw_structure_set_face_pressure(sw_structure* s, const double* pressure, size_t count) {
    return guarded([&] {
        requirePointer(s, "structure");
        if (count != s->definition.triangleCount()) throw std::out_of_range("Triangle index is out of range");
        if (count) requirePointer(pressure, "pressure");
        s->wing->setFacePressure(std::vector<double>(pressure, pressure + count));
    });
}

sw_status sw_structure_step(sw_structure* s, const sw_step_settings* settings, sw_diagnostics* diagnostics) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(settings, "settings");
        runSteps(*s, *settings, 1, diagnostics);
    });
}

sw_status sw_structure_step_many(sw_structure* s, const sw_step_settings* settings, int count) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(settings, "settings");
        // Loads apply to every step of the batch; they are consumed at the end.
        runSteps(*s, *settings, count, nullptr);
    });
}

sw_status sw_structure_diagnostics(sw_structure* s, sw_diagnostics* diagnostics) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(diagnostics, "diagnostics");
        fillDiagnostics(*s, diagnostics);
    });
}

sw_status sw_structure_node_states(const sw_structure* s, double* position, double* previous, double* velocity) {
    return guarded([&] {
        requirePointer(s, "structure");
        const swx::DeviceWing::State st = s->wing->download();
        if (position) std::copy(st.position.begin(), st.position.end(), position);
        if (previous) std::copy(st.previous.begin(), st.previous.end(), previous);
        if (velocity) std::copy(st.velocity.begin(), st.velocity.end(), velocity);
    });
}

sw_status sw_structure_frame_fields(const sw_structure* s, double* vertex_position, double* vertex_velocity,
                                    double* vertex_pending_force, double* line_length, double* line_violation) {
    return guarded([&] {
        requirePointer(s, "structure");
        const swx::DeviceWing::State st = s->wing->download();
        const swx::DeviceWing::Frame f = s->wing->frame();
        const std::size_t n = s->definition.nodeCount();
        const std::size_t h = f.harnessPosition.size() / 3;
        if (vertex_position) {
            std::copy(st.position.begin(), st.position.end(), vertex_position);
            std::copy(f.harnessPosition.begin(), f.harnessPosition.end(), vertex_position + 3 * n);
        }
        if (vertex_velocity) {
            std::copy(st.velocity.begin(), st.velocity.end(), vertex_velocity);
            std::copy(f.harnessVelocity.begin(), f.harnessVelocity.end(), vertex_velocity + 3 * n);
        }
        if (vertex_pending_force) {
            std::copy(s->pending.begin(), s->pending.end(), vertex_pending_force);
            std::fill(vertex_pending_force + 3 * n, vertex_pending_force + 3 * (n + h), 0.0);
        }
        if (line_length) std::copy(f.lineLength.begin(), f.lineLength.end(), line_length);
        if (line_violation) std::copy(f.lineViolation.begin(), f.lineViolation.end(), line_violation);
    });
}

sw_status sw_structure_frame_fields_noalloc(const sw_structure* s, double* vertex_position,
                                            double* vertex_velocity, double* vertex_pending_force,
                                            double* line_length, double* line_violation) {
    return guarded([&] {
        requirePointer(s, "structure");
        const std::size_t doubles = s->wing->frameFieldsDoubles();
        if (doubles > swx::kFrameScratchDoubles)
            throw std::invalid_argument("frame fields of this structure (" + std::to_string(doubles) +
                                        " doubles) exceed the static staging maximum of " +
                                        std::to_string(swx::kFrameScratchDoubles) + " doubles");
        double* staging = swx::frameScratchPointer(s->device);
        s->wing->frameFields(staging, vertex_position, vertex_velocity, line_length, line_violation);
        if (vertex_pending_force) {
            const std::size_t n = s->definition.nodeCount();
            const std::size_t h = s->wing->suspension().hangOrder.size();
            std::copy(s->pending.begin(), s->pending.end(), vertex_pending_force);
            std::fill(vertex_pending_force + 3 * n, vertex_pending_force + 3 * (n + h), 0.0);
        }
    });
}

namespace {

void emitFrame(const std::vector<std::uint8_t>& frame, std::uint8_t* bytes, std::size_t capacity, std::size_t* size) {
    *size = frame.size();
    if (capacity < frame.size()) throw std::invalid_argument("frame buffer is smaller than the frame");
    std::copy(frame.begin(), frame.end(), bytes);
}

}  // namespace

sw_status sw_structure_frame(sw_structure* s, const sw_frame_mapping* mapping, const sw_frame_context* context,
                             uint8_t* bytes, size_t capacity, size_t* size) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(mapping, "mapping");
        requirePointer(context, "context");
        requirePointer(size, "size");
        swx::frameLineVertices(s->definition, *mapping);
        *size = swx::frameBytes(s->definition, *mapping, *context);
        if (!bytes) return;
        if (capacity < *size) throw std::invalid_argument("frame buffer is smaller than the frame");
        const std::size_t v = mapping->vertex_count, lines = mapping->line_count;
        std::vector<double> position(3 * v), velocity(3 * v), pending(3 * v), length(lines), violation(lines);
        check(sw_structure_frame_fields(s, position.data(), velocity.data(), pending.data(), length.data(),
                                        violation.data()));
        sw_diagnostics diagnostics;
        check(sw_structure_diagnostics(s, &diagnostics));
        swx::FrameFields f;
        f.step = s->acceptedSteps;
        f.simulationTime = s->time;
        f.vertexPosition = position.data();
        f.vertexVelocity = velocity.data();
        f.vertexPendingForce = pending.data();
        f.lineLength = length.data();
        f.lineViolation = violation.data();
        f.diagnostics = &diagnostics;
        emitFrame(swx::encodeFrame(s->definition, *mapping, *context, f), bytes, capacity, size);
    });
}

sw_status sw_frame_encode(const sw_definition* definition, const sw_frame_mapping* mapping,
                          const sw_frame_context* context, uint64_t step, double simulation_time,
                          const double* vertex_position, const double* vertex_velocity,
                          const double* vertex_pending_force, const double* line_length,
                          const double* line_violation, const sw_diagnostics* diagnostics, uint8_t* bytes,
                          size_t capacity, size_t* size) {
    return guarded([&] {
        requirePointer(definition, "definition");
        requirePointer(mapping, "mapping");
        requirePointer(context, "context");
        requirePointer(size, "size");
        const swx::Definition d = swx::Definition::fromView(*definition);
        swx::frameLineVertices(d, *mapping);
        *size = swx::frameBytes(d, *mapping, *context);
        if (!bytes) return;
        swx::FrameFields f;
        f.step = step;
        f.simulationTime = simulation_time;
        f.vertexPosition = vertex_position;
        f.vertexVelocity = vertex_velocity;
        f.vertexPendingForce = vertex_pending_force;
        f.lineLength = line_length;
        f.lineViolation = line_violation;
        f.diagnostics = diagnostics;
        emitFrame(swx::encodeFrame(d, *mapping, *context, f), bytes, capacity, size);
    });
}

sw_status sw_structure_payload_state(const sw_structure* s, double* payload13) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(payload13, "payload");
        if (!s->definition.suspension) throw std::invalid_argument("structure has no suspension");
        const swx::DeviceWing::State st = s->wing->download();
        std::copy(st.payload.begin(), st.payload.begin() + 13, payload13);
    });
}

sw_status sw_structure_solver_state(const sw_structure* s, double* constraintLambda, double* membraneLambda,
                                    double* membraneResultant, double* membraneResidual, double* dihedralLambda,
                                    double* segmentLambda) {
    return guarded([&] {
        requirePointer(s, "structure");
        const swx::DeviceWing::State st = s->wing->download();
        if (constraintLambda) std::copy(st.constraintLambda.begin(), st.constraintLambda.end(), constraintLambda);
        if (membraneLambda) std::copy(st.membraneLambda.begin(), st.membraneLambda.end(), membraneLambda);
        if (membraneResultant) std::copy(st.membraneResultant.begin(), st.membraneResultant.end(), membraneResultant);
        if (membraneResidual) std::copy(st.membraneResidual.begin(), st.membraneResidual.end(), membraneResidual);
        if (dihedralLambda) std::copy(st.dihedralLambda.begin(), st.dihedralLambda.end(), dihedralLambda);
        if (segmentLambda) std::copy(st.segmentLambda.begin(), st.segmentLambda.end(), segmentLambda);
    });
}

sw_status sw_structure_checkpoint(const sw_structure* s, sw_checkpoint** out) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(out, "out");
        auto c = std::make_unique<sw_checkpoint>();
        c->fingerprint = s->fingerprint;
        c->acceptedSteps = s->acceptedSteps;
        c->time = s->time;
        c->state = s->wing->download();
        c->pending = s->pending;
        for (int k = 0; k < 3; ++k) c->lastApplied[k] = s->lastApplied[k];
        c->maxResidual = s->maxResidual;
        c->lastSubstepDt = s->lastSubstepDt;
        c->suspensionStepped = s->suspensionStepped;
        c->record = s->record;
        *out = c.release();
    });
}

sw_status sw_structure_restore(sw_structure* s, const sw_checkpoint* c) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(c, "checkpoint");
        if (c->fingerprint != s->fingerprint)
            throw std::invalid_argument("Structure checkpoint belongs to a different definition");
        if (c->pending.size() != s->pending.size() || c->state.position.size() != s->pending.size() ||
            !std::isfinite(c->time) || c->time < 0.0)
            throw std::invalid_argument("Invalid structure checkpoint");
        for (double v : c->state.position)
            if (!std::isfinite(v)) throw std::invalid_argument("Structure checkpoint is non-finite");
        s->wing->upload(c->state);
        s->pending = c->pending;
        s->pendingDirty = true;
        for (int k = 0; k < 3; ++k) s->lastApplied[k] = c->lastApplied[k];
        s->acceptedSteps = c->acceptedSteps;
        s->time = c->time;
        s->maxResidual = c->maxResidual;
        s->lastSubstepDt = c->lastSubstepDt;
        s->suspensionStepped = c->suspensionStepped;
        s->record = c->record;
    });
}

void sw_checkpoint_release(sw_checkpoint* checkpoint) { delete checkpoint; }

sw_status sw_checkpoint_serialize(const sw_structure* s, const sw_checkpoint* c, uint8_t* out, size_t capacity,
                                  size_t* size) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(c, "checkpoint");
        requirePointer(size, "size");
        if (c->fingerprint != s->fingerprint)
            throw std::invalid_argument("invalid Structure checkpoint or persistence limits");
        swx::CheckpointImage image;
        image.fingerprint = c->fingerprint;
        image.acceptedSteps = c->acceptedSteps;
        image.time = c->time;
        image.state = c->state;
        image.pending = c->pending;
        for (int k = 0; k < 3; ++k) image.lastApplied[k] = c->lastApplied[k];
        image.suspensionStepped = c->suspensionStepped;
        image.substepDt = c->lastSubstepDt;
        for (int k = 0; k < 3; ++k) image.gravity[k] = c->record.gravity[k];
        image.lineIterations = c->record.lineIterations;
        image.maxResidual = c->maxResidual;
        image.suspensionSection = c->record.section;
        const std::vector<std::uint8_t> bytes = swx::encodeStructureCheckpoint(s->definition, image);
        *size = bytes.size();
        if (!out) return;
        if (capacity < bytes.size()) throw std::invalid_argument("checkpoint buffer is smaller than the encoding");
        std::copy(bytes.begin(), bytes.end(), out);
    });
}

sw_status sw_checkpoint_deserialize(sw_structure* s, const uint8_t* bytes, size_t size, sw_checkpoint** out) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(bytes, "bytes");
        requirePointer(out, "out");
        *out = nullptr;
        swx::CheckpointImage image;
        image.state = s->wing->download();  // the shape of every state array
        swx::decodeStructureCheckpoint(s->definition, s->fingerprint, bytes, size, image);
        auto c = std::make_unique<sw_checkpoint>();
        c->fingerprint = image.fingerprint;
        c->acceptedSteps = image.acceptedSteps;
        c->time = image.time;
        // Contact-sphere multipliers restart from zero every substep.
        std::fill(image.state.sphereLambda.begin(), image.state.sphereLambda.end(), 0.0);
        c->state = std::move(image.state);
        c->pending = std::move(image.pending);
        for (int k = 0; k < 3; ++k) c->lastApplied[k] = image.lastApplied[k];
        c->suspensionStepped = image.suspensionStepped;
        c->maxResidual = image.maxResidual;
        for (int k = 0; k < 3; ++k) c->record.gravity[k] = image.gravity[k];
        c->record.lineIterations = image.lineIterations;
        c->record.section = std::move(image.suspensionSection);
        c->record.length = std::move(image.sectionLength);
        c->record.commanded = std::move(image.sectionCommanded);
        c->record.tension = std::move(image.sectionTension);
        *out = c.release();
    });
}

sw_status sw_checkpoint_info(const sw_checkpoint* c, uint32_t* version, uint64_t* fingerprint,
                             uint64_t* accepted_steps, double* simulation_time, size_t* node_count) {
    return guarded([&] {
        requirePointer(c, "checkpoint");
        if (version) *version = 2;  // structureCheckpointVersion (structure.h:248)
        if (fingerprint) *fingerprint = c->fingerprint;
        if (accepted_steps) *accepted_steps = c->acceptedSteps;
        if (simulation_time) *simulation_time = c->time;
        if (node_count) *node_count = c->pending.size() / 3;
    });
}

sw_status sw_checkpoint_nodes(const sw_checkpoint* c, double* position, double* previous, double* velocity,
                              double* pending_forces, double* last_applied_force) {
    return guarded([&] {
        requirePointer(c, "checkpoint");
        if (position) std::copy(c->state.position.begin(), c->state.position.end(), position);
        if (previous) std::copy(c->state.previous.begin(), c->state.previous.end(), previous);
        if (velocity) std::copy(c->state.velocity.begin(), c->state.velocity.end(), velocity);
        if (pending_forces) std::copy(c->pending.begin(), c->pending.end(), pending_forces);
        if (last_applied_force) std::copy(c->lastApplied, c->lastApplied + 3, last_applied_force);
    });
}

sw_status sw_structure_suspension_state(const sw_structure* s, int* present, double* payload13,
                                        double* harness_position, double* harness_velocity, uint64_t* segment_ids,
                                        double* segment_length, double* segment_commanded_rest,
                                        double* segment_tension) {
    return guarded([&] {
        requirePointer(s, "structure");
        const swx::Definition& d = s->definition;
        if (present) *present = d.suspension ? 1 : 0;
        if (!d.suspension) return;
        const swx::SuspensionDef& sd = *d.suspension;
        const std::size_t segments = sd.segmentIds.size();
        const swx::DeviceWing::State st = s->wing->download();
        if (payload13) {
            // com, q (w, x, y, z), v, w of the payload (State::payload holds
            // com, q, v, w, then the previous pose).
            std::copy(st.payload.begin(), st.payload.begin() + 13, payload13);
        }
        const swx::DeviceWing::Frame f = s->wing->frame();
        if (harness_position) std::copy(f.harnessPosition.begin(), f.harnessPosition.end(), harness_position);
        if (harness_velocity) std::copy(f.harnessVelocity.begin(), f.harnessVelocity.end(), harness_velocity);
        std::vector<double> lambda(segments, 0.0);
        const auto& order = s->wing->suspension().segmentOrder;  // runtime position -> definition index
        for (std::size_t r = 0; r < order.size(); ++r) lambda[order[r]] = st.segmentLambda[r];
        const std::size_t constraints = d.constraintCount();
        const double dt = s->lastSubstepDt;
        // A restored checkpoint section reports its own diagnostics (those of
        // the substep it was written after) until the next step.
        const bool restored = !s->record.section.empty();
        std::vector<double> length(segments), commanded(segments), tension(segments);
        for (std::size_t r = 0; r < order.size() && restored && !s->record.length.empty(); ++r) {
            length[order[r]] = s->record.length[r];
            commanded[order[r]] = s->record.commanded[r];
            tension[order[r]] = s->record.tension[r];
        }
        const bool stepped = restored ? !s->record.length.empty() : s->suspensionStepped;
        for (std::size_t j = 0; j < segments; ++j) {
            const double rest = sd.segmentRest[j];
            if (segment_ids) segment_ids[j] = sd.segmentIds[j];
            if (segment_commanded_rest) segment_commanded_rest[j] = restored && stepped ? commanded[j] : rest;
            if (segment_length)
                segment_length[j] = !stepped ? rest : restored ? length[j] : f.lineLength[constraints + j];
            if (segment_tension) segment_tension[j] = !stepped ? 0.0 : restored ? tension[j] : lambda[j] / (dt * dt);
        }
    });
}

sw_status sw_structure_colouring(const sw_structure* s, int family, uint64_t* order, size_t* orderCount,
                                 uint64_t* offsets, size_t* offsetCount) {
    return guarded([&] {
        requirePointer(s, "structure");
        const swx::ColourOrder* c = family == 0   ? &s->wing->distanceColours()
                                    : family == 1 ? &s->wing->membraneColours()
                                    : family == 2 ? &s->wing->hingeColours()
                                                  : nullptr;
        if (!c) throw std::out_of_range("colouring family must be 0, 1 or 2");
        if (orderCount) *orderCount = c->order.size();
        if (offsetCount) *offsetCount = c->offsets.size();
        if (order) std::copy(c->order.begin(), c->order.end(), order);
        if (offsets) std::copy(c->offsets.begin(), c->offsets.end(), offsets);
    });
}

sw_status sw_definition_fingerprint(const sw_definition* definition, uint64_t* fingerprint) {
    return guarded([&] {
        requirePointer(definition, "definition");
        requirePointer(fingerprint, "fingerprint");
        const swx::Definition d = swx::Definition::fromView(*definition);
        swx::validateDefinition(d);
        *fingerprint = swx::definitionFingerprint(d);
    });
}

sw_status sw_definition_colouring(const sw_definition* definition, int family, uint64_t* order, size_t* orderCount,
                                  uint64_t* offsets, size_t* offsetCount) {
    return guarded([&] {
        requirePointer(definition, "definition");
        const swx::Definition d = swx::Definition::fromView(*definition);
        swx::validateDefinition(d);
        swx::ColourOrder c;
        if (family == 0) c = swx::colourDistance(d);
        else if (family == 1) c = swx::colourMembranes(d);
        else if (family == 2) c = swx::colourDihedrals(d);
        else if (family == 3) {
            const swx::SuspensionLayout layout = swx::suspensionLayout(d);
            std::vector<std::vector<std::uint32_t>> writes(layout.from.size()), reads(layout.from.size());
            std::vector<std::uint8_t> payload(layout.from.size(), 0);
            for (std::size_t k = 0; k < layout.from.size(); ++k)
                for (int e : {layout.from[k], layout.to[k]}) {
                    if (e >= 0) {
                        writes[k].push_back(static_cast<std::uint32_t>(e));
                        reads[k].push_back(static_cast<std::uint32_t>(e));
                    } else {
                        payload[k] = 1;
                    }
                }
            const auto level = swx::dependencyLevels(writes, reads, payload, d.nodeCount());
            std::uint32_t levels = 0;
            for (auto l : level) levels = std::max(levels, l);
            c.offsets.assign(levels + 1, 0);
            for (auto l : level) ++c.offsets[l];
            for (std::uint32_t l = 1; l <= levels; ++l) c.offsets[l] += c.offsets[l - 1];
            std::vector<std::uint32_t> cursor(c.offsets.begin(), c.offsets.end() - 1);
            c.order.resize(level.size());
            for (std::size_t k = 0; k < level.size(); ++k) c.order[cursor[level[k] - 1]++] = layout.segmentOrder[k];
        } else {
            throw std::out_of_range("colouring family must be 0, 1, 2 or 3");
        }
        if (orderCount) *orderCount = c.order.size();
        if (offsetCount) *offsetCount = c.offsets.size();
        if (order) std::copy(c.order.begin(), c.order.end(), order);
        if (offsets) std::copy(c.offsets.begin(), c.offsets.end(), offsets);
    });
}

sw_status sw_device_math(int device, int function, const double* a, const double* b, double* out, size_t count) {
    return guarded([&] {
        if (count) {
            requirePointer(a, "a");
            requirePointer(b, "b");
            requirePointer(out, "out");
        }
        swx::deviceMath(device, function, a, b, out, count);
    });
}

sw_status sw_structure_set_launch(sw_structure* s, int workgroups, int threads) {
    return guarded([&] {
        requirePointer(s, "structure");
        s->wing->setLaunch(workgroups, threads);
    });
}

sw_status sw_structure_set_cooperative(sw_structure* s, int cooperative) {
    return guarded([&] {
        requirePointer(s, "structure");
        s->wing->setCooperative(cooperative != 0);
    });
}

sw_status sw_structure_set_xcd_local(sw_structure* s, int enabled, int* active) {
    return guarded([&] {
        requirePointer(s, "structure");
        const bool on = s->wing->setXcdLocal(enabled != 0);
        if (active) *active = on && enabled ? 1 : 0;
    });
}

sw_status sw_structure_set_kernel_timing(sw_structure* s, int enabled) {
    return guarded([&] {
        requirePointer(s, "structure");
        s->wing->setTiming(enabled != 0);
    });
}

sw_status sw_structure_kernel_time(sw_structure* s, double* milliseconds, long long* launches) {
    return guarded([&] {
        requirePointer(s, "structure");
        s->wing->kernelTime(milliseconds, launches);
    });
}

sw_status sw_structure_set_profiling(sw_structure* s, int enabled) {
    return guarded([&] {
        requirePointer(s, "structure");
        if (enabled && !s->profile.on) {
            s->profile = ProfileCounts{};
            s->wing->setProfiling(true);
            s->wing->clearProfile();
        } else if (!enabled) {
            s->wing->setProfiling(false);
        }
        s->profile.on = enabled != 0;
    });
}

sw_status sw_structure_reset_profile(sw_structure* s) {
    return guarded([&] {
        requirePointer(s, "structure");
        const bool on = s->profile.on;
        s->profile = ProfileCounts{};
        s->profile.on = on;
        s->wing->clearProfile();
    });
}

sw_status sw_structure_profile(sw_structure* s, sw_step_profile* out) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(out, "profile");
        double parts[9];
        long long launches = 0;
        s->wing->profileTotals(parts, &launches);
        const ProfileCounts& p = s->profile;
        const auto ns = [](double v) { return static_cast<std::uint64_t>(v < 0.0 ? 0.0 : v + 0.5); };
        *out = sw_step_profile{};
        out->soft_body_total_ns = ns(p.totalNs);
        out->prediction_ns = ns(parts[0]);
        out->distance_constraint_ns = ns(parts[1]);
        out->cable_constraint_ns = ns(parts[2]);
        out->membrane_constraint_ns = ns(parts[3]);
        out->bending_constraint_ns = ns(parts[4]);
        out->suspension_constraint_ns = ns(parts[5]);
        out->contact_constraint_ns = ns(parts[6]);
        out->membrane_diagnostics_ns = ns(parts[7]);
        out->finalization_ns = ns(parts[8]);
        double kernel = 0.0;
        for (double v : parts) kernel += v;
        out->transaction_snapshot_ns = ns(p.totalNs - kernel);
        out->structural_substeps = p.substeps;
        out->constraint_iterations = p.iterations;
        out->distance_constraint_visits = p.distance;
        out->cable_constraint_visits = p.cable;
        out->membrane_constraint_visits = p.membrane;
        out->bending_constraint_visits = p.bending;
        out->suspension_constraint_visits = p.suspension;
        out->contact_constraint_visits = p.contact;
    });
}

sw_status sw_structure_get_launch(const sw_structure* s, int* workgroups, int* threads) {
    return guarded([&] {
        requirePointer(s, "structure");
        if (workgroups) *workgroups = s->wing->workgroups();
        if (threads) *threads = s->wing->threads();
    });
}

}  // extern "C"

// ---- scene and assembly -------------------------------------------------

struct sw_scene {
    swx::GeneratedWing wing;
};

struct sw_assembly {
    swx::Assembly assembly;
    sw_suspension_definition suspensionView{};
    sw_definition view{};
    swx::V3 pilotCentre{};
};

extern "C" {

sw_status sw_scene_generate(const char* preset, sw_scene** out) {
    return guarded([&] {
        requirePointer(preset, "preset");
        requirePointer(out, "out");
        std::string name(preset);
        std::string overrides;
        if (const auto q = name.find('?'); q != std::string::npos) {
            overrides = name.substr(q + 1);
            name = name.substr(0, q);
        }
        swx::WingParams params;
        if (name.rfind("batch:", 0) == 0) {
            unsigned long long seed = 0;
            int index = 0, count = 0;
            if (std::sscanf(name.c_str(), "batch:%llu:%d:%d", &seed, &index, &count) != 3 || index < 0 ||
                index >= count)
                throw std::invalid_argument("batch preset must be batch:<seed>:<index>:<count>");
            params = swx::batchPresets(seed, count, 2000.0, 40000.0)[static_cast<std::size_t>(index)];
        } else if (name.rfind("adjbatch:", 0) == 0) {
            // The batched-adjoint wings: triangle counts log-uniform in
            // about 5k..70k (node counts 3.3k..42.7k), at the quantiles.
            unsigned long long seed = 0;
            int index = 0, count = 0;
            if (std::sscanf(name.c_str(), "adjbatch:%llu:%d:%d", &seed, &index, &count) != 3 || index < 0 ||
                index >= count)
                throw std::invalid_argument("adjoint batch preset must be adjbatch:<seed>:<index>:<count>");
            params = swx::batchPresets(seed, count, 3300.0, 42700.0, true)[static_cast<std::size_t>(index)];
        } else {
            params = swx::wingPreset(name);
        }
        swx::applyOverrides(params, overrides);
        auto s = std::make_unique<sw_scene>();
        s->wing = swx::generateWing(params);
        *out = s.release();
    });
}

void sw_scene_destroy(sw_scene* scene) { delete scene; }

sw_status sw_scene_read(const uint8_t* bytes, size_t size, sw_scene** out) {
    return guarded([&] {
        requirePointer(out, "out");
        auto s = std::make_unique<sw_scene>();
        s->wing.scene = swx::readSceneBinary(bytes, size);
        s->wing.facePressure.assign(s->wing.scene.triangles.size(), 0.0);
        *out = s.release();
    });
}

sw_status sw_scene_write(const sw_scene* scene, uint8_t* bytes, size_t capacity, size_t* size) {
    return guarded([&] {
        requirePointer(scene, "scene");
        requirePointer(size, "size");
        const std::vector<std::uint8_t> data = swx::writeSceneBinary(scene->wing.scene);
        *size = data.size();
        if (!bytes) return;
        if (capacity < data.size()) throw std::invalid_argument("scene buffer is smaller than the stream");
        std::copy(data.begin(), data.end(), bytes);
    });
}

sw_status sw_scene_face_pressure(const sw_scene* scene, const double** pressure, size_t* count) {
    return guarded([&] {
        requirePointer(scene, "scene");
        if (pressure) *pressure = scene->wing.facePressure.data();
        if (count) *count = scene->wing.facePressure.size();
    });
}

sw_status sw_scene_assemble(const sw_scene* scene, const sw_assembly_settings* settings, sw_assembly** out) {
    return guarded([&] {
        requirePointer(scene, "scene");
        requirePointer(out, "out");
        swx::AssemblySettings s;
        if (settings) {
            s.seamCoincidenceTolerance = settings->seam_coincidence_tolerance;
            s.suspensionIterations = settings->suspension_iterations;
            s.suspensionAttachmentTolerance = settings->suspension_attachment_tolerance;
            s.suspensionMinimumLineLength = settings->suspension_minimum_line_length;
            s.suspensionMaximumLineResidual = settings->suspension_maximum_line_residual;
            s.suspensionMaximumControlWork = settings->suspension_maximum_control_work;
            s.contactSphereCompliance = settings->contact_sphere_compliance;
        }
        auto a = std::make_unique<sw_assembly>();
        a->assembly = swx::assembleScene(scene->wing.scene, s, scene->wing.contactSpheres);
        a->view = a->assembly.definition.view(&a->suspensionView);
        if (!scene->wing.scene.pilots.empty()) {
            const auto& p = scene->wing.scene.pilots.front();
            a->pilotCentre = p.centerOfMass;
        }
        *out = a.release();
    });
}

void sw_assembly_destroy(sw_assembly* assembly) { delete assembly; }

sw_status sw_assembly_definition(const sw_assembly* assembly, sw_definition* out) {
    return guarded([&] {
        requirePointer(assembly, "assembly");
        requirePointer(out, "out");
        *out = assembly->view;
    });
}

sw_status sw_assembly_info(const sw_assembly* a, size_t* vertexNodes, size_t* junctionNodes, double* fabricMass,
                           double* seamMass) {
    return guarded([&] {
        requirePointer(a, "assembly");
        if (vertexNodes) *vertexNodes = a->assembly.nodeVertexIds.size();
        if (junctionNodes) *junctionNodes = a->assembly.nodeJunctionIds.size();
        if (fabricMass) *fabricMass = a->assembly.totalFabricMass;
        if (seamMass) *seamMass = a->assembly.totalSeamMass;
    });
}

sw_status sw_assembly_flight_loads(const sw_assembly* a, double* forces) {
    return guarded([&] {
        requirePointer(a, "assembly");
        requirePointer(forces, "forces");
        const swx::Definition& d = a->assembly.definition;
        double total = 0.0;
        for (double m : d.nodeMass) total += m;
        if (d.suspension) total += d.suspension->pilotMass;
        swx::V3 pilot{};
        if (d.suspension) pilot = {d.suspension->comWorld[0], d.suspension->comWorld[1], d.suspension->comWorld[2]};
        const swx::V3 centre = swx::flightLoadCentre(d.nodePosition.data(), d.nodeFixed.data(), d.nodeCount(),
                                                     d.suspension ? &pilot : nullptr);
        swx::flightLoads(d.nodePosition.data(), d.nodeMass.data(), d.nodeCount(), a->assembly.nodeVertexIds.size(),
                         total * 9.80665, centre, forces);
    });
}

}  // extern "C"

// ---- batched wings ---------------------------------------------------------

// Wings of uneven size stepped by one persistent launch. Each wing is a full
// structure (definition, validation, pending loads, accounting, checkpoint
// state); the batch shares only the launch. A wing whose step fails is
// restored to its own pre-step state; the others keep their accepted steps.
struct sw_batch {
    std::vector<std::unique_ptr<sw_structure>> wings;
    std::unique_ptr<swx::DeviceBatch> device;
};

namespace {

sw_structure& batchWing(const sw_batch* b, size_t wing) {
    requirePointer(b, "batch");
    if (wing >= b->wings.size()) throw std::out_of_range("wing index is outside the batch");
    return *b->wings[wing];
}

}  // namespace

extern "C" {

sw_status sw_batch_create(const sw_definition* const* definitions, size_t count, int device, sw_batch** out) {
    return guarded([&] {
        requirePointer(definitions, "definitions");
        requirePointer(out, "out");
        *out = nullptr;
        if (count == 0) throw std::invalid_argument("a batch needs at least one wing");
        auto b = std::make_unique<sw_batch>();
        std::vector<swx::DeviceWing*> devices;
        for (size_t w = 0; w < count; ++w) {
            sw_structure* s = nullptr;
            const sw_status status = sw_structure_create(definitions[w], device, &s);
            if (status != SW_OK) throw std::invalid_argument("wing " + std::to_string(w) + ": " + lastError);
            b->wings.emplace_back(s);
            devices.push_back(s->wing.get());
        }
        b->device = std::make_unique<swx::DeviceBatch>(devices, device);
        b->device->plan(64, 256);
        *out = b.release();
    });
}

void sw_batch_destroy(sw_batch* batch) { delete batch; }

sw_status sw_batch_set_face_pressure(sw_batch* b, size_t wing, const double* pressure, size_t count) {
    return guarded([&] { check(sw_structure_set_face_pressure(&batchWing(b, wing), pressure, count)); });
}

sw_status sw_batch_set_external_forces(sw_batch* b, size_t wing, const double* forces, size_t node_count) {
    return guarded([&] { check(sw_structure_set_external_forces(&batchWing(b, wing), forces, node_count)); });
}

sw_status sw_batch_step(sw_batch* b, const sw_step_settings* settings, int count) {
    return guarded([&] {
        requirePointer(b, "batch");
        requirePointer(settings, "settings");
        if (count < 1) throw std::invalid_argument("step count must be positive");
        const size_t n = b->wings.size();
        std::vector<swx::StepArgs> args(n);
        for (size_t w = 0; w < n; ++w) args[w] = prepare(*b->wings[w], *settings);
        b->device->ensurePlan();
        for (auto& s : b->wings) {
            if (s->pendingDirty) {
                s->wing->setExternalForces(s->pending);
                s->pendingDirty = false;
            }
            s->wing->clearError();
            s->wing->synchronize();
        }
        for (int step = 0; step < count; ++step) {
            b->device->snapshot();
            for (int sub = 0; sub < settings->substeps; ++sub) b->device->launchSubstep(args);
        }
        b->device->synchronize();
        std::vector<unsigned char> failed(n, 0);
        std::vector<swx::DeviceError> errors(n);
        std::vector<double> residual(n, 0.0);
        bool any = false;
        for (size_t w = 0; w < n; ++w) {
            errors[w] = b->wings[w]->wing->readError(&residual[w]);
            failed[w] = errors[w].code != 0;
            any = any || failed[w];
        }
        if (any) {
            b->device->restore(failed);
            for (size_t w = 0; w < n; ++w)
                if (failed[w]) b->wings[w]->wing->clearError();
        }
        for (size_t w = 0; w < n; ++w) {
            if (failed[w]) continue;
            sw_structure& s = *b->wings[w];
            double applied[3] = {0.0, 0.0, 0.0};
            for (size_t i = 0; i < s.pending.size() / 3; ++i)
                for (int k = 0; k < 3; ++k) applied[k] += s.pending[3 * i + k];
            for (int k = 0; k < 3; ++k) s.lastApplied[k] = applied[k];
            std::fill(s.pending.begin(), s.pending.end(), 0.0);
            s.pendingDirty = true;
            s.acceptedSteps += static_cast<std::uint64_t>(count);
            for (int step = 0; step < count; ++step) s.time += settings->time_step;
            s.maxResidual = residual[w];
            noteSubsteps(s, settings->time_step / settings->substeps, *settings);
        }
        for (size_t w = 0; w < n; ++w)
            if (failed[w]) throwDeviceError(*b->wings[w], errors[w]);
    });
}

sw_status sw_batch_node_states(const sw_batch* b, size_t wing, double* position, double* velocity) {
    return guarded([&] {
        sw_structure& s = batchWing(b, wing);
        check(sw_structure_node_states(&s, position, nullptr, velocity));
    });
}

sw_status sw_batch_set_launch(sw_batch* b, int workgroups_per_wing_cap, int threads) {
    return guarded([&] {
        requirePointer(b, "batch");
        b->device->plan(workgroups_per_wing_cap, threads);
    });
}

sw_status sw_batch_set_cooperative(sw_batch* b, int cooperative) {
    return guarded([&] {
        requirePointer(b, "batch");
        b->device->setCooperative(cooperative != 0);
    });
}

sw_status sw_batch_structure(sw_batch* b, size_t wing, sw_structure** out) {
    return guarded([&] {
        requirePointer(out, "out");
        *out = &batchWing(b, wing);
    });
}

sw_status sw_batch_workgroups(const sw_batch* b, int* per_wing, size_t count) {
    return guarded([&] {
        requirePointer(b, "batch");
        const auto& blocks = b->device->blocks();
        if (count != blocks.size()) throw std::invalid_argument("per_wing must hold one entry per wing");
        for (size_t w = 0; w < count; ++w) per_wing[w] = blocks[w];
    });
}

// ---- Batched adjoint ---------------------------------------------------

struct sw_batch_adjoint {
    sw_batch* batch = nullptr;
    std::vector<std::vector<double>> loads;  // per wing, pending at configuration
    std::unique_ptr<swx::DeviceBatchAdjoint> device;
};

sw_status sw_batch_adjoint_configure(sw_batch* b, const sw_step_settings* settings, uint64_t steps,
                                     uint32_t checkpoint_interval, uint32_t jacobian_chunk, sw_batch_adjoint** out) {
    return guarded([&] {
        requirePointer(b, "batch");
        requirePointer(settings, "settings");
        requirePointer(out, "out");
        if (steps < 1 || steps > (1ull << 40)) throw std::invalid_argument("adjoint step count must be positive");
        if (checkpoint_interval < 1) throw std::invalid_argument("checkpoint interval must be positive");
        auto a = std::make_unique<sw_batch_adjoint>();
        a->batch = b;
        std::vector<const swx::Definition*> definitions;
        std::vector<swx::StepArgs> args;
        for (auto& s : b->wings) {
            args.push_back(prepare(*s, *settings));
            requireDifferentiable(*s);
            definitions.push_back(&s->definition);
            a->loads.push_back(s->pending);
        }
        a->device = std::make_unique<swx::DeviceBatchAdjoint>(*b->device, definitions, args, settings->substeps,
                                                              static_cast<long>(steps),
                                                              static_cast<int>(checkpoint_interval),
                                                              static_cast<int>(jacobian_chunk));
        *out = a.release();
    });
}

void sw_batch_adjoint_destroy(sw_batch_adjoint* a) { delete a; }

sw_status sw_batch_adjoint_run(sw_batch_adjoint* a, const double* const* rest_lengths, const double* const* targets,
                               double* losses, double* const* grad_rest_lengths, double* const* final_positions) {
    return guarded([&] {
        requirePointer(a, "adjoint");
        requirePointer(targets, "targets");
        const size_t n = a->batch->wings.size();
        for (size_t w = 0; w < n; ++w) {
            sw_structure& s = *a->batch->wings[w];
            s.wing->setExternalForces(a->loads[w]);
            s.pendingDirty = true;
        }
        std::vector<const double*> rest, target(targets, targets + n);
        if (rest_lengths) rest.assign(rest_lengths, rest_lengths + n);
        std::vector<double*> grads, finals;
        if (grad_rest_lengths) grads.assign(grad_rest_lengths, grad_rest_lengths + n);
        if (final_positions) finals.assign(final_positions, final_positions + n);
        try {
            a->device->run(rest, target, losses, grads, finals);
        } catch (const swx::DeviceStepFailure& e) {
            throwDeviceError(*a->batch->wings[static_cast<size_t>(std::max(a->device->failedWing(), 0))], e.error);
        }
    });
}

sw_status sw_batch_adjoint_timing(const sw_batch_adjoint* a, double* forward_ms, double* reverse_ms,
                                  size_t* device_bytes) {
    return guarded([&] {
        requirePointer(a, "adjoint");
        if (forward_ms) *forward_ms = a->device->forwardMs();
        if (reverse_ms) *reverse_ms = a->device->reverseMs();
        if (device_bytes) *device_bytes = a->device->deviceBytes();
    });
}

sw_status sw_batch_adjoint_timing_detail(const sw_batch_adjoint* a, double* replay_ms, double* jacobian_ms,
                                         double* sweep_ms) {
    return guarded([&] {
        requirePointer(a, "adjoint");
        if (replay_ms) *replay_ms = a->device->replayMs();
        if (jacobian_ms) *jacobian_ms = a->device->jacobianMs();
        if (sweep_ms) *sweep_ms = a->device->sweepMs();
    });
}

sw_status sw_batch_adjoint_workgroups(const sw_batch_adjoint* a, int* per_wing, size_t count) {
    return guarded([&] {
        requirePointer(a, "adjoint");
        const auto& blocks = a->device->sweepWorkgroups();
        if (count != blocks.size()) throw std::invalid_argument("per_wing must hold one entry per wing");
        for (size_t w = 0; w < count; ++w) per_wing[w] = blocks[w];
    });
}

// ---- Adjoint -----------------------------------------------------------

struct sw_adjoint {
    sw_structure* structure = nullptr;
    std::vector<double> loads;  // pending at configuration, applied at every step
    std::unique_ptr<swx::DeviceAdjoint> device;
};

sw_status sw_adjoint_configure(sw_structure* s, const sw_step_settings* settings, uint64_t steps,
                               uint32_t checkpoint_interval, sw_adjoint** out) {
    return guarded([&] {
        requirePointer(s, "structure");
        requirePointer(settings, "settings");
        requirePointer(out, "out");
        const swx::StepArgs args = prepare(*s, *settings);
        if (steps < 1 || steps > (1ull << 40)) throw std::invalid_argument("adjoint step count must be positive");
        if (checkpoint_interval < 1) throw std::invalid_argument("checkpoint interval must be positive");
        requireDifferentiable(*s);
        auto a = std::make_unique<sw_adjoint>();
        a->structure = s;
        a->loads = s->pending;
        a->device = std::make_unique<swx::DeviceAdjoint>(*s->wing, s->definition, args, settings->substeps,
                                                         static_cast<long>(steps), static_cast<int>(checkpoint_interval));
        *out = a.release();
    });
}

void sw_adjoint_destroy(sw_adjoint* a) { delete a; }

sw_status sw_adjoint_run(sw_adjoint* a, const double* log_compliance, const double* rest_length,
                         const double* pressure, const double* target, double* loss, double* final_position,
                         double* grad_log_compliance, double* grad_rest_length, double* grad_pressure) {
    return guarded([&] {
        requirePointer(a, "adjoint");
        requirePointer(target, "target");
        sw_structure& s = *a->structure;
        s.wing->setExternalForces(a->loads);
        s.pendingDirty = true;  // the next regular step uploads its own pending loads
        try {
            const double value = a->device->run(log_compliance, rest_length, pressure, target, final_position,
                                                grad_log_compliance, grad_rest_length, grad_pressure);
            if (loss) *loss = value;
        } catch (const swx::DeviceStepFailure& e) {
            throwDeviceError(s, e.error);
        }
    });
}

sw_status sw_adjoint_set_workgroups(sw_adjoint* a, int workgroups) {
    return guarded([&] {
        requirePointer(a, "adjoint");
        a->device->setWorkgroups(workgroups);
    });
}

sw_status sw_adjoint_timing_detail(const sw_adjoint* a, double* replay_ms, double* jacobian_ms, double* sweep_ms) {
    return guarded([&] {
        requirePointer(a, "adjoint");
        if (replay_ms) *replay_ms = a->device->replayMs();
        if (jacobian_ms) *jacobian_ms = a->device->jacobianMs();
        if (sweep_ms) *sweep_ms = a->device->sweepMs();
    });
}

sw_status sw_adjoint_timing(const sw_adjoint* a, double* forward_ms, double* reverse_ms, size_t* device_bytes) {
    return guarded([&] {
        requirePointer(a, "adjoint");
        if (forward_ms) *forward_ms = a->device->forwardMs();
        if (reverse_ms) *reverse_ms = a->device->reverseMs();
        if (device_bytes) *device_bytes = a->device->deviceBytes();
    });
}

}  // extern "C"
