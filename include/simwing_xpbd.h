/*
 * simwing_xpbd.h -- C-ABI of the MI355X XPBD structural substep.
 *
 * This is the drop-in boundary for SimWing's structural adapter
 * simwing::fsi::Structure (reference src/fsi/structure.h:273-309) and the
 * SoftBody-level extensions it needs (per-face pressure, soft_body.h:401-402;
 * colouring view, soft_body.h:462-474). Everything is plain C: opaque
 * handles, SoA arrays of doubles/uint64, status codes. Caller arrays are
 * borrowed for the duration of the call only.
 *
 * Errors: every call returns an sw_status. The code mirrors the C++
 * exception class the reference throws (std::invalid_argument ->
 * SW_INVALID_ARGUMENT, std::out_of_range -> SW_OUT_OF_RANGE,
 * softwing::SuspensionError -> SW_SUSPENSION_ERROR with phase and entity,
 * std::logic_error -> SW_LOGIC_ERROR); SW_DEVICE_ERROR reports a HIP runtime
 * failure. The message of the last failure on the calling thread is
 * available from sw_last_error().
 *
 * Threading: a structure is single-owner and not thread-safe, like the
 * reference adapter. Each structure owns one HIP stream; sw_structure_step
 * is synchronous and transactional (a failed step restores the exact
 * pre-step state before returning the error, structure.cpp:807-831).
 */
#ifndef SIMWING_XPBD_H
#define SIMWING_XPBD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sw_status {
    SW_OK = 0,
    SW_INVALID_ARGUMENT = 1,
    SW_OUT_OF_RANGE = 2,
    SW_SUSPENSION_ERROR = 3,
    SW_LOGIC_ERROR = 4,
    SW_DEVICE_ERROR = 5,
    SW_BAD_ALLOC = 6
} sw_status;

/* Last failure on this thread: message, and for SW_SUSPENSION_ERROR the
 * softwing::SuspensionPhase value (suspension.h:29-40) and entity id. */
const char* sw_last_error(void);
int sw_last_suspension_phase(void);
const char* sw_last_suspension_entity(void);

/* ---- Definition (structure.h:30-230, as SoA views) ------------------- */

enum {
    SW_CONSTRAINT_DISTANCE = 1,
    SW_CONSTRAINT_CABLE = 2,
    SW_CONSTRAINT_SUSPENSION_TIE = 3,
    SW_CONSTRAINT_SEAM_STITCH = 4
};
enum { SW_ENDPOINT_ATTACHMENT = 1, SW_ENDPOINT_JUNCTION = 2, SW_ENDPOINT_HARNESS = 3 };
enum { SW_ROLE_BULK = 1, SW_ROLE_SEAM = 2, SW_ROLE_REINFORCEMENT = 3 };

/* StructureSuspensionDefinition (structure.h:147-166). Arrays are indexed
 * by entry; segment endpoint kinds use SW_ENDPOINT_*. */
typedef struct sw_suspension_definition {
    uint64_t pilot_stable_id;
    double pilot_mass;
    double pilot_com_local[3];
    double pilot_com_world[3];
    double pilot_velocity[3];
    double pilot_angular_velocity[3];
    double pilot_orientation[4]; /* w, x, y, z */
    double pilot_inertia[3];     /* principal, body frame */
    size_t attachment_count;
    const uint64_t* attachment_ids;
    const uint64_t* attachment_nodes;
    size_t junction_count;
    const uint64_t* junction_ids;
    const uint64_t* junction_nodes;
    size_t harness_count;
    const uint64_t* harness_ids;
    const double* harness_local; /* 3 per harness point */
    size_t segment_count;
    const uint64_t* segment_ids;
    const uint8_t* segment_from_kind;
    const uint64_t* segment_from_id;
    const uint8_t* segment_to_kind;
    const uint64_t* segment_to_id;
    const double* segment_rest;
    const double* segment_stiffness;
    const double* segment_damping;
    const uint32_t* segment_role;
    int solver_iterations;
    double attachment_tolerance;
    double minimum_line_length;
    double maximum_line_residual;
    double maximum_control_work;
} sw_suspension_definition;

/* StructureDefinition (structure.h:168-176). Node indices are uint64.
 * membrane_chart: 6 per membrane (three x,y corners); membrane_material: 8
 * per membrane (warp, weft, coupling, shear, warp pre-tension, weft
 * pre-tension, damping seconds, compression ratio).
 * Contact spheres are a build-defined family with no reference
 * counterpart: sphere k is centred on node sphere_centre[k] with radius
 * sphere_radius[k] and compliance sphere_compliance[k], and holds off the
 * nodes sphere_nodes[sphere_offsets[k] .. sphere_offsets[k+1]). */
typedef struct sw_definition {
    size_t node_count;
    const double* node_position; /* 3 per node */
    const double* node_mass;
    const uint8_t* node_fixed;
    size_t triangle_count;
    const uint64_t* triangle_nodes; /* 3 per triangle */
    size_t constraint_count;
    const uint8_t* constraint_kind;
    const uint64_t* constraint_nodes; /* 2 per constraint */
    const double* constraint_rest;
    const double* constraint_compliance;
    size_t membrane_count;
    const uint64_t* membrane_triangle;
    const double* membrane_chart;
    const double* membrane_material;
    const uint8_t* membrane_role;
    size_t dihedral_count;
    const uint64_t* dihedral_nodes; /* 4 per hinge */
    const double* dihedral_rest_angle;
    const double* dihedral_compliance;
    const sw_suspension_definition* suspension; /* NULL: no pilot */
    size_t sphere_count;
    const uint64_t* sphere_centre;
    const double* sphere_radius;
    const double* sphere_compliance;
    const uint64_t* sphere_offsets; /* sphere_count + 1 */
    const uint64_t* sphere_nodes;
} sw_definition;

/* StructureStepSettings (structure.h:208-219). worker_threads selects the
 * sweep order as in the reference: 0 the serial sweep (distance constraints
 * and membranes in definition order, soft_body.cpp:1316-1345), any other
 * count the coloured sweep (the count itself does not change results). The
 * device runs the serial order as dependency levels over the definition
 * order. Batched steps and adjoint runs take the coloured sweep only
 * (worker_threads = 0 is rejected there). */
typedef struct sw_step_settings {
    double time_step;
    int substeps;
    int constraint_iterations;
    int cable_sweep_pairs;
    double gravity[3];
    double velocity_damping;
    double damping_reference_velocity[3];
    unsigned worker_threads;
} sw_step_settings;

/* StructureDiagnostics (structure.h:221-246). */
typedef struct sw_diagnostics {
    size_t node_count;
    size_t dynamic_node_count;
    size_t triangle_count;
    size_t constraint_count;
    size_t membrane_count;
    size_t dihedral_count;
    size_t contact_pair_count;
    size_t active_contact_count;
    size_t suspension_segment_count;
    double total_dynamic_mass;
    double center_of_mass[3];
    double linear_momentum[3];
    double kinetic_energy;
    double maximum_distance_error;
    double maximum_cable_extension;
    double maximum_absolute_membrane_strain;
    double maximum_membrane_residual;
    double maximum_contact_penetration;
    double maximum_suspension_residual;
    double pending_external_force[3];
    double last_applied_external_force[3];
    int finite;
} sw_diagnostics;

/* ---- Structure (structure.h:273-309) ---------------------------------- */

typedef struct sw_structure sw_structure;
typedef struct sw_checkpoint sw_checkpoint;

/* Structure(StructureDefinition) (structure.cpp:682-685): validates, builds
 * the colour schedules and uploads the SoA tables to device `device`. */
sw_status sw_structure_create(const sw_definition* definition, int device, sw_structure** out);
void sw_structure_destroy(sw_structure* structure);

/* definitionFingerprint() (structure.cpp:326-431), acceptedStepCount(),
 * simulationTimeSeconds(). */
sw_status sw_structure_fingerprint(const sw_structure* s, uint64_t* fingerprint);
/* The same fingerprint computed on the host from a definition (validated as
 * sw_structure_create validates it; no device involved). */
sw_status sw_definition_fingerprint(const sw_definition* definition, uint64_t* fingerprint);
sw_status sw_structure_progress(const sw_structure* s, uint64_t* accepted_steps, double* time);

/* clearExternalForces / addExternalForce / setExternalForces
 * (structure.cpp:742-776): pending loads consumed by the next accepted step. */
sw_status sw_structure_clear_external_forces(sw_structure* s);
sw_status sw_structure_add_external_force(sw_structure* s, size_t node, const double force[3]);
sw_status sw_structure_set_external_forces(sw_structure* s, const double* forces, size_t node_count);

/* SoftBody::setFacePressureDifference for every triangle (soft_body.h:401). */
sw_status sw_structure_set_face_pressure(sw_structure* s, const double* pressure, size_t triangle_count);

/* step(settings) -> diagnostics (structure.cpp:807-831); diagnostics()
 * (structure.cpp:833-959). `diagnostics` may be NULL. */
sw_status sw_structure_step(sw_structure* s, const sw_step_settings* settings, sw_diagnostics* diagnostics);
sw_status sw_structure_diagnostics(sw_structure* s, sw_diagnostics* diagnostics);

/* Steps `count` times back to back on the device without reading back
 * diagnostics between steps (the benchmark loop). The loads pending at the
 * call apply to every one of the steps (as a caller setting the same loads
 * before each step) and are consumed at the end. Same transactional rule
 * per step: a failure in step k restores the state at the start of step k,
 * the k steps before it stay accepted (step count, time and last applied
 * load advance by k; with k > 0 the pending loads are consumed) and the
 * failure is reported. */
sw_status sw_structure_step_many(sw_structure* s, const sw_step_settings* settings, int count);

/* nodeStates() (structure.cpp:715-724): 3 doubles per node per array; any
 * pointer may be NULL. */
sw_status sw_structure_node_states(const sw_structure* s, double* position, double* previous, double* velocity);

/* Payload state (structure.h:195-206): com[3], orientation[4] (w,x,y,z),
 * linear velocity[3], angular velocity[3] = 13 doubles. */
sw_status sw_structure_payload_state(const sw_structure* s, double* payload13);

/* Diagnostic-frame fields of the committed state: the per-vertex and per-line
 * data buildStructureFrame reads (structure_frame.cpp:491-640), computed on
 * the device. Frame vertices are the nodes, then the pilot harness points in
 * definition order (positions and velocities of the rigid payload points,
 * suspension.cpp:1018-1032; pending force 0); lines are the constraints, then
 * the suspension segments, in definition order: current length
 * |second - first| and violation (|L - rest| for distance, tie and stitch,
 * max(0, L - rest) for cables and segments, structure_frame.cpp:60-73,
 * 577-598). Arrays: 3 (N + H), 3 (N + H), 3 (N + H), C + S, C + S doubles;
 * any may be NULL. The global fields come from sw_structure_diagnostics and
 * sw_structure_progress. */
sw_status sw_structure_frame_fields(const sw_structure* s, double* vertex_position, double* vertex_velocity,
                                    double* vertex_pending_force, double* line_length, double* line_violation);

/* The same fields as sw_structure_frame_fields, from an entry point that
 * allocates nothing: no hipMalloc, hipFree or other allocating routine on
 * any call. No counterpart in the reference. It is for callers that cannot
 * allocate device memory: a frame publisher inside a tight stepping loop,
 * code under HIP graph capture, hosts that forbid allocation after set-up.
 * The fields are staged in a static device array of 2,097,152 doubles
 * (16 MiB per GPU, held until the process ends); a structure needs
 * 6 (N + H) + 2 (C + S) doubles of it, and a larger one is rejected with
 * SW_INVALID_ARGUMENT. The array's address on a device is looked up on the
 * first call on that device. All calls on one GPU share the array: calls
 * that overlap on one GPU race and must be serialised by the caller. Every
 * staged value is written before it is read, so nothing carries over from
 * an earlier call. Results equal sw_structure_frame_fields bit for bit. */
sw_status sw_structure_frame_fields_noalloc(const sw_structure* s, double* vertex_position,
                                            double* vertex_velocity, double* vertex_pending_force,
                                            double* line_length, double* line_violation);

/* Stable identity of a frame: StructureFrameMappingDefinition
 * (structure_frame.h:35-39). vertex_count = N + H, triangle_count = T,
 * line_count = C + S. Ids nonzero and unique per kind, side regions
 * nonzero; a segment line's id is the segment's stable id. The line
 * endpoints are not passed: constraints use their nodes, segments the
 * endpoint rule of makeStructureFrameMapping (structure_frame.cpp:427-474;
 * harness point k is vertex N + k). */
typedef struct sw_frame_mapping {
    size_t vertex_count;
    const uint64_t* vertex_ids;
    size_t triangle_count;
    const uint64_t* triangle_ids;
    const uint64_t* triangle_negative_region;
    const uint64_t* triangle_positive_region;
    size_t line_count;
    const uint64_t* line_ids;
    const uint32_t* line_roles;
} sw_frame_mapping;

/* StructureFrameContext (structure_frame.h:86-93) with CouplingResiduals and
 * ConservationValues (viewer_protocol.h:127-145); strings NUL-terminated,
 * nonempty, at most 1024 bytes. */
typedef struct sw_frame_context {
    const char* scene_checksum;
    const char* solver_commit;
    double time_step;
    uint32_t coupling_iteration;
    double coupling_residuals[5]; /* displacement, traction, fluid, structure, interface power */
    double fluid_mass;
    double total_momentum[3];
    double total_energy;
    double interface_force_residual[3];
    double interface_moment_residual[3];
    double interface_power_residual;
} sw_frame_context;

/* buildStructureFrame (structure_frame.cpp:480-666) followed by
 * serializeFrame (viewer_protocol.cpp:838-882): the frame of the committed
 * state as SWVF bytes. Field arrays come from the device
 * (sw_structure_frame_fields), the global fields from
 * sw_structure_diagnostics and the step and time from
 * sw_structure_progress. bytes NULL: only *size is set. A capacity below
 * *size fails with SW_INVALID_ARGUMENT and *size set. Rejections follow
 * the reference (mapping checks, validateFrame). */
sw_status sw_structure_frame(sw_structure* s, const sw_frame_mapping* mapping, const sw_frame_context* context,
                             uint8_t* bytes, size_t capacity, size_t* size);

/* The same encoding from caller-held fields (host only; no device is used):
 * vertex arrays 3 (N + H) doubles, line arrays C + S doubles, the global
 * fields from *diagnostics. */
sw_status sw_frame_encode(const sw_definition* definition, const sw_frame_mapping* mapping,
                          const sw_frame_context* context, uint64_t step, double simulation_time,
                          const double* vertex_position, const double* vertex_velocity,
                          const double* vertex_pending_force, const double* line_length,
                          const double* line_violation, const sw_diagnostics* diagnostics, uint8_t* bytes,
                          size_t capacity, size_t* size);

/* Solver multipliers and residuals, in definition order (constraints,
 * membranes, dihedrals) and runtime segment order (string-id order of
 * "segment:<id>", suspension.cpp:996-1005). Any pointer may be NULL.
 * membrane_lambda, membrane_resultant: 3 per membrane. */
sw_status sw_structure_solver_state(const sw_structure* s, double* constraint_lambda,
                                    double* membrane_lambda, double* membrane_resultant,
                                    double* membrane_residual, double* dihedral_lambda,
                                    double* segment_lambda);

/* checkpoint() / restore() (structure.cpp:961-1025). A checkpoint is an
 * immutable value owned by the caller until released; it stays valid after
 * its structure is destroyed. */
sw_status sw_structure_checkpoint(const sw_structure* s, sw_checkpoint** out);
sw_status sw_structure_restore(sw_structure* s, const sw_checkpoint* checkpoint);
void sw_checkpoint_release(sw_checkpoint* checkpoint);

/* serializeStructureCheckpoint (structure_checkpoint_persistence.cpp:229-333)
 * of a checkpoint of `s`: the SWST bytes the reference writes for the same
 * state, with the SoftBody section of serializeSoftBodyCheckpoint
 * (soft_body_checkpoint_persistence.cpp:858-919) and, with a pilot, the
 * suspension section of serializeSuspensionCheckpoint
 * (suspension_checkpoint_persistence.cpp:752-800), whose diagnostics are
 * those SuspensionSystem::finishSubstep leaves after the last substep
 * (suspension.cpp:2481-2610), restated from the committed state (or, for a
 * checkpoint read from bytes and not stepped since, the bytes read). *size
 * receives the encoded length; with out == NULL nothing else is written.
 * Rejected with SW_INVALID_ARGUMENT: structures with contact spheres (no
 * reference counterpart) and, after a step, damped suspension lines (their
 * damping impulses are diagnostics the device does not record). */
sw_status sw_checkpoint_serialize(const sw_structure* s, const sw_checkpoint* checkpoint, uint8_t* out,
                                  size_t capacity, size_t* size);
/* deserializeStructureCheckpoint (structure_checkpoint_persistence.cpp:
 * 335-520): reads SWST bytes written for `s`'s definition into a new
 * checkpoint for sw_structure_restore; until the restored structure steps,
 * its suspensionState() and checkpoints report the diagnostics the bytes
 * hold. Malformed, truncated, mismatched (fingerprints, counts, checksums,
 * the suspension's state fingerprint) or inconsistent input, and suspension
 * sections commanding controls or ground contact, are rejected with
 * SW_INVALID_ARGUMENT and the reference's reason. */
sw_status sw_checkpoint_deserialize(sw_structure* s, const uint8_t* bytes, size_t size, sw_checkpoint** out);

/* StructureCheckpoint's public fields (structure.h:254-271): version
 * (structureCheckpointVersion = 2), definitionFingerprint,
 * acceptedStepCount, simulationTimeSeconds and the node count; then the
 * nodes (position, previousPosition, velocity: 3 doubles per node each),
 * pendingExternalForcesNewtons (3 per node) and
 * lastAppliedExternalForceNewtons (3). Any pointer may be NULL. */
sw_status sw_checkpoint_info(const sw_checkpoint* checkpoint, uint32_t* version, uint64_t* fingerprint,
                             uint64_t* accepted_steps, double* simulation_time, size_t* node_count);
sw_status sw_checkpoint_nodes(const sw_checkpoint* checkpoint, double* position, double* previous,
                              double* velocity, double* pending_forces, double* last_applied_force);

/* suspensionState() (structure.cpp:718-776): *present is 0 for a structure
 * without a suspension (std::nullopt) and nothing else is written. Else:
 * payload13 = centre of mass, body-to-world quaternion (w, x, y, z), linear
 * and angular velocity; harness positions and velocities (3 per harness
 * point, definition order); per segment in definition order its stable id,
 * current length, commanded rest length and tension (multiplier / dt^2 of
 * the last substep, suspension.cpp:2513-2517). Before the first accepted
 * substep the segments report their rest length as the current length and
 * zero tension, as the reference does. Any array pointer may be NULL. */
sw_status sw_structure_suspension_state(const sw_structure* s, int* present, double* payload13,
                                        double* harness_position, double* harness_velocity, uint64_t* segment_ids,
                                        double* segment_length, double* segment_commanded_rest,
                                        double* segment_tension);

/* Colour schedule of a family, colour-major: family 0 = distance
 * constraints (= SoftBody::constraintColouringView, soft_body.h:462-474),
 * 1 = membranes, 2 = dihedral hinges. Pass NULL arrays to query sizes. */
sw_status sw_structure_colouring(const sw_structure* s, int family, uint64_t* order,
                                 size_t* order_count, uint64_t* offsets, size_t* offset_count);

/* The same colour schedules computed on the host from a definition, with
 * no device involved (family 0, 1, 2 as above; family 3 = suspension
 * segments in runtime string-id order, reported as definition indices with
 * one "colour" per dependency level). */
sw_status sw_definition_colouring(const sw_definition* definition, int family, uint64_t* order,
                                  size_t* order_count, uint64_t* offsets, size_t* offset_count);

/* Device math conformance probe: evaluates `function` (0 sqrt, 1 divide
 * a/b, 2 atan2(a, b), 3 remainder(a, 2*pi), 4 sin, 5 cos, 6 acos) on the
 * device for `count` inputs (a[i], b[i]) and writes the results to out. */
sw_status sw_device_math(int device, int function, const double* a, const double* b, double* out, size_t count);

/* Device launch shape (workgroups used by one step) and the autotuner. */
sw_status sw_structure_set_launch(sw_structure* s, int workgroups, int threads);
sw_status sw_structure_get_launch(const sw_structure* s, int* workgroups, int* threads);

/* Multi-workgroup launches use hipLaunchCooperativeKernel (default, 1) or a
 * plain launch (0; resident in practice on an idle device, for tools that
 * do not follow cooperative launches). Every in-kernel wait is bounded and
 * reports SW_DEVICE_ERROR instead of hanging. No counterpart in the
 * reference (launch configuration). */
sw_status sw_structure_set_cooperative(sw_structure* s, int cooperative);

/* One-XCD synchronisation (no counterpart in the reference; launch
 * configuration). When enabled, a multi-workgroup substep launch carries 8
 * workgroups per participant and only the participants that land on one XCD
 * (chosen at run time from the hardware XCD id) run; they share that XCD's
 * L2, so a phase publishes its writes by draining them to the L2 and a
 * waiter invalidates only its own L1, instead of an agent-scope L2
 * write-back and invalidation per phase. Results are identical. *active is
 * 0 when the current launch shape does not allow it (single workgroup, or
 * 8 x the workgroups not resident at once, or more workgroups than one
 * XCD's compute units hold); the request is kept and re-checked by
 * sw_structure_set_launch. Enabled by default. */
sw_status sw_structure_set_xcd_local(sw_structure* s, int enabled, int* active);

/* Substep-kernel timing (no counterpart in the reference; measurement only).
 * While enabled, each substep kernel launch is bracketed by HIP events on the
 * structure's own stream. sw_structure_kernel_time waits for the stream and
 * returns the summed kernel time in milliseconds and the launch count since
 * timing was last enabled. */
sw_status sw_structure_set_kernel_timing(sw_structure* s, int enabled);
sw_status sw_structure_kernel_time(sw_structure* s, double* milliseconds, long long* launches);

/* Step profile: the soft-body fields of softwing::StepPerformanceProfile
 * (soft_body.h:136-170; filled by SoftBody::stepCoupled when
 * StepSettings::performanceProfile is set, soft_body.cpp:1075-1444),
 * accumulated over the steps taken while profiling is on. The visit counts
 * follow the reference's rules exactly (soft_body.cpp:1222-1245, 1389-1404:
 * per substep, iterations x constraints / membranes / dihedrals / segments,
 * 2 x cable pairs x load-path constraints, one more visit per seam stitch in
 * the closing pass, iterations once for a wing with contact pairs). The times
 * are the device's: soft_body_total_ns is the host's wall clock around each
 * step call (launch to completion); each part's time is the mean workgroup's
 * wall-clock time in that part of the substep kernel (waits included),
 * measured with the device's 100 MHz clock, so the parts add up to about the
 * kernel time; transaction_snapshot_ns is the remainder of the total. */
typedef struct sw_step_profile {
    uint64_t soft_body_total_ns;
    uint64_t transaction_snapshot_ns;
    uint64_t prediction_ns;
    uint64_t distance_constraint_ns;
    uint64_t cable_constraint_ns;
    uint64_t membrane_constraint_ns;
    uint64_t bending_constraint_ns;
    uint64_t suspension_constraint_ns;
    uint64_t contact_constraint_ns;
    uint64_t membrane_diagnostics_ns;
    uint64_t finalization_ns;
    uint64_t structural_substeps;
    uint64_t constraint_iterations;
    uint64_t distance_constraint_visits;
    uint64_t cable_constraint_visits;
    uint64_t membrane_constraint_visits;
    uint64_t bending_constraint_visits;
    uint64_t suspension_constraint_visits;
    uint64_t contact_constraint_visits;
} sw_step_profile;
/* Turns profiling on (the totals start from zero) or off (they are kept). */
sw_status sw_structure_set_profiling(sw_structure* s, int enabled);
sw_status sw_structure_profile(sw_structure* s, sw_step_profile* out);
sw_status sw_structure_reset_profile(sw_structure* s);

/* ---- Scene input and assembly (scene_structure.h:162-165) ------------- */

typedef struct sw_scene sw_scene;
typedef struct sw_assembly sw_assembly;

/* Seeded synthetic wing by preset name ("tiny", "small", "export",
 * "variant10k"), or a batch member (preset "batch:<seed>:<index>:<count>"). */
sw_status sw_scene_generate(const char* preset, sw_scene** out);
void sw_scene_destroy(sw_scene* scene);
/* Per-triangle pressure generated with the scene (ascending triangle id). */
sw_status sw_scene_face_pressure(const sw_scene* scene, const double** pressure, size_t* count);

/* readScene (scene.cpp:1225-1418): a scene from its compact binary stream
 * (binary versions 2 and 3, little-endian). Rejections carry readScene's
 * messages (magic header, binary version, count and string limits,
 * truncation, trailing data, junction fixed flag) as SW_INVALID_ARGUMENT;
 * the decoded scene is then checked with the structural subset of
 * validateScene that sw_scene_assemble applies. A read scene has zero face
 * pressure. */
sw_status sw_scene_read(const uint8_t* bytes, size_t size, sw_scene** out);

/* writeScene (scene.cpp:1051-1223): the scene's binary stream (version 3,
 * every collection in stable-ID order). bytes NULL: only *size is set; a
 * capacity below *size fails with SW_INVALID_ARGUMENT and *size set. */
sw_status sw_scene_write(const sw_scene* scene, uint8_t* bytes, size_t capacity, size_t* size);

/* SceneStructureSettings (scene_structure.h:27-39) subset. */
typedef struct sw_assembly_settings {
    double seam_coincidence_tolerance;
    int suspension_iterations;
    double suspension_attachment_tolerance;
    double suspension_minimum_line_length;
    double suspension_maximum_line_residual;
    double suspension_maximum_control_work;
    double contact_sphere_compliance;
} sw_assembly_settings;

/* assembleSceneStructure (scene_structure.cpp:894-1282): integer-exact node,
 * triangle, constraint, membrane, dihedral and suspension tables. */
sw_status sw_scene_assemble(const sw_scene* scene, const sw_assembly_settings* settings, sw_assembly** out);
void sw_assembly_destroy(sw_assembly* assembly);
/* Borrowed view of the assembled definition (valid while `assembly` lives). */
sw_status sw_assembly_definition(const sw_assembly* assembly, sw_definition* out);
/* Mappings: stable ids of vertex nodes and junction nodes, fabric and seam
 * masses (scene_structure.h:83-140). */
sw_status sw_assembly_info(const sw_assembly* assembly, size_t* vertex_nodes, size_t* junction_nodes,
                           double* fabric_mass, double* seam_mass);

/* Steady flight loads of the benchmark workload: the system weight carried
 * by the vertex nodes, mass-weighted, radially away from the pilot. */
sw_status sw_assembly_flight_loads(const sw_assembly* assembly, double* forces);

/* ---- Batched wings ----------------------------------------------------- */

typedef struct sw_batch sw_batch;

/* Many wings of uneven size stepped together by one persistent launch (no
 * counterpart in the reference, which steps one Structure at a time; each
 * wing behaves exactly as a Structure built from the same definition and
 * returns the same bits). Small wings run one LDS-resident workgroup each;
 * large wings share the remaining compute units in proportion to their
 * items. A wing whose step fails is restored to its pre-step state and the
 * call reports the first failing wing's error; the other wings keep their
 * accepted steps. */
sw_status sw_batch_create(const sw_definition* const* definitions, size_t count, int device, sw_batch** out);
void sw_batch_destroy(sw_batch* batch);
sw_status sw_batch_set_face_pressure(sw_batch* b, size_t wing, const double* pressure, size_t count);
sw_status sw_batch_set_external_forces(sw_batch* b, size_t wing, const double* forces, size_t node_count);
/* Steps every wing `count` times with the loads set before the call (as
 * Structure::step consumes them after an accepted step). */
sw_status sw_batch_step(sw_batch* b, const sw_step_settings* settings, int count);
sw_status sw_batch_node_states(const sw_batch* b, size_t wing, double* position, double* velocity);
/* Re-plans the launch: at most `workgroups_per_wing_cap` workgroups per
 * wing, `threads` lanes per workgroup; reports the chosen workgroups. */
sw_status sw_batch_set_launch(sw_batch* b, int workgroups_per_wing_cap, int threads);
sw_status sw_batch_workgroups(const sw_batch* b, int* per_wing, size_t count);
/* Cooperative (default) or plain launches, as sw_structure_set_cooperative. */
sw_status sw_batch_set_cooperative(sw_batch* b, int cooperative);
/* The wing's own structure handle, owned by the batch (valid until
 * sw_batch_destroy): every sw_structure_* call applies to it, including
 * stepping it alone with its own launch shape; sw_batch_set_launch restores
 * the batched plan afterwards. */
sw_status sw_batch_structure(sw_batch* b, size_t wing, sw_structure** out);

/* ---- Adjoint ------------------------------------------------------------ */

typedef struct sw_adjoint sw_adjoint;

/* Reverse-mode gradients of the shape loss
 *     L = sum over nodes of |x(end) - x_target|^2
 * after `steps` steps of `settings` from the structure's committed state at
 * the start of each run (face pressure included), with the loads that were
 * pending at configure applied at every step (no counterpart in the
 * reference, which has no differentiation). Parameters:
 *   - log-compliance per membrane element (definition order): the element's
 *     compliance matrix becomes C e^theta (its stiffness is scaled by
 *     e^-theta before the cofactor inverse of membrane.cpp:268-271);
 *   - rest length per suspension segment (definition order; a line's trim is
 *     the sum over its segments);
 *   - pressure per triangle.
 * The forward pass keeps a state checkpoint every `checkpoint_interval`
 * steps; the reverse pass replays each interval from its checkpoint with
 * the trajectory recorded and runs the reverse kernels substep by substep
 * in FP64, carrying the Lagrange-multiplier adjoints through every colour
 * sweep and iteration. The structure must outlive the adjoint; its state
 * (face pressure included), step count and parameters are what they were
 * before each run once the run returns, whether it succeeds or fails. Membrane damping is
 * rejected with SW_INVALID_ARGUMENT. */
sw_status sw_adjoint_configure(sw_structure* structure, const sw_step_settings* settings, uint64_t steps,
                               uint32_t checkpoint_interval, sw_adjoint** out);
void sw_adjoint_destroy(sw_adjoint* adjoint);
/* One forward + reverse run at the given parameters (null: the
 * definition's values; arrays of membrane_count, segment_count and
 * triangle_count). `target` holds 3 per node. `final_position` (3 per node)
 * and every gradient output are optional; with no gradient output only the
 * forward pass and the loss run. A failing step restores the structure and
 * reports the step's error. */
sw_status sw_adjoint_run(sw_adjoint* adjoint, const double* log_compliance, const double* rest_length,
                         const double* pressure, const double* target, double* loss, double* final_position,
                         double* grad_log_compliance, double* grad_rest_length, double* grad_pressure);
/* Workgroups of the reverse kernel (default: enough for the widest colour). */
sw_status sw_adjoint_set_workgroups(sw_adjoint* adjoint, int workgroups);
/* Device milliseconds of the last run: checkpointed forward pass, and
 * recorded replay + reverse kernels; device bytes the adjoint holds. */
sw_status sw_adjoint_timing(const sw_adjoint* adjoint, double* forward_ms, double* reverse_ms, size_t* device_bytes);
/* The last run's reverse pass split: recorded replay, Jacobian passes (every
 * item's columns, one parallel launch per family and substep) and the
 * cooperative sweeps that contract them. */
sw_status sw_adjoint_timing_detail(const sw_adjoint* adjoint, double* replay_ms, double* jacobian_ms,
                                   double* sweep_ms);

/* ---- Batched adjoint ---------------------------------------------------- */

typedef struct sw_batch_adjoint sw_batch_adjoint;

/* The adjoint of a batch (no counterpart in the reference): every wing's
 * shape-loss gradient with respect to its line rest lengths (definition
 * order), from the wings' current states over `steps` steps with the loads
 * pending now. The forward pass and the checkpoint replays are the batch's
 * own persistent launches; the reverse sweeps of all wings run in one
 * persistent launch per substep and Jacobian chunk (`jacobian_chunk`
 * iterations of Jacobian columns held at once, 0 = all: memory for launches).
 * Each wing's gradient equals the one its own sw_adjoint computes. The batch
 * must outlive the adjoint. */
sw_status sw_batch_adjoint_configure(sw_batch* b, const sw_step_settings* settings, uint64_t steps,
                                     uint32_t checkpoint_interval, uint32_t jacobian_chunk, sw_batch_adjoint** out);
void sw_batch_adjoint_destroy(sw_batch_adjoint* adjoint);
/* Per wing: rest lengths (rest_lengths NULL or an entry NULL: the
 * definition's), target positions; outputs per wing (NULL arrays or entries
 * skipped): loss, rest-length gradient, final positions. A failing step
 * restores every wing and reports the first failing wing's error. */
sw_status sw_batch_adjoint_run(sw_batch_adjoint* adjoint, const double* const* rest_lengths,
                               const double* const* targets, double* losses, double* const* grad_rest_lengths,
                               double* const* final_positions);
sw_status sw_batch_adjoint_timing(const sw_batch_adjoint* adjoint, double* forward_ms, double* reverse_ms,
                                  size_t* device_bytes);
/* The last run's reverse pass split, as sw_adjoint_timing_detail. */
sw_status sw_batch_adjoint_timing_detail(const sw_batch_adjoint* adjoint, double* replay_ms, double* jacobian_ms,
                                         double* sweep_ms);
/* Workgroups of each wing in the batched reverse sweep. */
sw_status sw_batch_adjoint_workgroups(const sw_batch_adjoint* adjoint, int* per_wing, size_t count);

#ifdef __cplusplus
}
#endif

#endif /* SIMWING_XPBD_H */
