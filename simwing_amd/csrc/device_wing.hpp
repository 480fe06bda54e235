// Device-resident XPBD wing: SoA tables in HBM, the phase program of one
// substep, and the launch entry points. Host-only header (no device code).
#pragma once

#include "definition.hpp"
#include "schedule.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace swx {

// One entry of the per-iteration phase program. Items of a phase are
// independent; consecutive phases are separated by a barrier.
enum PhaseType : int {
    kPhaseDistance = 0,   // distance-constraint colour [begin, end)
    kPhaseMembrane = 1,   // membrane colour
    kPhaseHinge = 2,      // dihedral colour
    kPhaseSegments = 3,   // one suspension dependency level
    kPhaseSegmentRun = 4, // consecutive single-item levels run by one lane
    kPhaseSpheres = 5,    // one contact-sphere dependency level
    kPhaseStitches = 6,   // one level of the seam-stitch closing pass
    // Serial sweep (workerThreads = 0): one dependency level over the
    // definition order of the distance constraints or the membranes; the
    // phase's [begin, end) indexes the level-ordered list of device slots.
    kPhaseDistanceList = 7,
    kPhaseMembraneList = 8,
    // An interleaved load-path cable pair's level (flags: the pair's index
    // within the iteration; [begin, end) indexes the pair's level list).
    kPhaseCableList = 9,
};

struct Phase {
    int type;
    int begin;
    int end;
    int flags;      // unused (0)
    int slotBegin;  // kPhaseSegmentRun: the run's staged nodes are
    int slotEnd;    // runNodes[slotBegin, slotEnd)
};

// Capacity of one staged serial segment run (kPhaseSegmentRun): items and
// distinct nodes. Longer runs are split into consecutive phases.
constexpr int kChainItems = 128;
constexpr int kChainSlots = 256;

// Rigid payload state and definition on the device.
struct PayloadDev {
    double com[3];
    double q[4];  // w, x, y, z
    double v[3];
    double w[3];
    double prevCom[3];
    double prevQ[4];
    double mass;
    double comLocal[3];
    double inertia[3];
    double iinv[9];
    double maxResidual;
    double minLineLength;
    double prevV[3];  // velocities at the start of the substep
    double prevW[3];  // (previousPayloadState, suspension.cpp:2085)
};

// First failure of a step: code (SW_* status), suspension phase, the
// failing item's kind (0 membrane, 1 segment, 2 payload) and index.
struct DeviceError {
    int code;
    int phase;
    int kind;
    int index;
};

struct WingView;     // device-side view (defined in the .hip file)
struct AdjointView;  // device-side adjoint buffers (defined with the adjoint kernels)

struct StepArgs {
    double dt;
    double invDt;
    double invDt2;
    double damping;  // exp(-c * dt), evaluated on the host
    double gravity[3];
    double reference[3];
    double maxLineResidual;
    int iterations;
    int cableBase, cableRem;  // interleaved cable pairs: per iteration, and one more before iteration cableRem
};

class DeviceBatch;
class DeviceAdjoint;

class DeviceWing {
public:
    DeviceWing(const Definition& definition, int device);
    ~DeviceWing();
    DeviceWing(const DeviceWing&) = delete;
    DeviceWing& operator=(const DeviceWing&) = delete;

    // Upload pending loads and per-triangle pressure (host arrays, 3N / T).
    void setExternalForces(const std::vector<double>& forces);
    void setFacePressure(const std::vector<double>& pressure);

    // One substep on the wing's stream (asynchronous); with `trace`, every
    // projection also records its inputs there (the adjoint's replay).
    void launchSubstep(const StepArgs& args, double* trace = nullptr);
    // Snapshot of the mutable state (taken only while no error is pending),
    // and restore from it.
    void snapshot();
    void restoreSnapshot();
    // Clears the device error word (and the snapshot count).
    void clearError();
    // Snapshots taken since the last clearError (one per step that started
    // without a pending failure).
    int snapshotsTaken();
    DeviceError readError(double* maxResidual);
    void synchronize();

    // Host copies of the mutable state (definition order).
    struct State {
        std::vector<double> position, previous, velocity;
        std::vector<double> constraintLambda, membraneLambda, membraneResultant, membraneResidual;
        std::vector<double> dihedralLambda, segmentLambda, sphereLambda, pressure;
        std::vector<double> payload;  // com, q, v, w, prevCom, prevQ, prevV, prevW (26)
    };
    State download() const;
    void upload(const State& state);

    // Diagnostic-frame fields of the committed state: harness point
    // positions and velocities (3 per point, definition order), and per line
    // (constraints, then suspension segments, definition order) the current
    // length and its violation.
    struct Frame {
        std::vector<double> harnessPosition, harnessVelocity, lineLength, lineViolation;
    };
    Frame frame();

    // The frame fields of sw_structure_frame_fields_noalloc, computed into
    // `staging` (device memory of at least frameFieldsDoubles() doubles that
    // the caller provides) in frame order and copied from there into the
    // non-null host outputs: positions and velocities 3 (N + H), line
    // lengths and violations C + S. Allocates nothing.
    std::size_t frameFieldsDoubles() const;
    void frameFields(double* staging, double* position, double* velocity, double* length, double* violation);

    // Reductions for StructureDiagnostics (see capi.cpp for the fields).
    struct Reductions {
        double mass, weighted[3], momentum[3], kinetic;
        double maxDistanceError, maxCableExtension, maxStrain, maxResidual;
        int finite;
        int dynamicNodes;
    };
    Reductions reduce();

    const ColourOrder& distanceColours() const { return distance_; }
    const ColourOrder& membraneColours() const { return membrane_; }
    const ColourOrder& hingeColours() const { return hinge_; }
    const SuspensionLayout& suspension() const { return layout_; }
    const std::vector<Phase>& iterationPhases() const { return iterPhases_; }

    // Substep-kernel timing: while enabled, every substep launch is
    // bracketed by a pair of HIP events recorded on the wing's stream.
    // kernelTime() waits for the stream and returns the summed elapsed time
    // of the pairs recorded since timing was enabled, and their count.
    void setTiming(bool enabled);
    void kernelTime(double* milliseconds, long long* launches);

    // Step profile: while enabled, every substep launch adds, per part of
    // the substep (prediction, distance, cable, membrane, bending,
    // suspension, contact, membrane diagnostics, finalisation), the
    // wall-clock ticks its workgroups spent there to device totals.
    // profileTotals() waits for the stream and returns the totals (10 MHz
    // ticks converted to nanoseconds, each divided by the workgroups per
    // launch, so each is the mean workgroup's time) and the launches since
    // the totals were last cleared.
    void setProfiling(bool enabled);
    void profileTotals(double nanoseconds[9], long long* launches);
    void clearProfile();

    void setLaunch(int workgroups, int threads);
    // Incremented by every setLaunch (batches re-plan when a wing changed).
    unsigned long long launchGeneration() const { return launchGeneration_; }
    // Multi-workgroup launches go through hipLaunchCooperativeKernel (the
    // runtime checks that every workgroup is resident at once) unless this is
    // off; then a plain launch is used, which is resident in practice on an
    // otherwise idle device and lets tools that do not follow cooperative
    // launches observe the kernel. Every wait in the kernel is bounded.
    void setCooperative(bool cooperative) { cooperative_ = cooperative; }
    // One-XCD synchronisation for multi-workgroup launches: the launch
    // carries 8 workgroups per participant and only those on one XCD take
    // part, so the participants share that XCD's L2 and synchronise through
    // it (no agent-scope L2 write-back or invalidation per phase). Returns
    // false (and stays off) when the launch shape cannot hold 8 x the
    // workgroups on the device at once. On by default wherever the launch
    // shape allows it; the request is kept across setLaunch.
    bool setXcdLocal(bool enabled);
    // The reference's serial sweep (StepSettings::workerThreads = 0:
    // distance constraints and membranes in definition order,
    // soft_body.cpp:1316-1345) or its coloured sweep (workerThreads >= 1).
    // The serial sweep runs as dependency levels over the definition order
    // (items of one level touch disjoint nodes, so the level's parallel
    // projection equals the serial one); hinges, segments and spheres are
    // the same in both.
    void setSerialSweep(bool serial);
    // Load-path cable sweep pairs (StepSettings::cableConstraintSweepPairs)
    // and the iteration count they are split over; rebuilds the programs
    // when either changes. interleavedCablePairs() is the part run inside
    // the iterations.
    void setCablePairs(int pairs, int iterations);
    int interleavedCablePairs() const { return cableInterleaved_; }
    bool serialSweep() const { return serialSweep_; }
    // A free membrane (momentum restored after each substep), damped
    // suspension lines and membranes with a compressionStiffnessRatio other
    // than 1: stepped, not differentiated.
    bool preservesMomentum() const { return preserveMomentum_; }
    std::size_t dampedSegments() const { return dampedSegments_; }
    bool compressedMembranes() const { return compressedMembranes_; }
    // Whether a substep needs the kernel variant compiled with the list
    // phases (serial sweep, interleaved cable pairs), compression softening,
    // momentum restoration and line damping.
    bool needsGeneralKernel() const {
        return serialSweep_ || cableInterleaved_ > 0 || preserveMomentum_ || dampedSegments_ > 0 ||
               compressedMembranes_;
    }
    bool xcdLocal() const { return xcdLocal_; }
    int workgroups() const { return workgroups_; }
    int threads() const { return threads_; }
    bool sharedNodes() const { return sharedNodes_; }
    hipStream_t stream() const { return stream_; }

    // The wing's device view with its current launch modes (for batched
    // launches), the dynamic LDS its substep needs, and its item count.
    WingView launchView() const;
    std::size_t dynamicLdsBytes() const;
    std::size_t itemCount() const { return membranes_ + hinges_ + constraints_ + segments_ + pairs_; }
    std::size_t nodeCount() const { return nodes_; }
    bool fitsInLds() const;

private:
    friend class DeviceBatch;
    friend class DeviceAdjoint;
    friend class DeviceBatchAdjoint;
    void build(const Definition& d);
    int device_;
    hipStream_t stream_ = nullptr;
    std::vector<void*> allocations_;
    std::unique_ptr<WingView> view_;
    ColourOrder distance_, membrane_, hinge_;
    SuspensionLayout layout_;
    std::vector<Phase> iterPhases_, closePhases_;
    std::vector<Phase> colouredPhases_, serialPhases_;  // the two iteration programs
    Phase* colouredPhasesDev_ = nullptr;
    Phase* serialPhasesDev_ = nullptr;
    std::vector<int> distList_, membList_;  // serial-sweep level order (device slots)
    bool serialSweep_ = false;
    bool preserveMomentum_ = false;
    std::size_t dampedSegments_ = 0;
    bool compressedMembranes_ = false;
    std::vector<Phase> colouredBase_, serialBase_;   // the programs without cable phases
    std::vector<int> stitchList_;                    // seam-stitch closing levels (device slots)
    std::vector<Phase> stitchLevels_;
    std::vector<int> cablePair_;                     // one cable pair's levels (device slots)
    std::vector<std::pair<int, int>> cablePairLevels_;
    int cablePairs_ = 0, cableIterations_ = 1, cableInterleaved_ = 0;
    std::vector<void*> programAlloc_;                // device copies of the current programs
    void applyCablePrograms(int pairs, int iterations);
    // Nodes of each item in device slot order (the payload is node nodes_),
    // kept on the host for the synchronisation analysis.
    struct ItemNodes {
        std::vector<int> dist, memb, hinge;  // 2, 3 and 4 per slot
        std::vector<int> segFrom, segTo;     // < 0: harness point on the payload
        std::vector<int> sphereNode, sphereCentre;
    } items_;
    void itemAccess(int type, int k, std::vector<int>& reads, std::vector<int>& writes) const;
    // Point-to-point dependency lists of the iteration program for the
    // current launch shape.
    void buildSync();
    std::size_t sharedNodeBudget() const;
    std::size_t stageBytes_ = 0;  // dynamic LDS of the staged point-to-point program
    bool hasRun_ = false, neighbourSync_ = false, dedicatedChain_ = false, cooperative_ = true;
    bool xcdLocal_ = false;      // one-XCD synchronisation active for the current shape
    bool xcdRequested_ = true;   // ... and wanted (the default: on where the shape allows it)
    unsigned epoch_ = 0;
    int participants_ = 0, maxParticipants_ = 0;
    unsigned long long launchGeneration_ = 0;
    void* depAlloc_[3] = {nullptr, nullptr, nullptr};
    bool timing_ = false;
    unsigned long long* profile_ = nullptr;  // device totals (10 slots), allocated on first use
    long long profiledLaunches_ = 0;
    std::vector<hipEvent_t> events_;  // start/stop pairs, grown on demand
    std::size_t eventsUsed_ = 0;
    double timedMs_ = 0.0;
    long long timedLaunches_ = 0;
    std::vector<std::uint32_t> segmentDevicePosition_;  // runtime order -> device slot
    std::vector<std::uint32_t> sphereDevicePosition_;   // definition pair -> device slot
    std::size_t nodes_ = 0, triangles_ = 0, constraints_ = 0, membranes_ = 0, hinges_ = 0, segments_ = 0,
                pairs_ = 0;
    void* partials_ = nullptr;  // diagnostics partial records
    int* frameLineSlot_ = nullptr;  // definition line -> device item (frameFieldsKernel)
    int* frameHangSlot_ = nullptr;  // definition harness point -> device hang index
    int workgroups_ = 1;
    int threads_ = 512;
    bool sharedNodes_ = false;
    std::size_t snapshotBytes_ = 0;
};

// Workgroups launched per participant in a one-XCD launch (the XCD count of
// an MI355X).
constexpr int kXcdLaunchFactor = 8;

// Capacity of the static device staging array of
// sw_structure_frame_fields_noalloc, in doubles (16 MiB per GPU).
constexpr std::size_t kFrameScratchDoubles = std::size_t{1} << 21;

// Device address of that array on `device` (hipGetSymbolAddress, looked up
// once per device with the device current, then cached).
double* frameScratchPointer(int device);

// A failed step inside an adjoint run (the device error word), thrown after
// the wing's state has been restored.
struct DeviceStepFailure : std::runtime_error {
    DeviceError error;
    explicit DeviceStepFailure(const DeviceError& e) : std::runtime_error("device step failed"), error(e) {}
};

// Reverse-mode gradient of L = sum_i |x_i(end) - target_i|^2 over `steps`
// steps (of `substeps` substeps each, the pending loads applied at every
// step) from the wing's state at construction, with respect to
//   * the log-compliance theta_m of every membrane (compliance C e^theta;
//     theta is applied by scaling the element's stiffness by e^-theta
//     before the cofactor inverse),
//   * the rest length of every suspension segment,
//   * the pressure of every triangle.
// The forward pass keeps a state checkpoint every `interval` steps; the
// reverse pass replays each interval from its checkpoint with the substep
// trajectories recorded, then runs the reverse kernels substep by substep,
// last first. Every run starts from the construction state and leaves the
// wing's state and parameters as they were.
class DeviceAdjoint {
public:
    // jacobianChunk: iterations whose Jacobian columns are held at once
    // (0: all of a substep; fewer trades launches for memory).
    DeviceAdjoint(DeviceWing& wing, const Definition& definition, const StepArgs& args, int substeps, long steps,
                  int interval, int jacobianChunk = 0);
    ~DeviceAdjoint();
    DeviceAdjoint(const DeviceAdjoint&) = delete;
    DeviceAdjoint& operator=(const DeviceAdjoint&) = delete;

    // Parameters in definition order (membranes, segments, triangles); a
    // null pointer keeps the definition's values. Gradient outputs may be
    // null (all null: forward and loss only). Returns the loss.
    double run(const double* logCompliance, const double* restLength, const double* pressure, const double* target,
               double* finalPosition, double* gradLogCompliance, double* gradRestLength, double* gradPressure);
    // Workgroups of the reverse kernel (cooperative launch).
    void setWorkgroups(int workgroups);
    int workgroups() const { return workgroups_; }
    // Device time of the last run: the checkpointed forward pass, and the
    // recorded replay plus the reverse kernels.
    double forwardMs() const { return forwardMs_; }
    double reverseMs() const { return reverseMs_; }
    // The reverse pass split: recorded replay, Jacobian passes, sweeps.
    double replayMs() const { return replayMs_; }
    double jacobianMs() const { return jacobianMs_; }
    double sweepMs() const { return sweepMs_; }
    std::size_t deviceBytes() const { return bytes_; }

private:
    friend class DeviceBatchAdjoint;
    AdjointView adjointView(const double* trace) const;
    WingView sweepView() const;
    void launchJacobians(const WingView& v, const AdjointView& A);
    void launchPressure(const WingView& v, const AdjointView& A);
    int jacobianChunk_ = 1;
    hipStream_t stream_ = nullptr;  // stream of the reverse launches (the wing's, or a batch's)
    void applyParameters(const double* logCompliance, const double* restLength, const double* pressure);
    void restoreBase();
    void launchRecording(double* trace);
    void launchReverse(const double* trace);
    void fail(const DeviceError& e);
    DeviceWing& wing_;
    Definition definition_;
    StepArgs args_;
    int substeps_;
    long steps_;
    int interval_;
    int workgroups_ = 1;
    long traceDoubles_ = 0;
    std::size_t bytes_ = 0;
    std::vector<void*> owned_;
    double *initial_ = nullptr, *checkpoints_ = nullptr, *trace_ = nullptr;
    double *xbar_ = nullptr, *vbar_ = nullptr, *qbar_ = nullptr, *fbar_ = nullptr, *lbar_ = nullptr, *pbar_ = nullptr;
    double *gTheta_ = nullptr, *gRest_ = nullptr, *gPressure_ = nullptr;
    double *jDist_ = nullptr, *jMemb_ = nullptr, *jHinge_ = nullptr, *jSeg_ = nullptr, *jSphere_ = nullptr,
           *jClose_ = nullptr;
    int firstPayloadSlot_ = -1;
    // Point-to-point schedule of the sweep (see buildReverseSync).
    void buildReverseSync();
    bool neighbourSync_ = false;
    int workers_ = 1, runOwner_ = 0, inertiaOwner_ = 0, maxParts_ = 1;
    void* depAlloc_[3] = {nullptr, nullptr, nullptr};
    unsigned* flags_ = nullptr;
    int lastPayloadPhase_ = -1;
    std::vector<double> baseCompliance_;  // 6 per membrane slot: xx yy zz xy xz yz
    std::vector<double> baseRest_;        // per segment slot
    double forwardMs_ = 0.0, reverseMs_ = 0.0, replayMs_ = 0.0, jacobianMs_ = 0.0, sweepMs_ = 0.0;
    hipEvent_t events_[3] = {nullptr, nullptr, nullptr};
    // Events marking the reverse pass's pieces (grown on demand), and which
    // pairs belong to which piece.
    std::vector<hipEvent_t> marks_;
    std::size_t marksUsed_ = 0;
    std::vector<std::pair<std::size_t, int>> spans_;  // first mark of a pair, piece (0 replay, 1 jacobian, 2 sweep)
    std::size_t mark();
    std::size_t mark(hipStream_t stream);
    // Pipelined reverse pass of a solo run: the replay of the next interval
    // on the wing's stream, beside the Jacobian passes and sweeps of the
    // current one on `reverseStream_`, over two trace buffers.
    bool pipelineFits() const;
    hipStream_t reverseStream_ = nullptr;
    double* traceB_ = nullptr;
    unsigned* barrier_ = nullptr;  // the sweep's own grid barrier words
    hipEvent_t replayDone_[2] = {nullptr, nullptr}, sweepDone_[2] = {nullptr, nullptr};
};

// The adjoint of many wings stepped together: the forward pass and the
// recorded replays are the batch's own launches, the reverse sweeps of all
// wings run in one persistent launch per substep (and Jacobian chunk); each
// wing's gradient equals that of its own DeviceAdjoint. Parameters: the
// wings' segment rest lengths (gradients of the shape loss per wing).
class DeviceBatchAdjoint {
public:
    DeviceBatchAdjoint(DeviceBatch& batch, const std::vector<const Definition*>& definitions,
                       const std::vector<StepArgs>& args, int substeps, long steps, int interval, int jacobianChunk);
    ~DeviceBatchAdjoint();
    DeviceBatchAdjoint(const DeviceBatchAdjoint&) = delete;
    DeviceBatchAdjoint& operator=(const DeviceBatchAdjoint&) = delete;

    // Per wing: rest lengths (definition order; empty vector or null entry:
    // the definition's), target positions (3 per node); outputs per wing
    // (null entries skipped): loss, rest-length gradient, final positions.
    void run(const std::vector<const double*>& rest, const std::vector<const double*>& targets, double* losses,
             const std::vector<double*>& gradRest, const std::vector<double*>& finalPositions);
    const std::vector<int>& sweepWorkgroups() const { return blocks_; }
    int failedWing() const { return failedWing_; }
    double forwardMs() const { return forwardMs_; }
    double reverseMs() const { return reverseMs_; }
    std::size_t deviceBytes() const { return bytes_; }

private:
    DeviceBatch& batch_;
    std::vector<StepArgs> args_;
    int substeps_;
    long steps_;
    int interval_;
    std::vector<std::unique_ptr<DeviceAdjoint>> wings_;
    std::vector<int> blocks_;
    int totalBlocks_ = 0, chunks_ = 1, failedWing_ = -1;
    DeviceError failure_{};
    std::vector<void*> owned_;
    void *views_ = nullptr, *sweepArgs_ = nullptr, *blockWing_ = nullptr, *blockLocal_ = nullptr,
         *blockCount_ = nullptr, *traceTable_ = nullptr, *adjointTable_ = nullptr;
    std::size_t bytes_ = 0;
    double forwardMs_ = 0.0, reverseMs_ = 0.0, replayMs_ = 0.0, jacobianMs_ = 0.0, sweepMs_ = 0.0;
    hipEvent_t events_[3] = {nullptr, nullptr, nullptr};
    std::vector<hipEvent_t> marks_;
    // The wings' Jacobian passes of one chunk are independent launches: they
    // run on kJacobianStreams streams of their own (default priority), wing w
    // on stream w mod kJacobianStreams, between a fork event on the batch
    // stream and one join event per stream that the batch stream waits on.
    static constexpr int kJacobianStreams = 4;
    hipStream_t jacobianStreams_[kJacobianStreams] = {};
    hipEvent_t jacobianJoin_[kJacobianStreams] = {};
    hipEvent_t jacobianFork_ = nullptr;

public:
    // The reverse pass split: recorded replays, Jacobian passes, sweeps.
    double replayMs() const { return replayMs_; }
    double jacobianMs() const { return jacobianMs_; }
    double sweepMs() const { return sweepMs_; }
};

// Device math conformance probe (function codes as sw_device_math).
void deviceMath(int device, int function, const double* a, const double* b, double* out, std::size_t count);

}  // namespace swx

namespace swx {

// Many wings stepped by one persistent launch: workgroup b runs workgroup
// blockLocal[b] of wing blockWing[b]. Each wing keeps its own tables, state,
// snapshot, error word and synchronisation words; only the launch is shared.
class DeviceBatch {
public:
    DeviceBatch(std::vector<DeviceWing*> wings, int device);
    ~DeviceBatch();
    DeviceBatch(const DeviceBatch&) = delete;
    DeviceBatch& operator=(const DeviceBatch&) = delete;

    // Chooses each wing's workgroups: one LDS-resident workgroup when its
    // nodes fit, else a share of the compute units in proportion to its items
    // (at least 3: workers and the chain workgroup; at most `cap`), and
    // uploads the launch tables. The total never exceeds the compute units,
    // so every workgroup of the launch is resident at once.
    void plan(int cap, int threads);
    // Re-plans with the last cap and threads when a wing's launch shape was
    // changed through its own handle since the plan.
    void ensurePlan();
    const std::vector<int>& blocks() const { return blocks_; }
    int totalBlocks() const { return total_; }

    void snapshot();                                 // wings without a pending error
    void restore(const std::vector<unsigned char>& which);
    // `traces` (a device array of one pointer per wing, or null): the
    // adjoint's recording of this substep for each wing.
    void launchSubstep(const std::vector<StepArgs>& args, double* const* traces = nullptr);
    void synchronize();
    void setCooperative(bool cooperative) { cooperative_ = cooperative; }

private:
    friend class DeviceBatchAdjoint;
    void release();
    std::vector<DeviceWing*> wings_;
    int device_;
    hipStream_t stream_ = nullptr;
    void *views_ = nullptr, *args_ = nullptr, *blockWing_ = nullptr, *blockLocal_ = nullptr, *mask_ = nullptr;
    std::vector<int> blocks_;
    std::vector<unsigned long long> generations_;
    int cap_ = 8;
    int total_ = 0;
    int threads_ = 256;
    std::size_t lds_ = 0;
    bool cooperative_ = true;
};

}  // namespace swx
