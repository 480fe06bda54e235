// gfx950 kernels of the XPBD structural substep and the device-side wing.
//
// One persistent launch runs a whole substep (soft_body.cpp:1175-1460):
//   A  prev = x, multiplier resets, payload prediction (beginSubstep)
//   B  pressure gather (ascending incident triangles) + predict
//   48 x [ distance colours | membrane colours | hinge colours |
//          suspension levels | contact-sphere levels ]
//   closing seam-stitch levels
//   membrane residual diagnostics + velocity finalisation + line residuals
//   payload velocity reconstruction + line-residual certification
// Phases are separated by a workgroup barrier (one workgroup) or a
// sense-reversing grid barrier (several workgroups, all resident).
// Node positions and inverse masses live in LDS when they fit (one
// workgroup), else in HBM; every constraint family is SoA in colour-major
// (or level-major) order, so each phase is one coalesced sweep.
#include "device_wing.hpp"
#include "xpbd_math.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>
#include <stdexcept>

namespace swx {

using namespace dev;

struct WingView {
    int nodeCount;
    double *px, *py, *pz;
    double *qx, *qy, *qz;
    double *vx, *vy, *vz;
    double *fx, *fy, *fz;
    double* w;
    int *incOffsets, *incTriangles;
    int *ta, *tb, *tc;
    double* pressure;
    int distCount;
    int *da, *db;
    double *drest, *dcomp, *dlam;
    unsigned char* dcable;
    int membCount;
    int *ma, *mb, *mc;
    double *m00, *m01, *m10, *m11, *mscale;
    double *cxx, *cyy, *czz, *cxy, *cxz, *cyz;
    double *preW, *preT, *mdamp;
    double *mlx, *mly, *mlz, *mrx, *mry, *mrz, *mres;
    int anyPreTension, anyDamping;
    int anyCompression;                  // some compressionStiffnessRatio != 1:
    double *mratio, *kxx, *kyy, *kzz, *kxy;  //   the ratio and the material stiffness
    int hingeCount;
    int *ha, *hb, *hc, *hd;
    double *hrest, *hcomp, *hlam;
    int segCount;
    int *sfrom, *sto;
    double *srest, *sstiff, *slam;
    unsigned char* slastPayload;
    int *sruntime;  // device slot -> runtime index (error reporting)
    int *sfromSlot, *stoSlot;  // staged-run endpoints: run-local node slot, or the hang-point code
    int* runNodes;             // node of each staged-run slot (see Phase::slotBegin)
    int hangCount;
    double* hangLocal;
    PayloadDev* payload;
    int pairCount;
    int *spNode, *spCentre;
    double *spRadius, *spComp, *splam;
    int stitchCount;
    int* stitchItems;
    int *distList, *membList;  // serial-sweep levels: device slots in level order
    int* cableList;            // one load-path cable pair's levels (device slots)
    int preserveMomentum;      // free membrane: restore the predicted momentum
    double* momentum;          // mass, centre, linear, angular after prediction
    int dampCount;             // segments with axial damping, runtime order:
    int* dampSlot;             //   device slot
    double* dampValue;         //   damping (N s/m)
    Phase* iterPhases;
    int iterPhaseCount;
    Phase* closePhases;
    int closePhaseCount;
    DeviceError* error;
    unsigned* barrier;
    // Point-to-point synchronisation of the iteration program: for
    // participant p and phase ph, entries [depOffsets[p*m+ph],
    // depOffsets[p*m+ph+1]) name a participant (depPart) whose progress must
    // have reached step - depLag + 1 before p starts that step; flags[32*q]
    // is participant q's progress (steps completed in this launch).
    int *depOffsets, *depPart, *depLag;
    unsigned* flags;
    int hasRun;  // the iteration program holds a serial segment run
    // Launch modes of this wing (set on the host before each launch):
    // point-to-point synchronisation, dedicated chain workgroup, nodes in
    // LDS, and the wing's workgroup count in a batched launch.
    int neighbourSync, dedicatedChain, sharedNodes, launchBlocks;
    // One-XCD synchronisation (set for launches whose participants all run
    // on one XCD, so they share one L2): published writes are drained to
    // that L2 and a waiter invalidates only its compute unit's L1, instead of
    // writing back and invalidating the L2 at agent scope. `epoch` tags the
    // launch's participant tickets (nonzero, different from the previous
    // launch's).
    int xcdLocal;
    unsigned epoch;
    // Trajectory recording for the adjoint (null otherwise): every item
    // writes the values its projection reads (its input vector), laid out as
    // traceLayout() describes; any launch mode records the same values.
    double* trace;
    // Step profile totals (null unless profiling): per part of the substep,
    // the wall-clock ticks (100 MHz) its workgroups spent there, then the
    // number of workgroup launches (kProfileParts entries).
    unsigned long long* profile;
    double* state;      // contiguous mutable state block
    double* snapshot;   // copy of `state`
    long stateDoubles;
};

// Layout of one substep's trace, in doubles:
//   prologue: x, v (3N each, node-major xyz), payload com, q, v, w (13);
//   per iteration, the input vector of every item before its projection,
//     by family in slot order: distance (pa, pb, lambda: 7), membrane
//     (p0, p1, p2, lambda xyz: 12), hinge (a, b, c, d, lambda: 13), segment
//     (A, B, lambda, rest, payload com, q, inverse inertia: 24; zero for a
//     harness endpoint), sphere (x, centre, lambda: 7);
//   closing: the input vector (7) of every closing-stitch entry;
//   end: payload com, q (7) after the iterations.
constexpr int kRecDist = 7, kRecMemb = 12, kRecHinge = 13, kRecSeg = 24, kRecSphere = 7;
struct TraceLayout {
    long nodes3, lambdaCount, distBase, membBase, hingeBase, segBase, sphereBase, iterStride, iterBase, closeBase,
        total;
};
__host__ __device__ inline TraceLayout traceLayout(const WingView& v, int iterations) {
    TraceLayout t;
    t.nodes3 = 3L * v.nodeCount;
    t.lambdaCount = static_cast<long>(v.distCount) + 3L * v.membCount + v.hingeCount + v.segCount + v.pairCount;
    t.distBase = 0;
    t.membBase = t.distBase + static_cast<long>(kRecDist) * v.distCount;
    t.hingeBase = t.membBase + static_cast<long>(kRecMemb) * v.membCount;
    t.segBase = t.hingeBase + static_cast<long>(kRecHinge) * v.hingeCount;
    t.sphereBase = t.segBase + static_cast<long>(kRecSeg) * v.segCount;
    t.iterStride = t.sphereBase + static_cast<long>(kRecSphere) * v.pairCount;
    t.iterBase = 2 * t.nodes3 + 13;
    t.closeBase = t.iterBase + static_cast<long>(iterations) * t.iterStride;
    t.total = t.closeBase + static_cast<long>(kRecDist) * v.stitchCount + 7;
    return t;
}

namespace {

constexpr int kMaxThreads = 256;

__device__ __forceinline__ void raise(const WingView& v, int code, int phase, int kind, int index) {
    if (atomicCAS(&v.error->code, 0, code) == 0) {
        v.error->phase = phase;
        v.error->kind = kind;
        v.error->index = index;
    }
}

// Bound on the polls of one wait (each poll is a device-scope round trip
// plus a short sleep, so this is several seconds). A wait that runs out
// raises SW_DEVICE_ERROR and lets the kernel finish instead of hanging.
constexpr unsigned kSpinLimit = 1u << 22;

// Spin bookkeeping shared by every wait: true when this wait has run out
// (it then flags the failure for all others) or another wait already has.
// The host clears the flag, the barrier words and the progress words before
// the next launch.
__device__ __forceinline__ bool syncTimedOut(const WingView& v, unsigned spins) {
    unsigned* failed = v.barrier + 64;
    if (spins > kSpinLimit) {
        __hip_atomic_store(failed, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        raise(v, SW_DEVICE_ERROR, -1, 3, -1);
        return true;
    }
    return (spins & 255u) == 0 && __hip_atomic_load(failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// Grid barrier over all resident workgroups of the launch (generation
// counter). One agent-scope release before arriving and one acquire after
// leaving make every workgroup's earlier global writes visible to every
// workgroup's later reads; the counter and generation words are touched only
// with relaxed atomics in between.
// Release half of a publication without an L2 write-back: every wave waits
// until its own earlier global stores are complete (in the XCD's L2 for a
// one-XCD launch; at the device's coherence point for the node stores of a
// device-wide substep, which are device-scope stores, see NodeRefs). Called
// by all threads before the workgroup barrier that precedes a publication.
// kLocal: the launch is a one-XCD launch (compiled per kernel variant);
// kCoherent: a device-wide substep whose cross-workgroup stores are
// device-scope.
template <bool kLocal = false, bool kCoherent = false>
__device__ __forceinline__ void drainStores(const WingView& v) {
    (void)v;
    if constexpr (kLocal || kCoherent) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// Acquire after a wait: the agent-scope acquire (buffer_inv sc1: this
// compute unit's L1, and the L2 lines other XCDs may have changed). With
// every participant on one XCD the writers did not write their L2 lines
// back, which the invalidation keeps (dirty lines are not dropped), so the
// same invalidation is the acquire; measured on gfx950, the workgroup-scope
// form (sc0) leaves stale L1 lines of other compute units' writes.
__device__ __forceinline__ void acquireAfterWait(const WingView& v) {
    (void)v;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

template <bool kLocal = false>
__device__ __forceinline__ void gridSync(const WingView& v, int workers) {
    if (workers == 1) {
        __syncthreads();
        return;
    }
    drainStores<kLocal>(v);
    __syncthreads();
    if (threadIdx.x == 0) {
        constexpr bool local = kLocal;
        unsigned* count = v.barrier;
        unsigned* generation = v.barrier + 32;
        const unsigned gen = __hip_atomic_load(generation, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // The generation must be read before this workgroup arrives: once it
        // has arrived, the last arrival may already advance it.
        if constexpr (!local) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        const unsigned arrived = __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == static_cast<unsigned>(workers) - 1u) {
            // Every other workgroup has arrived: reset the counter, then
            // publish the next generation (release orders the reset first;
            // one XCD: the reset is an earlier atomic at the same L2).
            acquireAfterWait(v);
            __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if constexpr (local) {
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                __hip_atomic_store(generation, gen + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                __hip_atomic_store(generation, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else {
            unsigned spins = 0;
            while (__hip_atomic_load(generation, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(1);
                if (syncTimedOut(v, ++spins)) break;
            }
            acquireAfterWait(v);
        }
    }
    __syncthreads();
}

// XCD (accelerator complex die) of the calling wave: hardware register
// XCC_ID (20), bits [3:0].
__device__ __forceinline__ unsigned xccId() { return __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 0xfu; }

// Participant index of this workgroup in a one-XCD launch (thread 0 only),
// or -1 if it does not take part. Each XCD's workgroups draw tickets from
// the XCD's own counter (barrier word 96 + XCD); the workgroup drawing
// ticket `participants - 1` claims the launch for its XCD (word 112) unless
// another XCD did first; workgroups of the claiming XCD with a ticket below
// `participants` take part under that index. Counters and the claim carry
// the launch's epoch in their upper 20 bits, so words left by an earlier
// launch count as empty. Participants therefore all run on one XCD,
// whatever the dispatcher's placement, provided one XCD receives at least
// `participants` workgroups (the host launches 8 x participants).
__device__ __forceinline__ int xcdParticipant(const WingView& v, int participants) {
    const unsigned xcc = xccId();
    unsigned* counter = v.barrier + 96 + xcc;
    unsigned* claim = v.barrier + 112;
    const unsigned tag = (v.epoch & 0xfffffu) << 12;
    unsigned old = __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned ticket = 0;
    for (;;) {
        const bool current = (old & ~0xfffu) == tag;
        const unsigned next = current ? old + 1u : (tag | 1u);
        if (__hip_atomic_compare_exchange_strong(counter, &old, next, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
            ticket = (next & 0xfffu) - 1u;
            break;
        }
    }
    if (ticket >= static_cast<unsigned>(participants)) return -1;
    if (ticket == static_cast<unsigned>(participants) - 1u) {
        unsigned seen = __hip_atomic_load(claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while ((seen & ~0xfffu) != tag) {
            if (__hip_atomic_compare_exchange_strong(claim, &seen, tag | (xcc + 1u), __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                break;
        }
    }
    unsigned spins = 0;
    unsigned seen = 0;
    while (((seen = __hip_atomic_load(claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) & ~0xfffu) != tag) {
        __builtin_amdgcn_s_sleep(1);
        if (syncTimedOut(v, ++spins)) return -1;
    }
    return (seen & 0xfffu) == xcc + 1u ? static_cast<int>(ticket) : -1;
}

// Node position/inverse-mass store: LDS copies or the HBM arrays. A staged
// segment run uses the same interface over its run-local slots.
struct NodeRefs {
    double *x, *y, *z, *w;
    const double* hangLocal;  // harness points in the payload body frame (3 per point)
    __device__ __forceinline__ d3 get(int i) const {
        if (coherent)
            return d3{__hip_atomic_load(x + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), __hip_atomic_load(y + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                      __hip_atomic_load(z + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
        return d3{x[i], y[i], z[i]};
    }
    __device__ __forceinline__ d3 hang(int h) const {
        return d3{hangLocal[3 * h], hangLocal[3 * h + 1], hangLocal[3 * h + 2]};
    }
    __device__ __forceinline__ double invMass(int i) const { return w[i]; }
    // coherent: nodes in HBM under device-wide point-to-point publication
    // (publishCounter with kCoherent): loads and stores are device-scope
    // relaxed atomics (global_load / global_store sc1). The stores are
    // complete at the device's coherence point once drained, so no L2
    // write-back is needed before a publication, and the loads read that
    // point, so no cache invalidation is needed after a wait.
    bool coherent = false;
    __device__ __forceinline__ void put(int i, d3 p) const {
        if (coherent) {
            __hip_atomic_store(x + i, p.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(y + i, p.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(z + i, p.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        x[i] = p.x;
        y[i] = p.y;
        z[i] = p.z;
    }
};

// Distance / cable / tie / stitch projection (soft_body.cpp:1598-1640).
__device__ __forceinline__ void put3(double* r, const d3& p) {
    r[0] = p.x;
    r[1] = p.y;
    r[2] = p.z;
}

// Per-item values of a projection that no other item of the iteration
// changes: indices, inverse masses, constants and the item's own multiplier
// (only the thread that projects the item writes it). The point-to-point
// loop loads them before it waits for its dependencies.
struct DistanceItem {
    int a, b;
    double wa, wb, rest, compliance, lambda;
    bool cable;
};

__device__ __forceinline__ DistanceItem distanceItem(const WingView& v, const NodeRefs& N, int k) {
    DistanceItem it;
    it.a = v.da[k];
    it.b = v.db[k];
    it.wa = N.w[it.a];
    it.wb = N.w[it.b];
    it.rest = v.drest[k];
    it.compliance = v.dcomp[k];
    it.lambda = v.dlam[k];
    it.cable = v.dcable[k] != 0;
    return it;
}

// `rec` (recording only): the item's input vector is written there first.
__device__ __forceinline__ void projectDistance(const WingView& v, const NodeRefs& N, int k, const DistanceItem& it,
                                                double invDt2, double* rec) {
    d3 pa = N.get(it.a), pb = N.get(it.b);
    if (rec) {
        put3(rec, pa);
        put3(rec + 3, pb);
        rec[6] = it.lambda;
    }
    const double wa = it.wa, wb = it.wb;
    const d3 diff = sub(pb, pa);
    const double length = len(diff);
    if (length <= 1.0e-12) return;
    const double c = length - it.rest;
    const bool cable = it.cable;
    double lambda = it.lambda;
    if (cable && c <= 0.0 && lambda == 0.0) return;
    const double wsum = wa + wb;
    const double alpha = it.compliance * invDt2;
    if (wsum + alpha <= 0.0) return;
    const double dl = (-c - alpha * lambda) / (wsum + alpha);
    double applied = dl;
    if (cable) {
        const double next = stdmin(0.0, lambda + dl);
        applied = next - lambda;
        lambda = next;
    } else {
        lambda += dl;
    }
    v.dlam[k] = lambda;
    const double s = applied / length;
    const double sa = wa * s, sb = wb * s;
    pa = sub(pa, scale(diff, sa));
    pb = add(pb, scale(diff, sb));
    N.put(it.a, pa);
    N.put(it.b, pb);
}

__device__ __forceinline__ void projectDistance(const WingView& v, const NodeRefs& N, int k, double invDt2,
                                                double* rec = nullptr) {
    projectDistance(v, N, k, distanceItem(v, N, k), invDt2, rec);
}

// Membrane strain and its gradients (soft_body.cpp:204-282, StrainOnly).
struct MembraneEval {
    d3 strain;
    d3 grad[3][3];  // [component][corner]
};

__device__ __forceinline__ MembraneEval evaluateMembrane(d3 p0, d3 p1, d3 p2, double m00, double m01, double m10,
                                                         double m11, double preW, double preT) {
    MembraneEval e;
    const d3 e1 = sub(p1, p0);
    const d3 e2 = sub(p2, p0);
    const d3 fw = add(scale(e1, m00), scale(e2, m10));
    const d3 ft = add(scale(e1, m01), scale(e2, m11));
    e.strain = d3{0.5 * (dot(fw, fw) - 1.0) + preW, 0.5 * (dot(ft, ft) - 1.0) + preT, dot(fw, ft)};
    const d3 zero{0.0, 0.0, 0.0};
    const d3 byWarp[3] = {fw, zero, ft};
    const d3 byWeft[3] = {zero, ft, fw};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const d3 g1 = add(scale(byWarp[k], m00), scale(byWeft[k], m01));
        const d3 g2 = add(scale(byWarp[k], m10), scale(byWeft[k], m11));
        e.grad[k][0] = neg(add(g1, g2));
        e.grad[k][1] = g1;
        e.grad[k][2] = g2;
    }
    return e;
}

// One membrane element (soft_body.cpp:1822-1929).
struct MembraneItem {
    int i0, i1, i2;
    double w[3];
    double preW, preT, m00, m01, m10, m11, areaScale;
    double cxx, cyy, czz, cxy, cxz, cyz;
    double damping;
    d3 lambda;
    d3 prev[3];  // start-of-substep positions (damped membranes only)
};

__device__ __forceinline__ MembraneItem membraneItem(const WingView& v, const NodeRefs& N, int k) {
    MembraneItem it;
    it.i0 = v.ma[k];
    it.i1 = v.mb[k];
    it.i2 = v.mc[k];
    it.w[0] = N.w[it.i0];
    it.w[1] = N.w[it.i1];
    it.w[2] = N.w[it.i2];
    it.preW = v.anyPreTension ? v.preW[k] : 0.0;
    it.preT = v.anyPreTension ? v.preT[k] : 0.0;
    it.m00 = v.m00[k];
    it.m01 = v.m01[k];
    it.m10 = v.m10[k];
    it.m11 = v.m11[k];
    it.areaScale = v.mscale[k];
    it.cxx = v.cxx[k];
    it.cyy = v.cyy[k];
    it.czz = v.czz[k];
    it.cxy = v.cxy[k];
    it.cxz = v.cxz[k];
    it.cyz = v.cyz[k];
    it.lambda = d3{v.mlx[k], v.mly[k], v.mlz[k]};
    it.damping = v.anyDamping ? v.mdamp[k] : 0.0;
    if (v.anyDamping) {
        const int idx[3] = {it.i0, it.i1, it.i2};
#pragma unroll
        for (int corner = 0; corner < 3; ++corner)
            it.prev[corner] = d3{v.qx[idx[corner]], v.qy[idx[corner]], v.qz[idx[corner]]};
    }
    return it;
}

// The compliance of a membrane whose compressionStiffnessRatio is not 1 for
// its current Green strain: effectiveMembraneStiffness (membrane.cpp:273-296;
// the material stiffness has no xz and yz entries) inverted by
// checkedInverse (membrane.cpp:208-235). False where the reference throws.
__device__ __forceinline__ bool compressedCompliance(const WingView& v, int k, double ratio, d3 strain, s3& out) {
    const double retained = sqrt(ratio);
    const double warp = strain.x < 0.0 ? retained : 1.0;
    const double weft = strain.y < 0.0 ? retained : 1.0;
    const double shear = sqrt(warp * weft);
    const s3 effective{v.kxx[k] * warp * warp, v.kyy[k] * weft * weft, v.kzz[k] * shear * shear,
                       v.kxy[k] * warp * weft, 0.0, 0.0};
    return checkedInverse(effective, out);
}

// kGeneral: compiled with the compressionStiffnessRatio branch.
template <bool kGeneral = true>
__device__ __forceinline__ void projectMembrane(const WingView& v, const NodeRefs& N, int k, const MembraneItem& it,
                                                const StepArgs& a, double* rec) {
    const int i0 = it.i0, i1 = it.i1, i2 = it.i2;
    const d3 p[3] = {N.get(i0), N.get(i1), N.get(i2)};
    if (rec) {
        put3(rec, p[0]);
        put3(rec + 3, p[1]);
        put3(rec + 6, p[2]);
        rec[9] = it.lambda.x;
        rec[10] = it.lambda.y;
        rec[11] = it.lambda.z;
    }
    const double* w = it.w;
    const MembraneEval e = evaluateMembrane(p[0], p[1], p[2], it.m00, it.m01, it.m10, it.m11, it.preW, it.preT);
    const double areaScale = it.areaScale;
    d3 J[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int corner = 0; corner < 3; ++corner) J[c][corner] = scale(e.grad[c][corner], areaScale);
    s3 M{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int corner = 0; corner < 3; ++corner) {
        const double wc = w[corner];
        M.xx += wc * dot(J[0][corner], J[0][corner]);
        M.yy += wc * dot(J[1][corner], J[1][corner]);
        M.zz += wc * dot(J[2][corner], J[2][corner]);
        M.xy += wc * dot(J[0][corner], J[1][corner]);
        M.xz += wc * dot(J[0][corner], J[2][corner]);
        M.yz += wc * dot(J[1][corner], J[2][corner]);
    }
    s3 at{a.invDt2 * it.cxx, a.invDt2 * it.cyy, a.invDt2 * it.czz,
          a.invDt2 * it.cxy, a.invDt2 * it.cxz, a.invDt2 * it.cyz};
    if (kGeneral && v.anyCompression) {
        // soft_body.cpp:1869-1873
        const double ratio = v.mratio[k];
        if (ratio != 1.0) {
            s3 state;
            if (!compressedCompliance(v, k, ratio, e.strain, state)) {
                raise(v, SW_INVALID_ARGUMENT, -1, 0, k);
                return;
            }
            at = s3{a.invDt2 * state.xx, a.invDt2 * state.yy, a.invDt2 * state.zz,
                    a.invDt2 * state.xy, a.invDt2 * state.xz, a.invDt2 * state.yz};
        }
    }
    const d3 lambda = it.lambda;
    const d3 cv = scale(e.strain, areaScale);
    const d3 atl = mul(at, lambda);
    s3 S;
    d3 rhs;
    if (v.anyDamping) {
        const double gamma = it.damping / a.dt;
        d3 dm{0.0, 0.0, 0.0};
#pragma unroll
        for (int corner = 0; corner < 3; ++corner) {
            const d3 disp = sub(p[corner], it.prev[corner]);
            dm.x += dot(J[0][corner], disp);
            dm.y += dot(J[1][corner], disp);
            dm.z += dot(J[2][corner], disp);
        }
        const double f = 1.0 + gamma;
        S = s3{f * M.xx + at.xx, f * M.yy + at.yy, f * M.zz + at.zz,
               f * M.xy + at.xy, f * M.xz + at.xz, f * M.yz + at.yz};
        rhs = sub(sub(neg(cv), atl), scale(dm, gamma));
    } else {
        // dampingTime == 0: gamma == 0, (1 + 0) * M == M exactly and the
        // damping term contributes a signed zero only.
        S = s3{M.xx + at.xx, M.yy + at.yy, M.zz + at.zz, M.xy + at.xy, M.xz + at.xz, M.yz + at.yz};
        rhs = sub(neg(cv), atl);
    }
    d3 dl;
    if (!checkedSolve(S, rhs, dl)) {
        raise(v, SW_INVALID_ARGUMENT, -1, 0, k);
        return;
    }
    d3 corr[3];
#pragma unroll
    for (int corner = 0; corner < 3; ++corner) {
        const d3 c = add(add(scale(J[0][corner], dl.x), scale(J[1][corner], dl.y)), scale(J[2][corner], dl.z));
        corr[corner] = scale(c, w[corner]);
    }
    v.mlx[k] = lambda.x + dl.x;
    v.mly[k] = lambda.y + dl.y;
    v.mlz[k] = lambda.z + dl.z;
    N.put(i0, add(p[0], corr[0]));
    N.put(i1, add(p[1], corr[1]));
    N.put(i2, add(p[2], corr[2]));
}

template <bool kGeneral = true>
__device__ __forceinline__ void projectMembrane(const WingView& v, const NodeRefs& N, int k, const StepArgs& a,
                                                double* rec = nullptr) {
    projectMembrane<kGeneral>(v, N, k, membraneItem(v, N, k), a, rec);
}

struct HingeItem {
    int idx[4];
    double w[4];
    double compliance, rest, lambda;
};

__device__ __forceinline__ HingeItem hingeItem(const WingView& v, const NodeRefs& N, int k) {
    HingeItem it;
    it.idx[0] = v.ha[k];
    it.idx[1] = v.hb[k];
    it.idx[2] = v.hc[k];
    it.idx[3] = v.hd[k];
#pragma unroll
    for (int corner = 0; corner < 4; ++corner) it.w[corner] = N.w[it.idx[corner]];
    it.compliance = v.hcomp[k];
    it.rest = v.hrest[k];
    it.lambda = v.hlam[k];
    return it;
}

// The atan table of the portable atan2 (pm::kAtanTable), copied into LDS at
// the start of every substep (substepBody): a lookup in global memory would
// miss in the caches each hinge phase, after the acquire that precedes it has
// invalidated them.
__shared__ double atanLds[65][2];

// Step profile (the device side of StepPerformanceProfile, soft_body.h:136-170):
// thread 0 of each workgroup adds the wall-clock ticks (100 MHz) it spends in
// each part of the substep to an LDS tally, which it adds to the wing's
// totals at the end of the launch. Parts: kProfilePrediction ..
// kProfileFinalization; the last slot counts workgroup launches.
constexpr int kProfilePrediction = 0, kProfileDistance = 1, kProfileCable = 2, kProfileMembrane = 3,
              kProfileBending = 4, kProfileSuspension = 5, kProfileContact = 6, kProfileDiagnostics = 7,
              kProfileFinalization = 8, kProfileParts = 10;
__shared__ unsigned long long profileTicks[kProfileParts];
__device__ __forceinline__ void profileStart(const WingView& v, unsigned long long& since) {
    if (v.profile && threadIdx.x == 0) {
        for (int k = 0; k < kProfileParts; ++k) profileTicks[k] = 0;
        since = wall_clock64();
    }
}
__device__ __forceinline__ void profileAdd(const WingView& v, int part, unsigned long long& since) {
    if (v.profile && threadIdx.x == 0) {
        const unsigned long long now = wall_clock64();
        profileTicks[part] += now - since;
        since = now;
    }
}
__device__ __forceinline__ int profilePart(int type) {
    switch (type) {
    case kPhaseCableList: return kProfileCable;
    case kPhaseMembrane:
    case kPhaseMembraneList: return kProfileMembrane;
    case kPhaseHinge: return kProfileBending;
    case kPhaseSegments:
    case kPhaseSegmentRun: return kProfileSuspension;
    case kPhaseSpheres: return kProfileContact;
    default: return kProfileDistance;
    }
}
__device__ __forceinline__ void profileFlush(const WingView& v) {
    if (v.profile && threadIdx.x == 0) {
        profileTicks[kProfileParts - 1] = 1;
        for (int k = 0; k < kProfileParts; ++k) atomicAdd(v.profile + k, profileTicks[k]);
    }
}

// Dihedral hinge (soft_body.cpp:1648-1711 and 1754-1792).
// remainder(x, 2π) (the double nearest 2π): for |x| < 3π the quotient is
// -1, 0 or 1, and |x| - 2π is exact (Sterbenz), so the result equals the
// library's, ties included (|x| = π keeps n = 0, the even quotient) and a
// zero result carrying the sign of x; other arguments (and NaN) take the
// library call.
__device__ __forceinline__ double remainderTwoPi(double x) {
    constexpr double y = 6.283185307179586476925286766559;
    const double ax = fabs(x);
    if (ax <= 0.5 * y) return x;
    if (ax < 1.5 * y) {
        const double r = ax - y;
        return x < 0.0 ? -r : r;
    }
    return remainder(x, y);
}

__device__ __forceinline__ void projectHinge(const WingView& v, const NodeRefs& N, int k, const HingeItem& it,
                                             double invDt2, double* rec) {
    const int* idx = it.idx;
    const d3 a = N.get(idx[0]), b = N.get(idx[1]), c = N.get(idx[2]), d = N.get(idx[3]);
    if (rec) {
        put3(rec, a);
        put3(rec + 3, b);
        put3(rec + 6, c);
        put3(rec + 9, d);
        rec[12] = it.lambda;
    }
    const d3 edge = sub(b, a);
    const d3 toC = sub(c, a);
    const d3 toD = sub(d, a);
    const d3 area1 = cross(edge, toC);
    const d3 area2 = cross(toD, edge);
    const double edgeLength = len(edge);
    const double l1 = len(area1);
    const double l2 = len(area2);
    if (!(edgeLength > 1.0e-12) || !(l1 > 1.0e-12 * edgeLength) || !(l2 > 1.0e-12 * edgeLength)) return;
    const d3 eu = divide(edge, edgeLength);
    const d3 n1 = divide(area1, l1);
    const d3 n2 = divide(area2, l2);
    const double cosine = stdclamp(dot(n1, n2), -1.0, 1.0);
    const double sine = dot(eu, cross(n1, n2));
    const d3 n1Adj = sub(scale(cross(n2, eu), cosine), scale(n2, sine));
    const d3 n2Adj = sub(scale(cross(eu, n1), cosine), scale(n1, sine));
    const d3 euAdj = scale(cross(n1, n2), cosine);
    const d3 a1Adj = divide(sub(n1Adj, scale(n1, dot(n1Adj, n1))), l1);
    const d3 a2Adj = divide(sub(n2Adj, scale(n2, dot(n2Adj, n2))), l2);
    d3 edgeAdj = divide(sub(euAdj, scale(eu, dot(euAdj, eu))), edgeLength);
    edgeAdj = add(edgeAdj, cross(toC, a1Adj));
    const d3 cAdj = cross(a1Adj, edge);
    edgeAdj = add(edgeAdj, cross(a2Adj, toD));
    const d3 dAdj = cross(edge, a2Adj);
    const double angle = pm::atan2(sine, cosine, atanLds);
    const d3 g[4] = {scale(add(add(edgeAdj, cAdj), dAdj), -1.0), edgeAdj, cAdj, dAdj};
    if (!isfinite(angle) || !fin3(g[0]) || !fin3(g[1]) || !fin3(g[2]) || !fin3(g[3])) return;
    double wgrad = 0.0;
    const double* w = it.w;
#pragma unroll
    for (int corner = 0; corner < 4; ++corner) wgrad += w[corner] * dot(g[corner], g[corner]);
    const double alpha = it.compliance * invDt2;
    const double den = wgrad + alpha;
    if (!(den > 0.0) || !isfinite(den)) return;
    const double value = remainderTwoPi(angle - it.rest);
    const double lambda = it.lambda;
    const double dl = (-value - alpha * lambda) / den;
    if (!isfinite(dl)) return;
    const d3 p[4] = {a, b, c, d};
#pragma unroll
    for (int corner = 0; corner < 4; ++corner) N.put(idx[corner], add(p[corner], scale(g[corner], w[corner] * dl)));
    v.hlam[k] = lambda + dl;
}

__device__ __forceinline__ void projectHinge(const WingView& v, const NodeRefs& N, int k, double invDt2,
                                             double* rec = nullptr) {
    projectHinge(v, N, k, hingeItem(v, N, k), invDt2, rec);
}

// Payload state held in registers by the lane running a chain of
// payload segments.
struct Payload {
    d3 com;
    quat q;
};

struct Endpoint {
    d3 position;
    d3 offset;
    double invMass;
    bool rigid;
    bool fixed;
};

template <class Store>
__device__ __forceinline__ bool endpointState(const WingView& v, const Store& N, int e, const Payload& P,
                                              const PayloadDev& def, Endpoint& out) {
    if (e >= 0) {
        out.position = N.get(e);
        out.invMass = N.invMass(e);
        out.fixed = out.invMass == 0.0;
        out.rigid = false;
        out.offset = d3{0.0, 0.0, 0.0};
        return true;
    }
    const d3 local = N.hang(-e - 1);
    m3 r;
    if (!rotationMatrix(P.q, r)) return false;
    const d3 rel = sub(local, d3{def.comLocal[0], def.comLocal[1], def.comLocal[2]});
    out.position = add(P.com, mulM(r, rel));
    out.offset = sub(out.position, P.com);
    out.rigid = true;
    out.fixed = false;
    out.invMass = 1.0 / def.mass;
    return true;
}

// generalizedLineEffectiveInverseMass (suspension.cpp:1104-1123).
// `iinvFinite` is the finiteness of every entry of `iinv`, which the
// reference checks on each call; callers evaluate it once per inverse.
__device__ __forceinline__ bool finiteM(const m3& m) {
    bool ok = true;
    for (int k = 0; k < 9; ++k) ok = ok && isfinite(m.a[k]);
    return ok;
}
__device__ __forceinline__ bool effectiveInverseMass(const Endpoint& e, bool rigid, d3 g, const m3& iinv,
                                                     bool iinvFinite, double& out) {
    if (!(e.invMass >= 0.0) || !isfinite(e.invMass) || !fin3(g) || !fin3(e.offset)) return false;
    if (!iinvFinite) return false;
    double r = e.invMass;
    if (rigid && e.invMass > 0.0) {
        const d3 rot = cross(e.offset, g);
        r += dot(rot, mulM(iinv, rot));
    }
    if (!(r >= 0.0) || !isfinite(r)) return false;
    out = r;
    return true;
}

// One suspension segment of solveLineIteration (suspension.cpp:2197-2274)
// with projectUnilateralLine (suspension.cpp:1125-1163): endpoints fe/te
// (node or run slot of `N`, or a hang-point code), its multiplier updated
// in place. Returns false on a LineSolve failure (already raised).
template <class Store>
__device__ bool projectSegment(const WingView& v, const Store& N, int fe, int te, double rest, double stiffness,
                               double& lambda, int runtime, Payload& P, const PayloadDev& def, const m3& iinv,
                               bool iinvFinite, const StepArgs& a) {
    Endpoint from, to;
    if (!endpointState(v, N, fe, P, def, from) || !endpointState(v, N, te, P, def, to)) {
        raise(v, SW_SUSPENSION_ERROR, 1, 2, runtime);
        return false;
    }
    const d3 diff = sub(to.position, from.position);
    const double length = len(diff);
    const double minLength = def.minLineLength;
    if (!isfinite(length) || length <= minLength) {
        raise(v, SW_SUSPENSION_ERROR, 6, 1, runtime);
        return false;
    }
    const d3 dir = divide(diff, length);
    const d3 toGrad = neg(dir);
    const double gradLength = len(dir);
    if (!(length > 0.0) || !isfinite(length) || !(rest > 0.0) || !isfinite(rest) || !(stiffness > 0.0) ||
        !isfinite(stiffness) || !(a.dt > 0.0) || !isfinite(a.dt) || !(lambda >= 0.0) || !isfinite(lambda) ||
        fabs(gradLength - 1.0) > 1.0e-10) {
        raise(v, SW_SUSPENSION_ERROR, 6, 1, runtime);
        return false;
    }
    double effFrom, effTo;
    if (!effectiveInverseMass(from, from.rigid && !from.fixed, dir, iinv, iinvFinite, effFrom) ||
        !effectiveInverseMass(to, to.rigid && !to.fixed, neg(dir), iinv, iinvFinite, effTo)) {
        raise(v, SW_SUSPENSION_ERROR, 6, 1, runtime);
        return false;
    }
    const double eff = effFrom + effTo;
    const double alphaTilde = 1.0 / (stiffness * a.dt * a.dt);
    if (!(eff + alphaTilde > 0.0)) {
        raise(v, SW_SUSPENSION_ERROR, 6, 1, runtime);
        return false;
    }
    const double gap = rest - length;
    const double candidate = lambda + (-gap - alphaTilde * lambda) / (eff + alphaTilde);
    const double next = stdmax(0.0, candidate);
    const double increment = next - lambda;
    lambda = next;
    // Corrections: from endpoint first, then to endpoint.
    const Endpoint* ends[2] = {&from, &to};
    const int ids[2] = {fe, te};
    const d3 grads[2] = {dir, toGrad};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const Endpoint& e = *ends[s];
        if (e.fixed) continue;
        const d3 step = scale(grads[s], e.invMass * increment);
        if (ids[s] >= 0) {
            // The endpoint's position as read above: the two endpoints are
            // distinct nodes (a zero length was rejected), so the first
            // correction's store does not change the second's node.
            N.put(ids[s], add(e.position, step));
        } else {
            P.com = add(P.com, step);
            const d3 rotation = scale(mulM(iinv, cross(e.offset, grads[s])), increment);
            quat inc, rotated;
            if (!rotationIncrement(rotation, inc) || !canonical(qmul(inc, P.q), rotated)) {
                raise(v, SW_SUSPENSION_ERROR, 5, 2, 0);
                return false;
            }
            P.q = rotated;
        }
    }
    return true;
}

// normalizedCanonical with the failure returned as a flag after the values
// are computed (the same values as canonical where it succeeds), so no
// branch waits on the normalisation.
__device__ __forceinline__ bool canonicalDeferred(const quat& q, quat& out) {
    const double n2 = qnorm2(q);
    const bool ok = finQ(q) && (n2 > 1.0e-28);
    const double inv = 1.0 / sqrt(n2);
    const quat r{q.w * inv, q.x * inv, q.y * inv, q.z * inv};
    const bool flip = r.w < 0.0 || (r.w == 0.0 && (r.x < 0.0 || (r.x == 0.0 && (r.y < 0.0 || (r.y == 0.0 && r.z < 0.0)))));
    out = quat{flip ? -r.w : r.w, flip ? -r.x : r.x, flip ? -r.y : r.y, flip ? -r.z : r.z};
    return ok;
}

__device__ __forceinline__ bool rotationMatrixDeferred(const quat& orientation, m3& r) {
    quat q;
    const bool ok = canonicalDeferred(orientation, q);
    const double xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const double xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    const double wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    r.a[0] = 1.0 - 2.0 * (yy + zz);
    r.a[1] = 2.0 * (xy - wz);
    r.a[2] = 2.0 * (xz + wy);
    r.a[3] = 2.0 * (xy + wz);
    r.a[4] = 1.0 - 2.0 * (xx + zz);
    r.a[5] = 2.0 * (yz - wx);
    r.a[6] = 2.0 * (xz - wy);
    r.a[7] = 2.0 * (yz + wx);
    r.a[8] = 1.0 - 2.0 * (xx + yy);
    return ok;
}

__device__ __forceinline__ bool rotationIncrementDeferred(d3 v, quat& out) {
    const bool finite = fin3(v);
    const double angle = len(v);
    if (angle < 1.0e-12) return canonicalDeferred(quat{1.0, 0.5 * v.x, 0.5 * v.y, 0.5 * v.z}, out) && finite;
    const double half = 0.5 * angle;
    double sine, cosine;
    pm::sincos(half, &sine, &cosine);
    const double s = sine / angle;
    out = quat{cosine, v.x * s, v.y * s, v.z * s};
    return finite;
}

__device__ __forceinline__ bool refreshInverseInertia(PayloadDev& def, const quat& q) {
    m3 world, inv;
    if (!worldInertia(q, def.inertia, world) || !inverseM(world, inv)) return false;
    for (int k = 0; k < 9; ++k) __hip_atomic_store(&def.iinv[k], inv.a[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // see storePayload
    return true;
}

// Device-scope loads (see storePayload).
__device__ __forceinline__ Payload loadPayload(const PayloadDev& p) {
    return Payload{d3{__hip_atomic_load(&p.com[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), __hip_atomic_load(&p.com[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                      __hip_atomic_load(&p.com[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)},
                   quat{__hip_atomic_load(&p.q[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), __hip_atomic_load(&p.q[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                        __hip_atomic_load(&p.q[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), __hip_atomic_load(&p.q[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)}};
}
// Device-scope stores: the payload pose is read by other workgroups after
// a publication without an L2 write-back (publishCounter, kCoherent).
__device__ __forceinline__ void storePayload(PayloadDev& p, const Payload& s) {
    __hip_atomic_store(&p.com[0], s.com.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&p.com[1], s.com.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&p.com[2], s.com.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&p.q[0], s.q.w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&p.q[1], s.q.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&p.q[2], s.q.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&p.q[3], s.q.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ m3 loadInverseInertia(const PayloadDev& p) {
    m3 r;
    for (int k = 0; k < 9; ++k) r.a[k] = __hip_atomic_load(&p.iinv[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return r;
}

constexpr int kChainHang = 32;  // harness points staged with a serial run

// Staging of a serial run: distinct LDS objects addressed directly, so the
// serial loop reads and writes them with LDS instructions and its loads of
// item values never wait on its stores of node slots.
__shared__ double chainX[kChainSlots], chainY[kChainSlots], chainZ[kChainSlots], chainW[kChainSlots];
__shared__ double chainHang[3 * kChainHang];
__shared__ double chainLambda[kChainItems], chainRest[kChainItems], chainStiffness[kChainItems];
__shared__ int chainItem[4 * kChainItems];  // from, to, runtime index, last-payload flag
__shared__ double chainAlpha[kChainItems];   // 1 / (stiffness dt^2)
__shared__ int chainStatic[kChainItems];     // rest, stiffness and dt pass projectSegment's checks
__shared__ d3 chainRel[kChainHang];          // harness point local - comLocal

struct SlotRefs {
    __device__ __forceinline__ d3 get(int i) const { return d3{chainX[i], chainY[i], chainZ[i]}; }
    __device__ __forceinline__ double invMass(int i) const { return chainW[i]; }
    __device__ __forceinline__ d3 hang(int h) const {
        return d3{chainHang[3 * h], chainHang[3 * h + 1], chainHang[3 * h + 2]};
    }
    __device__ __forceinline__ void put(int i, d3 p) const {
        chainX[i] = p.x;
        chainY[i] = p.y;
        chainZ[i] = p.z;
    }
};

// One segment of a staged serial run: projectSegment's operations in the
// same order, specialised on the endpoint kinds (FromHang / ToHang: a
// harness point, whose offset from the centre of mass local - comLocal is
// staged in `rel`; otherwise a run slot). The per-segment constants are
// staged: alphaTilde = 1 / (stiffness dt^2) and `staticOk`, the checks of
// rest, stiffness and dt. Every failure becomes a flag tested once before
// the corrections (and once after the payload rotation), so no branch waits
// on the projection's arithmetic; the error raised is the one
// projectSegment raises.
struct ChainEnd {
    d3 position, offset;
    double invMass;
};

template <bool Hang>
__device__ __forceinline__ ChainEnd chainEnd(int e, const m3& r, const Payload& P, const d3* rel,
                                             double invPayloadMass) {
    ChainEnd out;
    if constexpr (Hang) {
        out.position = add(P.com, mulM(r, rel[-e - 1]));
        out.offset = sub(out.position, P.com);
        out.invMass = invPayloadMass;
    } else {
        out.position = d3{chainX[e], chainY[e], chainZ[e]};
        out.offset = d3{0.0, 0.0, 0.0};
        out.invMass = chainW[e];
    }
    return out;
}

template <bool Hang>
__device__ __forceinline__ bool chainEffectiveMass(const ChainEnd& e, d3 g, const m3& iinv, bool iinvFinite,
                                                   double& out) {
    const bool pre = (e.invMass >= 0.0) && isfinite(e.invMass) && fin3(g) && fin3(e.offset) && iinvFinite;
    double r = e.invMass;
    if constexpr (Hang) {
        if (e.invMass > 0.0) {
            const d3 rot = cross(e.offset, g);
            r += dot(rot, mulM(iinv, rot));
        }
    }
    out = r;
    return pre && (r >= 0.0) && isfinite(r);
}

template <bool Hang>
__device__ __forceinline__ bool chainCorrect(int e, const ChainEnd& end, d3 grad, double increment, Payload& P,
                                             const m3& iinv) {
    if constexpr (Hang) {
        const d3 step = scale(grad, end.invMass * increment);
        P.com = add(P.com, step);
        const d3 rotation = scale(mulM(iinv, cross(end.offset, grad)), increment);
        quat inc, rotated;
        const bool ok = rotationIncrementDeferred(rotation, inc);
        const bool normal = canonicalDeferred(qmul(inc, P.q), rotated);
        P.q = rotated;
        return ok && normal;
    } else {
        if (end.invMass == 0.0) return true;  // fixed node
        const d3 step = scale(grad, end.invMass * increment);
        const d3 p{chainX[e], chainY[e], chainZ[e]};
        const d3 q = add(p, step);
        chainX[e] = q.x;
        chainY[e] = q.y;
        chainZ[e] = q.z;
        return true;
    }
}

template <bool FromHang, bool ToHang>
__device__ __forceinline__ bool chainSegment(const WingView& v, int fe, int te, double rest, double alphaTilde,
                                             bool staticOk, double& lambda, int runtime, Payload& P, const d3* rel,
                                             double invPayloadMass, double minLength, const m3& iinv,
                                             bool iinvFinite) {
    m3 r{};
    bool endsOk = true;
    if constexpr (FromHang || ToHang) endsOk = rotationMatrixDeferred(P.q, r);
    const ChainEnd from = chainEnd<FromHang>(fe, r, P, rel, invPayloadMass);
    const ChainEnd to = chainEnd<ToHang>(te, r, P, rel, invPayloadMass);
    const d3 diff = sub(to.position, from.position);
    const double length = len(diff);
    const d3 dir = divide(diff, length);
    const d3 toGrad = neg(dir);
    const double gradLength = len(dir);
    double effFrom, effTo;
    const bool effOk = chainEffectiveMass<FromHang>(from, dir, iinv, iinvFinite, effFrom) &
                       chainEffectiveMass<ToHang>(to, toGrad, iinv, iinvFinite, effTo);
    const double eff = effFrom + effTo;
    const bool lineOk = isfinite(length) && !(length <= minLength) && (length > 0.0) && staticOk &&
                        (lambda >= 0.0) && isfinite(lambda) && !(fabs(gradLength - 1.0) > 1.0e-10) && effOk &&
                        (eff + alphaTilde > 0.0);
    const double gap = rest - length;
    const double candidate = lambda + (-gap - alphaTilde * lambda) / (eff + alphaTilde);
    const double next = stdmax(0.0, candidate);
    const double increment = next - lambda;
    if (!(endsOk && lineOk)) {
        if (!endsOk) raise(v, SW_SUSPENSION_ERROR, 1, 2, runtime);
        else raise(v, SW_SUSPENSION_ERROR, 6, 1, runtime);
        return false;
    }
    lambda = next;
    const bool fromOk = chainCorrect<FromHang>(fe, from, dir, increment, P, iinv);
    const bool toOk = chainCorrect<ToHang>(te, to, toGrad, increment, P, iinv);
    if (!(fromOk && toOk)) {
        raise(v, SW_SUSPENSION_ERROR, 5, 2, 0);
        return false;
    }
    return true;
}

// A serial segment run: consecutive single-item dependency levels, i.e.
// segments that share the payload or a node, in the reference's string-id
// order (suspension.cpp:2197-2274). The workgroup stages the run's nodes,
// per-segment values and the harness points in LDS; lane 0 then projects the
// segments one after another with the payload pose, its definition and the
// inverse inertia in registers, so the serial chain waits on no global
// memory; the workgroup finally writes nodes, multipliers and the payload
// back. Identical arithmetic to projecting the segments in place.
// Input vector of a segment projection: endpoint positions (zero for a
// harness endpoint), multiplier, rest length, payload com, q, inverse inertia.
__device__ __forceinline__ void recordSegment(double* r, const d3& pa, const d3& pb, double lambda, double rest,
                                              const Payload& state, const m3& iinv) {
    put3(r, pa);
    put3(r + 3, pb);
    r[6] = lambda;
    r[7] = rest;
    put3(r + 8, state.com);
    r[11] = state.q.w;
    r[12] = state.q.x;
    r[13] = state.q.y;
    r[14] = state.q.z;
    for (int c = 0; c < 9; ++c) r[15 + c] = iinv.a[c];
}

// `rec`: the iteration's segment records (recording only).
__device__ __forceinline__ void runChain(const WingView& v, const NodeRefs& N, const Phase P, PayloadDev& payload,
                                      const StepArgs& a, double* rec) {
    const int slots = P.slotEnd - P.slotBegin;
    const int items = P.end - P.begin;
    for (int s = threadIdx.x; s < slots; s += blockDim.x) {
        const int node = v.runNodes[P.slotBegin + s];
        const d3 x = N.get(node);
        chainX[s] = x.x;
        chainY[s] = x.y;
        chainZ[s] = x.z;
        chainW[s] = N.w[node];
    }
    for (int j = threadIdx.x; j < items; j += blockDim.x) {
        const int k = P.begin + j;
        const double rest = v.srest[k], stiffness = v.sstiff[k];
        chainLambda[j] = v.slam[k];
        chainRest[j] = rest;
        chainStiffness[j] = stiffness;
        chainAlpha[j] = 1.0 / (stiffness * a.dt * a.dt);
        chainStatic[j] = (rest > 0.0) && isfinite(rest) && (stiffness > 0.0) && isfinite(stiffness) && (a.dt > 0.0) &&
                         isfinite(a.dt);
        chainItem[4 * j] = v.sfromSlot[k];
        chainItem[4 * j + 1] = v.stoSlot[k];
        chainItem[4 * j + 2] = v.sruntime[k];
        chainItem[4 * j + 3] = v.slastPayload[k];
    }
    for (int h = threadIdx.x; h < 3 * v.hangCount; h += blockDim.x) chainHang[h] = v.hangLocal[h];
    for (int h = threadIdx.x; h < v.hangCount; h += blockDim.x)
        chainRel[h] = sub(d3{v.hangLocal[3 * h], v.hangLocal[3 * h + 1], v.hangLocal[3 * h + 2]},
                          d3{payload.comLocal[0], payload.comLocal[1], payload.comLocal[2]});
    __syncthreads();
    if (threadIdx.x == 0) {
        const SlotRefs S{};
        const PayloadDev def = payload;
        const double invPayloadMass = 1.0 / def.mass;
        const double minLength = def.minLineLength;
        Payload state = loadPayload(def);
        m3 iinv = loadInverseInertia(def);
        bool iinvFinite = finiteM(iinv);
        bool touched = false;
        for (int j = 0; j < items; ++j) {
            const int fe = chainItem[4 * j], te = chainItem[4 * j + 1];
            double lambda = chainLambda[j];
            if (rec)
                recordSegment(rec + kRecSeg * static_cast<long>(P.begin + j), fe >= 0 ? S.get(fe) : d3{0.0, 0.0, 0.0},
                              te >= 0 ? S.get(te) : d3{0.0, 0.0, 0.0}, lambda, chainRest[j], state, iinv);
            const double rest = chainRest[j], alpha = chainAlpha[j];
            const bool fixedOk = chainStatic[j] != 0;
            const int runtime = chainItem[4 * j + 2];
            bool ok;
            if (fe >= 0 && te >= 0)
                ok = chainSegment<false, false>(v, fe, te, rest, alpha, fixedOk, lambda, runtime, state, chainRel,
                                                invPayloadMass, minLength, iinv, iinvFinite);
            else if (fe >= 0)
                ok = chainSegment<false, true>(v, fe, te, rest, alpha, fixedOk, lambda, runtime, state, chainRel,
                                               invPayloadMass, minLength, iinv, iinvFinite);
            else if (te >= 0)
                ok = chainSegment<true, false>(v, fe, te, rest, alpha, fixedOk, lambda, runtime, state, chainRel,
                                               invPayloadMass, minLength, iinv, iinvFinite);
            else
                ok = chainSegment<true, true>(v, fe, te, rest, alpha, fixedOk, lambda, runtime, state, chainRel,
                                              invPayloadMass, minLength, iinv, iinvFinite);
            if (!ok) break;
            chainLambda[j] = lambda;
            if (fe >= 0 && te >= 0) continue;
            touched = true;
            if (chainItem[4 * j + 3]) {
                m3 world, inv;
                if (!worldInertia(state.q, def.inertia, world) || !inverseM(world, inv)) {
                    raise(v, SW_SUSPENSION_ERROR, 1, 2, 0);
                } else {
                    iinv = inv;
                    iinvFinite = finiteM(iinv);
                    for (int c = 0; c < 9; ++c) __hip_atomic_store(&payload.iinv[c], inv.a[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        if (touched) storePayload(payload, state);
    }
    __syncthreads();
    for (int s = threadIdx.x; s < slots; s += blockDim.x) {
        const int node = v.runNodes[P.slotBegin + s];
        N.put(node, d3{chainX[s], chainY[s], chainZ[s]});
    }
    for (int j = threadIdx.x; j < items; j += blockDim.x) v.slam[P.begin + j] = chainLambda[j];
}

// Waits for the progress words named by dependency entries [e0, e1) of
// (participant, phase) `index` at `step`, then acquires. Lanes of the first
// wave poll one entry each.
__device__ __forceinline__ void waitDeps(const WingView& v, int index, int step) {
    const int e0 = v.depOffsets[index], e1 = v.depOffsets[index + 1];
    if (e0 == e1) return;
    if (threadIdx.x < 64) {
        for (int e = e0 + static_cast<int>(threadIdx.x); e < e1; e += 64) {
            const int need = step - v.depLag[e] + 1;
            if (need <= 0) continue;
            unsigned* flag = v.flags + 32 * v.depPart[e];
            unsigned spins = 0;
            while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < static_cast<unsigned>(need)) {
                __builtin_amdgcn_s_sleep(1);
                if (syncTimedOut(v, ++spins)) break;
            }
        }
        acquireAfterWait(v);
    }
    __syncthreads();
}

// A participant's iteration program staged in dynamic LDS (after the
// reduction scratch) at the start of the point-to-point loop: the phase
// table, the participant's dependency offsets and its entries (progress
// word, lag), so the loop reads no phase descriptor and no dependency list
// from global memory. Layout for m phases and d entries: m Phase records,
// then m + 1 offsets, then d int2 entries, each part 8-byte aligned
// (stagedProgramBytes).
struct StagedProgram {
    Phase* phase;
    int* offset;
    int2* dep;
    int2* share;  // this participant's items [x, y) of every phase (forward loop only)
    // Acquire after each wait (buffer_inv: drops stale cached lines). Off when
    // every value read across workgroups is read with device-scope loads.
    bool acquire = true;
};

__host__ __device__ constexpr std::size_t alignTo8(std::size_t bytes) { return (bytes + 7) & ~std::size_t{7}; }
__host__ __device__ constexpr std::size_t stagedProgramBytes(int m, int deps) {
    return alignTo8(static_cast<std::size_t>(m) * sizeof(Phase)) + alignTo8(static_cast<std::size_t>(m + 1) * sizeof(int)) +
           static_cast<std::size_t>(m + deps) * sizeof(int2);
}
constexpr std::size_t kMaxStagedProgramBytes = 48 * 1024;

__device__ __forceinline__ StagedProgram stagedProgram(char* base, int m) {
    StagedProgram s;
    s.phase = reinterpret_cast<Phase*>(base);
    s.offset = reinterpret_cast<int*>(base + alignTo8(static_cast<std::size_t>(m) * sizeof(Phase)));
    s.share = reinterpret_cast<int2*>(base + alignTo8(static_cast<std::size_t>(m) * sizeof(Phase)) +
                                      alignTo8(static_cast<std::size_t>(m + 1) * sizeof(int)));
    s.dep = s.share + m;
    return s;
}

// waitDeps over the staged entries of phase `ph`.
__device__ __forceinline__ void waitStagedDeps(const WingView& v, const StagedProgram& S, int ph, int step) {
    const int e0 = S.offset[ph], e1 = S.offset[ph + 1];
    if (e0 == e1) return;
    if (threadIdx.x < 64) {
        for (int e = e0 + static_cast<int>(threadIdx.x); e < e1; e += 64) {
            const int2 dep = S.dep[e];
            const int need = step - dep.y + 1;
            if (need <= 0) continue;
            unsigned* flag = v.flags + 32 * dep.x;
            unsigned spins = 0;
            while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < static_cast<unsigned>(need)) {
                __builtin_amdgcn_s_sleep(1);
                if (syncTimedOut(v, ++spins)) break;
            }
        }
        if (S.acquire) acquireAfterWait(v);
    }
    __syncthreads();
}

// A phase of the point-to-point loop whose first item's fixed values are
// loaded before the dependency wait, so that load overlaps the wait and the
// projection after it waits only on node positions.
template <class Item, class Load, class Project>
__device__ __forceinline__ void prefetchedPhase(const WingView& v, const StagedProgram& S, int ph, int step,
                                                int first, int end, int stride, Load load, Project project) {
    Item item{};
    const bool have = first < end;
    if (have) item = load(first);
    waitStagedDeps(v, S, ph, step);
    if (have) project(first, item);
    for (int k = first + stride; k < end; k += stride) project(k, load(k));
}

// Publishes `value` on a progress word after releasing this workgroup's
// earlier writes (call after a workgroup barrier).
// Device-wide, general case: an agent-scope release (it writes the XCD's
// dirty L2 lines back, buffer_wbl2, so that other XCDs read them).
// One XCD (kLocal), or a device-wide substep (kCoherent): every wave drained
// its stores before that barrier (drainStores), so only the progress word
// itself is written. For kCoherent the only values other workgroups read
// during the iterations are node positions and the payload pose, and those
// are written with device-scope stores (NodeRefs::put, storePayload), which
// are complete at the device's coherence point once drained; the readers'
// acquire (acquireAfterWait) drops their stale cached lines. Each item's
// multipliers are read and written by one thread only. This saves the
// per-phase L2 write-back (2.1 ms of the export wing's substep, measured).
template <bool kLocal = false, bool kCoherent = false>
__device__ __forceinline__ void publishCounter(const WingView& v, unsigned* counter, unsigned value) {
    (void)v;
    if (threadIdx.x == 0) {
        if constexpr (kLocal || kCoherent) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        else __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(counter, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// A run of consecutive phases of one item type [ph0, ph1) of iteration `it`
// for participant p of `workers`: each thread's first item of the next phase
// is loaded (fixed values and own multiplier, as prefetchedPhase does) while
// this phase's writes drain and its progress is published. A participant
// with no item in a phase neither waits nor publishes for it.
template <bool kLocal, class Item, class Load, class Project>
__device__ __forceinline__ void pipelinedRun(const WingView& v, const StagedProgram& S, int it, int ph0, int ph1,
                                             int p, Load load, Project project) {
    const int m = v.iterPhaseCount;
    const int stride = static_cast<int>(blockDim.x);
    const auto share = [&](int ph, int& b, int& e) {
        const int2 x = S.share[ph];
        b = x.x, e = x.y;
    };
    int b, e;
    share(ph0, b, e);
    int first = b + static_cast<int>(threadIdx.x);
    Item cur{};
    if (first < e) cur = load(first);
    for (int ph = ph0; ph < ph1; ++ph) {
        const int step = it * m + ph;
        if (b < e) {
            waitStagedDeps(v, S, ph, step);
            if (first < e) project(first, cur);
            for (int k = first + stride; k < e; k += stride) project(k, load(k));
        }
        int nb = 0, ne = 0;
        if (ph + 1 < ph1) share(ph + 1, nb, ne);
        const int nfirst = nb + static_cast<int>(threadIdx.x);
        Item next{};
        if (nfirst < ne) next = load(nfirst);
        if (b < e) {
            drainStores<kLocal, true>(v);
            __syncthreads();
            publishCounter<kLocal, true>(v, v.flags + 32 * p, static_cast<unsigned>(step + 1));
        }
        b = nb, e = ne, first = nfirst, cur = next;
    }
}

// Contact sphere pair (build-defined): node held off the centre node to the
// radius; the centre is not moved by contact.
__device__ __forceinline__ void projectSphere(const WingView& v, const NodeRefs& N, int k, double invDt2,
                                              double* rec = nullptr) {
    const int node = v.spNode[k];
    const d3 c = N.get(v.spCentre[k]);
    d3 x = N.get(node);
    if (rec) {
        put3(rec, x);
        put3(rec + 3, c);
        rec[6] = v.splam[k];
    }
    const d3 diff = sub(x, c);
    const double length = len(diff);
    if (length <= 1.0e-12) return;
    const double gap = length - v.spRadius[k];
    const double lambda = v.splam[k];
    if (gap >= 0.0 && lambda == 0.0) return;
    const double wn = N.w[node];
    const double alpha = v.spComp[k] * invDt2;
    if (wn + alpha <= 0.0) return;
    const double dl = (-gap - alpha * lambda) / (wn + alpha);
    const double next = stdmax(0.0, lambda + dl);
    const double applied = next - lambda;
    v.splam[k] = next;
    const double s = applied / length;
    x = add(x, scale(diff, wn * s));
    N.put(node, x);
}

// Payload prediction of beginSubstep (suspension.cpp:2072-2182 with
// integratePayloadTorqueFree, suspension.cpp:1075-1102).
__device__ __forceinline__ bool beginPayload(PayloadDev& p, const StepArgs& a) {
    for (int k = 0; k < 3; ++k) p.prevCom[k] = p.com[k];
    for (int k = 0; k < 4; ++k) p.prevQ[k] = p.q[k];
    for (int k = 0; k < 3; ++k) p.prevV[k] = p.v[k];
    for (int k = 0; k < 3; ++k) p.prevW[k] = p.w[k];
    const d3 force{0.0 + p.mass * a.gravity[0], 0.0 + p.mass * a.gravity[1], 0.0 + p.mass * a.gravity[2]};
    const d3 moment{0.0, 0.0, 0.0};
    if (!(a.dt > 0.0) || !isfinite(a.dt) || !fin3(force)) return false;
    quat q;
    if (!canonical(quat{p.q[0], p.q[1], p.q[2], p.q[3]}, q)) return false;
    m3 inertia, inv;
    if (!worldInertia(q, p.inertia, inertia)) return false;
    d3 angular = mulM(inertia, d3{p.w[0], p.w[1], p.w[2]});
    d3 vel = add(d3{p.v[0], p.v[1], p.v[2]}, scale(force, a.dt / p.mass));
    angular = add(angular, scale(moment, a.dt));
    if (!inverseM(inertia, inv)) return false;
    d3 omega = mulM(inv, angular);
    d3 com = add(d3{p.com[0], p.com[1], p.com[2]}, scale(vel, a.dt));
    quat inc, rotated;
    if (!rotationIncrement(scale(omega, a.dt), inc) || !canonical(qmul(inc, q), rotated)) return false;
    m3 world2, inv2;
    if (!worldInertia(rotated, p.inertia, world2) || !inverseM(world2, inv2)) return false;
    omega = mulM(inv2, angular);
    if (!fin3(com) || !fin3(vel) || !fin3(omega) || !finQ(rotated)) return false;
    p.com[0] = com.x;
    p.com[1] = com.y;
    p.com[2] = com.z;
    p.q[0] = rotated.w;
    p.q[1] = rotated.x;
    p.q[2] = rotated.y;
    p.q[3] = rotated.z;
    p.v[0] = vel.x;
    p.v[1] = vel.y;
    p.v[2] = vel.z;
    p.w[0] = omega.x;
    p.w[1] = omega.y;
    p.w[2] = omega.z;
    return true;
}

// reconstructPayloadVelocity (suspension.cpp:2386-2411). False when the
// pose delta does not normalise.
__device__ __forceinline__ bool reconstructPayloadVelocity(PayloadDev& p, const StepArgs& a) {
    const d3 vel{(p.com[0] - p.prevCom[0]) / a.dt, (p.com[1] - p.prevCom[1]) / a.dt,
                 (p.com[2] - p.prevCom[2]) / a.dt};
    const quat q{p.q[0], p.q[1], p.q[2], p.q[3]};
    quat delta;
    if (!canonical(qmul(q, quat{p.prevQ[0], -p.prevQ[1], -p.prevQ[2], -p.prevQ[3]}), delta)) return false;
    const double half = pm::acos(stdclamp(delta.w, -1.0, 1.0));
    const double sine = pm::sin(half);
    d3 omega;
    if (fabs(sine) < 1.0e-12) {
        omega = scale(d3{delta.x, delta.y, delta.z}, 2.0 / a.dt);
    } else {
        const d3 axis{delta.x / sine, delta.y / sine, delta.z / sine};
        omega = scale(axis, 2.0 * half / a.dt);
    }
    p.v[0] = vel.x;
    p.v[1] = vel.y;
    p.v[2] = vel.z;
    p.w[0] = omega.x;
    p.w[1] = omega.y;
    p.w[2] = omega.z;
    return true;
}

// The payload part of certifySubstep (suspension.cpp:2324-2333), after the
// line damping has changed the payload's velocities.
__device__ bool payloadCertified(const PayloadDev& p) {
    const quat q{p.q[0], p.q[1], p.q[2], p.q[3]};
    const double n2 = qnorm2(q);
    return fin3(d3{p.com[0], p.com[1], p.com[2]}) && finQ(q) && fin3(d3{p.v[0], p.v[1], p.v[2]}) &&
           fin3(d3{p.w[0], p.w[1], p.w[2]}) && !(fabs(n2 - 1.0) > 1.0e-12);
}

// Lane j's value of `x` for a wave-uniform j.
__device__ __forceinline__ double laneValue(double x, int j) {
    const long long bits = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), j);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), j);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// measureMomentum (soft_body.cpp:118-149) over the dynamic nodes in index
// order. Run by one whole wave: the lanes load 64 nodes at a time and every
// lane adds them up in node order, so the sums are the serial ones.
struct MomentumDev {
    double mass;
    d3 centre, linear, angular;
};

template <class Store>
__device__ __forceinline__ MomentumDev measureMomentum(const WingView& v, const Store& N) {
    const int lane = static_cast<int>(threadIdx.x) & 63;
    MomentumDev r{0.0, d3{0.0, 0.0, 0.0}, d3{0.0, 0.0, 0.0}, d3{0.0, 0.0, 0.0}};
    const int n = v.nodeCount;
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane;
        const bool live = i < n;
        const double w = live ? N.w[i] : 0.0;
        const d3 x = live ? N.get(i) : d3{0.0, 0.0, 0.0};
        const d3 u = live ? d3{v.vx[i], v.vy[i], v.vz[i]} : d3{0.0, 0.0, 0.0};
        const int count = n - base < 64 ? n - base : 64;
        for (int j = 0; j < count; ++j) {
            const double wj = laneValue(w, j);
            if (wj <= 0.0) continue;
            const double m = 1.0 / wj;
            r.mass += m;
            r.centre = add(r.centre, scale(d3{laneValue(x.x, j), laneValue(x.y, j), laneValue(x.z, j)}, m));
            r.linear = add(r.linear, scale(d3{laneValue(u.x, j), laneValue(u.y, j), laneValue(u.z, j)}, m));
        }
    }
    if (!(r.mass > 0.0)) return r;
    r.centre = divide(r.centre, r.mass);
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane;
        const bool live = i < n;
        const double w = live ? N.w[i] : 0.0;
        const d3 x = live ? N.get(i) : d3{0.0, 0.0, 0.0};
        const d3 u = live ? d3{v.vx[i], v.vy[i], v.vz[i]} : d3{0.0, 0.0, 0.0};
        const int count = n - base < 64 ? n - base : 64;
        for (int j = 0; j < count; ++j) {
            const double wj = laneValue(w, j);
            if (wj <= 0.0) continue;
            const double m = 1.0 / wj;
            const d3 xj{laneValue(x.x, j), laneValue(x.y, j), laneValue(x.z, j)};
            const d3 uj{laneValue(u.x, j), laneValue(u.y, j), laneValue(u.z, j)};
            r.angular = add(r.angular, scale(cross(sub(xj, r.centre), uj), m));
        }
    }
    return r;
}

// restoreMomentum (soft_body.cpp:151-190): a uniform velocity shift that
// restores the linear momentum, then a rigid rotation about the centre of
// mass that restores the angular momentum. One whole wave; false when the
// inertia solve fails (the reference throws from checkedSolve).
template <class Store>
__device__ __forceinline__ bool restoreMomentum(const WingView& v, const Store& N, const MomentumDev& target) {
    if (!(target.mass > 0.0)) return true;
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const int n = v.nodeCount;
    MomentumDev current = measureMomentum(v, N);
    const d3 translation = divide(sub(target.linear, current.linear), target.mass);
    for (int i = lane; i < n; i += 64) {
        if (!(N.w[i] > 0.0)) continue;
        v.vx[i] = v.vx[i] + translation.x;
        v.vy[i] = v.vy[i] + translation.y;
        v.vz[i] = v.vz[i] + translation.z;
    }
    current = measureMomentum(v, N);
    s3 inertia{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane;
        const bool live = i < n;
        const double w = live ? N.w[i] : 0.0;
        const d3 x = live ? N.get(i) : d3{0.0, 0.0, 0.0};
        const int count = n - base < 64 ? n - base : 64;
        for (int j = 0; j < count; ++j) {
            const double wj = laneValue(w, j);
            if (wj <= 0.0) continue;
            const double m = 1.0 / wj;
            const d3 r = sub(d3{laneValue(x.x, j), laneValue(x.y, j), laneValue(x.z, j)}, current.centre);
            inertia.xx += m * (r.y * r.y + r.z * r.z);
            inertia.yy += m * (r.x * r.x + r.z * r.z);
            inertia.zz += m * (r.x * r.x + r.y * r.y);
            inertia.xy -= m * r.x * r.y;
            inertia.xz -= m * r.x * r.z;
            inertia.yz -= m * r.y * r.z;
        }
    }
    if (!positiveDefinite(inertia)) return true;
    d3 omega;
    if (!checkedSolve(inertia, sub(target.angular, current.angular), omega)) return false;
    for (int i = lane; i < n; i += 64) {
        if (!(N.w[i] > 0.0)) continue;
        const d3 spin = cross(omega, sub(N.get(i), current.centre));
        v.vx[i] = v.vx[i] + spin.x;
        v.vy[i] = v.vy[i] + spin.y;
        v.vz[i] = v.vz[i] + spin.z;
    }
    return true;
}

// boundedAxialDampingImpulse (suspension.cpp:1165-1185). False on an
// invalid input (the reference throws).
__device__ __forceinline__ bool boundedAxialDampingImpulse(double relative, double damping, double dt,
                                                           double effective, double& out) {
    if (!isfinite(relative) || !(damping >= 0.0) || !isfinite(damping) || !(dt > 0.0) || !isfinite(dt) ||
        !(effective >= 0.0) || !isfinite(effective))
        return false;
    out = 0.0;
    if (damping == 0.0 || effective == 0.0 || relative == 0.0) return true;
    const double requested = damping * fabs(relative) * dt;
    const double stopping = fabs(relative) / effective;
    const double bounded = stdmin(requested, stopping);
    out = relative < 0.0 ? -bounded : bounded;
    return true;
}

// applyLineDamping for a free payload (suspension.cpp:2413-2479): an axial
// impulse along every taut, damped segment, in runtime (string-id) order,
// each seeing the velocities the earlier ones left. One lane; false after
// raising a failure.
template <class Store>
__device__ __forceinline__ bool dampLines(const WingView& v, const Store& N, PayloadDev& payload, const StepArgs& a) {
    bool any = false;
    for (int j = 0; j < v.dampCount; ++j) any = any || v.slam[v.dampSlot[j]] > 0.0;
    if (!any) return true;
    const Payload state = loadPayload(payload);
    m3 world, inverseInertia;
    if (!worldInertia(state.q, payload.inertia, world) || !inverseM(world, inverseInertia)) {
        raise(v, SW_SUSPENSION_ERROR, 6, 2, 0);
        return false;
    }
    m3 rot;
    if (!rotationMatrix(state.q, rot)) {
        raise(v, SW_SUSPENSION_ERROR, 6, 2, 0);
        return false;
    }
    const d3 comLocal{payload.comLocal[0], payload.comLocal[1], payload.comLocal[2]};
    d3 pv{payload.v[0], payload.v[1], payload.v[2]};
    d3 pw{payload.w[0], payload.w[1], payload.w[2]};
    const double mass = payload.mass;
    for (int j = 0; j < v.dampCount; ++j) {
        const int k = v.dampSlot[j];
        if (!(v.slam[k] > 0.0)) continue;
        const int ends[2] = {v.sfrom[k], v.sto[k]};
        d3 point[2], velocity[2];
        double effective = 0.0;
        Endpoint e[2];
        for (int s = 0; s < 2; ++s) {
            if (!endpointState(v, N, ends[s], state, payload, e[s])) {
                raise(v, SW_SUSPENSION_ERROR, 6, 2, 0);
                return false;
            }
            point[s] = e[s].position;
            if (ends[s] >= 0) {
                velocity[s] = d3{v.vx[ends[s]], v.vy[ends[s]], v.vz[ends[s]]};
            } else {
                const d3 offset = mulM(rot, sub(N.hang(-ends[s] - 1), comLocal));
                velocity[s] = add(pv, cross(pw, offset));
            }
        }
        const d3 d = sub(point[1], point[0]);
        const double magnitude = len(d);
        const d3 direction = magnitude > 0.0 ? divide(d, magnitude) : d3{0.0, 0.0, 0.0};
        const double relative = dot(sub(velocity[1], velocity[0]), direction);
        for (int s = 0; s < 2; ++s) {
            if (ends[s] >= 0) {
                effective += N.w[ends[s]];
            } else {
                const d3 rxn = cross(sub(point[s], state.com), direction);
                effective += 1.0 / mass + dot(rxn, mulM(inverseInertia, rxn));
            }
        }
        if (!(effective > 0.0) || relative == 0.0) continue;
        double impulse;
        if (!boundedAxialDampingImpulse(relative, v.dampValue[j], a.dt, effective, impulse)) {
            raise(v, SW_SUSPENSION_ERROR, 6, 1, v.sruntime[k]);
            return false;
        }
        const d3 fromImpulse = scale(direction, impulse);
        for (int s = 0; s < 2; ++s) {
            const d3 J = s == 0 ? fromImpulse : neg(fromImpulse);
            if (ends[s] >= 0) {
                const double w = N.w[ends[s]];
                if (w > 0.0) {
                    const d3 dv = scale(J, w);
                    v.vx[ends[s]] = v.vx[ends[s]] + dv.x;
                    v.vy[ends[s]] = v.vy[ends[s]] + dv.y;
                    v.vz[ends[s]] = v.vz[ends[s]] + dv.z;
                }
            } else {
                pv = add(pv, divide(J, mass));
                pw = add(pw, mulM(inverseInertia, cross(sub(point[s], state.com), J)));
            }
        }
    }
    payload.v[0] = pv.x;
    payload.v[1] = pv.y;
    payload.v[2] = pv.z;
    payload.w[0] = pw.x;
    payload.w[1] = pw.y;
    payload.w[2] = pw.z;
    return true;
}

// Wavefront reductions over the 64 lanes of a wave: a butterfly of
// cross-lane exchanges (__shfl_xor, lowered to DPP / ds_swizzle), no LDS.
// Every lane ends with the same value (a + b == b + a exactly), and the
// fixed exchange pattern makes sums reproducible run to run.
__device__ __forceinline__ double waveSum(double x) {
    for (int offset = 32; offset > 0; offset >>= 1) x += __shfl_xor(x, offset, 64);
    return x;
}
__device__ __forceinline__ double waveMax(double x) {
    for (int offset = 32; offset > 0; offset >>= 1) x = stdmax(x, __shfl_xor(x, offset, 64));
    return x;
}
__device__ __forceinline__ int waveSumInt(int x) {
    for (int offset = 32; offset > 0; offset >>= 1) x += __shfl_xor(x, offset, 64);
    return x;
}

__device__ __forceinline__ void atomicMaxNonNegative(double* address, double value) {
    // Non-negative doubles order like their bit patterns.
    atomicMax(reinterpret_cast<unsigned long long*>(address), static_cast<unsigned long long>(__double_as_longlong(value)));
}

// Items [first, end) of a parallel level of suspension segments.
__device__ __forceinline__ void segmentLevel(const WingView& v, const NodeRefs& N, int first, int end, int stride,
                                             PayloadDev& payload, const StepArgs& a, double* rec,
                                             const TraceLayout& L) {
    for (int k = first; k < end; k += stride) {
        Payload state = loadPayload(payload);
        const m3 iinv = loadInverseInertia(payload);
        const int fe = v.sfrom[k], te = v.sto[k];
        double lambda = v.slam[k];
        if (rec)
            recordSegment(rec + L.segBase + kRecSeg * static_cast<long>(k), fe >= 0 ? N.get(fe) : d3{0.0, 0.0, 0.0},
                          te >= 0 ? N.get(te) : d3{0.0, 0.0, 0.0}, lambda, v.srest[k], state, iinv);
        if (!projectSegment(v, N, fe, te, v.srest[k], v.sstiff[k], lambda, v.sruntime[k], state, payload, iinv,
                            finiteM(iinv), a))
            break;
        v.slam[k] = lambda;
        if (fe < 0 || te < 0) {
            storePayload(payload, state);
            if (v.slastPayload[k] && !refreshInverseInertia(payload, state.q))
                raise(v, SW_SUSPENSION_ERROR, 1, 2, 0);
        }
    }
}

// Items [first, end) with the given stride of one phase of the iteration
// program; a serial run (kPhaseSegmentRun) is staged and projected by the
// calling workgroup as a whole. `rec`: the iteration's trace records
// (recording only).
// Interleaved load-path cable pairs of iteration `it`
// (soft_body.cpp:1308-1315): interleaved / iterations, one more for the
// first interleaved % iterations iterations.
__device__ __forceinline__ int cablePairsOf(const StepArgs& a, int it) {
    return a.cableBase + (it < a.cableRem ? 1 : 0);
}

// kGeneral: compiled with the list phases (serial sweep, interleaved cable
// pairs) and the compressionStiffnessRatio branch; the host launches the
// general kernel variant whenever a wing's programs or definition use them.
template <bool kGeneral>
__device__ __forceinline__ void runPhase(const WingView& v, const NodeRefs& N, const Phase& P, int first, int end, int stride,
                         PayloadDev& payload, const StepArgs& a, double* rec, const TraceLayout& L) {
    switch (P.type) {
    case kPhaseDistance:
        for (int k = first; k < end; k += stride)
            projectDistance(v, N, k, a.invDt2, rec ? rec + L.distBase + kRecDist * static_cast<long>(k) : nullptr);
        break;
    case kPhaseMembrane:
        for (int k = first; k < end; k += stride)
            projectMembrane<kGeneral>(v, N, k, a, rec ? rec + L.membBase + kRecMemb * static_cast<long>(k) : nullptr);
        break;
    case kPhaseHinge:
        for (int k = first; k < end; k += stride)
            projectHinge(v, N, k, a.invDt2, rec ? rec + L.hingeBase + kRecHinge * static_cast<long>(k) : nullptr);
        break;
    case kPhaseSegments:
        segmentLevel(v, N, first, end, stride, payload, a, rec, L);
        break;
    case kPhaseSegmentRun:
        runChain(v, N, P, payload, a, rec ? rec + L.segBase : nullptr);
        break;
    case kPhaseSpheres:
        for (int k = first; k < end; k += stride)
            projectSphere(v, N, k, a.invDt2, rec ? rec + L.sphereBase + kRecSphere * static_cast<long>(k) : nullptr);
        break;
    case kPhaseDistanceList:
        if constexpr (kGeneral) {
            for (int j = first; j < end; j += stride) {
                const int k = v.distList[j];
                projectDistance(v, N, k, a.invDt2, rec ? rec + L.distBase + kRecDist * static_cast<long>(k) : nullptr);
            }
        }
        break;
    case kPhaseMembraneList:
        if constexpr (kGeneral) {
            for (int j = first; j < end; j += stride) {
                const int k = v.membList[j];
                projectMembrane<true>(v, N, k, a, rec ? rec + L.membBase + kRecMemb * static_cast<long>(k) : nullptr);
            }
        }
        break;
    case kPhaseCableList:  // the caller skips pairs beyond the iteration's share; never recorded
        if constexpr (kGeneral)
            for (int j = first; j < end; j += stride) projectDistance(v, N, v.cableList[j], a.invDt2, nullptr);
        break;
    default:
        break;
    }
}

// One substep of one wing, run by `blocks` workgroups of which this is
// `block`. kMode: 0 nodes in HBM, 1 nodes resident in LDS, 2 either, as the
// wing's view says (batched launches mix both kinds of wing).
// kLocal: a one-XCD launch; kGeneral: compiled with the list phases, the
// compressionStiffnessRatio branch, momentum restoration and line damping.
template <int kMode, bool kLocal, bool kGeneral>
// `trace`: the adjoint's recording of this substep, or null.
__device__ __forceinline__ void substepBody(const WingView& v, const StepArgs& a, int block, int blocks, double* trace) {
    extern __shared__ double lds[];
    // A pending failure stops later substeps of a queued batch. Only a
    // single-workgroup wing may leave early: with several workgroups the
    // error word can change between their starts, and a workgroup that left
    // would strand the others at the grid barrier.
    if (blocks == 1 && __hip_atomic_load(&v.error->code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    const int n = v.nodeCount;
    const bool shared = kMode == 1 || (kMode == 2 && v.sharedNodes != 0);
    unsigned long long since = 0;
    profileStart(v, since);
    for (int i = threadIdx.x; i < 65; i += blockDim.x) {
        atanLds[i][0] = pm::kAtanTable[i][0];
        atanLds[i][1] = pm::kAtanTable[i][1];
    }
    __syncthreads();
    const bool dedicated = !shared && v.neighbourSync != 0 && v.dedicatedChain != 0 && blocks > 1 && v.hasRun;
    const int workers = dedicated ? blocks - 1 : blocks;
    // The chain workgroup takes part in the grid barriers around the
    // iterations but owns no item of the other passes.
    const bool chainOnly = dedicated && block == workers;
    const int gtid = chainOnly ? 0x3fffffff : static_cast<int>(block * blockDim.x + threadIdx.x);
    const int gstride = workers * blockDim.x;
    NodeRefs N;
    if (shared) {
        N = NodeRefs{lds, lds + n, lds + 2 * n, lds + 3 * n, v.hangLocal};
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            N.x[i] = v.px[i];
            N.y[i] = v.py[i];
            N.z[i] = v.pz[i];
            N.w[i] = v.w[i];
        }
        __syncthreads();
    } else {
        N = NodeRefs{v.px, v.py, v.pz, v.w, v.hangLocal};
        N.coherent = !kLocal;  // device-wide publication without L2 write-back
    }
    PayloadDev& payload = *v.payload;
    const TraceLayout layout = traceLayout(v, a.iterations);
    if (trace) {
        // Step-start state. Phase A writes no position and no velocity, and
        // only workgroup 0's thread 0 touches the payload.
        for (int i = gtid; i < n; i += gstride) {
            trace[3L * i] = v.px[i];
            trace[3L * i + 1] = v.py[i];
            trace[3L * i + 2] = v.pz[i];
            trace[layout.nodes3 + 3L * i] = v.vx[i];
            trace[layout.nodes3 + 3L * i + 1] = v.vy[i];
            trace[layout.nodes3 + 3L * i + 2] = v.vz[i];
        }
        if (gtid == 0) {
            double* t = trace + 2 * layout.nodes3;
            for (int k = 0; k < 3; ++k) t[k] = payload.com[k];
            for (int k = 0; k < 4; ++k) t[3 + k] = payload.q[k];
            for (int k = 0; k < 3; ++k) t[7 + k] = payload.v[k];
            for (int k = 0; k < 3; ++k) t[10 + k] = payload.w[k];
        }
    }

    // A: prev = x, multiplier resets, payload prediction.
    for (int i = gtid; i < n; i += gstride) {
        v.qx[i] = N.x[i];
        v.qy[i] = N.y[i];
        v.qz[i] = N.z[i];
    }
    for (int k = gtid; k < v.distCount; k += gstride) v.dlam[k] = 0.0;
    for (int k = gtid; k < v.membCount; k += gstride) {
        v.mlx[k] = 0.0;
        v.mly[k] = 0.0;
        v.mlz[k] = 0.0;
        v.mrx[k] = 0.0;
        v.mry[k] = 0.0;
        v.mrz[k] = 0.0;
        v.mres[k] = 0.0;
    }
    for (int k = gtid; k < v.hingeCount; k += gstride) v.hlam[k] = 0.0;
    for (int k = gtid; k < v.segCount; k += gstride) v.slam[k] = 0.0;
    for (int k = gtid; k < v.pairCount; k += gstride) v.splam[k] = 0.0;
    if (gtid == 0 && v.segCount > 0) {
        payload.maxResidual = 0.0;
        if (!beginPayload(payload, a) ||
            !refreshInverseInertia(payload, quat{payload.q[0], payload.q[1], payload.q[2], payload.q[3]}))
            raise(v, SW_SUSPENSION_ERROR, 5, 2, 0);
    }
    gridSync<kLocal>(v, blocks);

    // B: pressure gather from the start-of-substep positions, then predict
    // (soft_body.cpp:1492-1566).
    for (int i = gtid; i < n; i += gstride) {
        d3 force{v.fx[i], v.fy[i], v.fz[i]};
        for (int s = v.incOffsets[i]; s < v.incOffsets[i + 1]; ++s) {
            const int t = v.incTriangles[s];
            const int ia = v.ta[t], ib = v.tb[t], ic = v.tc[t];
            const d3 pa{v.qx[ia], v.qy[ia], v.qz[ia]};
            const d3 pb{v.qx[ib], v.qy[ib], v.qz[ib]};
            const d3 pc{v.qx[ic], v.qy[ic], v.qz[ic]};
            const d3 areaVector = scale(cross(sub(pb, pa), sub(pc, pa)), 0.5);
            const double p = v.pressure[t];
            force = add(force, divide(scale(areaVector, p), 3.0));
        }
        const double wi = N.w[i];
        if (wi == 0.0) {
            v.vx[i] = 0.0;
            v.vy[i] = 0.0;
            v.vz[i] = 0.0;
            continue;
        }
        const d3 acc = add(d3{a.gravity[0], a.gravity[1], a.gravity[2]}, scale(force, wi));
        const d3 ref{a.reference[0], a.reference[1], a.reference[2]};
        const d3 vel = add(ref, scale(sub(add(d3{v.vx[i], v.vy[i], v.vz[i]}, scale(acc, a.dt)), ref), a.damping));
        v.vx[i] = vel.x;
        v.vy[i] = vel.y;
        v.vz[i] = vel.z;
        N.put(i, add(N.get(i), scale(vel, a.dt)));
    }
    gridSync<kLocal>(v, blocks);
    if (kGeneral && v.preserveMomentum) {
        // The free membrane's momentum after prediction (soft_body.cpp:1196-1205),
        // restored after finalisation.
        if (block == 0 && threadIdx.x < 64) {
            const MomentumDev m = measureMomentum(v, N);
            if (threadIdx.x == 0) {
                v.momentum[0] = m.mass;
                v.momentum[1] = m.centre.x;
                v.momentum[2] = m.centre.y;
                v.momentum[3] = m.centre.z;
                v.momentum[4] = m.linear.x;
                v.momentum[5] = m.linear.y;
                v.momentum[6] = m.linear.z;
                v.momentum[7] = m.angular.x;
                v.momentum[8] = m.angular.y;
                v.momentum[9] = m.angular.z;
            }
        }
        gridSync<kLocal>(v, blocks);
    }

    profileAdd(v, kProfilePrediction, since);

    // Constraint iterations.
    if (v.neighbourSync && !shared && blocks > 1) {
        // Point-to-point synchronisation: workgroup p owns a contiguous share
        // of every phase's items (the dedicated chain workgroup owns the
        // serial runs). Before a phase it waits only for the workgroups whose
        // earlier accesses to its nodes the in-order program puts first
        // (host-computed), and after it publishes its progress.
        const int m = v.iterPhaseCount;
        const int p = block;
        const bool chainGroup = chainOnly;
        // The host sizes the dynamic LDS for this participant's program
        // (DeviceWing::buildSync); it follows the reduction scratch.
        StagedProgram S = stagedProgram(reinterpret_cast<char*>(lds + kMaxThreads), m);
        S.acquire = !N.coherent;
        const int depBase = v.depOffsets[p * m];
        const int depCount = v.depOffsets[p * m + m] - depBase;
        for (int k = threadIdx.x; k < m; k += blockDim.x) {
            const Phase Q = v.iterPhases[k];
            S.phase[k] = Q;
            int b = Q.begin, e = Q.begin;
            if (Q.type == kPhaseSegmentRun) {
                if (p == (dedicated ? workers : 0)) e = Q.end;
            } else {
                const long count = Q.end - Q.begin;
                b = Q.begin + static_cast<int>(count * p / workers);
                e = Q.begin + static_cast<int>(count * (p + 1) / workers);
            }
            S.share[k] = int2{b, e};
        }
        for (int k = threadIdx.x; k <= m; k += blockDim.x) S.offset[k] = v.depOffsets[p * m + k] - depBase;
        for (int e = threadIdx.x; e < depCount; e += blockDim.x)
            S.dep[e] = int2{v.depPart[depBase + e], v.depLag[depBase + e]};
        __syncthreads();
        if (chainGroup) {
            // The dedicated chain workgroup owns only the serial runs: a loop
            // of its own, away from the code of the other phases.
            for (int it = 0; it < a.iterations; ++it) {
                double* rec = trace ? trace + layout.iterBase + it * layout.iterStride : nullptr;
                for (int ph = 0; ph < m; ++ph) {
                    const Phase P = S.phase[ph];
                    if (P.type != kPhaseSegmentRun) continue;
                    const int step = it * m + ph;
                    waitStagedDeps(v, S, ph, step);
                    runChain(v, N, P, payload, a, rec ? rec + layout.segBase : nullptr);
                    drainStores<kLocal, true>(v);
                    __syncthreads();
                    publishCounter<kLocal, true>(v, v.flags + 32 * p, static_cast<unsigned>(step + 1));
                    profileAdd(v, kProfileSuspension, since);
                }
            }
        }
        for (int it = 0; it < (chainGroup ? 0 : a.iterations); ++it) {
            double* const irec = trace ? trace + layout.iterBase + it * layout.iterStride : nullptr;
            for (int ph = 0; ph < m; ++ph) {
                const Phase P = S.phase[ph];
                if (P.type == kPhaseDistance || P.type == kPhaseMembrane || P.type == kPhaseHinge) {
                    int end = ph + 1;
                    while (end < m && S.phase[end].type == P.type) ++end;
                    if (P.type == kPhaseDistance) {
                        pipelinedRun<kLocal, DistanceItem>(
                            v, S, it, ph, end, p, [&](int k) { return distanceItem(v, N, k); },
                            [&](int k, const DistanceItem& item) {
                                projectDistance(v, N, k, item, a.invDt2,
                                                irec ? irec + layout.distBase + kRecDist * static_cast<long>(k)
                                                     : nullptr);
                            });
                    } else if (P.type == kPhaseMembrane) {
                        pipelinedRun<kLocal, MembraneItem>(
                            v, S, it, ph, end, p, [&](int k) { return membraneItem(v, N, k); },
                            [&](int k, const MembraneItem& item) {
                                projectMembrane<kGeneral>(v, N, k, item, a,
                                                irec ? irec + layout.membBase + kRecMemb * static_cast<long>(k)
                                                     : nullptr);
                            });
                    } else {
                        pipelinedRun<kLocal, HingeItem>(
                            v, S, it, ph, end, p, [&](int k) { return hingeItem(v, N, k); },
                            [&](int k, const HingeItem& item) {
                                projectHinge(v, N, k, item, a.invDt2,
                                             irec ? irec + layout.hingeBase + kRecHinge * static_cast<long>(k)
                                                  : nullptr);
                            });
                    }
                    ph = end - 1;
                    profileAdd(v, profilePart(P.type), since);
                    continue;
                }
                const int b = S.share[ph].x, e = S.share[ph].y;
                if (b >= e) continue;
                const int step = it * m + ph;
                double* rec = trace ? trace + layout.iterBase + it * layout.iterStride : nullptr;
                const int first = b + static_cast<int>(threadIdx.x), stride = static_cast<int>(blockDim.x);
                switch (P.type) {
                case kPhaseDistance:
                    prefetchedPhase<DistanceItem>(
                        v, S, ph, step, first, e, stride, [&](int k) { return distanceItem(v, N, k); },
                        [&](int k, const DistanceItem& item) {
                            projectDistance(v, N, k, item, a.invDt2,
                                            rec ? rec + layout.distBase + kRecDist * static_cast<long>(k) : nullptr);
                        });
                    break;
                case kPhaseMembrane:
                    prefetchedPhase<MembraneItem>(
                        v, S, ph, step, first, e, stride, [&](int k) { return membraneItem(v, N, k); },
                        [&](int k, const MembraneItem& item) {
                            projectMembrane<kGeneral>(v, N, k, item, a,
                                            rec ? rec + layout.membBase + kRecMemb * static_cast<long>(k) : nullptr);
                        });
                    break;
                case kPhaseHinge:
                    prefetchedPhase<HingeItem>(
                        v, S, ph, step, first, e, stride, [&](int k) { return hingeItem(v, N, k); },
                        [&](int k, const HingeItem& item) {
                            projectHinge(v, N, k, item, a.invDt2,
                                         rec ? rec + layout.hingeBase + kRecHinge * static_cast<long>(k) : nullptr);
                        });
                    break;
                case kPhaseSegments:
                    waitStagedDeps(v, S, ph, step);
                    segmentLevel(v, N, first, e, stride, payload, a, rec, layout);
                    break;
                case kPhaseSegmentRun:  // no dedicated chain workgroup
                    waitStagedDeps(v, S, ph, step);
                    runChain(v, N, P, payload, a, rec ? rec + layout.segBase : nullptr);
                    break;
                case kPhaseSpheres:
                    waitStagedDeps(v, S, ph, step);
                    for (int k = first; k < e; k += stride)
                        projectSphere(v, N, k, a.invDt2,
                                      rec ? rec + layout.sphereBase + kRecSphere * static_cast<long>(k) : nullptr);
                    break;
                case kPhaseDistanceList:
                case kPhaseMembraneList:
                    waitStagedDeps(v, S, ph, step);
                    if constexpr (kGeneral) runPhase<true>(v, N, P, first, e, stride, payload, a, rec, layout);
                    break;
                case kPhaseCableList:
                    waitStagedDeps(v, S, ph, step);
                    if constexpr (kGeneral)
                        if (P.flags < cablePairsOf(a, it)) runPhase<true>(v, N, P, first, e, stride, payload, a, rec, layout);
                    break;
                default:
                    waitStagedDeps(v, S, ph, step);
                    break;
                }
                drainStores<kLocal, true>(v);
                __syncthreads();
                publishCounter<kLocal, true>(v, v.flags + 32 * p, static_cast<unsigned>(step + 1));
                profileAdd(v, profilePart(P.type), since);
            }
        }
        gridSync<kLocal>(v, blocks);
        // Every participant has left the loop: nobody polls the progress
        // words any more, so each resets its own for the next launch.
        if (threadIdx.x == 0) __hip_atomic_store(v.flags + 32 * p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (chainGroup) {
            profileAdd(v, kProfileSuspension, since);
            profileFlush(v);
            return;
        }
    } else {
        for (int it = 0; it < a.iterations; ++it) {
            double* rec = trace ? trace + layout.iterBase + it * layout.iterStride : nullptr;
            for (int ph = 0; ph < v.iterPhaseCount; ++ph) {
                const Phase P = v.iterPhases[ph];
                if (P.type == kPhaseSegmentRun) {
                    if (block == 0) runPhase<kGeneral>(v, N, P, P.begin, P.end, 1, payload, a, rec, layout);
                } else if (!kGeneral || P.type != kPhaseCableList || P.flags < cablePairsOf(a, it)) {
                    runPhase<kGeneral>(v, N, P, P.begin + gtid, P.end, gstride, payload, a, rec, layout);
                }
                gridSync<kLocal>(v, workers);
                profileAdd(v, profilePart(P.type), since);
            }
        }
    }
    // Closing seam-stitch pass (soft_body.cpp:1389-1405).
    if (trace && gtid == 0) {
        double* t = trace + layout.total - 7;
        for (int k = 0; k < 3; ++k) t[k] = payload.com[k];
        for (int k = 0; k < 4; ++k) t[3 + k] = payload.q[k];
    }
    for (int ph = 0; ph < v.closePhaseCount; ++ph) {
        const Phase P = v.closePhases[ph];
        for (int k = P.begin + gtid; k < P.end; k += gstride)
            projectDistance(v, N, v.stitchItems[k], a.invDt2,
                            trace ? trace + layout.closeBase + kRecDist * static_cast<long>(k) : nullptr);
        gridSync<kLocal>(v, workers);
    }
    // The wait for the last workgroups of the iterations and the closing
    // stitches count as distance-constraint time (soft_body.cpp:1389-1408).
    profileAdd(v, kProfileDistance, since);

    // Membrane solver diagnostics (soft_body.cpp:2366-2404), velocity
    // finalisation (soft_body.cpp:1568-1585) and line residuals
    // (suspension.cpp:2334-2349).
    for (int k = gtid; k < v.membCount; k += gstride) {
        const int i0 = v.ma[k], i1 = v.mb[k], i2 = v.mc[k];
        const d3 p[3] = {N.get(i0), N.get(i1), N.get(i2)};
        const double preW = v.anyPreTension ? v.preW[k] : 0.0;
        const double preT = v.anyPreTension ? v.preT[k] : 0.0;
        const MembraneEval e = evaluateMembrane(p[0], p[1], p[2], v.m00[k], v.m01[k], v.m10[k], v.m11[k], preW, preT);
        const double areaScale = v.mscale[k];
        const s3 at{a.invDt2 * v.cxx[k], a.invDt2 * v.cyy[k], a.invDt2 * v.czz[k],
                    a.invDt2 * v.cxy[k], a.invDt2 * v.cxz[k], a.invDt2 * v.cyz[k]};
        const double gamma = (v.anyDamping ? v.mdamp[k] : 0.0) / a.dt;
        const int idx[3] = {i0, i1, i2};
        d3 dm{0.0, 0.0, 0.0};
        for (int corner = 0; corner < 3; ++corner) {
            const d3 disp = sub(p[corner], d3{v.qx[idx[corner]], v.qy[idx[corner]], v.qz[idx[corner]]});
            dm.x += areaScale * dot(e.grad[0][corner], disp);
            dm.y += areaScale * dot(e.grad[1][corner], disp);
            dm.z += areaScale * dot(e.grad[2][corner], disp);
        }
        const d3 lambda{v.mlx[k], v.mly[k], v.mlz[k]};
        const d3 residual = add(add(scale(e.strain, areaScale), mul(at, lambda)), scale(dm, gamma));
        v.mres[k] = len(residual) / areaScale;
        const double f = a.invDt2 / areaScale;
        v.mrx[k] = -lambda.x * f;
        v.mry[k] = -lambda.y * f;
        v.mrz[k] = -lambda.z * f;
    }
    profileAdd(v, kProfileDiagnostics, since);
    for (int i = gtid; i < n; i += gstride) {
        const double wi = N.w[i];
        if (wi == 0.0) {
            v.vx[i] = 0.0;
            v.vy[i] = 0.0;
            v.vz[i] = 0.0;
        } else {
            v.vx[i] = (N.x[i] - v.qx[i]) * a.invDt;
            v.vy[i] = (N.y[i] - v.qy[i]) * a.invDt;
            v.vz[i] = (N.z[i] - v.qz[i]) * a.invDt;
        }
    }
    if (v.segCount > 0) {
        double worst = 0.0;
        const Payload state = loadPayload(payload);
        for (int k = gtid; k < v.segCount; k += gstride) {
            Endpoint from, to;
            if (!endpointState(v, N, v.sfrom[k], state, payload, from) ||
                !endpointState(v, N, v.sto[k], state, payload, to)) {
                raise(v, SW_SUSPENSION_ERROR, 8, 2, 0);
                continue;
            }
            const double length = len(sub(to.position, from.position));
            if (!isfinite(length)) {
                raise(v, SW_SUSPENSION_ERROR, 8, 1, v.sruntime[k]);
                continue;
            }
            const double stretch = stdmax(0.0, length - v.srest[k]);
            const double compliant = v.slam[k] / (v.sstiff[k] * a.dt * a.dt);
            worst = stdmax(worst, fabs(stretch - compliant));
        }
        // Wave maxima by lane exchanges, then one atomic per wave (the max
        // is order-free, so the result does not depend on arrival order).
        worst = waveMax(worst);
        if ((threadIdx.x & 63) == 0) atomicMaxNonNegative(&payload.maxResidual, worst);
    }
    gridSync<kLocal>(v, workers);
    if (kGeneral && v.preserveMomentum && block == 0 && threadIdx.x < 64) {
        // soft_body.cpp:1446-1448
        const MomentumDev target{v.momentum[0], d3{v.momentum[1], v.momentum[2], v.momentum[3]},
                                 d3{v.momentum[4], v.momentum[5], v.momentum[6]},
                                 d3{v.momentum[7], v.momentum[8], v.momentum[9]}};
        if (!restoreMomentum(v, N, target) && threadIdx.x == 0) raise(v, SW_INVALID_ARGUMENT, 9, 2, 0);
    }
    if (gtid == 0 && v.segCount > 0) {
        // Payload velocity, line damping, then the payload and line-residual
        // certification (suspension.cpp:2324-2411, 2413-2479).
        if (!reconstructPayloadVelocity(payload, a)) raise(v, SW_SUSPENSION_ERROR, 8, 2, 0);
        else if (kGeneral && v.dampCount > 0 && !dampLines(v, N, payload, a)) {
        } else if (!payloadCertified(payload)) raise(v, SW_SUSPENSION_ERROR, 8, 2, 0);
        else if (payload.maxResidual > a.maxLineResidual) raise(v, SW_SUSPENSION_ERROR, 8, 3, 0);
    }
    if (shared) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            v.px[i] = N.x[i];
            v.py[i] = N.y[i];
            v.pz[i] = N.z[i];
        }
    }
    profileAdd(v, kProfileFinalization, since);
    profileFlush(v);
}

template <bool kShared, bool kLocal, bool kGeneral>
__global__ void __launch_bounds__(kMaxThreads) substepKernel(WingView v, StepArgs a) {
    int block = static_cast<int>(blockIdx.x), blocks = static_cast<int>(gridDim.x);
    if constexpr (kLocal) {
        // One-XCD launch: the grid holds several workgroups per XCD; those
        // holding the first launchBlocks tickets of the XCD that fills them
        // first are the participants, the rest leave.
        __shared__ int role;
        if (threadIdx.x == 0) role = xcdParticipant(v, v.launchBlocks);
        __syncthreads();
        if (role < 0) return;
        block = role;
        blocks = v.launchBlocks;
    }
    substepBody<kShared ? 1 : 0, kLocal, kGeneral>(v, a, block, blocks, v.trace);
}

// Batched substep: workgroup b runs workgroup blockLocal[b] of wing
// blockWing[b]; each wing's view holds its own workgroup count and modes.
// `traces` (recording only): each wing's trace for this substep.
template <bool kGeneral>
__global__ void __launch_bounds__(kMaxThreads) batchSubstepKernel(const WingView* views, const StepArgs* args,
                                                                   const int* blockWing, const int* blockLocal,
                                                                   double* const* traces) {
    const int wing = blockWing[blockIdx.x];
    // The wing's view in LDS: the substep reads its fields after every
    // acquire, which would otherwise send each read back to global memory.
    __shared__ WingView sView;
    static_assert(sizeof(WingView) % 8 == 0, "the view copies as 8-byte words");
    {
        const auto* from = reinterpret_cast<const unsigned long long*>(&views[wing]);
        auto* to = reinterpret_cast<unsigned long long*>(&sView);
        for (int k = threadIdx.x; k < static_cast<int>(sizeof(WingView) / 8); k += blockDim.x) to[k] = from[k];
    }
    __syncthreads();
    substepBody<2, false, kGeneral>(sView, args[wing], blockLocal[blockIdx.x], sView.launchBlocks,
                                    traces ? traces[wing] : nullptr);
}

// The substep kernel variants: nodes in LDS or HBM, a one-XCD launch, and
// the general features (list phases, compression softening, momentum
// restoration, line damping) compiled in or out.
const void* substepKernelFor(bool shared, bool local, bool general) {
    if (shared)
        return general ? reinterpret_cast<const void*>(&substepKernel<true, false, true>)
                       : reinterpret_cast<const void*>(&substepKernel<true, false, false>);
    if (local)
        return general ? reinterpret_cast<const void*>(&substepKernel<false, true, true>)
                       : reinterpret_cast<const void*>(&substepKernel<false, true, false>);
    return general ? reinterpret_cast<const void*>(&substepKernel<false, false, true>)
                   : reinterpret_cast<const void*>(&substepKernel<false, false, false>);
}
const void* batchKernelFor(bool general) {
    return general ? reinterpret_cast<const void*>(&batchSubstepKernel<true>)
                   : reinterpret_cast<const void*>(&batchSubstepKernel<false>);
}

// Copies the mutable state block into the snapshot while no error is
// pending (so a failing step keeps its own pre-step snapshot).
// Each snapshot taken increments barrier word 120 (cleared with the error
// word), so the host can tell how many steps of a queued run started before
// a failure.
__global__ void snapshotKernel(WingView v, int restore) {
    if (!restore && __hip_atomic_load(&v.error->code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    if (!restore && blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_fetch_add(v.barrier + 120, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < v.stateDoubles; i += stride) {
        if (restore) v.state[i] = v.snapshot[i];
        else v.snapshot[i] = v.state[i];
    }
}

// Snapshot (restore = 0) of every wing without a pending error, or restore
// (restore = 1) of the wings with mask[wing] set; blockIdx.y is the wing.
__global__ void batchSnapshotKernel(const WingView* views, const unsigned char* mask, int restore) {
    const WingView& v = views[blockIdx.y];
    if (restore ? mask[blockIdx.y] == 0
                : __hip_atomic_load(&v.error->code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
        return;
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < v.stateDoubles; i += stride) {
        if (restore) v.state[i] = v.snapshot[i];
        else v.snapshot[i] = v.state[i];
    }
}

// StructureDiagnostics reductions (structure.cpp:833-959): one partial
// record per workgroup, combined on the host in workgroup order.
struct Partial {
    double mass, wx, wy, wz, mx, my, mz, kinetic;
    double maxDistance, maxCable, maxStrain, maxResidual;
    int finite, dynamicNodes, pad0, pad1;
};

__global__ void __launch_bounds__(256) diagnosticsKernel(WingView v, Partial* out) {
    __shared__ Partial block[4];  // one per wave
    Partial p{};
    p.finite = 1;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    for (int i = tid; i < v.nodeCount; i += stride) {
        const d3 x{v.px[i], v.py[i], v.pz[i]};
        const d3 q{v.qx[i], v.qy[i], v.qz[i]};
        const d3 vel{v.vx[i], v.vy[i], v.vz[i]};
        const double wi = v.w[i];
        p.finite &= (fin3(x) && fin3(q) && fin3(vel) && isfinite(wi) && wi >= 0.0) ? 1 : 0;
        if (wi > 0.0) {
            const double m = 1.0 / wi;
            p.dynamicNodes += 1;
            p.mass += m;
            p.wx += x.x * m;
            p.wy += x.y * m;
            p.wz += x.z * m;
            p.mx += vel.x * m;
            p.my += vel.y * m;
            p.mz += vel.z * m;
            p.kinetic += 0.5 * m * dot(vel, vel);
        }
    }
    for (int k = tid; k < v.distCount; k += stride) {
        const d3 a{v.px[v.da[k]], v.py[v.da[k]], v.pz[v.da[k]]};
        const d3 b{v.px[v.db[k]], v.py[v.db[k]], v.pz[v.db[k]]};
        const double error = len(sub(b, a)) - v.drest[k];
        if (v.dcable[k]) p.maxCable = stdmax(p.maxCable, stdmax(0.0, error));
        else p.maxDistance = stdmax(p.maxDistance, fabs(error));
    }
    for (int k = tid; k < v.membCount; k += stride) {
        const int i0 = v.ma[k], i1 = v.mb[k], i2 = v.mc[k];
        const MembraneEval e = evaluateMembrane(d3{v.px[i0], v.py[i0], v.pz[i0]}, d3{v.px[i1], v.py[i1], v.pz[i1]},
                                                d3{v.px[i2], v.py[i2], v.pz[i2]}, v.m00[k], v.m01[k], v.m10[k],
                                                v.m11[k], v.anyPreTension ? v.preW[k] : 0.0,
                                                v.anyPreTension ? v.preT[k] : 0.0);
        p.maxStrain = stdmax(p.maxStrain, stdmax(stdmax(fabs(e.strain.x), fabs(e.strain.y)), fabs(e.strain.z)));
        p.maxResidual = stdmax(p.maxResidual, fabs(v.mres[k]));
        p.finite &= (fin3(e.strain) && isfinite(v.mres[k])) ? 1 : 0;
    }
    // Each wave reduces its lanes by exchanges; the waves' results are then
    // combined in wave order through LDS.
    p.mass = waveSum(p.mass);
    p.wx = waveSum(p.wx);
    p.wy = waveSum(p.wy);
    p.wz = waveSum(p.wz);
    p.mx = waveSum(p.mx);
    p.my = waveSum(p.my);
    p.mz = waveSum(p.mz);
    p.kinetic = waveSum(p.kinetic);
    p.maxDistance = waveMax(p.maxDistance);
    p.maxCable = waveMax(p.maxCable);
    p.maxStrain = waveMax(p.maxStrain);
    p.maxResidual = waveMax(p.maxResidual);
    p.finite = __all(p.finite) ? 1 : 0;
    p.dynamicNodes = waveSumInt(p.dynamicNodes);
    if ((threadIdx.x & 63) == 0) block[threadIdx.x >> 6] = p;
    __syncthreads();
    if (threadIdx.x == 0) {
        Partial r{};
        r.finite = 1;
        for (int t = 0; t < static_cast<int>(blockDim.x + 63) / 64; ++t) {
            const Partial& s = block[t];
            r.mass += s.mass;
            r.wx += s.wx;
            r.wy += s.wy;
            r.wz += s.wz;
            r.mx += s.mx;
            r.my += s.my;
            r.mz += s.mz;
            r.kinetic += s.kinetic;
            r.maxDistance = stdmax(r.maxDistance, s.maxDistance);
            r.maxCable = stdmax(r.maxCable, s.maxCable);
            r.maxStrain = stdmax(r.maxStrain, s.maxStrain);
            r.maxResidual = stdmax(r.maxResidual, s.maxResidual);
            r.finite &= s.finite;
            r.dynamicNodes += s.dynamicNodes;
        }
        out[blockIdx.x] = r;
    }
}

// Diagnostic-frame fields of the committed state (structure_frame.cpp:
// 491-640): harness point positions and velocities
// (payloadPointPosition/Velocity, suspension.cpp:1018-1032), and per line
// (distance slots, then segment slots) the current length |second - first|
// and its violation (structure_frame.cpp:60-73: |L - rest| for distance,
// tie and stitch, max(0, L - rest) for cables and suspension segments).
// out: harness position 3H, velocity 3H (runtime hang order), then length
// and violation per distance slot, then per segment slot.
__global__ void frameKernel(WingView v, double* out) {
    const int h = v.hangCount;
    double* hpos = out;
    double* hvel = out + 3L * h;
    double* length = out + 6L * h;
    double* violation = length + v.distCount + v.segCount;
    const PayloadDev& p = *v.payload;
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    const long first = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    m3 r{};
    bool rigid = false;
    if (v.segCount > 0) rigid = rotationMatrix(quat{p.q[0], p.q[1], p.q[2], p.q[3]}, r);
    const d3 com{p.com[0], p.com[1], p.com[2]};
    const auto hang = [&](int i, d3& position, d3& offset) {
        const d3 rel = sub(d3{v.hangLocal[3 * i], v.hangLocal[3 * i + 1], v.hangLocal[3 * i + 2]},
                           d3{p.comLocal[0], p.comLocal[1], p.comLocal[2]});
        offset = mulM(r, rel);
        position = add(com, offset);
    };
    for (long i = first; i < h; i += stride) {
        d3 position, offset;
        hang(static_cast<int>(i), position, offset);
        const d3 velocity = add(d3{p.v[0], p.v[1], p.v[2]}, cross(d3{p.w[0], p.w[1], p.w[2]}, offset));
        const double nan = __builtin_nan("");
        hpos[3 * i] = rigid ? position.x : nan;
        hpos[3 * i + 1] = rigid ? position.y : nan;
        hpos[3 * i + 2] = rigid ? position.z : nan;
        hvel[3 * i] = rigid ? velocity.x : nan;
        hvel[3 * i + 1] = rigid ? velocity.y : nan;
        hvel[3 * i + 2] = rigid ? velocity.z : nan;
    }
    for (long k = first; k < v.distCount; k += stride) {
        const int a = v.da[k], b = v.db[k];
        const double l = len(sub(d3{v.px[b], v.py[b], v.pz[b]}, d3{v.px[a], v.py[a], v.pz[a]}));
        const double extension = l - v.drest[k];
        length[k] = l;
        violation[k] = v.dcable[k] ? stdmax(0.0, extension) : fabs(extension);
    }
    for (long k = first; k < v.segCount; k += stride) {
        d3 ends[2], offset;
        const int codes[2] = {v.sfrom[k], v.sto[k]};
        for (int e = 0; e < 2; ++e) {
            if (codes[e] >= 0) ends[e] = d3{v.px[codes[e]], v.py[codes[e]], v.pz[codes[e]]};
            else hang(-codes[e] - 1, ends[e], offset);
        }
        const double l = len(sub(ends[1], ends[0]));
        length[v.distCount + k] = l;
        violation[v.distCount + k] = stdmax(0.0, l - v.srest[k]);
    }
}

// The diagnostic-frame fields of the committed state directly in frame
// order: positions, velocities (3 per vertex: the nodes, then the harness
// points in definition order), then per line in definition order the
// current length and its violation (the quantities frameKernel computes).
// lineSlot maps a definition line (constraints, then segments) to its
// device item (distance slot, or distCount + segment slot); hangSlot maps a
// definition harness point to its device index.
__global__ void frameFieldsKernel(WingView v, const int* __restrict__ lineSlot, const int* __restrict__ hangSlot,
                                  double* out) {
    const long n = v.nodeCount;
    const int h = v.hangCount;
    const long vertices = n + h;
    const long lines = static_cast<long>(v.distCount) + v.segCount;
    double* position = out;
    double* velocity = out + 3 * vertices;
    double* length = out + 6 * vertices;
    double* violation = length + lines;
    const PayloadDev& p = *v.payload;
    const long stride = static_cast<long>(gridDim.x) * blockDim.x;
    const long first = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    m3 r{};
    bool rigid = false;
    if (v.segCount > 0) rigid = rotationMatrix(quat{p.q[0], p.q[1], p.q[2], p.q[3]}, r);
    const d3 com{p.com[0], p.com[1], p.com[2]};
    const auto hang = [&](int i, d3& point, d3& offset) {
        const d3 rel = sub(d3{v.hangLocal[3 * i], v.hangLocal[3 * i + 1], v.hangLocal[3 * i + 2]},
                           d3{p.comLocal[0], p.comLocal[1], p.comLocal[2]});
        offset = mulM(r, rel);
        point = add(com, offset);
    };
    for (long i = first; i < n; i += stride) {
        position[3 * i] = v.px[i];
        position[3 * i + 1] = v.py[i];
        position[3 * i + 2] = v.pz[i];
        velocity[3 * i] = v.vx[i];
        velocity[3 * i + 1] = v.vy[i];
        velocity[3 * i + 2] = v.vz[i];
    }
    for (long j = first; j < h; j += stride) {
        d3 point, offset;
        hang(hangSlot[j], point, offset);
        const d3 pv = add(d3{p.v[0], p.v[1], p.v[2]}, cross(d3{p.w[0], p.w[1], p.w[2]}, offset));
        const double nan = __builtin_nan("");
        const long o = 3 * (n + j);
        position[o] = rigid ? point.x : nan;
        position[o + 1] = rigid ? point.y : nan;
        position[o + 2] = rigid ? point.z : nan;
        velocity[o] = rigid ? pv.x : nan;
        velocity[o + 1] = rigid ? pv.y : nan;
        velocity[o + 2] = rigid ? pv.z : nan;
    }
    for (long l = first; l < lines; l += stride) {
        const int k = lineSlot[l];
        if (k < v.distCount) {
            const int a = v.da[k], b = v.db[k];
            const double len_ = len(sub(d3{v.px[b], v.py[b], v.pz[b]}, d3{v.px[a], v.py[a], v.pz[a]}));
            const double extension = len_ - v.drest[k];
            length[l] = len_;
            violation[l] = v.dcable[k] ? stdmax(0.0, extension) : fabs(extension);
        } else {
            const int s = k - v.distCount;
            d3 ends[2], offset;
            const int codes[2] = {v.sfrom[s], v.sto[s]};
            for (int e = 0; e < 2; ++e) {
                if (codes[e] >= 0) ends[e] = d3{v.px[codes[e]], v.py[codes[e]], v.pz[codes[e]]};
                else hang(-codes[e] - 1, ends[e], offset);
            }
            const double len_ = len(sub(ends[1], ends[0]));
            length[l] = len_;
            violation[l] = stdmax(0.0, len_ - v.srest[s]);
        }
    }
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

// Staging for sw_structure_frame_fields_noalloc: a file-scope device array,
// one copy per GPU, held until the process ends. Host code never uses its
// name as an address; it uses the per-device pointer from
// hipGetSymbolAddress (frameScratchPointer).
static __device__ __align__(256) double scratch[kFrameScratchDoubles];

}  // namespace

double* frameScratchPointer(int device) {
    static std::mutex guard;
    static std::vector<double*> pointers;
    std::lock_guard<std::mutex> lock(guard);
    if (device < 0) throw std::invalid_argument("device index must not be negative");
    if (pointers.size() <= static_cast<std::size_t>(device)) pointers.resize(static_cast<std::size_t>(device) + 1);
    double*& ptr = pointers[static_cast<std::size_t>(device)];
    if (!ptr) {
        check(hipSetDevice(device), "hipSetDevice");
        void* address = nullptr;
        check(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        ptr = static_cast<double*>(address);
    }
    return ptr;
}

// ---------------------------------------------------------------------------
// DeviceWing host side.

namespace {

template <typename T>
T* deviceCopy(std::vector<void*>& owned, const std::vector<T>& host, std::size_t minimum = 1) {
    const std::size_t count = std::max(host.size(), minimum);
    void* p = nullptr;
    check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    owned.push_back(p);
    if (!host.empty()) check(hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
    else check(hipMemset(p, 0, count * sizeof(T)), "hipMemset");
    return static_cast<T*>(p);
}

}  // namespace

// Membrane compliance from the orthotropic material: the checkedInverse
// cofactors of the stiffness {xx warp, yy weft, zz shear, xy coupling,
// xz 0, yz 0} (membrane.cpp:268-271), with the four stiffness entries scaled
// by `stiffnessScale` first (1 for the definition's own material).
void membraneCompliance(const double* mat, double stiffnessScale, double out[6]) {
    const double xx = mat[0] * stiffnessScale, yy = mat[1] * stiffnessScale, zz = mat[3] * stiffnessScale,
                 xy = mat[2] * stiffnessScale, xz = 0.0, yz = 0.0;
    const double sMax = std::max({std::abs(xx), std::abs(yy), std::abs(zz), std::abs(xy), std::abs(xz), std::abs(yz)});
    const double sDet = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
    const double sTol = 128.0 * 2.220446049250313080847263336181640625e-16 * sMax * sMax * sMax;
    if (!(sMax > 0.0) || !std::isfinite(sDet) || std::abs(sDet) <= sTol)
        throw std::invalid_argument("SymmetricMatrix3 is singular or ill-conditioned");
    const double r = 1.0 / sDet;
    out[0] = (yy * zz - yz * yz) * r;
    out[1] = (xx * zz - xz * xz) * r;
    out[2] = (xx * yy - xy * xy) * r;
    out[3] = (xz * yz - xy * zz) * r;
    out[4] = (xy * yz - xz * yy) * r;
    out[5] = (xy * xz - xx * yz) * r;
}

namespace {

template <typename T>
std::vector<T> permute(const std::vector<T>& values, const std::vector<std::uint32_t>& order, std::size_t width = 1) {
    std::vector<T> out(order.size() * width);
    for (std::size_t k = 0; k < order.size(); ++k)
        for (std::size_t c = 0; c < width; ++c) out[k * width + c] = values[order[k] * width + c];
    return out;
}

std::vector<int> column(const std::vector<std::uint64_t>& values, const std::vector<std::uint32_t>& order,
                        std::size_t width, std::size_t col) {
    std::vector<int> out(order.size());
    for (std::size_t k = 0; k < order.size(); ++k) out[k] = static_cast<int>(values[order[k] * width + col]);
    return out;
}

std::vector<double> columnD(const std::vector<double>& values, const std::vector<std::uint32_t>& order,
                            std::size_t width, std::size_t col) {
    std::vector<double> out(order.size());
    for (std::size_t k = 0; k < order.size(); ++k) out[k] = values[order[k] * width + col];
    return out;
}

}  // namespace

DeviceWing::DeviceWing(const Definition& d, int device) : device_(device), view_(std::make_unique<WingView>()) {
    check(hipSetDevice(device_), "hipSetDevice");
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    try {
        build(d);
    } catch (...) {
        for (void* p : allocations_) (void)hipFree(p);
        allocations_.clear();
        (void)hipStreamDestroy(stream_);
        throw;
    }
}

DeviceWing::~DeviceWing() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : allocations_) (void)hipFree(p);
    for (void* p : programAlloc_) (void)hipFree(p);
    for (void* p : depAlloc_)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void DeviceWing::setTiming(bool enabled) {
    check(hipSetDevice(device_), "hipSetDevice");
    if (enabled && !timing_) {
        check(hipStreamSynchronize(stream_), "timing");
        eventsUsed_ = 0;
        timedMs_ = 0.0;
        timedLaunches_ = 0;
    }
    timing_ = enabled;
}

void DeviceWing::setProfiling(bool enabled) {
    check(hipSetDevice(device_), "hipSetDevice");
    check(hipStreamSynchronize(stream_), "profiling");
    if (enabled && !profile_) {
        void* p = nullptr;
        check(hipMalloc(&p, kProfileParts * sizeof(unsigned long long)), "hipMalloc profile");
        allocations_.push_back(p);
        profile_ = static_cast<unsigned long long*>(p);
        clearProfile();
    }
    view_->profile = enabled ? profile_ : nullptr;
    ++launchGeneration_;  // batched views copy the wing's view
}

void DeviceWing::clearProfile() {
    if (!profile_) return;
    check(hipSetDevice(device_), "hipSetDevice");
    check(hipMemsetAsync(profile_, 0, kProfileParts * sizeof(unsigned long long), stream_), "clear profile");
    check(hipStreamSynchronize(stream_), "clear profile");
    profiledLaunches_ = 0;
}

void DeviceWing::profileTotals(double nanoseconds[9], long long* launches) {
    unsigned long long ticks[kProfileParts] = {};
    if (profile_) {
        check(hipSetDevice(device_), "hipSetDevice");
        check(hipStreamSynchronize(stream_), "profile");
        check(hipMemcpy(ticks, profile_, sizeof(ticks), hipMemcpyDeviceToHost), "profile");
    }
    // ticks[9]: workgroup launches; profiledLaunches_: substep launches.
    const double perLaunch = profiledLaunches_ > 0 && ticks[kProfileParts - 1] > 0
                                 ? static_cast<double>(ticks[kProfileParts - 1]) / static_cast<double>(profiledLaunches_)
                                 : 1.0;
    for (int k = 0; k < 9; ++k) nanoseconds[k] = 10.0 * static_cast<double>(ticks[k]) / perLaunch;  // 100 MHz clock
    if (launches) *launches = profiledLaunches_;
}

void DeviceWing::kernelTime(double* milliseconds, long long* launches) {
    check(hipSetDevice(device_), "hipSetDevice");
    check(hipStreamSynchronize(stream_), "timing");
    for (std::size_t i = 0; i + 1 < eventsUsed_; i += 2) {
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, events_[i], events_[i + 1]), "hipEventElapsedTime");
        timedMs_ += ms;
        ++timedLaunches_;
    }
    eventsUsed_ = 0;
    if (milliseconds) *milliseconds = timedMs_;
    if (launches) *launches = timedLaunches_;
}

void DeviceWing::build(const Definition& d) {
    WingView& v = *view_;
    std::memset(&v, 0, sizeof(WingView));
    nodes_ = d.nodeCount();
    triangles_ = d.triangleCount();
    constraints_ = d.constraintCount();
    membranes_ = d.membraneCount();
    hinges_ = d.dihedralCount();
    if (nodes_ > 0x3fffffff || 3 * triangles_ > 0x3fffffff || constraints_ > 0x3fffffff)
        throw std::invalid_argument("structure exceeds the device index range");

    distance_ = colourDistance(d);
    membrane_ = colourMembranes(d);
    hinge_ = colourDihedrals(d);
    items_ = ItemNodes{};
    for (std::size_t k = 0; k < distance_.order.size(); ++k)
        for (int c = 0; c < 2; ++c) items_.dist.push_back(static_cast<int>(d.constraintNodes[2 * distance_.order[k] + c]));
    for (std::size_t k = 0; k < membrane_.order.size(); ++k) {
        const std::size_t t = d.membraneTriangle[membrane_.order[k]];
        for (int c = 0; c < 3; ++c) items_.memb.push_back(static_cast<int>(d.triangleNodes[3 * t + c]));
    }
    for (std::size_t k = 0; k < hinge_.order.size(); ++k)
        for (int c = 0; c < 4; ++c) items_.hinge.push_back(static_cast<int>(d.dihedralNodes[4 * hinge_.order[k] + c]));
    layout_ = suspensionLayout(d);
    segments_ = layout_.segmentOrder.size();

    // --- mutable state block layout ---------------------------------------
    // positions, prev, velocity (3N each), distance lambda (C), membrane
    // lambda + resultant (6M) + residual (M), hinge lambda (D), segment
    // lambda (S), sphere lambda (P), face pressure (T), payload.
    std::size_t pairCount = d.sphereNodes.size();
    pairs_ = pairCount;
    const std::size_t payloadDoubles = sizeof(PayloadDev) / sizeof(double);
    const std::size_t total = 9 * nodes_ + constraints_ + 7 * membranes_ + hinges_ + segments_ + pairCount +
                              triangles_ + payloadDoubles;
    snapshotBytes_ = total * sizeof(double);
    double* state = nullptr;
    double* snap = nullptr;
    check(hipMalloc(&state, snapshotBytes_), "hipMalloc state");
    allocations_.push_back(state);
    check(hipMalloc(&snap, snapshotBytes_), "hipMalloc snapshot");
    allocations_.push_back(snap);
    check(hipMemset(state, 0, snapshotBytes_), "memset state");
    v.state = state;
    v.snapshot = snap;
    v.stateDoubles = static_cast<long>(total);
    double* cursor = state;
    const auto take = [&](std::size_t count) {
        double* p = cursor;
        cursor += count;
        return p;
    };
    v.px = take(nodes_);
    v.py = take(nodes_);
    v.pz = take(nodes_);
    v.qx = take(nodes_);
    v.qy = take(nodes_);
    v.qz = take(nodes_);
    v.vx = take(nodes_);
    v.vy = take(nodes_);
    v.vz = take(nodes_);
    v.dlam = take(constraints_);
    v.mlx = take(membranes_);
    v.mly = take(membranes_);
    v.mlz = take(membranes_);
    v.mrx = take(membranes_);
    v.mry = take(membranes_);
    v.mrz = take(membranes_);
    v.mres = take(membranes_);
    v.hlam = take(hinges_);
    v.slam = take(segments_);
    v.splam = take(pairCount);
    v.pressure = take(triangles_);
    v.payload = reinterpret_cast<PayloadDev*>(take(payloadDoubles));

    // Nodes.
    v.nodeCount = static_cast<int>(nodes_);
    {
        std::vector<double> x(nodes_), y(nodes_), z(nodes_), w(nodes_);
        for (std::size_t i = 0; i < nodes_; ++i) {
            x[i] = d.nodePosition[3 * i];
            y[i] = d.nodePosition[3 * i + 1];
            z[i] = d.nodePosition[3 * i + 2];
            w[i] = d.nodeFixed[i] ? 0.0 : 1.0 / d.nodeMass[i];
        }
        check(hipMemcpy(v.px, x.data(), nodes_ * 8, hipMemcpyHostToDevice), "upload x");
        check(hipMemcpy(v.py, y.data(), nodes_ * 8, hipMemcpyHostToDevice), "upload y");
        check(hipMemcpy(v.pz, z.data(), nodes_ * 8, hipMemcpyHostToDevice), "upload z");
        check(hipMemcpy(v.qx, x.data(), nodes_ * 8, hipMemcpyHostToDevice), "upload qx");
        check(hipMemcpy(v.qy, y.data(), nodes_ * 8, hipMemcpyHostToDevice), "upload qy");
        check(hipMemcpy(v.qz, z.data(), nodes_ * 8, hipMemcpyHostToDevice), "upload qz");
        v.w = deviceCopy(allocations_, w);
        // An isolated free membrane (no suspension, no contact, every node
        // dynamic) has its predicted momentum restored after each substep
        // (soft_body.cpp:1196-1205).
        preserveMomentum_ = !d.suspension && (membranes_ > 0 || hinges_ > 0) && pairCount == 0 &&
                            std::all_of(w.begin(), w.end(), [](double x) { return x > 0.0; });
        v.preserveMomentum = preserveMomentum_ ? 1 : 0;
        if (preserveMomentum_) v.momentum = deviceCopy(allocations_, std::vector<double>(16, 0.0));
        std::vector<double> zeros(nodes_, 0.0);
        v.fx = deviceCopy(allocations_, zeros);
        v.fy = deviceCopy(allocations_, zeros);
        v.fz = deviceCopy(allocations_, zeros);
    }
    // Triangles and the node -> incident triangle CSR in ascending order.
    {
        std::vector<int> ta(triangles_), tb(triangles_), tc(triangles_);
        std::vector<int> offsets(nodes_ + 1, 0);
        for (std::size_t t = 0; t < triangles_; ++t) {
            ta[t] = static_cast<int>(d.triangleNodes[3 * t]);
            tb[t] = static_cast<int>(d.triangleNodes[3 * t + 1]);
            tc[t] = static_cast<int>(d.triangleNodes[3 * t + 2]);
            ++offsets[ta[t] + 1];
            ++offsets[tb[t] + 1];
            ++offsets[tc[t] + 1];
        }
        for (std::size_t i = 0; i < nodes_; ++i) offsets[i + 1] += offsets[i];
        std::vector<int> list(3 * triangles_);
        std::vector<int> fill(offsets.begin(), offsets.end() - 1);
        for (std::size_t t = 0; t < triangles_; ++t) {
            list[fill[ta[t]]++] = static_cast<int>(t);
            list[fill[tb[t]]++] = static_cast<int>(t);
            list[fill[tc[t]]++] = static_cast<int>(t);
        }
        v.ta = deviceCopy(allocations_, ta);
        v.tb = deviceCopy(allocations_, tb);
        v.tc = deviceCopy(allocations_, tc);
        v.incOffsets = deviceCopy(allocations_, offsets);
        v.incTriangles = deviceCopy(allocations_, list);
    }
    // Distance constraints, colour-major.
    {
        const auto& o = distance_.order;
        v.distCount = static_cast<int>(o.size());
        v.da = deviceCopy(allocations_, column(d.constraintNodes, o, 2, 0));
        v.db = deviceCopy(allocations_, column(d.constraintNodes, o, 2, 1));
        v.drest = deviceCopy(allocations_, permute(d.constraintRest, o));
        v.dcomp = deviceCopy(allocations_, permute(d.constraintCompliance, o));
        std::vector<unsigned char> cable(o.size());
        for (std::size_t k = 0; k < o.size(); ++k) cable[k] = d.constraintKind[o[k]] == SW_CONSTRAINT_CABLE ? 1 : 0;
        v.dcable = deviceCopy(allocations_, cable);
    }
    // Membranes, colour-major, with the element construction of
    // addMembraneElements (soft_body.cpp:663-690): reference area, inverse
    // reference matrix (membrane.cpp:131-151) and compliance matrix
    // (membrane.cpp:268-271).
    {
        const auto& o = membrane_.order;
        v.membCount = static_cast<int>(o.size());
        std::vector<int> a(o.size()), b(o.size()), c(o.size());
        std::vector<double> m00(o.size()), m01(o.size()), m10(o.size()), m11(o.size()), sc(o.size());
        std::vector<double> cxx(o.size()), cyy(o.size()), czz(o.size()), cxy(o.size()), cxz(o.size()), cyz(o.size());
        std::vector<double> pw(o.size()), pt(o.size()), damp(o.size());
        std::vector<double> ratio(o.size()), kxx(o.size()), kyy(o.size()), kzz(o.size()), kxy(o.size());
        int anyPre = 0, anyDamp = 0, anyCompression = 0;
        for (std::size_t k = 0; k < o.size(); ++k) {
            const std::size_t m = o[k];
            const std::uint64_t t = d.membraneTriangle[m];
            a[k] = static_cast<int>(d.triangleNodes[3 * t]);
            b[k] = static_cast<int>(d.triangleNodes[3 * t + 1]);
            c[k] = static_cast<int>(d.triangleNodes[3 * t + 2]);
            const double* chart = &d.membraneChart[6 * m];
            const double e1x = chart[2] - chart[0], e1y = chart[3] - chart[1];
            const double e2x = chart[4] - chart[0], e2y = chart[5] - chart[1];
            // Matrix2{e1.x, e2.x, e1.y, e2.y}
            const double r00 = e1x, r01 = e2x, r10 = e1y, r11 = e2y;
            const double det = r00 * r11 - r01 * r10;
            const double referenceArea = 0.5 * det;
            const double scaleValue = std::max({std::abs(r00), std::abs(r01), std::abs(r10), std::abs(r11)});
            const double tol = 64.0 * 2.220446049250313080847263336181640625e-16 * scaleValue * scaleValue;
            if (!(scaleValue > 0.0) || !std::isfinite(det) || std::abs(det) <= tol)
                throw std::invalid_argument("Matrix2 is singular or ill-conditioned");
            const double inv = 1.0 / det;
            m00[k] = r11 * inv;
            m01[k] = -r01 * inv;
            m10[k] = -r10 * inv;
            m11[k] = r00 * inv;
            sc[k] = std::sqrt(referenceArea);
            const double* mat = &d.membraneMaterial[8 * m];
            double cm[6];
            membraneCompliance(mat, 1.0, cm);
            cxx[k] = cm[0], cyy[k] = cm[1], czz[k] = cm[2], cxy[k] = cm[3], cxz[k] = cm[4], cyz[k] = cm[5];
            pw[k] = mat[4];
            pt[k] = mat[5];
            damp[k] = mat[6];
            anyPre |= (mat[4] != 0.0 || mat[5] != 0.0) ? 1 : 0;
            anyDamp |= mat[6] != 0.0 ? 1 : 0;
            // stiffnessMatrix() (membrane.cpp:259-266): warp, weft, shear, coupling.
            ratio[k] = mat[7];
            kxx[k] = mat[0];
            kyy[k] = mat[1];
            kzz[k] = mat[3];
            kxy[k] = mat[2];
            anyCompression |= mat[7] != 1.0 ? 1 : 0;
        }
        v.ma = deviceCopy(allocations_, a);
        v.mb = deviceCopy(allocations_, b);
        v.mc = deviceCopy(allocations_, c);
        v.m00 = deviceCopy(allocations_, m00);
        v.m01 = deviceCopy(allocations_, m01);
        v.m10 = deviceCopy(allocations_, m10);
        v.m11 = deviceCopy(allocations_, m11);
        v.mscale = deviceCopy(allocations_, sc);
        v.cxx = deviceCopy(allocations_, cxx);
        v.cyy = deviceCopy(allocations_, cyy);
        v.czz = deviceCopy(allocations_, czz);
        v.cxy = deviceCopy(allocations_, cxy);
        v.cxz = deviceCopy(allocations_, cxz);
        v.cyz = deviceCopy(allocations_, cyz);
        v.preW = deviceCopy(allocations_, pw);
        v.preT = deviceCopy(allocations_, pt);
        v.mdamp = deviceCopy(allocations_, damp);
        v.anyPreTension = anyPre;
        v.anyDamping = anyDamp;
        v.anyCompression = anyCompression;
        compressedMembranes_ = anyCompression != 0;
        if (anyCompression) {
            v.mratio = deviceCopy(allocations_, ratio);
            v.kxx = deviceCopy(allocations_, kxx);
            v.kyy = deviceCopy(allocations_, kyy);
            v.kzz = deviceCopy(allocations_, kzz);
            v.kxy = deviceCopy(allocations_, kxy);
        }
    }
    // Hinges, colour-major.
    {
        const auto& o = hinge_.order;
        v.hingeCount = static_cast<int>(o.size());
        v.ha = deviceCopy(allocations_, column(d.dihedralNodes, o, 4, 0));
        v.hb = deviceCopy(allocations_, column(d.dihedralNodes, o, 4, 1));
        v.hc = deviceCopy(allocations_, column(d.dihedralNodes, o, 4, 2));
        v.hd = deviceCopy(allocations_, column(d.dihedralNodes, o, 4, 3));
        v.hrest = deviceCopy(allocations_, permute(d.dihedralRestAngle, o));
        v.hcomp = deviceCopy(allocations_, permute(d.dihedralCompliance, o));
    }
    // Phase program.
    iterPhases_.clear();
    for (std::size_t c = 0; c + 1 < distance_.offsets.size(); ++c)
        iterPhases_.push_back({kPhaseDistance, static_cast<int>(distance_.offsets[c]),
                               static_cast<int>(distance_.offsets[c + 1]), 0});
    for (std::size_t c = 0; c + 1 < membrane_.offsets.size(); ++c)
        iterPhases_.push_back({kPhaseMembrane, static_cast<int>(membrane_.offsets[c]),
                               static_cast<int>(membrane_.offsets[c + 1]), 0});
    for (std::size_t c = 0; c + 1 < hinge_.offsets.size(); ++c)
        iterPhases_.push_back({kPhaseHinge, static_cast<int>(hinge_.offsets[c]), static_cast<int>(hinge_.offsets[c + 1]), 0});

    // Suspension: dependency levels over the runtime (string-id) order.
    {
        std::vector<std::vector<std::uint32_t>> writes(segments_), reads(segments_);
        std::vector<std::uint8_t> touchesPayload(segments_, 0);
        for (std::size_t k = 0; k < segments_; ++k) {
            for (int e : {layout_.from[k], layout_.to[k]}) {
                if (e >= 0) {
                    writes[k].push_back(static_cast<std::uint32_t>(e));
                    reads[k].push_back(static_cast<std::uint32_t>(e));
                } else {
                    touchesPayload[k] = 1;
                }
            }
        }
        const auto level = dependencyLevels(writes, reads, touchesPayload, nodes_);
        std::vector<std::uint32_t> slots(segments_);
        std::iota(slots.begin(), slots.end(), 0u);
        std::stable_sort(slots.begin(), slots.end(),
                         [&](std::uint32_t x, std::uint32_t y) { return level[x] < level[y]; });
        segmentDevicePosition_.assign(segments_, 0);
        std::vector<int> from(segments_), to(segments_), runtime(segments_);
        std::vector<double> rest(segments_), stiff(segments_);
        std::vector<unsigned char> last(segments_, 0);
        int lastPayloadSlot = -1;
        for (std::size_t s = 0; s < segments_; ++s) {
            const std::uint32_t k = slots[s];
            segmentDevicePosition_[k] = static_cast<std::uint32_t>(s);
            from[s] = layout_.from[k];
            to[s] = layout_.to[k];
            runtime[s] = static_cast<int>(k);
            const std::uint32_t def = layout_.segmentOrder[k];
            rest[s] = d.suspension->segmentRest[def];
            stiff[s] = d.suspension->segmentStiffness[def];
            if (touchesPayload[k]) lastPayloadSlot = static_cast<int>(s);
        }
        if (lastPayloadSlot >= 0) last[lastPayloadSlot] = 1;
        v.segCount = static_cast<int>(segments_);
        v.sfrom = deviceCopy(allocations_, from);
        v.sto = deviceCopy(allocations_, to);
        v.sruntime = deviceCopy(allocations_, runtime);
        v.srest = deviceCopy(allocations_, rest);
        v.sstiff = deviceCopy(allocations_, stiff);
        v.slastPayload = deviceCopy(allocations_, last);
        // Axial line damping (suspension.cpp:2413-2479) visits the damped
        // segments in runtime order.
        std::vector<int> dampSlot;
        std::vector<double> dampValue;
        for (std::size_t k = 0; k < segments_; ++k) {
            const double damping = d.suspension->segmentDamping[layout_.segmentOrder[k]];
            if (!(damping > 0.0)) continue;
            dampSlot.push_back(static_cast<int>(segmentDevicePosition_[k]));
            dampValue.push_back(damping);
        }
        v.dampCount = static_cast<int>(dampSlot.size());
        dampedSegments_ = dampSlot.size();
        if (!dampSlot.empty()) {
            v.dampSlot = deviceCopy(allocations_, dampSlot);
            v.dampValue = deviceCopy(allocations_, dampValue);
        }
        // Levels -> phases; consecutive single-item levels merge into one
        // serial run, staged in LDS in chunks of at most kChainItems
        // segments and kChainSlots distinct nodes.
        std::vector<int> fromSlot(segments_, 0), toSlot(segments_, 0), runNodes;
        std::vector<int> nodeSlot(nodes_, -1);
        const auto addRun = [&](std::size_t b, std::size_t e) {
            std::size_t c = b;
            while (c < e) {
                Phase ph{kPhaseSegmentRun, static_cast<int>(c), 0, 0, static_cast<int>(runNodes.size()), 0};
                std::size_t j = c;
                for (; j < e && j - c < static_cast<std::size_t>(kChainItems); ++j) {
                    int fresh = 0;
                    for (int endpoint : {from[j], to[j]})
                        if (endpoint >= 0 && nodeSlot[endpoint] < 0) ++fresh;
                    if (from[j] >= 0 && from[j] == to[j] && nodeSlot[from[j]] < 0) fresh = 1;
                    if (runNodes.size() - ph.slotBegin + fresh > static_cast<std::size_t>(kChainSlots)) break;
                    int* out[2] = {&fromSlot[j], &toSlot[j]};
                    const int ends[2] = {from[j], to[j]};
                    for (int q = 0; q < 2; ++q) {
                        if (ends[q] < 0) {
                            *out[q] = ends[q];
                            continue;
                        }
                        if (nodeSlot[ends[q]] < 0) {
                            nodeSlot[ends[q]] = static_cast<int>(runNodes.size() - ph.slotBegin);
                            runNodes.push_back(ends[q]);
                        }
                        *out[q] = nodeSlot[ends[q]];
                    }
                }
                ph.end = static_cast<int>(j);
                ph.slotEnd = static_cast<int>(runNodes.size());
                for (int q = ph.slotBegin; q < ph.slotEnd; ++q) nodeSlot[runNodes[q]] = -1;
                iterPhases_.push_back(ph);
                c = j;
            }
        };
        std::size_t s = 0;
        while (s < segments_) {
            std::size_t e = s;
            while (e < segments_ && level[slots[e]] == level[slots[s]]) ++e;
            if (e - s == 1) {
                std::size_t r = e;
                while (r < segments_) {
                    std::size_t r2 = r;
                    while (r2 < segments_ && level[slots[r2]] == level[slots[r]]) ++r2;
                    if (r2 - r != 1) break;
                    r = r2;
                }
                if (r - s > 1) addRun(s, r);
                else iterPhases_.push_back({kPhaseSegments, static_cast<int>(s), static_cast<int>(r), 0});
                s = r;
            } else {
                iterPhases_.push_back({kPhaseSegments, static_cast<int>(s), static_cast<int>(e), 0});
                s = e;
            }
        }
        items_.segFrom = from;
        items_.segTo = to;
        v.sfromSlot = deviceCopy(allocations_, fromSlot);
        v.stoSlot = deviceCopy(allocations_, toSlot);
        v.runNodes = deviceCopy(allocations_, runNodes);
        if (layout_.hangOrder.size() > static_cast<std::size_t>(kChainHang))
            throw std::invalid_argument("the device suspension sweep supports at most 32 harness points");
        std::vector<double> hang(3 * layout_.hangOrder.size());
        for (std::size_t r = 0; r < layout_.hangOrder.size(); ++r)
            for (int c = 0; c < 3; ++c) hang[3 * r + c] = d.suspension->harnessLocal[3 * layout_.hangOrder[r] + c];
        v.hangCount = static_cast<int>(layout_.hangOrder.size());
        v.hangLocal = deviceCopy(allocations_, hang);
    }
    // Contact-sphere pairs: dependency levels over definition order
    // (writes: the held-off node; reads: it and the centre).
    {
        std::vector<std::vector<std::uint32_t>> writes, reads;
        std::vector<int> node, centre;
        std::vector<double> radius, comp;
        for (std::size_t s = 0; s < d.sphereCount(); ++s)
            for (std::uint64_t j = d.sphereOffsets[s]; j < d.sphereOffsets[s + 1]; ++j) {
                node.push_back(static_cast<int>(d.sphereNodes[j]));
                centre.push_back(static_cast<int>(d.sphereCentre[s]));
                radius.push_back(d.sphereRadius[s]);
                comp.push_back(d.sphereCompliance[s]);
                writes.push_back({static_cast<std::uint32_t>(node.back())});
                reads.push_back({static_cast<std::uint32_t>(node.back()), static_cast<std::uint32_t>(centre.back())});
            }
        const auto level = dependencyLevels(writes, reads, std::vector<std::uint8_t>(writes.size(), 0), nodes_);
        std::vector<std::uint32_t> slots(writes.size());
        std::iota(slots.begin(), slots.end(), 0u);
        std::stable_sort(slots.begin(), slots.end(),
                         [&](std::uint32_t x, std::uint32_t y) { return level[x] < level[y]; });
        sphereDevicePosition_.assign(slots.size(), 0);
        for (std::size_t s = 0; s < slots.size(); ++s) sphereDevicePosition_[slots[s]] = static_cast<std::uint32_t>(s);
        v.pairCount = static_cast<int>(slots.size());
        items_.sphereNode = permute(node, slots);
        items_.sphereCentre = permute(centre, slots);
        v.spNode = deviceCopy(allocations_, items_.sphereNode);
        v.spCentre = deviceCopy(allocations_, items_.sphereCentre);
        v.spRadius = deviceCopy(allocations_, permute(radius, slots));
        v.spComp = deviceCopy(allocations_, permute(comp, slots));
        std::size_t s = 0;
        while (s < slots.size()) {
            std::size_t e = s;
            while (e < slots.size() && level[slots[e]] == level[slots[s]]) ++e;
            iterPhases_.push_back({kPhaseSpheres, static_cast<int>(s), static_cast<int>(e), 0});
            s = e;
        }
    }
    // Closing stitch pass: stitches in constraint order, as dependency
    // levels, addressed by their colour-major slot.
    {
        std::vector<std::uint32_t> slotOf(constraints_);
        for (std::size_t k = 0; k < distance_.order.size(); ++k) slotOf[distance_.order[k]] = static_cast<std::uint32_t>(k);
        std::vector<std::vector<std::uint32_t>> writes;
        std::vector<int> items;
        for (std::size_t c = 0; c < constraints_; ++c) {
            if (d.constraintKind[c] != SW_CONSTRAINT_SEAM_STITCH) continue;
            writes.push_back({static_cast<std::uint32_t>(d.constraintNodes[2 * c]),
                              static_cast<std::uint32_t>(d.constraintNodes[2 * c + 1])});
            items.push_back(static_cast<int>(slotOf[c]));
        }
        const auto level = dependencyLevels(writes, writes, std::vector<std::uint8_t>(writes.size(), 0), nodes_);
        std::vector<std::uint32_t> slots(items.size());
        std::iota(slots.begin(), slots.end(), 0u);
        std::stable_sort(slots.begin(), slots.end(),
                         [&](std::uint32_t x, std::uint32_t y) { return level[x] < level[y]; });
        stitchList_ = permute(items, slots);
        stitchLevels_.clear();
        std::size_t s = 0;
        while (s < slots.size()) {
            std::size_t e = s;
            while (e < slots.size() && level[slots[e]] == level[slots[s]]) ++e;
            stitchLevels_.push_back({kPhaseStitches, static_cast<int>(s), static_cast<int>(e), 0});
            s = e;
        }
    }
    // One load-path cable pair (soft_body.cpp:1265-1288): the load-path
    // order backwards, then forwards, as dependency levels of device slots.
    {
        std::vector<std::uint32_t> slotOf(constraints_);
        for (std::size_t k = 0; k < distance_.order.size(); ++k) slotOf[distance_.order[k]] = static_cast<std::uint32_t>(k);
        const std::vector<std::uint32_t> path = loadPathOrder(d);
        std::vector<std::uint32_t> sequence(path.rbegin(), path.rend());
        sequence.insert(sequence.end(), path.begin(), path.end());
        std::vector<std::vector<std::uint32_t>> writes;
        for (std::uint32_t c : sequence)
            writes.push_back({static_cast<std::uint32_t>(d.constraintNodes[2 * c]),
                              static_cast<std::uint32_t>(d.constraintNodes[2 * c + 1])});
        const auto level = dependencyLevels(writes, writes, std::vector<std::uint8_t>(writes.size(), 0), nodes_);
        std::vector<std::uint32_t> byLevel(sequence.size());
        std::iota(byLevel.begin(), byLevel.end(), 0u);
        std::stable_sort(byLevel.begin(), byLevel.end(),
                         [&](std::uint32_t x, std::uint32_t y) { return level[x] < level[y]; });
        cablePair_.clear();
        cablePairLevels_.clear();
        for (std::uint32_t i : byLevel) cablePair_.push_back(static_cast<int>(slotOf[sequence[i]]));
        std::size_t s = 0;
        while (s < byLevel.size()) {
            std::size_t e = s;
            while (e < byLevel.size() && level[byLevel[e]] == level[byLevel[s]]) ++e;
            cablePairLevels_.push_back({static_cast<int>(s), static_cast<int>(e)});
            s = e;
        }
        v.cableList = deviceCopy(allocations_, cablePair_);
    }
    hasRun_ = false;
    for (const Phase& P : iterPhases_) hasRun_ = hasRun_ || P.type == kPhaseSegmentRun;
    v.hasRun = hasRun_ ? 1 : 0;
    // The serial sweep's program: dependency levels over the definition
    // order of the distance constraints, then of the membranes, then the
    // coloured program's hinge, segment and sphere phases.
    {
        const auto levelList = [&](const ColourOrder& colours, std::size_t count, int width,
                                   const std::function<std::uint32_t(std::size_t, int)>& node, int type,
                                   std::vector<int>& list) {
            std::vector<std::uint32_t> slotOf(count);
            for (std::size_t k = 0; k < colours.order.size(); ++k) slotOf[colours.order[k]] = static_cast<std::uint32_t>(k);
            std::vector<std::vector<std::uint32_t>> writes(count);
            for (std::size_t i = 0; i < count; ++i)
                for (int c = 0; c < width; ++c) writes[i].push_back(node(i, c));
            const auto level = dependencyLevels(writes, std::vector<std::vector<std::uint32_t>>(count),
                                                std::vector<std::uint8_t>(count, 0), nodes_);
            std::vector<std::uint32_t> byLevel(count);
            std::iota(byLevel.begin(), byLevel.end(), 0u);
            std::stable_sort(byLevel.begin(), byLevel.end(),
                             [&](std::uint32_t x, std::uint32_t y) { return level[x] < level[y]; });
            list.clear();
            for (std::uint32_t i : byLevel) list.push_back(static_cast<int>(slotOf[i]));
            std::size_t s = 0;
            while (s < byLevel.size()) {
                std::size_t e = s;
                while (e < byLevel.size() && level[byLevel[e]] == level[byLevel[s]]) ++e;
                serialPhases_.push_back({type, static_cast<int>(s), static_cast<int>(e), 0});
                s = e;
            }
        };
        serialPhases_.clear();
        levelList(distance_, constraints_, 2,
                  [&](std::size_t i, int c) { return static_cast<std::uint32_t>(d.constraintNodes[2 * i + c]); },
                  kPhaseDistanceList, distList_);
        levelList(membrane_, membranes_, 3,
                  [&](std::size_t i, int c) {
                      return static_cast<std::uint32_t>(d.triangleNodes[3 * d.membraneTriangle[i] + c]);
                  },
                  kPhaseMembraneList, membList_);
        for (const Phase& P : iterPhases_)
            if (P.type != kPhaseDistance && P.type != kPhaseMembrane) serialPhases_.push_back(P);
        v.distList = deviceCopy(allocations_, distList_);
        v.membList = deviceCopy(allocations_, membList_);
    }
    colouredBase_ = iterPhases_;
    serialBase_ = serialPhases_;
    applyCablePrograms(0, 1);

    // Payload.
    {
        PayloadDev p{};
        if (d.suspension) {
            const SuspensionDef& s = *d.suspension;
            for (int k = 0; k < 3; ++k) {
                p.com[k] = s.comWorld[k];
                p.v[k] = s.velocity[k];
                p.w[k] = s.angularVelocity[k];
                p.comLocal[k] = s.comLocal[k];
                p.inertia[k] = s.inertia[k];
            }
            // validateAndNormalizeSuspensionDefinition normalises the
            // initial orientation (suspension.cpp:766-767).
            const double n2 = s.orientation[0] * s.orientation[0] + s.orientation[1] * s.orientation[1] +
                              s.orientation[2] * s.orientation[2] + s.orientation[3] * s.orientation[3];
            if (!(n2 > 1.0e-28)) throw SuspensionFailure(kPhaseValidation, "orientation", "quaternion is near zero");
            const double inv = 1.0 / std::sqrt(n2);
            double q[4] = {s.orientation[0] * inv, s.orientation[1] * inv, s.orientation[2] * inv, s.orientation[3] * inv};
            const bool flip = q[0] < 0.0 ||
                              (q[0] == 0.0 && (q[1] < 0.0 || (q[1] == 0.0 && (q[2] < 0.0 || (q[2] == 0.0 && q[3] < 0.0)))));
            for (int k = 0; k < 4; ++k) p.q[k] = flip ? -q[k] : q[k];
            for (int k = 0; k < 3; ++k) p.prevCom[k] = p.com[k];
            for (int k = 0; k < 4; ++k) p.prevQ[k] = p.q[k];
            for (int k = 0; k < 3; ++k) p.prevV[k] = p.v[k];
            for (int k = 0; k < 3; ++k) p.prevW[k] = p.w[k];
            p.mass = s.pilotMass;
            p.minLineLength = s.minimumLineLength;
        }
        check(hipMemcpy(v.payload, &p, sizeof(PayloadDev), hipMemcpyHostToDevice), "upload payload");
    }
    {
        DeviceError e{};
        std::vector<DeviceError> errors(1, e);
        v.error = deviceCopy(allocations_, errors);
        std::vector<unsigned> barrier(128, 0);  // count, generation, sync-failure flag
        v.barrier = deviceCopy(allocations_, barrier);
    }
    // Launch shape: one workgroup with LDS-resident nodes for a small wing
    // whose nodes fit beside the kernel's static LDS (the serial-run
    // staging); else nodes stay in HBM and a cooperative grid of up to one
    // workgroup per compute unit runs the phases with point-to-point
    // synchronisation: 16 workgroups up to ~50k items, 64 beyond (measured
    // on the 1.7k-, 6.6k- and 43k-node flight wings).
    const std::size_t dynamicLds = (4 * nodes_ + kMaxThreads) * sizeof(double);
    sharedNodes_ = dynamicLds <= sharedNodeBudget() && itemCount() <= 4096;
    workgroups_ = 1;
    threads_ = kMaxThreads;
    if (!sharedNodes_) {
        int cus = 0;
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_), "attribute");
        workgroups_ = std::max(1, std::min(cus, itemCount() < 50000 ? 16 : 64));
    }
    if (dynamicLds <= sharedNodeBudget())  // also for a later single-workgroup setLaunch
        for (bool general : {false, true})
            check(hipFuncSetAttribute(substepKernelFor(true, false, general), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(dynamicLds)),
                  "shared memory attribute");
    {
        int cus = 0;
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_), "attribute");
        maxParticipants_ = std::max(cus, 1);
        v.flags = deviceCopy(allocations_, std::vector<unsigned>(32 * static_cast<std::size_t>(maxParticipants_), 0u));
    }
    buildSync();
    setXcdLocal(xcdRequested_);
    check(hipMalloc(&partials_, 64 * sizeof(Partial)), "diagnostics partials");
    allocations_.push_back(partials_);
    {
        // Frame-order maps of frameFieldsKernel.
        std::vector<int> lineSlot(constraints_ + segments_, 0);
        std::vector<int> hangSlot(layout_.hangOrder.size(), 0);
        for (std::size_t k = 0; k < constraints_; ++k) lineSlot[distance_.order[k]] = static_cast<int>(k);
        for (std::size_t r = 0; r < segments_; ++r)
            lineSlot[constraints_ + layout_.segmentOrder[r]] =
                static_cast<int>(constraints_ + segmentDevicePosition_[r]);
        for (std::size_t r = 0; r < layout_.hangOrder.size(); ++r)
            hangSlot[layout_.hangOrder[r]] = static_cast<int>(r);
        frameLineSlot_ = deviceCopy(allocations_, lineSlot);
        frameHangSlot_ = deviceCopy(allocations_, hangSlot);
    }
    check(hipStreamSynchronize(stream_), "sync build");
    check(hipDeviceSynchronize(), "sync build");
}

// Dynamic LDS available to an LDS-resident wing: the per-workgroup maximum
// less the static LDS of the kernels that run it (the solo LDS kernel and the
// batched kernel).
std::size_t DeviceWing::sharedNodeBudget() const {
    int maxShared = 0;
    check(hipDeviceGetAttribute(&maxShared, hipDeviceAttributeMaxSharedMemoryPerBlock, device_), "attribute");
    std::size_t fixed = 0;
    for (bool general : {false, true}) {
        hipFuncAttributes solo{}, batch{};
        check(hipFuncGetAttributes(&solo, substepKernelFor(true, false, general)), "kernel attributes");
        check(hipFuncGetAttributes(&batch, batchKernelFor(general)), "kernel attributes");
        fixed = std::max({fixed, solo.sharedSizeBytes, batch.sharedSizeBytes});
    }
    const std::size_t total = static_cast<std::size_t>(maxShared);
    return total > fixed ? total - fixed : 0;
}

void DeviceWing::setLaunch(int workgroups, int threads) {
    ++launchGeneration_;
    if (workgroups < 1 || threads < 64 || threads > kMaxThreads || (threads & (threads - 1)) != 0)
        throw std::invalid_argument("launch shape must be >= 1 workgroup of 64..256 (power of two) threads");
    int cus = 0;
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_), "attribute");
    if (workgroups > cus) throw std::invalid_argument("launch shape exceeds one workgroup per compute unit");
    workgroups_ = workgroups;
    threads_ = threads;
    const std::size_t dynamicLds = (4 * nodes_ + kMaxThreads) * sizeof(double);
    sharedNodes_ = workgroups == 1 && dynamicLds <= sharedNodeBudget();
    if (sharedNodes_) {
        for (bool general : {false, true})
            check(hipFuncSetAttribute(substepKernelFor(true, false, general), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(dynamicLds)),
                  "shared memory attribute");
    }
    buildSync();
    setXcdLocal(xcdRequested_);
}

// The programs for `pairs` load-path cable sweep pairs and `iterations`
// iterations (soft_body.cpp:1296-1315, 1380-1388): the interleaved pairs as
// cable phases at the start of the iteration program (pair j of an
// iteration runs when j < the iteration's share), the closing pairs and
// then the seam-stitch levels as the closing program (with no iterations,
// every pair closes and no stitch pass runs).
void DeviceWing::applyCablePrograms(int pairs, int iterations) {
    WingView& v = *view_;
    cablePairs_ = pairs;
    cableIterations_ = iterations;
    int closing = pairs > 0 ? pairs - pairs / 4 : 0;
    int interleaved = pairs - closing;
    if (iterations == 0) {
        closing = pairs;
        interleaved = 0;
    }
    if (cablePair_.empty()) closing = interleaved = 0;
    for (void* p : programAlloc_) (void)hipFree(p);
    programAlloc_.clear();
    // Closing program.
    std::vector<int> closeList;
    closePhases_.clear();
    for (int c = 0; c < closing; ++c) {
        const int base = static_cast<int>(closeList.size());
        closeList.insert(closeList.end(), cablePair_.begin(), cablePair_.end());
        for (const auto& L : cablePairLevels_) closePhases_.push_back({kPhaseStitches, base + L.first, base + L.second, 0});
    }
    if (iterations > 0) {
        const int base = static_cast<int>(closeList.size());
        closeList.insert(closeList.end(), stitchList_.begin(), stitchList_.end());
        for (const Phase& P : stitchLevels_) closePhases_.push_back({kPhaseStitches, base + P.begin, base + P.end, 0});
    }
    v.stitchCount = static_cast<int>(closeList.size());
    v.stitchItems = deviceCopy(programAlloc_, closeList);
    v.closePhases = deviceCopy(programAlloc_, closePhases_);
    v.closePhaseCount = static_cast<int>(closePhases_.size());
    // Iteration programs.
    std::vector<Phase> prefix;
    const int iters = std::max(1, iterations);
    const int share = interleaved / iters + (interleaved % iters > 0 ? 1 : 0);
    for (int j = 0; j < share; ++j)
        for (const auto& L : cablePairLevels_) prefix.push_back({kPhaseCableList, L.first, L.second, j});
    cableInterleaved_ = interleaved;
    colouredPhases_ = prefix;
    colouredPhases_.insert(colouredPhases_.end(), colouredBase_.begin(), colouredBase_.end());
    serialPhases_ = prefix;
    serialPhases_.insert(serialPhases_.end(), serialBase_.begin(), serialBase_.end());
    colouredPhasesDev_ = deviceCopy(programAlloc_, colouredPhases_);
    serialPhasesDev_ = deviceCopy(programAlloc_, serialPhases_);
    iterPhases_ = serialSweep_ ? serialPhases_ : colouredPhases_;
    v.iterPhases = serialSweep_ ? serialPhasesDev_ : colouredPhasesDev_;
    v.iterPhaseCount = static_cast<int>(iterPhases_.size());
}

void DeviceWing::setCablePairs(int pairs, int iterations) {
    if (cablePair_.empty()) pairs = 0;  // no load-path constraint: the pairs do nothing
    if (pairs == cablePairs_ && (iterations > 0) == (cableIterations_ > 0) &&
        (pairs == 0 || iterations == cableIterations_))
        return;
    check(hipSetDevice(device_), "hipSetDevice");
    check(hipStreamSynchronize(stream_), "synchronise before rebuilding the programs");
    applyCablePrograms(pairs, iterations);
    ++launchGeneration_;
    buildSync();
    setXcdLocal(xcdRequested_);
}

void DeviceWing::setSerialSweep(bool serial) {
    if (serial == serialSweep_) return;
    check(hipSetDevice(device_), "hipSetDevice");
    check(hipStreamSynchronize(stream_), "synchronise before switching the sweep");
    serialSweep_ = serial;
    iterPhases_ = serial ? serialPhases_ : colouredPhases_;
    view_->iterPhases = serial ? serialPhasesDev_ : colouredPhasesDev_;
    view_->iterPhaseCount = static_cast<int>(iterPhases_.size());
    ++launchGeneration_;
    buildSync();
    setXcdLocal(xcdRequested_);
}

bool DeviceWing::setXcdLocal(bool enabled) {
    xcdRequested_ = enabled;
    xcdLocal_ = false;
    if (!enabled || sharedNodes_ || workgroups_ < 2) return !enabled;
    check(hipSetDevice(device_), "hipSetDevice");
    int cus = 0, perCu = 0;
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_), "attribute");
    const std::size_t scratch = kMaxThreads * sizeof(double) + (neighbourSync_ ? stageBytes_ : 0);
    perCu = 1 << 30;
    for (bool general : {false, true}) {
        int blocks = 0;
        check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, substepKernelFor(false, true, general), threads_,
                                                           scratch),
              "occupancy");
        perCu = std::min(perCu, blocks);
    }
    // Every workgroup of the grid resident at once, and one XCD's share of
    // the compute units able to hold all participants.
    const long grid = static_cast<long>(kXcdLaunchFactor) * workgroups_;
    if (grid > static_cast<long>(perCu) * cus || workgroups_ > perCu * (cus / kXcdLaunchFactor)) return false;
    xcdLocal_ = true;
    return true;
}

void DeviceWing::setExternalForces(const std::vector<double>& forces) {
    WingView& v = *view_;
    std::vector<double> x(nodes_), y(nodes_), z(nodes_);
    for (std::size_t i = 0; i < nodes_; ++i) {
        x[i] = forces[3 * i];
        y[i] = forces[3 * i + 1];
        z[i] = forces[3 * i + 2];
    }
    check(hipMemcpyAsync(v.fx, x.data(), nodes_ * 8, hipMemcpyHostToDevice, stream_), "forces");
    check(hipMemcpyAsync(v.fy, y.data(), nodes_ * 8, hipMemcpyHostToDevice, stream_), "forces");
    check(hipMemcpyAsync(v.fz, z.data(), nodes_ * 8, hipMemcpyHostToDevice, stream_), "forces");
    check(hipStreamSynchronize(stream_), "forces");
}

void DeviceWing::setFacePressure(const std::vector<double>& pressure) {
    check(hipMemcpyAsync(view_->pressure, pressure.data(), triangles_ * 8, hipMemcpyHostToDevice, stream_), "pressure");
    check(hipStreamSynchronize(stream_), "pressure");
}

void DeviceWing::launchSubstep(const StepArgs& stepArgs, double* trace) {
    WingView v = *view_;
    v.trace = trace;
    if (v.profile) ++profiledLaunches_;
    v.neighbourSync = neighbourSync_ ? 1 : 0;
    v.dedicatedChain = dedicatedChain_ ? 1 : 0;
    v.sharedNodes = sharedNodes_ ? 1 : 0;
    v.launchBlocks = sharedNodes_ ? 1 : workgroups_;
    StepArgs args = stepArgs;
    if (timing_) {
        while (events_.size() < eventsUsed_ + 2) {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            events_.push_back(e);
        }
        check(hipEventRecord(events_[eventsUsed_], stream_), "hipEventRecord");
    }
    const std::size_t scratch = kMaxThreads * sizeof(double) + (neighbourSync_ ? stageBytes_ : 0);
    const bool general = needsGeneralKernel();
    void* params[] = {&v, &args};
    if (sharedNodes_) {
        const std::size_t bytes = 4 * nodes_ * sizeof(double) + scratch;
        check(hipLaunchKernel(substepKernelFor(true, false, general), dim3(1), dim3(threads_), params, bytes, stream_),
              "substep launch");
    } else if (workgroups_ == 1) {
        check(hipLaunchKernel(substepKernelFor(false, false, general), dim3(1), dim3(threads_), params, scratch,
                              stream_),
              "substep launch");
    } else {
        int grid = workgroups_;
        if (xcdLocal_) {
            // One-XCD launch: 8 workgroups per participant, so that one XCD
            // receives `workgroups_` of them (xcdParticipant).
            v.xcdLocal = 1;
            epoch_ = (epoch_ + 1u) & 0xfffffu;
            if (epoch_ == 0) epoch_ = 1;
            v.epoch = epoch_;
            grid = kXcdLaunchFactor * workgroups_;
        }
        const void* kernel = substepKernelFor(false, xcdLocal_, general);
        if (!cooperative_) {
            check(hipLaunchKernel(kernel, dim3(grid), dim3(threads_), params, scratch, stream_), "substep launch");
        } else {
            check(hipLaunchCooperativeKernel(kernel, dim3(grid), dim3(threads_), params, scratch, stream_),
                  "cooperative launch");
        }
    }
    check(hipGetLastError(), "substep launch");
    if (timing_) {
        check(hipEventRecord(events_[eventsUsed_ + 1], stream_), "hipEventRecord");
        eventsUsed_ += 2;
    }
}

void DeviceWing::snapshot() {
    hipLaunchKernelGGL(snapshotKernel, dim3(64), dim3(256), 0, stream_, *view_, 0);
    check(hipGetLastError(), "snapshot");
}

void DeviceWing::restoreSnapshot() {
    hipLaunchKernelGGL(snapshotKernel, dim3(64), dim3(256), 0, stream_, *view_, 1);
    check(hipGetLastError(), "restore");
    check(hipStreamSynchronize(stream_), "restore");
}

int DeviceWing::snapshotsTaken() {
    unsigned taken = 0;
    check(hipMemcpyAsync(&taken, view_->barrier + 120, sizeof(unsigned), hipMemcpyDeviceToHost, stream_),
          "snapshot count");
    check(hipStreamSynchronize(stream_), "snapshot count");
    return static_cast<int>(taken);
}

void DeviceWing::clearError() {
    // Also resets the synchronisation words (they are zero after every
    // completed launch, but not after a launch whose waits timed out).
    check(hipMemsetAsync(view_->error, 0, sizeof(DeviceError), stream_), "clear error");
    check(hipMemsetAsync(view_->barrier, 0, 128 * sizeof(unsigned), stream_), "clear barrier");
    check(hipMemsetAsync(view_->flags, 0, 32 * sizeof(unsigned) * maxParticipants_, stream_), "clear progress");
}

void DeviceWing::itemAccess(int type, int k, std::vector<int>& reads, std::vector<int>& writes) const {
    reads.clear();
    writes.clear();
    const int payloadNode = static_cast<int>(nodes_);
    switch (type) {
    case kPhaseDistance:
        writes.assign(items_.dist.begin() + 2 * k, items_.dist.begin() + 2 * k + 2);
        break;
    case kPhaseMembrane:
        writes.assign(items_.memb.begin() + 3 * k, items_.memb.begin() + 3 * k + 3);
        break;
    case kPhaseDistanceList:
        writes.assign(items_.dist.begin() + 2 * distList_[k], items_.dist.begin() + 2 * distList_[k] + 2);
        break;
    case kPhaseCableList:
        writes.assign(items_.dist.begin() + 2 * cablePair_[k], items_.dist.begin() + 2 * cablePair_[k] + 2);
        break;
    case kPhaseMembraneList:
        writes.assign(items_.memb.begin() + 3 * membList_[k], items_.memb.begin() + 3 * membList_[k] + 3);
        break;
    case kPhaseHinge:
        writes.assign(items_.hinge.begin() + 4 * k, items_.hinge.begin() + 4 * k + 4);
        break;
    case kPhaseSegments:
    case kPhaseSegmentRun:
        for (int e : {items_.segFrom[k], items_.segTo[k]}) writes.push_back(e >= 0 ? e : payloadNode);
        break;
    case kPhaseSpheres:
        writes.push_back(items_.sphereNode[k]);
        reads.push_back(items_.sphereCentre[k]);
        break;
    default:
        throw std::logic_error("phase type outside the iteration program");
    }
}

// Dependency lists of the point-to-point synchronisation. Workgroup p owns
// items [n p / W, n (p + 1) / W) of every phase of n items (W workers); the
// serial runs belong to the dedicated chain workgroup (participant W) when
// there is one, else to workgroup 0. Replaying two iterations of the program
// on the host, every access of p at step s to node x (the payload counts as a
// node) records the last other participant that wrote x and, for a write,
// every other participant that read x since; p must see their progress
// reach that step before it starts. The second iteration gives the periodic
// lists, stored as lags (steps back); lags reaching before the first step
// refer to the prologue, which a grid barrier already orders.
void DeviceWing::buildSync() {
    WingView& v = *view_;
    check(hipSetDevice(device_), "hipSetDevice");
    check(hipStreamSynchronize(stream_), "synchronise before rebuilding the schedule");
    const int m = static_cast<int>(iterPhases_.size());
    neighbourSync_ = !sharedNodes_ && workgroups_ > 1 && m > 0 && workgroups_ <= maxParticipants_;
    dedicatedChain_ = neighbourSync_ && hasRun_ && workgroups_ > 2;
    for (void*& p : depAlloc_) {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    v.depOffsets = v.depPart = v.depLag = nullptr;
    participants_ = 0;
    stageBytes_ = 0;
    if (!neighbourSync_) return;
    const int parts = workgroups_;
    const int workers = dedicatedChain_ ? parts - 1 : parts;
    const int runOwner = dedicatedChain_ ? workers : 0;
    struct Access {
        int p, s;
    };
    const std::size_t slots = nodes_ + 1;
    std::vector<Access> lastWrite(slots, Access{-1, -1});
    std::vector<std::vector<Access>> readers(slots);
    std::vector<int> needStep(static_cast<std::size_t>(parts) * m * parts, -1);
    std::vector<int> reads, writes;
    struct Pending {
        int node, p;
        bool write;
    };
    std::vector<Pending> pending;
    for (int s = 0; s < 2 * m; ++s) {
        const int ph = s % m;
        const Phase& P = iterPhases_[ph];
        pending.clear();
        for (int p = 0; p < parts; ++p) {
            int b = P.begin, e = P.begin;
            if (P.type == kPhaseSegmentRun) {
                if (p == runOwner) e = P.end;
            } else if (p < workers) {
                const long long count = P.end - P.begin;
                b = P.begin + static_cast<int>(count * p / workers);
                e = P.begin + static_cast<int>(count * (p + 1) / workers);
            }
            int* need = needStep.data() + (static_cast<std::size_t>(p) * m + ph) * parts;
            const auto depend = [&](const Access& a) {
                if (s >= m && a.p >= 0 && a.p != p) need[a.p] = std::max(need[a.p], a.s);
            };
            for (int k = b; k < e; ++k) {
                itemAccess(P.type, k, reads, writes);
                for (int x : writes) {
                    depend(lastWrite[x]);
                    for (const Access& r : readers[x]) depend(r);
                    pending.push_back({x, p, true});
                }
                for (int x : reads) {
                    depend(lastWrite[x]);
                    pending.push_back({x, p, false});
                }
            }
        }
        for (const Pending& a : pending) {
            if (a.write) {
                lastWrite[a.node] = Access{a.p, s};
                readers[a.node].clear();
            }
        }
        for (const Pending& a : pending) {
            if (a.write) continue;
            auto& list = readers[a.node];
            bool found = false;
            for (Access& r : list)
                if (r.p == a.p) {
                    r.s = s;
                    found = true;
                }
            if (!found) list.push_back(Access{a.p, s});
        }
    }
    std::vector<int> offsets(static_cast<std::size_t>(parts) * m + 1, 0), part, lag;
    int maxDeps = 0;
    for (int p = 0; p < parts; ++p) {
        const std::size_t first = part.size();
        for (int ph = 0; ph < m; ++ph) {
            const int* need = needStep.data() + (static_cast<std::size_t>(p) * m + ph) * parts;
            for (int q = 0; q < parts; ++q)
                if (need[q] >= 0) {
                    const int l = (m + ph) - need[q];
                    if (l < 1 || l > 2 * m) throw std::logic_error("synchronisation analysis produced an invalid lag");
                    part.push_back(q);
                    lag.push_back(l);
                }
            offsets[static_cast<std::size_t>(p) * m + ph + 1] = static_cast<int>(part.size());
        }
        maxDeps = std::max(maxDeps, static_cast<int>(part.size() - first));
    }
    if (stagedProgramBytes(m, maxDeps) > kMaxStagedProgramBytes) {
        // The loop stages each participant's program in LDS; a program
        // beyond that capacity runs with grid barriers instead.
        neighbourSync_ = false;
        dedicatedChain_ = false;
        return;
    }
    stageBytes_ = stagedProgramBytes(m, maxDeps);
    const auto upload = [&](const std::vector<int>& host, int index) {
        void* ptr = nullptr;
        const std::size_t bytes = std::max<std::size_t>(host.size(), 1) * sizeof(int);
        check(hipMalloc(&ptr, bytes), "hipMalloc");
        depAlloc_[index] = ptr;
        if (!host.empty()) check(hipMemcpy(ptr, host.data(), host.size() * sizeof(int), hipMemcpyHostToDevice), "upload");
        return static_cast<int*>(ptr);
    };
    v.depOffsets = upload(offsets, 0);
    v.depPart = upload(part, 1);
    v.depLag = upload(lag, 2);
    participants_ = parts;
}

DeviceError DeviceWing::readError(double* maxResidual) {
    DeviceError e{};
    check(hipMemcpyAsync(&e, view_->error, sizeof(DeviceError), hipMemcpyDeviceToHost, stream_), "read error");
    double residual = 0.0;
    check(hipMemcpyAsync(&residual, &view_->payload->maxResidual, sizeof(double), hipMemcpyDeviceToHost, stream_),
          "read residual");
    check(hipStreamSynchronize(stream_), "read error");
    if (maxResidual) *maxResidual = residual;
    return e;
}

void DeviceWing::synchronize() { check(hipStreamSynchronize(stream_), "synchronize"); }

DeviceWing::State DeviceWing::download() const {
    const WingView& v = *view_;
    std::vector<double> block(v.stateDoubles);
    check(hipMemcpyAsync(block.data(), v.state, snapshotBytes_, hipMemcpyDeviceToHost, stream_), "download");
    check(hipStreamSynchronize(stream_), "download");
    State s;
    const double* base = block.data();
    const auto at = [&](const double* dev) { return base + (dev - v.state); };
    s.position.resize(3 * nodes_);
    s.previous.resize(3 * nodes_);
    s.velocity.resize(3 * nodes_);
    for (std::size_t i = 0; i < nodes_; ++i) {
        s.position[3 * i] = at(v.px)[i];
        s.position[3 * i + 1] = at(v.py)[i];
        s.position[3 * i + 2] = at(v.pz)[i];
        s.previous[3 * i] = at(v.qx)[i];
        s.previous[3 * i + 1] = at(v.qy)[i];
        s.previous[3 * i + 2] = at(v.qz)[i];
        s.velocity[3 * i] = at(v.vx)[i];
        s.velocity[3 * i + 1] = at(v.vy)[i];
        s.velocity[3 * i + 2] = at(v.vz)[i];
    }
    s.constraintLambda.resize(constraints_);
    for (std::size_t k = 0; k < constraints_; ++k) s.constraintLambda[distance_.order[k]] = at(v.dlam)[k];
    s.membraneLambda.resize(3 * membranes_);
    s.membraneResultant.resize(3 * membranes_);
    s.membraneResidual.resize(membranes_);
    for (std::size_t k = 0; k < membranes_; ++k) {
        const std::size_t m = membrane_.order[k];
        s.membraneLambda[3 * m] = at(v.mlx)[k];
        s.membraneLambda[3 * m + 1] = at(v.mly)[k];
        s.membraneLambda[3 * m + 2] = at(v.mlz)[k];
        s.membraneResultant[3 * m] = at(v.mrx)[k];
        s.membraneResultant[3 * m + 1] = at(v.mry)[k];
        s.membraneResultant[3 * m + 2] = at(v.mrz)[k];
        s.membraneResidual[m] = at(v.mres)[k];
    }
    s.dihedralLambda.resize(hinges_);
    for (std::size_t k = 0; k < hinges_; ++k) s.dihedralLambda[hinge_.order[k]] = at(v.hlam)[k];
    s.segmentLambda.resize(segments_);
    for (std::size_t k = 0; k < segments_; ++k) s.segmentLambda[k] = at(v.slam)[segmentDevicePosition_[k]];
    s.sphereLambda.resize(pairs_);
    for (std::size_t k = 0; k < pairs_; ++k) s.sphereLambda[k] = at(v.splam)[sphereDevicePosition_[k]];
    s.pressure.assign(at(v.pressure), at(v.pressure) + triangles_);
    const double* p = at(reinterpret_cast<double*>(v.payload));
    const PayloadDev* pd = reinterpret_cast<const PayloadDev*>(p);
    s.payload.assign(26, 0.0);
    for (int k = 0; k < 3; ++k) {
        s.payload[k] = pd->com[k];
        s.payload[7 + k] = pd->v[k];
        s.payload[10 + k] = pd->w[k];
        s.payload[13 + k] = pd->prevCom[k];
        s.payload[20 + k] = pd->prevV[k];
        s.payload[23 + k] = pd->prevW[k];
    }
    for (int k = 0; k < 4; ++k) {
        s.payload[3 + k] = pd->q[k];
        s.payload[16 + k] = pd->prevQ[k];
    }
    return s;
}

void DeviceWing::upload(const State& s) {
    const WingView& v = *view_;
    std::vector<double> block(v.stateDoubles);
    check(hipMemcpyAsync(block.data(), v.state, snapshotBytes_, hipMemcpyDeviceToHost, stream_), "upload read");
    check(hipStreamSynchronize(stream_), "upload read");
    double* base = block.data();
    const auto at = [&](double* dev) { return base + (dev - v.state); };
    for (std::size_t i = 0; i < nodes_; ++i) {
        at(v.px)[i] = s.position[3 * i];
        at(v.py)[i] = s.position[3 * i + 1];
        at(v.pz)[i] = s.position[3 * i + 2];
        at(v.qx)[i] = s.previous[3 * i];
        at(v.qy)[i] = s.previous[3 * i + 1];
        at(v.qz)[i] = s.previous[3 * i + 2];
        at(v.vx)[i] = s.velocity[3 * i];
        at(v.vy)[i] = s.velocity[3 * i + 1];
        at(v.vz)[i] = s.velocity[3 * i + 2];
    }
    for (std::size_t k = 0; k < constraints_; ++k) at(v.dlam)[k] = s.constraintLambda[distance_.order[k]];
    for (std::size_t k = 0; k < membranes_; ++k) {
        const std::size_t m = membrane_.order[k];
        at(v.mlx)[k] = s.membraneLambda[3 * m];
        at(v.mly)[k] = s.membraneLambda[3 * m + 1];
        at(v.mlz)[k] = s.membraneLambda[3 * m + 2];
        at(v.mrx)[k] = s.membraneResultant[3 * m];
        at(v.mry)[k] = s.membraneResultant[3 * m + 1];
        at(v.mrz)[k] = s.membraneResultant[3 * m + 2];
        at(v.mres)[k] = s.membraneResidual[m];
    }
    for (std::size_t k = 0; k < hinges_; ++k) at(v.hlam)[k] = s.dihedralLambda[hinge_.order[k]];
    for (std::size_t k = 0; k < segments_; ++k) at(v.slam)[segmentDevicePosition_[k]] = s.segmentLambda[k];
    for (std::size_t k = 0; k < pairs_; ++k) at(v.splam)[sphereDevicePosition_[k]] = s.sphereLambda[k];
    std::copy(s.pressure.begin(), s.pressure.end(), at(v.pressure));
    PayloadDev* pd = reinterpret_cast<PayloadDev*>(at(reinterpret_cast<double*>(v.payload)));
    for (int k = 0; k < 3; ++k) {
        pd->com[k] = s.payload[k];
        pd->v[k] = s.payload[7 + k];
        pd->w[k] = s.payload[10 + k];
        pd->prevCom[k] = s.payload[13 + k];
        pd->prevV[k] = s.payload[20 + k];
        pd->prevW[k] = s.payload[23 + k];
    }
    for (int k = 0; k < 4; ++k) {
        pd->q[k] = s.payload[3 + k];
        pd->prevQ[k] = s.payload[16 + k];
    }
    check(hipMemcpyAsync(v.state, block.data(), snapshotBytes_, hipMemcpyHostToDevice, stream_), "upload");
    check(hipStreamSynchronize(stream_), "upload");
}

DeviceWing::Reductions DeviceWing::reduce() {
    constexpr int blocks = 64;
    Partial* partial = static_cast<Partial*>(partials_);
    hipLaunchKernelGGL(diagnosticsKernel, dim3(blocks), dim3(256), 0, stream_, *view_, partial);
    check(hipGetLastError(), "diagnostics");
    std::vector<Partial> host(blocks);
    check(hipMemcpyAsync(host.data(), partial, blocks * sizeof(Partial), hipMemcpyDeviceToHost, stream_), "diagnostics");
    check(hipStreamSynchronize(stream_), "diagnostics");
    Reductions r{};
    r.finite = 1;
    for (const Partial& p : host) {
        r.mass += p.mass;
        r.weighted[0] += p.wx;
        r.weighted[1] += p.wy;
        r.weighted[2] += p.wz;
        r.momentum[0] += p.mx;
        r.momentum[1] += p.my;
        r.momentum[2] += p.mz;
        r.kinetic += p.kinetic;
        r.maxDistanceError = std::max(r.maxDistanceError, p.maxDistance);
        r.maxCableExtension = std::max(r.maxCableExtension, p.maxCable);
        r.maxStrain = std::max(r.maxStrain, p.maxStrain);
        r.maxResidual = std::max(r.maxResidual, p.maxResidual);
        r.finite &= p.finite;
        r.dynamicNodes += p.dynamicNodes;
    }
    return r;
}

DeviceWing::Frame DeviceWing::frame() {
    const WingView& v = *view_;
    const std::size_t h = static_cast<std::size_t>(v.hangCount);
    const std::size_t doubles = 6 * h + 2 * (constraints_ + segments_);
    double* out = nullptr;
    check(hipMalloc(&out, std::max<std::size_t>(doubles, 1) * sizeof(double)), "hipMalloc frame");
    std::vector<double> host(doubles);
    try {
        const long work = static_cast<long>(std::max({h, constraints_, segments_, std::size_t{1}}));
        const int blocks = static_cast<int>(std::min<long>(1024, (work + 255) / 256));
        hipLaunchKernelGGL(frameKernel, dim3(blocks), dim3(256), 0, stream_, v, out);
        check(hipGetLastError(), "frame");
        check(hipMemcpyAsync(host.data(), out, doubles * sizeof(double), hipMemcpyDeviceToHost, stream_), "frame");
        check(hipStreamSynchronize(stream_), "frame");
    } catch (...) {
        (void)hipFree(out);
        throw;
    }
    (void)hipFree(out);
    Frame f;
    // Harness points in definition order; lines in definition order.
    f.harnessPosition.assign(3 * h, 0.0);
    f.harnessVelocity.assign(3 * h, 0.0);
    for (std::size_t r = 0; r < h; ++r)
        for (int c = 0; c < 3; ++c) {
            f.harnessPosition[3 * layout_.hangOrder[r] + c] = host[3 * r + c];
            f.harnessVelocity[3 * layout_.hangOrder[r] + c] = host[3 * h + 3 * r + c];
        }
    const double* length = host.data() + 6 * h;
    const double* violation = length + constraints_ + segments_;
    f.lineLength.assign(constraints_ + segments_, 0.0);
    f.lineViolation.assign(constraints_ + segments_, 0.0);
    for (std::size_t k = 0; k < constraints_; ++k) {
        f.lineLength[distance_.order[k]] = length[k];
        f.lineViolation[distance_.order[k]] = violation[k];
    }
    for (std::size_t r = 0; r < segments_; ++r) {
        const std::size_t slot = segmentDevicePosition_[r];
        f.lineLength[constraints_ + layout_.segmentOrder[r]] = length[constraints_ + slot];
        f.lineViolation[constraints_ + layout_.segmentOrder[r]] = violation[constraints_ + slot];
    }
    return f;
}

std::size_t DeviceWing::frameFieldsDoubles() const {
    const std::size_t vertices = nodes_ + static_cast<std::size_t>(view_->hangCount);
    return 6 * vertices + 2 * (constraints_ + segments_);
}

void DeviceWing::frameFields(double* staging, double* position, double* velocity, double* length,
                             double* violation) {
    const WingView& v = *view_;
    const std::size_t vertices = nodes_ + static_cast<std::size_t>(v.hangCount);
    const std::size_t lines = constraints_ + segments_;
    check(hipSetDevice(device_), "hipSetDevice");
    const long work = static_cast<long>(std::max({vertices, lines, std::size_t{1}}));
    const int blocks = static_cast<int>(std::min<long>(1024, (work + 255) / 256));
    hipLaunchKernelGGL(frameFieldsKernel, dim3(blocks), dim3(256), 0, stream_, v, frameLineSlot_, frameHangSlot_,
                       staging);
    check(hipGetLastError(), "frame fields");
    const struct {
        double* host;
        std::size_t offset, count;
    } copies[] = {{position, 0, 3 * vertices},
                  {velocity, 3 * vertices, 3 * vertices},
                  {length, 6 * vertices, lines},
                  {violation, 6 * vertices + lines, lines}};
    for (const auto& c : copies)
        if (c.host && c.count)
            check(hipMemcpyAsync(c.host, staging + c.offset, c.count * sizeof(double), hipMemcpyDeviceToHost,
                                 stream_),
                  "frame fields");
    check(hipStreamSynchronize(stream_), "frame fields");
}

WingView DeviceWing::launchView() const {
    WingView v = *view_;
    v.neighbourSync = neighbourSync_ ? 1 : 0;
    v.dedicatedChain = dedicatedChain_ ? 1 : 0;
    v.sharedNodes = sharedNodes_ ? 1 : 0;
    v.launchBlocks = sharedNodes_ ? 1 : workgroups_;
    return v;
}

std::size_t DeviceWing::dynamicLdsBytes() const {
    return (sharedNodes_ ? 4 * nodes_ * sizeof(double) : 0) + kMaxThreads * sizeof(double) +
           (neighbourSync_ ? stageBytes_ : 0);
}

bool DeviceWing::fitsInLds() const { return (4 * nodes_ + kMaxThreads) * sizeof(double) <= sharedNodeBudget(); }

// ---- batched wings ----------------------------------------------------------

DeviceBatch::DeviceBatch(std::vector<DeviceWing*> wings, int device) : wings_(std::move(wings)), device_(device) {
    if (wings_.empty()) throw std::invalid_argument("a batch needs at least one wing");
    check(hipSetDevice(device_), "hipSetDevice");
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
}

DeviceBatch::~DeviceBatch() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    release();
    if (stream_) (void)hipStreamDestroy(stream_);
}

void DeviceBatch::release() {
    for (void** p : {&views_, &args_, &blockWing_, &blockLocal_, &mask_}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
}

void DeviceBatch::plan(int cap, int threads) {
    check(hipSetDevice(device_), "hipSetDevice");
    check(hipStreamSynchronize(stream_), "synchronise before planning");
    if (threads < 64 || threads > kMaxThreads || (threads & (threads - 1)) != 0)
        throw std::invalid_argument("batch threads must be a power of two in 64..256");
    int cus = 0;
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_), "attribute");
    const int count = static_cast<int>(wings_.size());
    if (count > cus) throw std::invalid_argument("a batch holds at most one wing per compute unit");
    cap = std::max(cap, 1);
    // One LDS-resident workgroup per small wing. The large wings first get
    // a floor of three workgroups (workers plus the chain workgroup); the
    // remaining compute units are then dealt out one at a time, longest-
    // processing-time first: each goes to the large wing with the most items
    // per workgroup so far (ties to the lower wing index), until none are
    // left or every large wing is at the cap. The batched step lasts as long
    // as its slowest wing, and a wing's phases take longer the more items
    // each of its threads holds, so this levels the wings' phase loads.
    blocks_.assign(count, 1);
    std::vector<int> large;
    for (int w = 0; w < count; ++w)
        if (!wings_[w]->fitsInLds() || cap == 1) large.push_back(w);
    const int floor = std::min(3, cap);
    int spare = cus - count;
    for (int w : large) {
        const int add = std::max(0, std::min(floor - 1, spare));
        blocks_[w] += add;
        spare -= add;
    }
    const auto load = [&](int w) { return static_cast<double>(wings_[w]->itemCount()) / blocks_[w]; };
    while (spare > 0) {
        int best = -1;
        for (int w : large)
            if (blocks_[w] < cap && (best < 0 || load(w) > load(best))) best = w;
        if (best < 0) break;
        ++blocks_[best];
        --spare;
    }
    lds_ = 0;
    std::vector<WingView> views(count);
    std::vector<int> blockWing, blockLocal;
    for (int w = 0; w < count; ++w) {
        wings_[w]->setLaunch(blocks_[w], threads);
        views[w] = wings_[w]->launchView();
        views[w].profile = nullptr;  // the step profile covers a wing's own launches
        lds_ = std::max(lds_, wings_[w]->dynamicLdsBytes());
        for (int b = 0; b < views[w].launchBlocks; ++b) {
            blockWing.push_back(w);
            blockLocal.push_back(b);
        }
    }
    if (lds_ > kMaxThreads * sizeof(double))
        for (bool general : {false, true})
            check(hipFuncSetAttribute(batchKernelFor(general), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(lds_)),
                  "shared memory attribute");
    total_ = static_cast<int>(blockWing.size());
    threads_ = threads;
    cap_ = cap;
    generations_.clear();
    for (DeviceWing* w : wings_) generations_.push_back(w->launchGeneration());
    release();
    check(hipMalloc(&views_, views.size() * sizeof(WingView)), "hipMalloc");
    check(hipMemcpy(views_, views.data(), views.size() * sizeof(WingView), hipMemcpyHostToDevice), "upload");
    check(hipMalloc(&args_, count * sizeof(StepArgs)), "hipMalloc");
    check(hipMalloc(&blockWing_, blockWing.size() * sizeof(int)), "hipMalloc");
    check(hipMemcpy(blockWing_, blockWing.data(), blockWing.size() * sizeof(int), hipMemcpyHostToDevice), "upload");
    check(hipMalloc(&blockLocal_, blockLocal.size() * sizeof(int)), "hipMalloc");
    check(hipMemcpy(blockLocal_, blockLocal.data(), blockLocal.size() * sizeof(int), hipMemcpyHostToDevice), "upload");
    check(hipMalloc(&mask_, count), "hipMalloc");
}

void DeviceBatch::ensurePlan() {
    for (std::size_t w = 0; w < wings_.size(); ++w)
        if (generations_.size() != wings_.size() || wings_[w]->launchGeneration() != generations_[w]) {
            plan(cap_, threads_);
            return;
        }
}

void DeviceBatch::snapshot() {
    const dim3 grid(16, static_cast<unsigned>(wings_.size()));
    hipLaunchKernelGGL(batchSnapshotKernel, grid, dim3(256), 0, stream_, static_cast<const WingView*>(views_),
                       static_cast<const unsigned char*>(mask_), 0);
    check(hipGetLastError(), "batch snapshot");
}

void DeviceBatch::restore(const std::vector<unsigned char>& which) {
    check(hipMemcpyAsync(mask_, which.data(), which.size(), hipMemcpyHostToDevice, stream_), "restore mask");
    const dim3 grid(16, static_cast<unsigned>(wings_.size()));
    hipLaunchKernelGGL(batchSnapshotKernel, grid, dim3(256), 0, stream_, static_cast<const WingView*>(views_),
                       static_cast<const unsigned char*>(mask_), 1);
    check(hipGetLastError(), "batch restore");
    check(hipStreamSynchronize(stream_), "batch restore");
}

void DeviceBatch::launchSubstep(const std::vector<StepArgs>& args, double* const* traces) {
    if (!views_) throw std::logic_error("the batch has no launch plan");
    check(hipMemcpyAsync(args_, args.data(), args.size() * sizeof(StepArgs), hipMemcpyHostToDevice, stream_),
          "batch arguments");
    const WingView* views = static_cast<const WingView*>(views_);
    const StepArgs* stepArgs = static_cast<const StepArgs*>(args_);
    const int* blockWing = static_cast<const int*>(blockWing_);
    const int* blockLocal = static_cast<const int*>(blockLocal_);
    bool general = false;
    for (const DeviceWing* w : wings_) general = general || w->needsGeneralKernel();
    void* params[] = {&views, &stepArgs, &blockWing, &blockLocal, &traces};
    if (!cooperative_) {
        check(hipLaunchKernel(batchKernelFor(general), dim3(total_), dim3(threads_), params, lds_, stream_),
              "batch substep launch");
    } else {
        check(hipLaunchCooperativeKernel(batchKernelFor(general), dim3(total_), dim3(threads_), params, lds_, stream_),
              "cooperative batch launch");
    }
    check(hipGetLastError(), "batch substep launch");
}

void DeviceBatch::synchronize() { check(hipStreamSynchronize(stream_), "batch synchronize"); }

}  // namespace swx

// ---------------------------------------------------------------------------
// Device math conformance probe (sw_device_math).

namespace swx {
namespace {

__global__ void mathKernel(int function, const double* a, const double* b, double* out, long count) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const double x = a[i], y = b[i];
    double r = 0.0;
    switch (function) {
    case 0: r = sqrt(x); break;
    case 1: r = x / y; break;
    case 2: r = pm::atan2(x, y); break;
    case 3: r = remainderTwoPi(x); break;
    case 4: r = pm::sin(x); break;
    case 5: r = pm::cos(x); break;
    case 6: r = pm::acos(x); break;
    default: r = 0.0; break;
    }
    out[i] = r;
}

}  // namespace

void deviceMath(int device, int function, const double* a, const double* b, double* out, std::size_t count) {
    if (function < 0 || function > 6) throw std::out_of_range("math function must be 0..6");
    check(hipSetDevice(device), "hipSetDevice");
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    const std::size_t bytes = std::max<std::size_t>(1, count) * sizeof(double);
    check(hipMalloc(&da, bytes), "hipMalloc");
    check(hipMalloc(&db, bytes), "hipMalloc");
    check(hipMalloc(&dout, bytes), "hipMalloc");
    try {
        check(hipMemcpy(da, a, count * sizeof(double), hipMemcpyHostToDevice), "upload");
        check(hipMemcpy(db, b, count * sizeof(double), hipMemcpyHostToDevice), "upload");
        const int threads = 256;
        const int blocks = static_cast<int>((count + threads - 1) / threads);
        if (blocks > 0) {
            hipLaunchKernelGGL(mathKernel, dim3(blocks), dim3(threads), 0, nullptr, function, da, db, dout,
                               static_cast<long>(count));
            check(hipGetLastError(), "math launch");
        }
        check(hipMemcpy(out, dout, count * sizeof(double), hipMemcpyDeviceToHost), "download");
    } catch (...) {
        (void)hipFree(da);
        (void)hipFree(db);
        (void)hipFree(dout);
        throw;
    }
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dout);
}

}  // namespace swx

#include "adjoint_device.inc"
