#!/usr/bin/env python3
"""Throughput of the XPBD structural substep on one MI355X.

Workload: the seeded 70k-triangle FP64 export-scale wing (43k nodes, 70.8k
membrane triangles, 98.5k dihedral hinges, 42.4k seam constraints, 300
suspension segments with a rigid 90 kg pilot, 8 contact spheres), flight
settings dt = 1/120 s, one substep of 48 iterations per step. Inputs are
resident in HBM before the timed region; the timed region is K whole steps
(`sw_structure_step_many`), bracketed by device synchronisation.

Printed: one JSON line. ``value`` is constraint projections per second over
the timed steps; ``roofline`` prices the substep kernel's algorithmic bytes
(per-family byte model below) against its HIP-event duration on the
structure's stream; ``cpu_baseline`` times the reference's own CPU solver
(``oracle/_ref/ref_driver``, built from the reference sources) on a bounded
sample of the same wing, or the serial oracle when that binary is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "constraint projections/s @70k-tri FP64 wing"
UNIT = "constraint projections/s"

# Algorithmic HBM bytes per projection visit (FP64 values, 32-bit indices,
# SoA). Node traffic per visit is counted once per visit because the items of
# one colour are node-disjoint.
BYTES = {
    "membrane": 316,        # idx 12, inverse chart 32, sqrt(A) 8, compliance 48, lambda r+w 48, x r+w 144, w 24
    "membrane_damped": 72,  # + previous positions when the membrane damping is non-zero
    "hinge": 272,           # idx 16, rest angle 8, compliance 8, lambda r+w 16, x r+w 192, w 32
    "rail": 152,            # idx 8, rest 8, compliance 8, lambda r+w 16, x r+w 96, w 16
    "stitch": 136,          # rail without the per-constraint rest/compliance reads
    "segment": 152,
    "sphere": 120,          # idx 8, radius 8, compliance 8, lambda r+w 16, x r+w 48, centre x 24, w 8
    "predict": 152,         # per node: x, v, f, w read; prev, x, v written
    "finalize": 80,         # per node: x, prev, w read; v written
    "pressure": 96,         # per node-triangle incidence (3 per triangle)
    "residual": 300,        # per membrane element diagnostic pass
}


def family_counts(definition) -> dict:
    a = definition.arrays
    kinds = np.asarray(a["constraint_kind"])
    suspension = definition.suspension
    return {
        "nodes": int(definition.node_count),
        "membranes": int(np.asarray(a["membrane_triangle"]).size),
        "hinges": int(np.asarray(a["dihedral_rest_angle"]).size),
        "stitches": int(np.count_nonzero(kinds == 4)),
        "rails": int(np.count_nonzero(kinds != 4)),
        "segments": 0 if suspension is None else int(np.asarray(suspension["segment_ids"]).size),
        "sphere_pairs": int(np.asarray(a["sphere_nodes"]).size),
        "triangles": int(definition.triangle_count),
    }


def projections_per_substep(c: dict, iterations: int) -> int:
    per_iteration = c["membranes"] + c["hinges"] + c["stitches"] + c["rails"] + c["segments"] + c["sphere_pairs"]
    return iterations * per_iteration + c["stitches"]  # + the seam-stitch closing pass


def bytes_per_substep(c: dict, iterations: int, membrane_damped: bool) -> int:
    b = BYTES
    per_iteration = (c["membranes"] * (b["membrane"] + (b["membrane_damped"] if membrane_damped else 0))
                     + c["hinges"] * b["hinge"] + c["rails"] * b["rail"] + c["stitches"] * b["stitch"]
                     + c["segments"] * b["segment"] + c["sphere_pairs"] * b["sphere"])
    once = (c["nodes"] * (b["predict"] + b["finalize"]) + 3 * c["triangles"] * b["pressure"]
            + c["membranes"] * b["residual"] + c["stitches"] * b["stitch"])
    return iterations * per_iteration + once


def launch_candidates(cus: int) -> list[tuple[int, int]]:
    shapes = []
    for wg in (cus // 16, cus // 8, cus // 4, cus // 2, cus):
        for threads in (128, 256):
            if wg >= 1:
                shapes.append((wg, threads))
    return shapes


def tune(structure, loads, settings, cus: int) -> tuple[tuple[int, int, bool], dict]:
    """Launch autotuner: one timed step per candidate, keep the fastest. A
    candidate is a workgroup count (which sets each colour's block of items
    per workgroup: the colour's size over the workgroups), a tile of threads
    per workgroup (items per lane within the block) and, where every
    participant fits on one XCD, the one-XCD synchronisation mode (no L2
    write-back per phase) or the device-wide one. The integer tables (colour
    orders, levels) do not depend on the candidate, nor does the trajectory."""
    results = {}
    for shape in launch_candidates(cus):
        try:
            structure.set_launch(*shape)
        except Exception:  # shape outside the device's limits
            continue
        for xcd in (True, False):
            if structure.set_xcd_local(xcd) != xcd:
                continue  # the one-XCD mode is not available for this shape
            structure.set_kernel_timing(True)
            structure.set_external_forces(loads)
            structure.step_many(1, settings)
            ms, launches = structure.kernel_time()
            structure.set_kernel_timing(False)
            results[f"{shape[0]}x{shape[1]}" + ("/xcd" if xcd else "")] = ms / max(launches, 1)
    best = min(results, key=results.get)
    xcd = best.endswith("/xcd")
    wg, threads = (int(v) for v in best.split("/")[0].split("x"))
    structure.set_launch(wg, threads)
    structure.set_xcd_local(xcd)
    return (wg, threads, xcd), results


def measured_traffic(workload: str, launch: str) -> int | None:
    """HBM bytes per substep-kernel launch from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (scaled by a copy kernel of known size),
    when they were taken on this workload and launch shape."""
    path = ROOT / "profiles" / f"substepKernel_{workload}_counters.json"
    if not path.exists():
        return None
    record = json.loads(path.read_text()).get("launches", {}).get(launch)
    return None if record is None else int(record["traffic_bytes_per_launch"])


def cpu_share() -> int:
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(cores, 16))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload: str, counts_no_spheres: dict, iterations: int, budget_s: float) -> dict | None:
    """The reference's own CPU solver (oracle/_ref/ref_driver --bench, built
    from the reference sources) on the same wing, in the softwing-bench
    manner: warm-up steps, steady_clock around each step, best of several
    runs, and the reference's per-family StepPerformanceProfile split. Timed
    with workerThreads = 1, 12 (the surface-flight setting) and P, the
    reference's hardwarePhysicalCoreCount() (parallel.cpp:359-398, on Linux
    std::thread::hardware_concurrency()), run on at most this process's CPU
    share; the P run is the reported value."""
    driver = ROOT / "oracle" / "_ref" / "ref_driver"
    share = cpu_share()
    if driver.exists():
        def run(workers: int, seconds: float, warmup: int, runs: int) -> dict:
            cmd = [str(driver), "--bench", workload, str(workers), str(seconds), str(warmup), str(runs)]
            out = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
            if out.returncode != 0:
                raise RuntimeError(f"reference CPU baseline failed: {out.stderr.strip()}")
            return json.loads(out.stdout.strip().splitlines()[-1])
        p_workers = max(1, min(os.cpu_count() or 1, share))  # hardware_concurrency, capped at the share
        legs = {}
        for workers, warmup, runs, share_of_budget in ((p_workers, 5, 5, 0.5), (12, 2, 3, 0.25), (1, 1, 3, 0.25)):
            if workers in legs:
                continue
            legs[workers] = run(workers, budget_s * share_of_budget, warmup, runs)
        best = legs[p_workers]
        physical = int(best["physical_cores"])
        ms = best["ms_per_step_best_run"]
        per_leg = {str(w): {"ms_per_step_best_run": round(l["ms_per_step_best_run"], 3),
                            "ms_per_step_mean": round(l["ms_per_step_mean"], 3), "warmup": l["warmup"],
                            "runs": l["runs"], "steps": l["steps"]} for w, l in legs.items()}
        return {
            "value": projections_per_substep(counts_no_spheres, iterations) / (ms * 1e-3),
            "unit": UNIT,
            "cores": p_workers,
            "kind": "reference",
            "sample": (f"{workload} wing without contact spheres (the reference has no sphere family); "
                       f"workerThreads={p_workers} (P={physical} from hardwarePhysicalCoreCount, capped at this "
                       f"process's {share}-core share), {best['warmup']} warm-up steps, best of {best['runs']} "
                       f"runs ({best['steps']} timed steps): {ms:.1f} ms/step"),
            "cpu_model": cpu_model(),
            "physical_cores": physical,
            "workers": per_leg,
            "split_ms_per_step": best["split_ms"],
        }
    # Serial oracle (bit-identical to the reference path), one core.
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    import simwing_amd as sw
    definition, pressure, loads = make_workload(workload)
    body = orc.OracleBody(definition)
    body.set_face_pressure(pressure)
    settings = sw.flight_step_settings(iterations)
    body.set_external_forces(loads)
    body.step(settings)
    steps, start = 0, time.perf_counter()
    while time.perf_counter() - start < budget_s or steps < 2:
        body.set_external_forces(loads)
        body.step(settings)
        steps += 1
    ms = 1e3 * (time.perf_counter() - start) / steps
    return {
        "value": projections_per_substep(family_counts(definition), iterations) / (ms * 1e-3),
        "unit": UNIT,
        "cores": 1,
        "kind": "port",
        "sample": f"{workload} wing, serial oracle, 1 warm-up + {steps} timed steps, mean {ms:.1f} ms/step",
    }


def batched_leg(count: int, steps: int, seed: int = 0x5EED0002) -> dict:
    """64 wings (node counts log-uniform 2k-40k, span 6-12 m, per-wing
    compliance, pressure 20-100 Pa and line rest scale 0.97-1.03 from the
    seed) stepped by one persistent launch, against the same wings stepped
    one after another with their own launches (one LDS-resident workgroup
    for wings whose nodes fit, else 64 workgroups)."""
    import simwing_amd as sw
    settings = sw.flight_step_settings()
    t0 = time.perf_counter()
    cases = [make_workload(f"batch:{seed}:{i}:{count}") for i in range(count)]
    batch = sw.Batch([c[0] for c in cases])
    for w, (_, pressure, loads) in enumerate(cases):
        batch.set_face_pressure(w, pressure)
        batch.set_external_forces(w, loads)
    setup = time.perf_counter() - t0
    batch.step(1, settings)  # first launch
    for w, (_, _, loads) in enumerate(cases):
        batch.set_external_forces(w, loads)
    start = time.perf_counter()
    batch.step(steps, settings)  # the loads apply to every step of the call
    batched_s = time.perf_counter() - start
    blocks = batch.workgroups()
    # Per-wing launches: each wing alone with its own solo launch shape.
    solo_steps = 3
    solo_s = 0.0
    for w, (d, _, loads) in enumerate(cases):
        s = batch.structure(w)
        s.set_launch(1, 256) if blocks[w] == 1 else s.set_launch(64, 256)
        s.set_external_forces(loads)
        s.step_many(1, settings)
        s.set_external_forces(loads)
        start = time.perf_counter()
        s.step_many(solo_steps, settings)
        solo_s += (time.perf_counter() - start) / solo_steps
    batch.set_launch(64, 256)  # the default plan again after the solo runs
    nodes = [c[0].node_count for c in cases]
    batched_rate = count * steps / batched_s
    solo_rate = count / solo_s
    batch.close()
    # Roofline of the batched step: the wings' algorithmic bytes per substep
    # (the per-family model above, summed over the wings) over its duration.
    step_bytes = sum(bytes_per_substep(family_counts(c[0]), 48, False) for c in cases)
    achieved = step_bytes / (batched_s / steps) / 1e9
    return {
        "metric": "batched wings/s",
        "value": batched_rate,
        "unit": "wing-steps/s",
        "per_wing_launches": solo_rate,
        "speedup": batched_rate / solo_rate,
        "ms_per_batched_step": 1e3 * batched_s / steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_step": step_bytes},
        "steps": steps,
        "workload": (f"{count} wings, {min(nodes)}-{max(nodes)} nodes (total {sum(nodes)}), seed {seed:#x}, "
                     f"dt 1/120 s, 48 iterations; {int(blocks.sum())} workgroups "
                     f"({int((blocks == 1).sum())} LDS-resident wings); setup {setup:.1f} s"),
    }


def adjoint_leg(wing: str, steps: int, interval: int, seed: int = 0x5EED0003) -> dict:
    """Shape-loss gradient (membrane log-compliance, line rest length,
    triangle pressure) over `steps` steps of `wing`, checkpoints every
    `interval` steps, against the forward pass alone. The target is a hidden
    run with compliance x U(0.8, 1.2) and rest lengths x U(0.98, 1.02)."""
    import simwing_amd as sw
    definition, pressure, loads = make_workload(wing)
    s = sw.Structure(definition)
    s.set_face_pressure(pressure)
    s.set_external_forces(loads)
    adj = sw.Adjoint(s, steps, interval)
    rng = np.random.default_rng(seed)
    rest = np.asarray(definition.suspension["segment_rest"], dtype=np.float64)
    hidden_theta = np.log(rng.uniform(0.8, 1.2, adj.membrane_count))
    hidden_rest = rest * rng.uniform(0.98, 1.02, rest.size)
    x0 = np.asarray(definition.arrays["node_position"]).reshape(-1, 3)
    target = adj.run(x0, hidden_theta, hidden_rest, gradients=False)["final_position"]
    start = time.perf_counter()
    forward = adj.run(target, gradients=False)
    forward_s = time.perf_counter() - start
    forward_ms, _, _ = adj.timing()
    start = time.perf_counter()
    result = adj.run(target)
    gradient_s = time.perf_counter() - start
    fwd_ms, rev_ms, device_bytes = adj.timing()
    replay_ms, jacobian_ms, sweep_ms = adj.timing_detail()
    finite = all(np.all(np.isfinite(result[k])) for k in ("grad_log_compliance", "grad_rest_length", "grad_pressure"))
    if not finite or result["loss"] != forward["loss"]:
        raise RuntimeError("adjoint run is not finite or not reproducible")
    adj.close()
    s.close()
    return {
        "metric": "adjoint/fwd cost",
        "value": gradient_s / forward_s,
        "unit": "x (gradient run / forward run, wall clock)",
        "higher_is_better": False,
        "device_ratio": (fwd_ms + rev_ms) / forward_ms,
        "forward_ms": forward_ms,
        "gradient_forward_ms": fwd_ms,
        "gradient_reverse_ms": rev_ms,
        "gradient_replay_ms": replay_ms,
        "gradient_jacobian_ms": jacobian_ms,
        "gradient_sweep_ms": sweep_ms,
        "device_bytes": device_bytes,
        "workload": (f"{wing} wing: {definition.node_count} nodes, {adj.membrane_count} membranes, "
                     f"{rest.size} line segments; {steps} steps x 48 iterations, checkpoints every {interval} steps; "
                     f"{adj.membrane_count + rest.size + adj.triangle_count} parameters"),
    }


def batch_adjoint_leg(count: int, steps: int, interval: int, chunk: int, seed: int = 0x5EED0004) -> dict:
    """Rest-length gradients of `count` wings (triangle counts log-uniform
    5k-70k at the quantiles) stepped together, against the batched forward
    pass alone. Each wing's target is a hidden batched run with its membrane
    compliances x U(0.8, 1.2) and line rest lengths x U(0.98, 1.02)."""
    import math
    import simwing_amd as sw
    settings = sw.flight_step_settings()
    t0 = time.perf_counter()
    cases = [make_workload(f"adjbatch:{seed}:{i}:{count}") for i in range(count)]
    rng = np.random.default_rng(seed)
    hidden = []
    for d, _, _ in cases:
        h = d.copy()
        mat = h.arrays["membrane_material"].reshape(-1, 8)
        for m, scale in enumerate(rng.uniform(0.8, 1.2, mat.shape[0])):
            mat[m, :4] = mat[m, :4] / scale  # compliance x scale
        h.suspension["segment_rest"] = np.asarray(h.suspension["segment_rest"]) * rng.uniform(0.98, 1.02, len(
            h.suspension["segment_rest"]))
        hidden.append(h)
    target_batch = sw.Batch(hidden)
    for w, (_, pressure, loads) in enumerate(cases):
        target_batch.set_face_pressure(w, pressure)
        target_batch.set_external_forces(w, loads)
    target_batch.step(steps, settings)
    targets = [target_batch.node_states(w)[0] for w in range(count)]
    target_batch.close()
    batch = sw.Batch([d for d, _, _ in cases])
    for w, (_, pressure, loads) in enumerate(cases):
        batch.set_face_pressure(w, pressure)
        batch.set_external_forces(w, loads)
    adj = sw.BatchAdjoint(batch, steps, interval, chunk, settings)
    setup = time.perf_counter() - t0
    start = time.perf_counter()
    forward = adj.run(targets, gradients=False)
    forward_s = time.perf_counter() - start
    start = time.perf_counter()
    result = adj.run(targets)
    gradient_s = time.perf_counter() - start
    fwd_ms, rev_ms, device_bytes = adj.timing()
    replay_ms, jacobian_ms, sweep_ms = adj.timing_detail()
    # Large wings released from the flat state fold violently in the first
    # steps; where a hinge angle crosses its +-pi wrap the loss is
    # discontinuous and the exact derivative overflows over long horizons.
    # The run's cost does not depend on it; the count is reported.
    finite = [bool(np.all(np.isfinite(g))) for g in result["grad_rest_length"]]
    if not np.array_equal(forward["loss"], result["loss"]):
        raise RuntimeError("batched adjoint forward pass is not reproducible")
    triangles = [d.triangle_count for d, _, _ in cases]
    workgroups = adj.workgroups()
    adj.close()
    batch.close()
    return {
        "metric": "adjoint/fwd cost (batched)",
        "value": gradient_s / forward_s,
        "unit": "x (gradient run / forward run, wall clock)",
        "higher_is_better": False,
        "forward_ms": 1e3 * forward_s,
        "gradient_ms": 1e3 * gradient_s,
        "device_forward_ms": fwd_ms,
        "device_reverse_ms": rev_ms,
        "device_replay_ms": replay_ms,
        "device_jacobian_ms": jacobian_ms,
        "device_sweep_ms": sweep_ms,
        "device_bytes": device_bytes,
        "wing_gradients_per_s": count / gradient_s,
        "wings_with_finite_gradients": int(sum(finite)),
        "workload": (f"{count} wings, {min(triangles)}-{max(triangles)} triangles (total {sum(triangles)}), seed "
                     f"{seed:#x}; {steps} steps x 48 iterations, checkpoints every {interval} steps, Jacobian "
                     f"columns {chunk} iterations at a time, {int(workgroups.sum())} sweep workgroups; line "
                     f"rest-length gradients; setup {setup:.1f} s"),
    }


def make_workload(name: str):
    import simwing_amd as sw
    scene = sw.generate_scene(name)
    assembly = scene.assemble()
    return assembly.definition, scene.face_pressure(), assembly.flight_loads()


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--workload", default="export")
    parser.add_argument("--iterations", type=int, default=48)
    parser.add_argument("--cpu-seconds", type=float, default=15.0)
    parser.add_argument("--no-cpu", action="store_true")
    parser.add_argument("--no-tune", action="store_true")
    parser.add_argument("--no-batch", action="store_true")
    parser.add_argument("--batch-wings", type=int, default=64)
    parser.add_argument("--batch-steps", type=int, default=120)
    parser.add_argument("--no-adjoint", action="store_true")
    parser.add_argument("--adjoint-steps", type=int, default=120)
    parser.add_argument("--adjoint-wings", type=int, default=32)
    parser.add_argument("--launch", default=None, help="fixed launch shape WGxTHREADS, /xcd for the one-XCD mode")
    parser.add_argument("--plain-launch", action="store_true",
                        help="plain instead of cooperative launches (for profilers that do not follow the latter)")
    args = parser.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the structural substep of one wing runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        return 2

    import simwing_amd as sw

    t0 = time.perf_counter()
    definition, pressure, loads = make_workload(args.workload)
    counts = family_counts(definition)
    print(f"# workload {args.workload}: {counts} (assembled in {time.perf_counter() - t0:.1f} s)",
          file=sys.stderr, flush=True)
    structure = sw.Structure(definition)
    structure.set_face_pressure(pressure)
    if args.plain_launch:
        structure.set_cooperative(False)
    settings = sw.flight_step_settings(args.iterations)
    cus = 256  # MI355X compute units; shapes beyond the device's limits are skipped by the tuner

    structure.set_external_forces(loads)
    structure.step_many(1, settings)  # first launch: code object load
    tuning = {}
    if args.launch:
        wg, threads = (int(v) for v in args.launch.split("/")[0].split("x"))
        structure.set_launch(wg, threads)
        shape = (wg, threads, structure.set_xcd_local(args.launch.endswith("/xcd")))
    elif args.no_tune:
        shape = (*structure.launch(), structure.set_xcd_local(True))  # the default request
    else:
        shape, tuning = tune(structure, loads, settings, cus)
    print(f"# launch shape {shape} tuning(ms/launch)={tuning}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        structure.set_external_forces(loads)
        structure.step_many(1, settings)

    structure.set_kernel_timing(True)
    structure.set_external_forces(loads)
    start = time.perf_counter()
    structure.step_many(args.steps, settings)  # returns after the stream has drained
    elapsed = time.perf_counter() - start
    kernel_ms, launches = structure.kernel_time()
    structure.set_kernel_timing(False)
    diagnostics = structure.diagnostics()
    if diagnostics["finite"] != 1:
        raise RuntimeError("non-finite state after the timed steps")
    # The device's per-part split of a substep (StepPerformanceProfile's
    # fields), from a few profiled steps after the timed region.
    structure.set_profiling(True)
    profiled = 3
    for _ in range(profiled):
        structure.set_external_forces(loads)
        structure.step_many(1, settings)
    prof = structure.profile()
    structure.set_profiling(False)
    device_split = {k[:-3]: prof[k] * 1e-6 / profiled for k in prof if k.endswith("_ns")}

    substeps = settings.substeps
    proj_step = projections_per_substep(counts, args.iterations) * substeps
    ms_per_step = 1e3 * elapsed / args.steps
    value = proj_step * args.steps / elapsed
    avg_kernel_ms = kernel_ms / max(launches, 1)
    algorithmic = bytes_per_substep(counts, args.iterations, membrane_damped=False)
    achieved = algorithmic / (avg_kernel_ms * 1e-3) / 1e9

    structure.close()
    batched = None if args.no_batch else batched_leg(args.batch_wings, args.batch_steps)
    adjoint = None if args.no_adjoint else [adjoint_leg(w, args.adjoint_steps, 8) for w in ("small", "variant10k")]
    batched_adjoint = None if args.no_adjoint else batch_adjoint_leg(args.adjoint_wings, 240, 16, 8)

    cpu = None
    if not args.no_cpu:
        no_spheres = dict(counts, sphere_pairs=0)
        cpu = cpu_baseline(args.workload, no_spheres, args.iterations, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload} wing: {counts['nodes']} nodes, {counts['membranes']} membrane triangles, "
                        f"{counts['hinges']} hinges, {counts['rails'] + counts['stitches']} seam constraints, "
                        f"{counts['segments']} line segments + rigid pilot, {counts['sphere_pairs']} sphere pairs; "
                        f"dt 1/120 s, {args.iterations} iterations, {substeps} substep",
            "projections_per_step": proj_step,
            "launch": f"{shape[0]}x{shape[1]}" + ("/xcd" if shape[2] else ""),
            "tuning_ms_per_launch": tuning,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.workload, f"{shape[0]}x{shape[1]}" + ("/xcd" if shape[2] else "")),
            "kernel": "substepKernel",
            "kernel_ms": avg_kernel_ms,
            "algorithmic_bytes": algorithmic,
            "device_split_ms_per_step": device_split,
        },
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adjoint,
        "batched_adjoint": batched_adjoint,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
