"""Batched adjoint (needs an MI355X): the rest-length gradients of wings
stepped together in one persistent launch per substep equal, wing for wing,
those of each wing's own adjoint (Jacobian columns held in chunks here, all
at once there), and the run leaves the batch untouched."""
import numpy as np
import pytest

import simwing_amd as sw
from helpers import flight_case

pytestmark = pytest.mark.gpu

PRESETS = ["tiny", "small", "batch:4242:0:8", "batch:4242:5:8"]
STEPS, INTERVAL = 5, 2


def cases():
    out = []
    rng = np.random.default_rng(44)
    for name in PRESETS:
        d, pressure, loads = flight_case(name)
        rest = np.asarray(d.suspension["segment_rest"], dtype=np.float64)
        trim = rest * rng.uniform(0.99, 1.01, rest.size)
        target = np.asarray(d.arrays["node_position"]).reshape(-1, 3) + rng.normal(0.0, 0.02, (d.node_count, 3))
        out.append((name, d, pressure, loads, trim, target))
    return out


def test_batch_adjoint_equals_solo_adjoints():
    data = cases()
    solo = []
    for name, d, pressure, loads, trim, target in data:
        s = sw.Structure(d)
        s.set_face_pressure(pressure)
        s.set_external_forces(loads)
        adj = sw.Adjoint(s, STEPS, INTERVAL)
        r = adj.run(target, None, trim, None)
        solo.append((r["loss"], r["grad_rest_length"], r["final_position"]))
        adj.close()
        s.close()
    batch = sw.Batch([d for _, d, _, _, _, _ in data])
    for w, (_, _, pressure, loads, _, _) in enumerate(data):
        batch.set_face_pressure(w, pressure)
        batch.set_external_forces(w, loads)
    before = [batch.node_states(w) for w in range(len(data))]
    adj = sw.BatchAdjoint(batch, STEPS, INTERVAL, jacobian_chunk=16)
    print("sweep workgroups per wing:", adj.workgroups().tolist(), flush=True)
    r = adj.run([t for *_, t in data], [tr for *_, tr, _ in data])
    for w, (name, *_rest) in enumerate(data):
        loss, grad, final = solo[w]
        assert np.array_equal(r["final_position"][w], final), f"{name}: batched forward differs"
        assert r["loss"][w] == loss, name
        scale = np.abs(grad).max()
        err = np.abs(r["grad_rest_length"][w] - grad).max() / scale
        print(f"{name}: loss {loss:.6e}, max |grad| {scale:.3e}, batch vs solo {err:.1e}", flush=True)
        assert err <= 1e-12, name
    for w in range(len(data)):
        after = batch.node_states(w)
        assert np.array_equal(before[w][0], after[0]) and np.array_equal(before[w][1], after[1])
    fwd, rev, nbytes = adj.timing()
    print(f"device ms: forward {fwd:.1f}, reverse {rev:.1f}; {nbytes / 2**20:.0f} MiB", flush=True)
    adj.close()
    batch.close()


CFG4_SEED = 0x5EED0004
CFG4_WINGS = (0, 15, 31)  # smallest, median and largest of the 32 (5.4k-67.7k triangles)


def test_batch_adjoint_at_cfg4_settings():
    """Three of the 32 batched-adjoint benchmark wings (smallest, median,
    largest) at the benchmark's settings: recomputation every 16 steps,
    Jacobian columns in chunks of 8 iterations, 32 steps (two checkpoint
    intervals). Each wing's loss is its solo adjoint's exactly and its
    rest-length gradient is within 1e-12 of the solo one, non-finite entries
    in the same places. The smallest wing's gradient is finite; the larger
    wings fold under the flight loads within these steps (the solver's 48
    iterations cannot carry the load across their finer meshes), so their
    exact derivatives grow without bound and may overflow."""
    steps, interval, chunk = 32, 16, 8
    rng = np.random.default_rng(45)
    data = []
    for i in CFG4_WINGS:
        d, pressure, loads = flight_case(f"adjbatch:{CFG4_SEED}:{i}:32")
        rest = np.asarray(d.suspension["segment_rest"], dtype=np.float64)
        target = np.asarray(d.arrays["node_position"]).reshape(-1, 3) + rng.normal(0.0, 0.02, (d.node_count, 3))
        data.append((i, d, pressure, loads, rest, target))
    solo = []
    for i, d, pressure, loads, rest, target in data:
        s = sw.Structure(d)
        s.set_face_pressure(pressure)
        s.set_external_forces(loads)
        adj = sw.Adjoint(s, steps, interval)
        r = adj.run(target, None, rest, None)
        solo.append((r["loss"], r["grad_rest_length"]))
        adj.close()
        s.close()
    batch = sw.Batch([d for _, d, *_ in data])
    for w, (_, _, pressure, loads, _, _) in enumerate(data):
        batch.set_face_pressure(w, pressure)
        batch.set_external_forces(w, loads)
    adj = sw.BatchAdjoint(batch, steps, interval, jacobian_chunk=chunk)
    r = adj.run([t for *_, t in data], [rest for *_, rest, _ in data])
    for w, (i, d, *_rest) in enumerate(data):
        loss, grad = solo[w]
        got = r["grad_rest_length"][w]
        finite = np.isfinite(grad)
        assert np.array_equal(np.isfinite(got), finite), f"cfg4 wing {i}: non-finite entries differ"
        if w == 0:
            assert finite.all(), f"cfg4 wing {i}: non-finite gradient"
        assert r["loss"][w] == loss, i
        scale = np.abs(grad[finite]).max() if finite.any() else 1.0
        err = np.abs(got[finite] - grad[finite]).max() / scale if finite.any() else 0.0
        print(f"cfg4 wing {i} ({d.triangle_count} triangles): loss {loss:.6e}, {int(finite.sum())}/{finite.size} "
              f"finite, max |grad| {scale:.3e}, batch vs solo {err:.1e}", flush=True)
        assert err <= 1e-12, i
    adj.close()
    batch.close()
