"""The Structure adapter surface on the device (needs an MI355X).

Against the reference's own simwing::fsi::Structure (oracle/ref_driver
--structure, tests/golden/<preset>.structure.w<N>.swfx) and the portable-math
oracle: the serial definition-order sweep (workerThreads 0) and the coloured
sweep, suspensionState() before the first step and after the last, the
checkpoint's public fields, and the accounting of a multi-step call that
fails part-way.
"""
import numpy as np
import pytest

import oracle as orc
import simwing_amd as sw
from conftest import bits_equal, relative_error
from simwing_amd import native

pytestmark = pytest.mark.gpu

STEPS = 20
TOL = 1e-10


def stepped(preset, workers, steps=STEPS):
    assembly = sw.generate_scene(preset).assemble()
    d, loads = assembly.definition, assembly.flight_loads()
    structure = sw.Structure(d)
    settings = sw.flight_step_settings()
    settings.worker_threads = workers
    initial = structure.suspension_state()
    for _ in range(steps):
        structure.set_external_forces(loads)
        structure.step(settings)
    structure.set_external_forces(loads)  # pending again, as the fixture's run
    return d, loads, structure, initial


@pytest.mark.parametrize("preset", ["tiny", "small"])
@pytest.mark.parametrize("workers", [0, 1])
def test_sweep_orders_match_reference_structure(golden, preset, workers):
    print(f"structure {preset}: workerThreads={workers}, {STEPS} steps", flush=True)
    ref = golden(f"{preset}.structure.w{workers}.swfx")
    d, loads, structure, _ = stepped(preset, workers)
    body = orc.OracleBody(d)
    body.use_portable_math(True)
    settings = sw.flight_step_settings()
    settings.worker_threads = workers
    for _ in range(STEPS):
        body.set_external_forces(loads)
        body.step(settings)
    expected = body.state()
    x, prev, v = structure.node_states()
    for key, got in (("position", x), ("previous", prev), ("velocity", v)):
        assert bits_equal(got.reshape(-1), expected[key]), f"{key} differs from the portable oracle"
        err = relative_error(got.reshape(-1), ref[f"checkpoint/{key}"])
        print(f"  {key}: {err:.2e} vs the reference Structure", flush=True)
        assert err <= TOL


@pytest.mark.parametrize("preset", ["tiny", "small"])
def test_suspension_state_matches_reference(golden, preset):
    ref = golden(f"{preset}.structure.w1.swfx")
    d, _, structure, initial = stepped(preset, 1)
    # Before the first step: the definition's pose, rest lengths, no tension.
    assert bits_equal(initial["payload"], ref["initial/payload"])
    assert np.array_equal(initial["segment_ids"], ref["initial/segment_ids"])
    for key in ("segment_length", "segment_commanded_rest", "segment_tension"):
        assert bits_equal(initial[key], ref[f"initial/{key}"]), f"initial {key}"
    assert relative_error(initial["harness_position"], ref["initial/harness_position"]) <= 1e-15
    final = structure.suspension_state()
    assert np.array_equal(final["segment_ids"], ref["final/segment_ids"])
    assert bits_equal(final["segment_commanded_rest"], ref["final/segment_commanded_rest"])
    for key in ("payload", "harness_position", "harness_velocity", "segment_length", "segment_tension"):
        err = relative_error(final[key], ref[f"final/{key}"])
        print(f"  {preset} final {key}: {err:.2e}", flush=True)
        assert err <= TOL, key
    assert (final["segment_tension"] > 0.0).any()
    # No suspension: suspensionState() is empty.
    free = d.copy()
    free.suspension = None
    free.arrays["node_fixed"] = np.ones_like(free.arrays["node_fixed"])
    assert sw.Structure(free).suspension_state() is None


@pytest.mark.parametrize("preset", ["tiny", "small"])
def test_checkpoint_public_fields_match_reference(golden, preset):
    ref = golden(f"{preset}.structure.w1.swfx")
    _, _, structure, _ = stepped(preset, 1)
    checkpoint = structure.checkpoint()
    info = checkpoint.info()
    assert info["version"] == int(ref["checkpoint/version"][0]) == 2
    assert info["accepted_steps"] == int(ref["checkpoint/accepted_steps"][0]) == STEPS
    assert info["time"] == float(ref["checkpoint/time"][0])
    assert info["fingerprint"] == structure.definition_fingerprint()
    nodes = checkpoint.nodes()
    assert bits_equal(nodes["pending_force"], ref["checkpoint/pending_force"])
    assert bits_equal(nodes["last_applied"], ref["checkpoint/last_applied"])
    for key in ("position", "previous", "velocity"):
        assert relative_error(nodes[key], ref[f"checkpoint/{key}"]) <= TOL, key
    # The checkpoint restores what it holds.
    structure.step(sw.flight_step_settings())
    structure.restore(checkpoint)
    x, _, _ = structure.node_states()
    assert bits_equal(x.reshape(-1), nodes["position"])
    assert structure.accepted_step_count() == STEPS


def test_failure_inside_a_multi_step_call_keeps_the_earlier_steps():
    """step_many(5) with a line-residual bound that the first steps meet and a
    later one exceeds: the structure keeps the steps before the failing one
    (state and accounting), as if they had been separate calls."""
    assembly = sw.generate_scene("tiny").assemble()
    d, loads = assembly.definition, assembly.flight_loads()
    settings = sw.flight_step_settings()
    # The oracle's line residual after each step picks a bound that step 1
    # meets and a later step exceeds.
    body = orc.OracleBody(d)
    body.use_portable_math(True)
    residuals, states = [], []
    for _ in range(5):
        body.set_external_forces(loads)
        body.step(settings)
        state = body.state()
        residuals.append(float(state["suspension_residual"][0]))
        states.append(state)
    later = max(residuals[1:])
    assert later > residuals[0]
    bound = 0.5 * (residuals[0] + later)
    fails_at = next(k for k in range(1, 5) if residuals[k] > bound)
    d = d.copy()
    d.suspension["maximum_line_residual"] = bound
    structure = sw.Structure(d)
    structure.set_external_forces(loads)
    with pytest.raises(sw.NativeError) as error:
        structure.step_many(5, settings)
    assert error.value.status == native.SW_SUSPENSION_ERROR
    assert structure.accepted_step_count() == fails_at
    x, _, _ = structure.node_states()
    assert bits_equal(x.reshape(-1), states[fails_at - 1]["position"])
