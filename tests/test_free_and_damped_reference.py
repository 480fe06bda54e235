"""Free-membrane momentum restoration and axial line damping against the
reference (CPU only).

* Free wings (tiny-free, small-free): the tiny and small wings with no rigid
  pilot, their lines ending at two dynamic riser junctions of half the
  pilot's mass each. No suspension, no contact and every node dynamic, so the
  reference measures the momentum after prediction and restores it after
  finalisation (soft_body.cpp:118-190, 1192-1205, 1446-1448).
  oracle/ref_driver ran the reference's SoftBody on them with the
  surface-flight settings: tests/golden holds the assembly, the colouring and
  the states (tiny) or their SHA-256 digests (small).
* Damped lines (<preset>.structure.damped.swfx): the reference's
  simwing::fsi::Structure stepped 20 times with every suspension segment's
  axial damping at 40 N s/m, which applies applyLineDamping's impulses after
  each substep (suspension.cpp:1165-1185, 2413-2479).
"""
import hashlib

import numpy as np
import pytest

import oracle as orc
import simwing_amd as sw
from conftest import bits_equal, relative_error
from test_assembly_and_schedules import ARRAYS, same_bytes

KEYS = ("position", "previous", "velocity", "constraint_lambda", "membrane_lambda", "membrane_resultant",
        "membrane_residual")
LINE_DAMPING = 40.0
STRUCTURE_STEPS = 20


def digest(array) -> np.ndarray:
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(array).tobytes()).digest(), np.uint64)


def run_oracle(preset, steps, settings=None):
    scene = sw.generate_scene(preset)
    assembly = scene.assemble()
    body = orc.OracleBody(assembly.definition)
    body.set_face_pressure(scene.face_pressure())
    loads = assembly.flight_loads()
    st = settings or sw.flight_step_settings()
    states = {}
    for step in range(1, max(steps) + 1):
        body.set_external_forces(loads)
        body.step(st)
        if step in steps:
            states[step] = body.state()
    return assembly.definition, states


def damped(definition):
    d = definition.copy()
    d.suspension["segment_damping"] = np.full_like(d.suspension["segment_damping"], LINE_DAMPING)
    return d


def test_free_assembly_matches_reference(golden):
    fixture = golden("tiny-free.assembly.swfx")
    d = sw.generate_scene("tiny-free").assemble().definition
    assert d.suspension is None and d.arrays["node_fixed"].sum() == 0
    for name in ARRAYS:
        assert same_bytes(d.arrays[name], fixture[name]), f"{name} differs from assembleSceneStructure"
    small = sw.generate_scene("small-free").assemble().definition
    reference = golden("small-free.digest.swfx")
    for name in ARRAYS:
        assert np.array_equal(digest(small.arrays[name]), reference[f"digest/{name}"]), f"small: {name}"


def test_oracle_tiny_free_bit_for_bit(golden):
    ref = golden("tiny-free.states.swfx")
    order = golden("tiny-free.assembly.swfx")["dihedral_colour_order"].astype(np.int64)
    _, states = run_oracle("tiny-free", (1, 5, 20))
    for step, state in states.items():
        for key in KEYS:
            assert bits_equal(state[key], ref[f"step{step}/{key}"]), f"step {step}: {key}"
        assert bits_equal(state["dihedral_lambda"][order], ref[f"step{step}/dihedral_lambda"])


def test_oracle_small_free_digests(golden):
    ref = golden("small-free.digest.swfx")
    d, states = run_oracle("small-free", (1, 20))
    order, _ = sw.definition_colouring(d, 2)
    for step, state in states.items():
        for key in ("position", "previous", "velocity", "membrane_residual", "membrane_lambda", "constraint_lambda"):
            assert np.array_equal(digest(state[key]), ref[f"step{step}/{key}/digest"]), f"step {step}: {key}"
        assert np.array_equal(digest(state["dihedral_lambda"][order.astype(np.int64)]),
                              ref[f"step{step}/dihedral_lambda/digest"])


def test_free_wing_momentum_is_restored_exactly():
    """An invariant beside the fixtures (which pin the restoration itself):
    with no gravity, load, pressure or velocity damping the predicted
    momentum is the step-start momentum, so each substep ends with the
    linear momentum it started with, to rounding."""
    d = sw.generate_scene("tiny-free").assemble().definition
    body = orc.OracleBody(d)
    st = sw.flight_step_settings()
    st.gravity[:] = (0.0, 0.0, 0.0)
    st.velocity_damping = 0.0
    st.substeps = 1
    masses = 1.0 / (1.0 / d.arrays["node_mass"])
    rng = np.random.default_rng(7)
    # A spinning, drifting start: two steps of a sideways load, then free.
    kick = rng.standard_normal(3 * d.node_count) * 0.05
    body.set_external_forces(kick)
    body.step(st)
    start = body.state()
    linear0 = (masses[:, None] * start["velocity"].reshape(-1, 3)).sum(axis=0)
    for _ in range(5):
        body.set_external_forces(np.zeros(3 * d.node_count))
        body.step(st)
    end = body.state()
    linear = (masses[:, None] * end["velocity"].reshape(-1, 3)).sum(axis=0)
    assert np.abs(linear - linear0).max() <= 1e-12 * max(1.0, np.abs(linear0).max())


@pytest.mark.parametrize("preset", ["tiny", "small"])
def test_oracle_damped_lines_match_reference_structure(golden, preset):
    ref = golden(f"{preset}.structure.damped.swfx")
    undamped = golden(f"{preset}.structure.w1.swfx")
    assembly = sw.generate_scene(preset).assemble()
    d, loads = damped(assembly.definition), assembly.flight_loads()
    body = orc.OracleBody(d)
    st = sw.flight_step_settings()
    st.worker_threads = 1
    for _ in range(STRUCTURE_STEPS):
        body.set_external_forces(loads)
        body.step(st)
    state = body.state()
    for key in ("position", "previous", "velocity"):
        assert bits_equal(state[key], ref[f"checkpoint/{key}"]), key
    # The damping is not a no-op on these wings.
    assert relative_error(ref["checkpoint/velocity"], undamped["checkpoint/velocity"]) > 1e-6
