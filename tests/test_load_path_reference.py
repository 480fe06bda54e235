"""Load-path cable pairs against the reference (CPU only).

The tethered wings (tiny-tethered, small-tethered: the tiny and small wings
with no rigid pilot, their lines ending at two fixed riser junctions) carry
their lines as Cable constraints, which the reference sweeps in load-path
order in cableConstraintSweepPairs reverse + forward pairs
(soft_body.cpp:528-629, 1265-1302, 1380-1388). oracle/ref_driver ran the
reference's own SoftBody on them with the surface-flight settings (3 pairs,
all closing); tests/golden holds the assembly, the colouring and the states
(tiny) or their SHA-256 digests and strided samples (small).
"""
import hashlib

import numpy as np
import pytest

import oracle as orc
import simwing_amd as sw
from conftest import bits_equal
from simwing_amd import native
from test_assembly_and_schedules import ARRAYS, same_bytes

KEYS = ("position", "previous", "velocity", "constraint_lambda", "membrane_lambda", "membrane_resultant",
        "membrane_residual")


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


def test_tethered_assembly_is_cables_and_matches_reference(golden):
    fixture = golden("tiny-tethered.assembly.swfx")
    assembly = sw.generate_scene("tiny-tethered").assemble()
    d = assembly.definition
    assert d.suspension is None
    kinds = d.arrays["constraint_kind"]
    assert (kinds == native.SW_CONSTRAINT_CABLE).sum() == 12 and d.arrays["node_fixed"].sum() == 2
    for name in ARRAYS:
        assert same_bytes(d.arrays[name], fixture[name]), f"{name} differs from assembleSceneStructure"
    colouring = golden("tiny-tethered.colouring.swfx")
    order, offsets = sw.definition_colouring(d, 0)
    assert np.array_equal(order, colouring["distance_order"]) and np.array_equal(offsets, colouring["distance_offsets"])
    small = sw.generate_scene("small-tethered").assemble().definition
    reference = golden("small-tethered.digest.swfx")
    for name in ARRAYS:
        assert np.array_equal(digest(small.arrays[name]), reference[f"digest/{name}"]), f"small: {name}"


def test_oracle_tiny_tethered_bit_for_bit(golden):
    ref = golden("tiny-tethered.states.swfx")
    order = golden("tiny-tethered.assembly.swfx")["dihedral_colour_order"].astype(np.int64)
    _, states = run_oracle("tiny-tethered", (1, 5, 20))
    for step, state in states.items():
        for key in KEYS:
            assert bits_equal(state[key], ref[f"step{step}/{key}"]), f"step {step}: {key}"
        assert bits_equal(state["dihedral_lambda"][order], ref[f"step{step}/dihedral_lambda"])


def test_oracle_small_tethered_digests(golden):
    ref = golden("small-tethered.digest.swfx")
    d, states = run_oracle("small-tethered", (1, 20))
    order, _ = sw.definition_colouring(d, 2)
    for step, state in states.items():
        for key in ("position", "previous", "velocity", "membrane_residual", "membrane_lambda", "constraint_lambda"):
            assert np.array_equal(digest(state[key]), ref[f"step{step}/{key}/digest"]), f"step {step}: {key}"
        assert np.array_equal(digest(state["dihedral_lambda"][order.astype(np.int64)]),
                              ref[f"step{step}/dihedral_lambda/digest"])


def test_cable_pairs_change_the_trajectory():
    """The pairs are not a no-op on a tethered wing: 0 and 3 pairs differ."""
    st0 = sw.flight_step_settings()
    st0.cable_sweep_pairs = 0
    _, a = run_oracle("tiny-tethered", (2,), st0)
    _, b = run_oracle("tiny-tethered", (2,))
    assert not bits_equal(a[2]["position"], b[2]["position"])
