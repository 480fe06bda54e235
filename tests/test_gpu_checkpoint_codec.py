"""The reference's checkpoint bytes (SWST around SWBC and, with a pilot,
SWSC) read and written by the device structure (needs an MI355X).

tests/golden/<preset>.structure.w1.swfx holds serializeStructureCheckpoint's
bytes of the reference's own Structure before its first step and after 20
steps (tiny wings: the bytes; small wings: SHA-256 digests). A fresh device
structure writes the reference's initial bytes exactly; the reference's
bytes after 20 steps, read into a checkpoint of this build, come out byte for
byte and restore the nodes, payload and suspension state they hold; a device
checkpoint written after its own 20 steps restates the suspension's
diagnostics within 1e-10 of the reference's; a device checkpoint written,
read and restored into a fresh structure steps on bit-identically; and
malformed or mismatched bytes are refused.
"""
import pytest

import hashlib

import numpy as np

import simwing_amd as sw
from conftest import bits_equal, relative_error
from simwing_amd import native
from test_structure_reference import colour_ordered

pytestmark = pytest.mark.gpu


def structure_of(preset):
    """A structure of the preset with its hinges in the order the reference's
    Structure was given them (the fixtures' definition fingerprint)."""
    assembly = sw.generate_scene(preset).assemble()
    return sw.Structure(colour_ordered(assembly.definition)), assembly.flight_loads()


def digest(data: bytes) -> np.ndarray:
    return np.frombuffer(hashlib.sha256(data).digest(), np.uint64)


@pytest.mark.parametrize("preset", ["tiny", "tiny-tethered", "tiny-free", "small", "small-tethered"])
def test_fresh_structure_writes_the_reference_initial_bytes(golden, preset):
    print(f"{preset}: checkpoint of a fresh structure", flush=True)
    ref = golden(f"{preset}.structure.w1.swfx")
    structure, _ = structure_of(preset)
    data = structure.serialize_checkpoint(structure.checkpoint())
    if preset.startswith("tiny"):
        assert data == ref["checkpoint_bytes_initial"].tobytes()
    else:
        assert len(data) == int(ref["checkpoint_bytes_initial_size"][0])
        assert np.array_equal(digest(data), ref["checkpoint_bytes_initial_digest"])


@pytest.mark.parametrize("preset", ["tiny", "tiny-tethered", "tiny-free"])
def test_reference_bytes_round_trip_exactly(golden, preset):
    print(f"{preset}: reference checkpoint bytes", flush=True)
    ref = golden(f"{preset}.structure.w1.swfx")
    data = bytes(ref["checkpoint_bytes"])
    structure, _ = structure_of(preset)
    checkpoint = structure.deserialize_checkpoint(data)
    info = checkpoint.info()
    assert info["accepted_steps"] == int(ref["checkpoint/accepted_steps"][0])
    assert info["time"] == float(ref["checkpoint/time"][0])
    nodes = checkpoint.nodes()
    for key in ("position", "previous", "velocity", "pending_force", "last_applied"):
        assert bits_equal(nodes[key], ref[f"checkpoint/{key}"]), key
    assert structure.serialize_checkpoint(checkpoint) == data
    # Restored, the structure holds the reference's nodes and writes the
    # same bytes from its own checkpoint.
    structure.restore(checkpoint)
    x, prev, v = structure.node_states()
    assert bits_equal(x.reshape(-1), ref["checkpoint/position"])
    assert bits_equal(v.reshape(-1), ref["checkpoint/velocity"])
    assert structure.serialize_checkpoint(structure.checkpoint()) == data
    if preset == "tiny":
        # suspensionState() of the restored structure is the reference's after
        # its 20th step: payload, harness points, and the per-segment length,
        # commanded rest length and tension its checkpoint's diagnostics hold.
        state = structure.suspension_state()
        for key in ("payload", "segment_length", "segment_commanded_rest", "segment_tension"):
            assert bits_equal(state[key], ref[f"final/{key}"]), key
        assert np.array_equal(state["segment_ids"], ref["final/segment_ids"])


def test_device_checkpoint_restates_the_suspension_diagnostics(golden):
    """20 device steps of the tiny pilot wing, written and read back into a
    fresh structure: its suspensionState() (from the written diagnostics)
    is the device's own and within 1e-10 of the reference's."""
    print("tiny: 20 steps, write, read into a fresh structure", flush=True)
    ref = golden("tiny.structure.w1.swfx")
    structure, loads = structure_of("tiny")
    st = sw.flight_step_settings()
    for _ in range(20):
        structure.set_external_forces(loads)
        structure.step(st)
    structure.set_external_forces(loads)
    data = structure.serialize_checkpoint(structure.checkpoint())
    assert len(data) == ref["checkpoint_bytes"].size
    own = structure.suspension_state()
    fresh, _ = structure_of("tiny")
    fresh.restore(fresh.deserialize_checkpoint(data))
    restored = fresh.suspension_state()
    for key in ("payload", "segment_length", "segment_commanded_rest", "segment_tension", "harness_position"):
        assert relative_error(restored[key], own[key]) <= 1e-15, key
        err = relative_error(restored[key], ref[f"final/{key}"])
        print(f"  {key}: {err:.2e} vs the reference", flush=True)
        assert err <= 1e-10, key
    assert fresh.serialize_checkpoint(fresh.checkpoint()) == data
    # Stepped on, both structures agree bit for bit.
    for s in (structure, fresh):
        s.step(st)
    for a, b in zip(structure.node_states(), fresh.node_states()):
        assert bits_equal(a.reshape(-1), b.reshape(-1))


def test_device_checkpoint_bytes_restore_bit_identically(golden):
    print("small-tethered: 5 steps, write, read into a fresh structure, 5 more steps", flush=True)
    structure, loads = structure_of("small-tethered")
    st = sw.flight_step_settings()
    for _ in range(5):
        structure.set_external_forces(loads)
        structure.step(st)
    structure.set_external_forces(loads)
    data = structure.serialize_checkpoint(structure.checkpoint())
    ref = golden("small-tethered.structure.w1.swfx")
    assert len(data) == int(ref["checkpoint_bytes_size"][0])
    fresh, _ = structure_of("small-tethered")
    fresh.restore(fresh.deserialize_checkpoint(data))
    assert fresh.accepted_step_count() == 5
    for s in (structure, fresh):
        for _ in range(5):
            s.step(st)  # the pending loads travel in the checkpoint
            s.set_external_forces(loads)
    for a, b in zip(structure.node_states(), fresh.node_states()):
        assert bits_equal(a.reshape(-1), b.reshape(-1))
    assert structure.serialize_checkpoint(structure.checkpoint()) == fresh.serialize_checkpoint(fresh.checkpoint())


def test_malformed_and_mismatched_bytes_are_refused(golden):
    data = bytearray(golden("tiny-tethered.structure.w1.swfx")["checkpoint_bytes"].tobytes())
    structure, _ = structure_of("tiny-tethered")
    cases = {
        "truncated": bytes(data[:-1]),
        "trailing": bytes(data) + b"\\0",
        "checksum": bytes(data[:100]) + bytes([data[100] ^ 1]) + bytes(data[101:]),
        "magic": b"SWSX" + bytes(data[4:]),
    }
    for name, bad in cases.items():
        with pytest.raises(sw.NativeError) as error:
            structure.deserialize_checkpoint(bad)
        assert error.value.status == native.SW_INVALID_ARGUMENT, name
    # Another definition: the free wing's bytes do not fit the tethered wing.
    other = golden("tiny-free.structure.w1.swfx")["checkpoint_bytes"].tobytes()
    with pytest.raises(sw.NativeError) as error:
        structure.deserialize_checkpoint(other)
    assert "different definition" in str(error.value)
    # Damped lines' impulses are not recorded by the device: after a step the
    # suspension section cannot be restated.
    assembly = sw.generate_scene("tiny").assemble()
    d = colour_ordered(assembly.definition)
    d.suspension["segment_damping"] = np.full_like(d.suspension["segment_damping"], 40.0)
    damped = sw.Structure(d)
    damped.set_external_forces(assembly.flight_loads())
    damped.step(sw.flight_step_settings())
    with pytest.raises(sw.NativeError) as error:
        damped.serialize_checkpoint(damped.checkpoint())
    assert error.value.status == native.SW_INVALID_ARGUMENT and "damping" in str(error.value)
