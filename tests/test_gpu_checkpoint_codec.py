"""The reference's checkpoint bytes (SWST around SWBC) read and written by
the device structure (needs an MI355X).

tests/golden/<preset>.structure.w1.swfx holds serializeStructureCheckpoint's
bytes of the reference's own Structure after 20 steps of the tethered and
free wings (no suspension). Read into a checkpoint of this build and written
back, they come out byte for byte; the checkpoint restores the nodes they
hold; a device checkpoint written, read and restored into a fresh structure
steps on bit-identically; and malformed or mismatched bytes are refused.
"""
import pytest

import simwing_amd as sw
from conftest import bits_equal
from simwing_amd import native
from test_structure_reference import colour_ordered

pytestmark = pytest.mark.gpu


def structure_of(preset):
    """A structure of the preset with its hinges in the order the reference's
    Structure was given them (the fixtures' definition fingerprint)."""
    assembly = sw.generate_scene(preset).assemble()
    return sw.Structure(colour_ordered(assembly.definition)), assembly.flight_loads()


@pytest.mark.parametrize("preset", ["tiny-tethered", "tiny-free"])
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
    # A suspension section is not written by this codec.
    pilot, _ = structure_of("tiny")
    with pytest.raises(sw.NativeError) as error:
        pilot.serialize_checkpoint(pilot.checkpoint())
    assert error.value.status == native.SW_INVALID_ARGUMENT
