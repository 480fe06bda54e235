"""Scene-v2 compact binary stream (sw_scene_read / sw_scene_write) against
the reference's own writeScene/readScene (CPU only).

tests/golden/tiny.scene.bin is the reference's writeScene stream of the
seeded tiny wing, and small.scene.digest.swfx the SHA-256 digest of the small
wing's stream (``oracle/_ref/ref_driver --scene``, which also checks that the
reference decodes and re-encodes the stream to itself). The library must
write the same bytes, read them back to a scene whose assembly equals the
generated one, and reject malformed streams with readScene's messages. The
layout is pinned independently by a field-by-field walk of the stream
(scene.cpp:1071-1204) whose entity counts must match the assembly's.
"""
import ctypes as C
import hashlib
import struct

import numpy as np
import pytest

import simwing_amd as sw
from conftest import GOLDEN
from simwing_amd import native

MAGIC = b"SIMWSCN2"


def reference_tiny() -> bytes:
    return (GOLDEN / "tiny.scene.bin").read_bytes()


def walk(data: bytes) -> dict:
    """Entity counts and field offsets of a version-3 stream."""
    at = 0

    def take(n):
        nonlocal at
        chunk = data[at:at + n]
        assert len(chunk) == n, "truncated"
        at += n
        return chunk

    u8 = lambda: take(1)[0]  # noqa: E731
    u32 = lambda: struct.unpack("<I", take(4))[0]  # noqa: E731
    u64 = lambda: struct.unpack("<Q", take(8))[0]  # noqa: E731
    real = lambda: struct.unpack("<d", take(8))[0]  # noqa: E731

    def string():
        return take(u32()).decode()

    assert take(8) == MAGIC
    out = {"version": u32(), "schema": (u32(), u32()), "units": (u8(), u8(), u8()), "scale": real(),
           "checksum": string(), "exporter": string()}
    counts = {}
    n = counts["regions"] = u32()
    for _ in range(n):
        u64(), u8(), string()
    n = counts["vertices"] = u32()
    ids, offsets = [], []
    for _ in range(n):
        offsets.append(at)
        ids.append(u64())
        take(24)
    out["vertex_ids"] = ids
    out["vertex_offsets"] = offsets
    n = counts["fabric"] = u32()
    for _ in range(n):
        u64(), string(), take(64)
    n = counts["seam_materials"] = u32()
    for _ in range(n):
        u64(), string(), take(16)
    n = counts["triangles"] = u32()
    out["triangle_offsets"] = []
    for _ in range(n):
        out["triangle_offsets"].append(at)
        take(8 + 24 + 48 + 32)
        u8()
    n = counts["openings"] = u32()
    for _ in range(n):
        u64()
        take(8 * u32())
        take(16)
        u8()
        take(24 * u32())
    n = counts["seams"] = u32()
    for _ in range(n):
        take(16)
        take(8 * u32())
        take(8 * u32())
    n = counts["line_materials"] = u32()
    for _ in range(n):
        u64(), string(), take(32)
    n = counts["pilots"] = u32()
    for _ in range(n):
        u64(), string(), take(8 + 24 + 24 + 32 + 24)
    n = counts["junctions"] = u32()
    out["junction_fixed_offsets"] = []
    for _ in range(n):
        take(8 + 24 + 8)
        out["junction_fixed_offsets"].append(at)
        assert u8() in (0, 1)
    n = counts["attachments"] = u32()
    for _ in range(n):
        u64(), u8(), take(8 + 8 + 24 + 8)
    n = counts["lines"] = u32()
    for _ in range(n):
        take(32 + 8)
        u8()
    out["counts"] = counts
    out["end"] = at
    return out


def read_status(data: bytes):
    lib = native.library()
    buf = np.frombuffer(data, np.uint8).copy()
    handle = C.c_void_p()
    status = lib.sw_scene_read(native.u8p(buf), buf.size, C.byref(handle))
    message = lib.sw_last_error().decode()
    if status == native.SW_OK:
        lib.sw_scene_destroy(handle)
    return status, message


def test_write_matches_reference_stream():
    assert sw.generate_scene("tiny").to_bytes() == reference_tiny()


def test_write_matches_reference_digest(golden):
    fixture = golden("small.scene.digest.swfx")
    data = sw.generate_scene("small").to_bytes()
    assert len(data) == int(fixture["size"][0])
    assert np.array_equal(np.frombuffer(hashlib.sha256(data).digest(), np.uint64), fixture["digest"])


def test_stream_layout_and_counts():
    data = reference_tiny()
    layout = walk(data)
    assert layout["end"] == len(data)
    assert layout["version"] == 3 and layout["schema"] == (2, 2) and layout["units"] == (1, 1, 1)
    assert layout["scale"] == 1.0 and layout["checksum"] == "synthetic-wing"
    assert layout["vertex_ids"] == sorted(layout["vertex_ids"])  # stable-ID order
    assembly = sw.generate_scene("tiny").assemble()
    counts = layout["counts"]
    assert counts["vertices"] == assembly.vertex_nodes
    assert counts["junctions"] == assembly.junction_nodes
    assert counts["triangles"] == assembly.definition.arrays["triangle_nodes"].size // 3


@pytest.mark.parametrize("preset", ["tiny", "small"])
def test_read_assembles_like_the_generated_scene(preset):
    generated = sw.generate_scene(preset)
    data = generated.to_bytes() if preset == "small" else reference_tiny()
    decoded = sw.read_scene(data)
    assert decoded.to_bytes() == data
    a = decoded.assemble().definition.arrays
    b = generated.assemble().definition.arrays
    assert a.keys() == b.keys()
    for key in b:
        assert np.array_equal(a[key].view(np.uint8), b[key].view(np.uint8)), key
    assert not decoded.face_pressure().any()


def test_read_assembles_like_the_reference_assembly(golden):
    reference = golden("tiny.assembly.swfx")
    arrays = sw.read_scene(reference_tiny()).assemble().definition.arrays
    assert np.array_equal(arrays["node_position"], reference["node_position"])
    assert np.array_equal(arrays["constraint_nodes"], reference["constraint_nodes"])


def test_version_two_stream():
    data = bytearray(reference_tiny())
    assert walk(bytes(data))["counts"]["openings"] == 0  # no cap-triangle counts to drop
    data[8:12] = struct.pack("<I", 2)
    assert sw.read_scene(bytes(data)).to_bytes() == reference_tiny()


def test_rejections():
    data = reference_tiny()
    layout = walk(data)
    cases = [
        (b"SIMWSCN3" + data[8:], "invalid magic header"),
        (data[:8] + struct.pack("<I", 4) + data[12:], "unsupported binary version"),
        (data[:-1], "truncated scene stream"),
        (data + b"\0", "trailing data"),
        (data[:5], "truncated scene stream"),
    ]
    offset = layout["junction_fixed_offsets"][0]
    cases.append((data[:offset] + b"\x02" + data[offset + 1:], "fixed flag is invalid"))
    for stream, text in cases:
        status, message = read_status(stream)
        assert status == native.SW_INVALID_ARGUMENT and text in message, (text, message)
    # A decodable stream whose scene fails the structural checks: the second
    # vertex given the first one's stable ID.
    second = layout["vertex_offsets"][1]
    duplicated = data[:second] + struct.pack("<Q", layout["vertex_ids"][0]) + data[second + 8:]
    status, message = read_status(duplicated)
    assert status == native.SW_INVALID_ARGUMENT and "vertex stable ID is duplicated" in message, message


def test_bending_topology_rejections():
    """UnsupportedBendingTopology (scene_structure.cpp:311-335): a triangle
    whose corners are listed in the opposite order shares its edges with its
    neighbours in the same direction, so no signed hinge exists; the scene is
    rejected with the reference's message for the first such edge."""
    data = reference_tiny()
    layout = walk(data)
    scene = sw.read_scene(data)
    assert scene is not None
    at = layout["triangle_offsets"][0]
    a, b, c = struct.unpack("<QQQ", data[at + 8:at + 32])
    charts = data[at + 32:at + 80]
    # Swap the second and third corners' vertices (the material chart keeps
    # its positive orientation).
    flipped = data[:at + 8] + struct.pack("<QQQ", a, c, b) + charts + data[at + 80:]
    flipped_scene = sw.read_scene(flipped)
    with pytest.raises(sw.NativeError) as error:
        flipped_scene.assemble()
    # The reference's assembleSceneStructure of the same stream reports this
    # first diagnostic (oracle/_ref/ref_driver --assemble on these bytes).
    assert str(error.value).endswith(
        "inconsistently oriented edge 1-18 in sheet 1 cannot define a signed fabric bending hinge")
