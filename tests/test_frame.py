"""Diagnostic frames on the CPU: the oracle's frame fields and the library's
SWVF encoder against the reference's own frames.

tests/golden/<preset>.frame.swfx were written by ``oracle/_ref/ref_driver
--frame`` (tests/golden/make_fixtures.py): the reference's
simwing::fsi::Structure stepped with the flight loads (no face pressure),
loads pending again after the last step, then
simwing::viewer::buildStructureFrame and serializeFrame
(structure_frame.cpp:480-666, viewer_protocol.cpp:838-882) under
sequential stable ids. The oracle (libm transcendentals, as the reference)
must reproduce the frame fields bit for bit, and sw_frame_encode must
reproduce the frame's bytes from the reference's own field values.
"""
import ctypes as C
import hashlib

import numpy as np
import pytest

import oracle as orc
import simwing_amd as sw
from conftest import bits_equal
from simwing_amd import native

FIELDS = ("vertex_position", "vertex_velocity", "vertex_pending_force", "line_length", "line_violation")
DT = 1.0 / 120.0


def oracle_frame(preset, steps):
    assembly = sw.generate_scene(preset).assemble()
    body = orc.OracleBody(assembly.definition)
    body.set_external_forces(assembly.flight_loads())
    settings = sw.flight_step_settings()
    for _ in range(steps):
        body.step(settings)
    return assembly.definition, body


def reference_diagnostics(frame):
    g = frame["global_scalars"]
    v = frame["global_vectors"].reshape(4, 3)
    return {"total_dynamic_mass": g[0], "kinetic_energy": g[1], "maximum_distance_error": g[2],
            "maximum_cable_extension": g[3], "maximum_absolute_membrane_strain": g[4],
            "maximum_membrane_residual": g[5], "maximum_contact_penetration": g[6],
            "maximum_suspension_residual": g[7], "pending_external_force": v[0],
            "last_applied_external_force": v[1], "center_of_mass": v[2], "linear_momentum": v[3]}


def accumulated_time(steps):
    t = 0.0
    for _ in range(steps):
        t += DT  # Structure::step adds the step's time (structure.cpp:826)
    return t


@pytest.mark.parametrize("preset", ["tiny", "small"])
def test_oracle_frame_fields_match_reference(golden, preset):
    frame = golden(f"{preset}.frame.swfx")
    definition, body = oracle_frame(preset, int(frame["steps"][0]))
    fields = body.frame_fields()
    for key in FIELDS:
        assert bits_equal(fields[key], frame[key]), f"{preset}: {key} differs from buildStructureFrame"
    assert bits_equal(body.state()["payload"], frame["payload"])
    # Independent of the restatement: the first N frame vertices are the
    # committed nodes, a harness point's pending force is zero, and every
    # violation follows from its length (|L - rest| or max(0, L - rest)).
    n = int(frame["node_count"][0])
    state = body.state()
    assert bits_equal(frame["vertex_position"][:3 * n], state["position"])
    assert bits_equal(frame["vertex_velocity"][:3 * n], state["velocity"])
    assert not frame["vertex_pending_force"][3 * n:].any()
    a = definition.arrays
    c = a["constraint_kind"].size
    extension = frame["line_length"][:c] - a["constraint_rest"]
    cable = a["constraint_kind"] == 2  # SW_CONSTRAINT_CABLE
    expected = np.where(cable, np.maximum(0.0, extension), np.abs(extension))
    assert bits_equal(frame["line_violation"][:c], expected)
    if definition.suspension is not None:
        seg = frame["line_length"][c:] - frame["segment_commanded_rest"]
        assert bits_equal(frame["line_violation"][c:], np.maximum(0.0, seg))
        assert bits_equal(frame["segment_commanded_rest"], definition.suspension["segment_rest"])


def test_frame_encoding_matches_reference_bytes(golden):
    frame = golden("tiny.frame.swfx")
    definition = sw.generate_scene("tiny").assemble().definition
    steps = int(frame["steps"][0])
    context = sw.frame_context("synthetic-wing", "reference", DT)
    data = sw.frame_encode(definition, sw.FrameMapping.sequential(definition), context, steps,
                           accumulated_time(steps), frame, reference_diagnostics(frame))
    assert data == frame["frame_bytes"].tobytes()
    # Line endpoints as the reference's mapping derived them.
    c = definition.arrays["constraint_kind"].size
    assert np.array_equal(frame["line_vertex"][:2 * c], definition.arrays["constraint_nodes"])


def test_frame_encoding_matches_reference_digest(golden):
    frame = golden("small.frame.swfx")
    definition = sw.generate_scene("small").assemble().definition
    steps = int(frame["steps"][0])
    context = sw.frame_context("synthetic-wing", "reference", DT)
    data = sw.frame_encode(definition, sw.FrameMapping.sequential(definition), context, steps,
                           accumulated_time(steps), frame, reference_diagnostics(frame))
    assert len(data) == int(frame["frame_bytes_size"][0])
    digest = np.frombuffer(hashlib.sha256(data).digest(), dtype=np.uint64)
    assert np.array_equal(digest, frame["frame_bytes_digest"])


def test_frame_envelope_layout(golden):
    """SWVF envelope: magic, version 1, reserved 0, payload size; then the
    scene checksum and solver commit strings (viewer_protocol.cpp:286-288,
    854-873)."""
    frame = golden("tiny.frame.swfx")
    definition = sw.generate_scene("tiny").assemble().definition
    context = sw.frame_context("abc", "xy", DT)
    data = sw.frame_encode(definition, sw.FrameMapping.sequential(definition), context, 3, 0.025, frame,
                           reference_diagnostics(frame))
    assert data[:4] == b"SWVF"
    assert int.from_bytes(data[4:6], "little") == 1 and int.from_bytes(data[6:8], "little") == 0
    assert int.from_bytes(data[8:16], "little") == len(data) - 16
    assert data[16:20] == (3).to_bytes(4, "little") and data[20:23] == b"abc"
    assert data[23:27] == (2).to_bytes(4, "little") and data[27:29] == b"xy"
    assert int.from_bytes(data[29:37], "little") == 3
    assert np.frombuffer(data[37:45], np.float64)[0] == 0.025
    # 3 fewer string bytes than the reference's context, everything else alike.
    assert len(data) == len(frame["frame_bytes"]) - (len("synthetic-wing") + len("reference") - 5)


def _encode_status(definition, mapping, context, fields, diagnostics, capacity=None):
    lib = native.library()
    view = definition.to_ctypes()
    m = mapping._ctypes()
    arrays = [np.ascontiguousarray(fields[k], np.float64) for k in FIELDS]
    d = sw._diagnostics_struct(diagnostics)
    size = C.c_size_t()
    out = np.zeros(capacity or 1, np.uint8)
    status = lib.sw_frame_encode(C.byref(view), C.byref(m), C.byref(context), 1, 0.0,
                                 *[native.f64p(x) for x in arrays], C.byref(d),
                                 native.u8p(out) if capacity is not None else None, capacity or 0, C.byref(size))
    return status, native.library().sw_last_error().decode(), size.value


def test_frame_rejections(golden):
    """The reference's rejections: mapping ids (structure_frame.cpp:17-31,
    205-264), segment identity (566-573) and validateFrame
    (viewer_protocol.cpp:690-836)."""
    frame = golden("tiny.frame.swfx")
    definition = sw.generate_scene("tiny").assemble().definition
    diagnostics = reference_diagnostics(frame)
    context = sw.frame_context("synthetic-wing", "reference", DT)
    good = sw.FrameMapping.sequential(definition)
    status, _, size = _encode_status(definition, good, context, frame, diagnostics)
    assert status == native.SW_OK and size == frame["frame_bytes"].size

    def variant(**changes):
        m = sw.FrameMapping.sequential(definition)
        for key, fn in changes.items():
            arr = getattr(m, key).copy()
            fn(arr)
            setattr(m, key, arr)
        return m

    cases = [
        (variant(vertex_ids=lambda a: a.__setitem__(3, a[2])), native.SW_INVALID_ARGUMENT, "vertex IDs must be"),
        (variant(triangle_ids=lambda a: a.__setitem__(0, 0)), native.SW_INVALID_ARGUMENT, "triangle IDs must be"),
        (variant(line_ids=lambda a: a.__setitem__(1, a[0])), native.SW_INVALID_ARGUMENT, "line IDs must be"),
        (variant(triangle_negative_region=lambda a: a.__setitem__(5, 0)), native.SW_INVALID_ARGUMENT,
         "nonzero side regions"),
        (variant(line_ids=lambda a: a.__setitem__(-1, a[-1] + 10_000)), native.SW_LOGIC_ERROR, "disagree"),
    ]
    for mapping, expected, text in cases:
        status, message, _ = _encode_status(definition, mapping, context, frame, diagnostics)
        assert status == expected and text in message, (status, message)
    short = sw.FrameMapping(good.vertex_ids[:-1], good.triangle_ids, good.triangle_negative_region,
                            good.triangle_positive_region, good.line_ids, good.line_roles)
    status, message, _ = _encode_status(definition, short, context, frame, diagnostics)
    assert status == native.SW_INVALID_ARGUMENT and "sizes do not match" in message
    bad = dict(frame)
    bad["line_length"] = frame["line_length"].copy()
    bad["line_length"][7] = np.nan
    status, message, _ = _encode_status(definition, good, context, bad, diagnostics, capacity=1 << 16)
    assert status == native.SW_INVALID_ARGUMENT and message.startswith("Invalid structure diagnostic frame")
    empty = sw.frame_context("", "reference", DT)
    status, message, _ = _encode_status(definition, good, empty, frame, diagnostics, capacity=1 << 16)
    assert status == native.SW_INVALID_ARGUMENT and "scene checksum must not be empty" in message
    status, message, size = _encode_status(definition, good, context, frame, diagnostics, capacity=100)
    assert status == native.SW_INVALID_ARGUMENT and size == frame["frame_bytes"].size
