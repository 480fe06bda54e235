"""Diagnostic frames from the device (needs an MI355X).

The device's frame fields (sw_structure_frame_fields) are bit-equal to the
oracle's under the device's portable transcendentals, within 1e-10 of the
reference's own buildStructureFrame run (libm transcendentals;
tests/golden/<preset>.frame.swfx), and sw_structure_frame's bytes are the
encoding of exactly those fields with the structure's own diagnostics and
progress. At the export wing's size the fields are checked against the
portable oracle after two steps.
"""
import numpy as np
import pytest

import oracle as orc
import simwing_amd as sw
from conftest import bits_equal, relative_error

pytestmark = pytest.mark.gpu

FIELDS = ("vertex_position", "vertex_velocity", "vertex_pending_force", "line_length", "line_violation")
TOL = 1e-10
DT = 1.0 / 120.0


def run_pair(preset, steps):
    """Device structure and portable-math oracle, no face pressure, flight
    loads before each step and pending again after the last."""
    assembly = sw.generate_scene(preset).assemble()
    d, loads = assembly.definition, assembly.flight_loads()
    structure = sw.Structure(d)
    body = orc.OracleBody(d)
    body.use_portable_math(True)
    body.set_external_forces(loads)
    settings = sw.flight_step_settings()
    for _ in range(steps):
        structure.set_external_forces(loads)
        structure.step(settings)
        body.step(settings)
    structure.set_external_forces(loads)
    return d, structure, body


def flat(fields):
    return {k: np.asarray(v, np.float64).reshape(-1) for k, v in fields.items()}


@pytest.mark.parametrize("preset", ["tiny", "small"])
def test_device_frame_matches_reference_frame(golden, preset):
    frame = golden(f"{preset}.frame.swfx")
    steps = int(frame["steps"][0])
    print(f"frame {preset}: {steps} steps", flush=True)
    d, structure, body = run_pair(preset, steps)
    got = flat(structure.frame_fields())
    portable = body.frame_fields()
    for key in FIELDS:
        assert bits_equal(got[key], portable[key]), f"{preset}: device {key} differs from the portable oracle"
        err = relative_error(got[key], frame[key])
        print(f"  {key}: {err:.2e} vs buildStructureFrame", flush=True)
        assert err <= TOL, f"{preset}: {key} relative error {err:.3e} vs the reference frame"

    # Global fields: the structure's diagnostics against the reference's.
    diagnostics = structure.diagnostics()
    g = frame["global_scalars"]
    v = frame["global_vectors"].reshape(4, 3)
    scalars = ("total_dynamic_mass", "kinetic_energy", "maximum_distance_error", "maximum_cable_extension",
               "maximum_absolute_membrane_strain", "maximum_membrane_residual", "maximum_contact_penetration",
               "maximum_suspension_residual")
    for k, name in enumerate(scalars):
        assert abs(diagnostics[name] - g[k]) <= 1e-8 * max(1.0, abs(g[k])), (name, diagnostics[name], g[k])
    vectors = ("pending_external_force", "last_applied_external_force", "center_of_mass", "linear_momentum")
    for k, name in enumerate(vectors):
        assert relative_error(np.asarray(diagnostics[name]), v[k]) <= 1e-9, (name, diagnostics[name], v[k])

    # The frame's bytes: exactly the encoding of the device's fields with the
    # structure's diagnostics and progress.
    context = sw.frame_context("synthetic-wing", "reference", DT)
    mapping = sw.FrameMapping.sequential(d)
    data = structure.frame(mapping, context)
    accepted, time = structure.accepted_step_count(), structure.simulation_time_seconds()
    assert accepted == steps
    expected = sw.frame_encode(d, mapping, context, accepted, time, got, diagnostics)
    assert data == expected
    if preset == "tiny":
        assert len(data) == frame["frame_bytes"].size


def test_device_frame_export_size():
    """The ~70k-triangle wing: frame fields bit-equal to the portable oracle
    after two steps."""
    print("frame export: 2 steps", flush=True)
    d, structure, body = run_pair("export", 2)
    got = flat(structure.frame_fields())
    expected = body.frame_fields()
    for key in FIELDS:
        assert bits_equal(got[key], expected[key]), f"export: device {key} differs from the portable oracle"
    assert np.isfinite(got["line_violation"]).all() and (got["line_violation"] >= 0.0).all()
    data = structure.frame()
    assert data[:4] == b"SWVF" and int.from_bytes(data[8:16], "little") == len(data) - 16
