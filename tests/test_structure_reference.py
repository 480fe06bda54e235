"""The Structure adapter surface against the reference's own Structure (CPU only).

`<preset>.structure.w<N>.swfx` were written by oracle/ref_driver --structure:
the reference's simwing::fsi::Structure (structure.cpp) stepped 20 times with
the flight loads and workerThreads N (0: the serial definition-order sweep,
1: the coloured sweep), its definitionFingerprint(), suspensionState() and
the checkpoint's public fields (tests/golden/make_fixtures.py).
"""
import numpy as np
import pytest

import oracle as orc
import simwing_amd as sw
from conftest import bits_equal

STEPS = 20


def colour_ordered(definition):
    """The definition with its hinges listed in the build's colour-major
    order, the order the reference's Structure was given them in."""
    order, _ = sw.definition_colouring(definition, 2)
    order = order.astype(np.int64)
    out = definition.copy()
    a = out.arrays
    a["dihedral_nodes"] = a["dihedral_nodes"].reshape(-1, 4)[order].reshape(-1).copy()
    a["dihedral_rest_angle"] = a["dihedral_rest_angle"][order].copy()
    a["dihedral_compliance"] = a["dihedral_compliance"][order].copy()
    return out


@pytest.mark.parametrize("preset", ["tiny", "small"])
def test_definition_fingerprint_matches_reference(golden, preset):
    d = sw.generate_scene(preset).assemble().definition
    want = int(golden(f"{preset}.structure.w1.swfx")["fingerprint"][0])
    assert sw.definition_fingerprint(colour_ordered(d)) == want
    # The fingerprint covers the hinge list order: the definition-order list differs.
    assert sw.definition_fingerprint(d) != want


@pytest.mark.parametrize("preset", ["tiny", "small"])
@pytest.mark.parametrize("workers", [0, 1])
def test_oracle_sweeps_match_reference_structure(golden, preset, workers):
    """The oracle's serial (workerThreads 0) and coloured sweeps against the
    reference Structure's committed nodes after 20 steps, bit for bit."""
    ref = golden(f"{preset}.structure.w{workers}.swfx")
    assembly = sw.generate_scene(preset).assemble()
    loads = assembly.flight_loads()
    body = orc.OracleBody(assembly.definition)
    settings = sw.flight_step_settings()
    settings.worker_threads = workers
    for _ in range(STEPS):
        body.set_external_forces(loads)
        body.step(settings)
    state = body.state()
    for key, field in (("position", "position"), ("previous", "previous"), ("velocity", "velocity")):
        assert bits_equal(state[key], ref[f"checkpoint/{field}"]), f"{preset} workers={workers}: {key}"
    assert bits_equal(state["payload"], ref["final/payload"])
    # The two sweep orders give different trajectories.
    other = golden(f"{preset}.structure.w{1 - workers}.swfx")
    assert not bits_equal(ref["checkpoint/position"], other["checkpoint/position"])
