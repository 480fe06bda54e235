"""Load-path cable pairs on the device (needs an MI355X).

The tethered wings (no pilot; lines as Cable constraints to two fixed riser
junctions) stepped with cable sweep pairs: bit-equal to the portable-math
oracle, within 1e-10 of the reference's own run (tests/golden), for the
surface-flight setting (3 pairs, all closing) and for pair counts that
interleave pairs inside the iterations, with no iterations, and with the
serial sweep.
"""
import numpy as np
import pytest

import oracle as orc
import simwing_amd as sw
from conftest import bits_equal, relative_error

pytestmark = pytest.mark.gpu

TOL = 1e-10


def pair(preset, steps, settings):
    scene = sw.generate_scene(preset)
    assembly = scene.assemble()
    d, pressure, loads = assembly.definition, scene.face_pressure(), assembly.flight_loads()
    structure = sw.Structure(d)
    structure.set_face_pressure(pressure)
    body = orc.OracleBody(d)
    body.use_portable_math(True)
    body.set_face_pressure(pressure)
    for _ in range(steps):
        structure.set_external_forces(loads)
        structure.step(settings)
        body.set_external_forces(loads)
        body.step(settings)
    return structure, body


def compare(structure, body, label):
    x, prev, v = structure.node_states()
    expected = body.state()
    for key, got in (("position", x), ("previous", prev), ("velocity", v)):
        assert bits_equal(got.reshape(-1), expected[key]), f"{label}: {key} differs from the portable oracle"
    solver = structure.solver_state()
    assert bits_equal(solver["constraint_lambda"], expected["constraint_lambda"]), f"{label}: cable multipliers"
    return x.reshape(-1), v.reshape(-1)


def test_tethered_wings_match_reference(golden):
    st = sw.flight_step_settings()
    print("tiny-tethered: 20 steps, 3 closing cable pairs", flush=True)
    structure, body = pair("tiny-tethered", 20, st)
    x, v = compare(structure, body, "tiny-tethered")
    ref = golden("tiny-tethered.states.swfx")
    assert relative_error(x, ref["step20/position"]) <= TOL
    assert relative_error(v, ref["step20/velocity"]) <= TOL
    print("small-tethered: 20 steps", flush=True)
    structure, body = pair("small-tethered", 20, st)
    x, v = compare(structure, body, "small-tethered")
    ref = golden("small-tethered.digest.swfx")
    for key, got in (("position", x), ("velocity", v)):
        rows = got.reshape(-1, 3)[ref[f"step20/{key}/rows"].astype(np.int64)].reshape(-1)
        err = float(np.abs(rows - ref[f"step20/{key}/samples"]).max()) / float(ref[f"step20/{key}/maxabs"][0])
        print(f"  small-tethered {key}: sampled rel err {err:.2e} vs the reference", flush=True)
        assert err <= TOL


@pytest.mark.parametrize("pairs,iterations,workers", [(0, 48, 1), (1, 48, 1), (4, 48, 1), (8, 48, 1), (5, 3, 1),
                                                      (3, 0, 1), (3, 48, 0), (4, 48, 0)])
def test_cable_pair_programs_match_oracle(pairs, iterations, workers):
    """Interleaved pairs (4 and more), more pairs than iterations, no
    iterations at all, and the serial sweep: bit-equal to the oracle."""
    st = sw.flight_step_settings()
    st.cable_sweep_pairs = pairs
    st.constraint_iterations = iterations
    st.worker_threads = workers
    print(f"tiny-tethered: pairs={pairs} iterations={iterations} workers={workers}", flush=True)
    structure, body = pair("tiny-tethered", 3, st)
    compare(structure, body, f"pairs={pairs} iterations={iterations} workers={workers}")
