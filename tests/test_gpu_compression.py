"""Membranes with a compressionStiffnessRatio below 1 on the device (needs an
MI355X): every membrane of the tiny and small wings at ratio 0.25, stepped 20
times, bit-equal to the portable-math oracle and within 1e-10 (velocities 1e-9) of the
reference's own Structure run (tests/golden/<preset>.structure.compressed.swfx),
for the default launch and a multi-workgroup one; the adjoints refuse it.
"""
import pytest

import oracle as orc
import simwing_amd as sw
from conftest import bits_equal, relative_error
from test_compression_reference import compressed

pytestmark = pytest.mark.gpu

TOL = 1e-10
# Velocities are difference quotients (x - prev) / dt, so they carry the
# position gap between the portable transcendentals (device, portable
# oracle) and glibc's (reference) magnified: the two oracle modes differ by
# 1.0e-12 in position and 1.2e-10 in velocity on the small compressed wing
# after 20 steps.
VELOCITY_TOL = 1e-9


@pytest.mark.parametrize("preset,launch", [("tiny", None), ("small", None), ("small", (16, 256))])
def test_compressed_membranes_match_reference_structure(golden, preset, launch):
    print(f"{preset}: compression ratio 0.25, 20 steps, launch {launch}", flush=True)
    ref = golden(f"{preset}.structure.compressed.swfx")
    assembly = sw.generate_scene(preset).assemble()
    d, loads = compressed(assembly.definition), assembly.flight_loads()
    structure = sw.Structure(d)
    if launch:
        structure.set_launch(*launch)
    body = orc.OracleBody(d)
    body.use_portable_math(True)
    st = sw.flight_step_settings()
    st.worker_threads = 1
    for _ in range(20):
        structure.set_external_forces(loads)
        structure.step(st)
        body.set_external_forces(loads)
        body.step(st)
    x, prev, v = structure.node_states()
    expected = body.state()
    for key, got in (("position", x), ("previous", prev), ("velocity", v)):
        assert bits_equal(got.reshape(-1), expected[key]), f"{key} differs from the portable oracle"
        err = relative_error(got.reshape(-1), ref[f"checkpoint/{key}"])
        print(f"  {preset} {key}: {err:.2e} vs the reference Structure", flush=True)
        assert err <= (VELOCITY_TOL if key == "velocity" else TOL)
    solver = structure.solver_state()
    assert bits_equal(solver["membrane_lambda"], expected["membrane_lambda"])
    assert bits_equal(solver["membrane_residual"], expected["membrane_residual"])


def test_adjoint_refuses_compressed_membranes():
    d = compressed(sw.generate_scene("tiny").assemble().definition)
    structure = sw.Structure(d)
    with pytest.raises(sw.NativeError) as error:
        sw.Adjoint(structure, 1)
    assert "does not differentiate" in str(error.value)
