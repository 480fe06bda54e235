"""Membranes with a compressionStiffnessRatio below 1 against the reference
(CPU only).

A ratio r < 1 softens each compressed material direction of a membrane by
sqrt(r) (shear by the geometric mean of the two) for the projection, whose
compliance is then the inverse of that state's stiffness
(membrane.cpp:273-296, soft_body.cpp:1869-1873); the solver diagnostics keep
the material compliance. The scene assembly always emits 1
(scene_structure.cpp:424-438), so oracle/ref_driver --structure set every
membrane's ratio to 0.25 on the tiny and small wings and stepped the
reference's own simwing::fsi::Structure 20 times
(<preset>.structure.compressed.swfx).
"""
import numpy as np
import pytest

import oracle as orc
import simwing_amd as sw
from conftest import bits_equal, relative_error
from test_structure_reference import colour_ordered

COMPRESSION_RATIO = 0.25


def compressed(definition, ratio=COMPRESSION_RATIO):
    d = definition.copy()
    material = d.arrays["membrane_material"].reshape(-1, 8).copy()
    material[:, 7] = ratio
    d.arrays["membrane_material"] = material.reshape(-1)
    return d


@pytest.mark.parametrize("preset", ["tiny", "small"])
def test_oracle_compressed_membranes_match_reference_structure(golden, preset):
    ref = golden(f"{preset}.structure.compressed.swfx")
    plain = golden(f"{preset}.structure.w1.swfx")
    assembly = sw.generate_scene(preset).assemble()
    d, loads = compressed(assembly.definition), assembly.flight_loads()
    assert sw.definition_fingerprint(colour_ordered(d)) == int(ref["fingerprint"][0])
    body = orc.OracleBody(d)
    st = sw.flight_step_settings()
    st.worker_threads = 1
    for _ in range(20):
        body.set_external_forces(loads)
        body.step(st)
    state = body.state()
    for key in ("position", "previous", "velocity"):
        assert bits_equal(state[key], ref[f"checkpoint/{key}"]), key
    # The ratio is not a no-op on these wings.
    assert relative_error(ref["checkpoint/position"], plain["checkpoint/position"]) > 1e-6


def test_ratio_outside_the_material_range_is_rejected():
    """validateOrthotropicMembraneMaterial (membrane.cpp:322-323): 0 < r <= 1
    (definition_fingerprint validates the definition as a structure does)."""
    d = sw.generate_scene("tiny").assemble().definition
    for ratio in (0.0, -0.5, 1.5, float("nan")):
        with pytest.raises(sw.NativeError):
            sw.definition_fingerprint(compressed(d, ratio))
