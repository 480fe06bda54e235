"""sw_structure_frame_fields_noalloc against sw_structure_frame_fields (needs
an MI355X).

The entry point that allocates nothing stages the frame fields in the
library's static device array; its outputs must equal the allocating entry
point's bit for bit, on two calls in a row (the second on a smaller
structure, so values left in the array by the first call are not read), and
a structure whose fields exceed the array is rejected.
"""
import numpy as np
import pytest

import simwing_amd as sw
from conftest import bits_equal
from helpers import definition

pytestmark = pytest.mark.gpu

FIELDS = ("vertex_position", "vertex_velocity", "vertex_pending_force", "line_length", "line_violation")


def stepped(preset, steps):
    assembly = sw.generate_scene(preset).assemble()
    d, loads = assembly.definition, assembly.flight_loads()
    structure = sw.Structure(d)
    settings = sw.flight_step_settings()
    for _ in range(steps):
        structure.set_external_forces(loads)
        structure.step(settings)
    structure.set_external_forces(loads)  # pending again: the frame reports them
    return structure


def test_noalloc_frame_fields_equal_allocating_entry_point():
    for preset, steps in (("small", 3), ("tiny", 2)):
        print(f"noalloc frame fields: {preset}, {steps} steps", flush=True)
        structure = stepped(preset, steps)
        want = structure.frame_fields()
        got = structure.frame_fields(noalloc=True)
        for key in FIELDS:
            assert bits_equal(got[key], want[key]), f"{preset}: {key} differs from sw_structure_frame_fields"
        assert np.isfinite(got["vertex_position"]).all()
        structure.close()


def test_noalloc_frame_fields_reject_structures_beyond_the_static_array():
    n = 350_000  # 6 n doubles > 2,097,152
    positions = np.zeros((n, 3))
    positions[:, 0] = np.arange(n, dtype=np.float64)
    structure = sw.Structure(definition(positions, np.ones(n), fixed=np.ones(n, np.uint8)))
    with pytest.raises(sw.NativeError) as failure:
        structure.frame_fields(noalloc=True)
    assert failure.value.status == 1 and "static staging maximum" in str(failure.value)
    structure.close()
