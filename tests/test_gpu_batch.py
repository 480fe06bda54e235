"""Batched wings on the device (needs an MI355X): many wings of uneven size
in one persistent launch return, wing for wing, the bits each wing returns
when stepped alone; a failing wing is rolled back on its own."""
import numpy as np
import pytest

import simwing_amd as sw
from conftest import bits_equal
from helpers import flight_case

pytestmark = pytest.mark.gpu

# LDS-resident wings (tiny, small), two generated batch wings and the
# ~70k-triangle export wing (multi-workgroup with a chain workgroup).
PRESETS = ["tiny", "small", "batch:4242:0:8", "batch:4242:5:8", "export"]


def solo_states(cases, steps):
    out = []
    for d, pressure, loads in cases:
        s = sw.Structure(d)
        s.set_face_pressure(pressure)
        for _ in range(steps):
            s.set_external_forces(loads)
            s.step(sw.flight_step_settings())
        x, _, v = s.node_states()
        out.append((x, v))
        s.close()
    return out


def test_batch_equals_solo_structures():
    cases = [flight_case(p) for p in PRESETS]
    batch = sw.Batch([d for d, _, _ in cases])
    blocks = batch.workgroups()
    print("workgroups per wing:", dict(zip(PRESETS, blocks.tolist())), flush=True)
    assert blocks[0] == 1 and blocks[1] == 1 and blocks[-1] >= 3
    for w, (_, pressure, _) in enumerate(cases):
        batch.set_face_pressure(w, pressure)
    steps = 4
    for _ in range(steps):
        for w, (_, _, loads) in enumerate(cases):
            batch.set_external_forces(w, loads)
        batch.step(1)
    expected = solo_states(cases, steps)
    for w, name in enumerate(PRESETS):
        x, v = batch.node_states(w)
        assert bits_equal(x, expected[w][0]) and bits_equal(v, expected[w][1]), f"{name}: batch differs from solo"
    batch.close()


def test_batch_wing_alone_and_replanned():
    """A wing of the batch stepped alone through its own structure handle,
    then the batch re-planned and stepped: still equal to solo stepping."""
    cases = [flight_case(p) for p in ("tiny", "batch:4242:1:8")]
    batch = sw.Batch([d for d, _, _ in cases])
    for w, (_, pressure, _) in enumerate(cases):
        batch.set_face_pressure(w, pressure)
    wing = batch.structure(1)
    wing.set_launch(32, 256)
    wing.set_external_forces(cases[1][2])
    wing.step(sw.flight_step_settings())
    batch.set_launch(8, 256)
    batch.set_external_forces(0, cases[0][2])
    batch.set_external_forces(1, cases[1][2])
    batch.step(1)
    assert batch.structure(0).accepted_step_count() == 1 and wing.accepted_step_count() == 2
    expected = solo_states(cases, 1)[0], solo_states(cases[1:], 2)[0]
    for w in range(2):
        x, v = batch.node_states(w)
        assert bits_equal(x, expected[w][0]) and bits_equal(v, expected[w][1])
    batch.close()


def test_failing_wing_is_rolled_back_alone():
    good = flight_case("tiny")
    bad_d, bad_p, bad_l = flight_case("tiny")
    bad_d = bad_d.copy()
    bad_d.suspension["maximum_line_residual"] = 1e-15
    batch = sw.Batch([good[0], bad_d])
    batch.set_face_pressure(0, good[1])
    batch.set_face_pressure(1, bad_p)
    before = batch.node_states(1)
    batch.set_external_forces(0, good[2])
    batch.set_external_forces(1, bad_l)
    with pytest.raises(sw.NativeError) as error:
        batch.step(1)
    assert error.value.entity == "line-residual"
    after = batch.node_states(1)
    assert bits_equal(before[0], after[0]) and bits_equal(before[1], after[1])
    assert batch.structure(0).accepted_step_count() == 1
    assert batch.structure(1).accepted_step_count() == 0
    x, v = batch.node_states(0)
    expected = solo_states([good], 1)[0]
    assert bits_equal(x, expected[0]) and bits_equal(v, expected[1])
    batch.close()


def test_config_batch_of_64_wings_equals_solo():
    """The benchmark's batched configuration: the 64 generated wings of seed
    0x5EED0002 (2k-40k nodes each) in one persistent launch, 3 steps, each
    wing's bits against the same wing stepped alone."""
    count, steps = 64, 3
    names = [f"batch:{0x5EED0002}:{i}:{count}" for i in range(count)]
    cases = [flight_case(name) for name in names]
    print(f"64 wings: {sum(d.node_count for d, _, _ in cases)} nodes", flush=True)
    batch = sw.Batch([d for d, _, _ in cases])
    for w, (_, pressure, _) in enumerate(cases):
        batch.set_face_pressure(w, pressure)
    for _ in range(steps):
        for w, (_, _, loads) in enumerate(cases):
            batch.set_external_forces(w, loads)
        batch.step(1)
    got = [batch.node_states(w) for w in range(count)]
    batch.close()
    print("batch stepped; stepping each wing alone", flush=True)
    expected = solo_states(cases, steps)
    for w, name in enumerate(names):
        assert bits_equal(got[w][0], expected[w][0]) and bits_equal(got[w][1], expected[w][1]), name
