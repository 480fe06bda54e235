"""Device path parity (needs an MI355X). Every check goes through the C-ABI.

Bar: integer tables bit-equal; FP64 states within 1e-10 max-norm relative
error of the reference's own CPU run after the stated step count (and
reported bit-equality counts, which are expected to be complete because the
device kernels keep the reference's operation order)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as orc
import simwing_amd as sw
from conftest import GOLDEN, bits_equal, relative_error
from helpers import definition, flight_case, settings
from simwing_amd import native

pytestmark = pytest.mark.gpu

TOL = 1e-10
STATE_KEYS = ("position", "previous", "velocity", "membrane_lambda", "membrane_resultant", "membrane_residual",
              "constraint_lambda", "dihedral_lambda", "segment_lambda", "payload")


def device_state(structure: sw.Structure) -> dict:
    x, p, v = structure.node_states()
    out = {"position": x.reshape(-1), "previous": p.reshape(-1), "velocity": v.reshape(-1)}
    out.update(structure.solver_state())
    if structure.definition.suspension is not None:
        out["payload"] = structure.payload_state()
    return out


def reference_state(golden, preset, step):
    states = golden(f"{preset}.states.swfx")
    order = golden(f"{preset}.assembly.swfx")["dihedral_colour_order"].astype(np.int64)
    out = {k: states[f"step{step}/{k}"] for k in STATE_KEYS if k != "dihedral_lambda"}
    hinge = np.zeros(order.size)
    hinge[order] = states[f"step{step}/dihedral_lambda"]
    out["dihedral_lambda"] = hinge
    return out


def compare(got, expected, label):
    report = {}
    for key, ref in expected.items():
        err = relative_error(got[key], ref)
        report[key] = (bits_equal(got[key], ref), err)
    print(label, {k: ("bit-equal" if eq else f"{err:.2e}") for k, (eq, err) in report.items()}, flush=True)
    for key, (eq, err) in report.items():
        assert eq or err <= TOL, f"{label}: {key} relative error {err:.3e} > {TOL}"
    return report


def make_structure(preset, strip_spheres=False):
    d, pressure, loads = flight_case(preset, strip_spheres)
    s = sw.Structure(d)
    s.set_face_pressure(pressure)
    return s, loads


def test_device_math_conformance():
    """sqrt and division are correctly rounded on the device (bit-equal to
    the host); transcendental mismatches against glibc are counted."""
    rng = np.random.default_rng(7)
    n = 200_000
    libm = C.CDLL("libm.so.6")
    for name in ("atan2", "remainder", "sin", "cos", "acos"):
        fn = getattr(libm, name)
        fn.restype = C.c_double
        fn.argtypes = [C.c_double, C.c_double] if name in ("atan2", "remainder") else [C.c_double]
    a = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
    b = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
    cases = {
        0: (np.abs(a), b, np.sqrt(np.abs(a))),
        1: (a, b, a / b),
    }
    for code, (x, y, expected) in cases.items():
        out = np.zeros(n)
        native.check(native.library().sw_device_math(0, code, native.f64p(x), native.f64p(y), native.f64p(out), n))
        assert bits_equal(out, expected), f"function {code} is not correctly rounded on the device"
    # Hinge-like atan2 inputs (unit sine/cosine pairs), remainder, small
    # rotation half-angles, and acos near 1.
    angle = rng.uniform(-math.pi, math.pi, n)
    s_, c_ = np.sin(angle), np.cos(angle)
    half = rng.standard_normal(n) * 10.0 ** rng.integers(-9, -1, n)
    near_one = 1.0 - np.abs(rng.standard_normal(n)) * 10.0 ** rng.integers(-16, -2, n)
    probes = {
        2: ("atan2", s_, c_, lambda i: libm.atan2(s_[i], c_[i])),
        3: ("remainder", angle * 3.0, c_, lambda i: libm.remainder(angle[i] * 3.0, 6.283185307179586)),
        4: ("sin", half, half, lambda i: libm.sin(half[i])),
        5: ("cos", half, half, lambda i: libm.cos(half[i])),
        6: ("acos", near_one, near_one, lambda i: libm.acos(near_one[i])),
    }
    for code, (name, x, y, host) in probes.items():
        x = np.ascontiguousarray(x)
        y = np.ascontiguousarray(y)
        out = np.zeros(n)
        native.check(native.library().sw_device_math(0, code, native.f64p(x), native.f64p(y), native.f64p(out), n))
        expected = np.array([host(i) for i in range(n)])
        mismatches = int(np.count_nonzero(out.view(np.uint64) != expected.view(np.uint64)))
        ulps = np.abs(out.view(np.int64) - expected.view(np.int64)).max()
        print(f"device {name}: {mismatches}/{n} differ from glibc, max {ulps} ulp", flush=True)
        assert ulps <= 2
        if name == "remainder":
            assert mismatches == 0


@pytest.mark.parametrize("preset,steps", [("tiny", (1, 5, 20)), ("small", (1, 120))])
def test_device_matches_reference(golden, preset, steps):
    structure, loads = make_structure(preset)
    settings_ = sw.flight_step_settings()
    for step in range(1, max(steps) + 1):
        structure.set_external_forces(loads)
        diagnostics = structure.step(settings_)
        assert diagnostics["finite"] == 1
        if step in steps:
            compare(device_state(structure), reference_state(golden, preset, step), f"{preset} step {step}")
    assert structure.accepted_step_count() == max(steps)


def test_device_one_step_bit_equal_to_oracle():
    d, pressure, loads = flight_case("small")
    structure = sw.Structure(d)
    structure.set_face_pressure(pressure)
    structure.set_external_forces(loads)
    structure.step(sw.flight_step_settings())
    body = orc.OracleBody(d)
    body.set_face_pressure(pressure)
    body.set_external_forces(loads)
    body.step(sw.flight_step_settings())
    expected = body.state()
    got = device_state(structure)
    report = compare(got, {k: expected[k] for k in STATE_KEYS}, "small 1 step vs oracle")
    assert all(eq for eq, _ in report.values()), "one device step is not bit-equal to the oracle"


def test_device_closed_forms():
    d = definition([[0.0, 0.0, 0.0]], [2.0])
    s = sw.Structure(d)
    s.set_external_forces(np.array([4.0, 0.0, 0.0]))
    s.step(settings(dt=0.5, iterations=1))
    x, _, v = s.node_states()
    assert v[0, 0] == 1.0 and x[0, 0] == 0.5
    d = definition([[0, 0, 0], [1.5, 0, 0]], [1.0, 1.0], constraints=[(1, 0, 1, 1.0, 0.0)])
    s = sw.Structure(d)
    s.step(settings(dt=0.1, iterations=1))
    x, _, _ = s.node_states()
    assert abs(np.linalg.norm(x[1] - x[0]) - 1.0) < 1e-15
    d = definition([[0, 0, 0], [0.02, 0, 0], [0, 0.3, 0]], [1.0, 1.0, 1.0], spheres=[(0, 0.05, 0.0, [1, 2])])
    s = sw.Structure(d)
    s.step(settings(dt=0.01, iterations=1))
    x, _, _ = s.node_states()
    assert math.isclose(x[1, 0], 0.05, rel_tol=1e-15) and np.array_equal(x[2], [0, 0.3, 0])


def test_checkpoint_restore_replays_bit_identically():
    structure, loads = make_structure("tiny")
    st = sw.flight_step_settings()
    for _ in range(3):
        structure.set_external_forces(loads)
        structure.step(st)
    checkpoint = structure.checkpoint()
    for _ in range(4):
        structure.set_external_forces(loads)
        structure.step(st)
    first = device_state(structure)
    structure.restore(checkpoint)
    assert structure.accepted_step_count() == 3
    for _ in range(4):
        structure.set_external_forces(loads)
        structure.step(st)
    second = device_state(structure)
    for key in first:
        assert bits_equal(first[key], second[key]), key


def test_failed_step_restores_committed_state():
    d, pressure, loads = flight_case("tiny")
    d = d.copy()
    d.suspension["maximum_line_residual"] = 1e-15
    structure = sw.Structure(d)
    structure.set_face_pressure(pressure)
    before = device_state(structure)
    structure.set_external_forces(loads)
    with pytest.raises(sw.NativeError) as error:
        structure.step(sw.flight_step_settings())
    assert error.value.status == native.SW_SUSPENSION_ERROR
    assert error.value.phase == 8 and error.value.entity == "line-residual"
    after = device_state(structure)
    for key in before:
        assert bits_equal(before[key], after[key]), key
    assert structure.accepted_step_count() == 0


def test_invalid_settings_are_rejected():
    structure, _ = make_structure("tiny")
    st = sw.flight_step_settings()
    st.constraint_iterations = 12  # != the suspension's lineIterations
    with pytest.raises(sw.NativeError) as error:
        structure.step(st)
    assert error.value.status == native.SW_INVALID_ARGUMENT
    with pytest.raises(sw.NativeError) as error:
        structure.add_external_force(10 ** 6, [0.0, 0.0, 1.0])
    assert error.value.status == native.SW_OUT_OF_RANGE


def test_device_schedules_equal_host_schedules(golden):
    structure, _ = make_structure("small")
    order, offsets = structure.colouring(0)
    colouring = golden("small.colouring.swfx")
    assert np.array_equal(order, colouring["distance_order"])
    assert np.array_equal(offsets, colouring["distance_offsets"])


def test_export_scale_matches_reference(golden):
    if not (GOLDEN / "export.digest.swfx").exists():
        pytest.skip("export fixture not generated")
    fixture = golden("export.digest.swfx")
    structure, loads = make_structure("export", strip_spheres=True)
    st = sw.flight_step_settings()
    structure.set_external_forces(loads)
    for step in range(240):
        structure.set_external_forces(loads)
        structure.step(st)
    got = device_state(structure)
    tag = "step240"
    for key in ("position", "previous", "velocity", "membrane_residual", "membrane_lambda", "dihedral_lambda",
                "constraint_lambda"):
        width = 3 if key in ("position", "previous", "velocity", "membrane_lambda") else 1
        rows = got[key].reshape(-1, width)[fixture[f"{tag}/{key}/rows"].astype(np.int64)].reshape(-1)
        expected = fixture[f"{tag}/{key}/samples"]
        scale = float(fixture[f"{tag}/{key}/maxabs"][0])
        err = float(np.abs(rows - expected).max()) / scale
        print(f"export step 240 {key}: sampled rel err {err:.2e} bit-equal {bits_equal(rows, expected)}", flush=True)
        assert err <= TOL
    for key in ("segment_lambda", "payload"):
        err = relative_error(got[key], fixture[f"{tag}/{key}"])
        assert err <= TOL, key
