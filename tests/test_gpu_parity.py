"""Device path parity (needs an MI355X). Every check goes through the C-ABI.

Bar: integer tables bit-equal; FP64 states bit-equal to the oracle run with
the device's portable transcendentals (the device keeps the reference's
operation order, and +, -, *, /, sqrt and fma are correctly rounded on both
sides); within 1e-10 max-norm relative error of the reference's own CPU run
(which uses the C library's atan2/sin/cos/acos) after the stated step count.
"""
import ctypes as C
import hashlib
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
    the host); the portable atan2, sin, cos and acos return the host's bits;
    remainder is exact."""
    rng = np.random.default_rng(7)
    n = 200_000
    libm = C.CDLL("libm.so.6")
    libm.remainder.restype = C.c_double
    libm.remainder.argtypes = [C.c_double, C.c_double]
    a = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
    b = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)

    def device(code, x, y):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        out = np.zeros(n)
        native.check(native.library().sw_device_math(0, code, native.f64p(x), native.f64p(y), native.f64p(out), n))
        return out

    assert bits_equal(device(0, np.abs(a), b), np.sqrt(np.abs(a))), "sqrt is not correctly rounded on the device"
    assert bits_equal(device(1, a, b), a / b), "division is not correctly rounded on the device"
    angle = rng.uniform(-math.pi, math.pi, n)
    r = 10.0 ** rng.uniform(-3, 3, n)
    # Quotients -1, 0, 1 and beyond, the ties at |x| = pi and 3 pi (half and
    # three halves of the double 2 pi), their neighbours, signed zeros and
    # non-finite values.
    two_pi = 6.283185307179586
    edges = [0.5 * two_pi, 1.5 * two_pi, two_pi, 0.0, 1e300, 1e-300]
    edges = [e2 for e in edges for e1 in (e, np.nextafter(e, 0.0), np.nextafter(e, np.inf)) for e2 in (e1, -e1)]
    edges += [np.inf, -np.inf, np.nan]
    x = np.concatenate([angle * 3.0, angle * 1.5, np.array(edges)])
    remainder = np.array([libm.remainder(v, two_pi) for v in x])
    got = np.empty_like(x)
    native.check(native.library().sw_device_math(0, 3, native.f64p(x), native.f64p(x), native.f64p(got), x.size))
    finite = np.isfinite(remainder)
    assert np.array_equal(np.isnan(got), ~finite)  # inf and nan give nan (its sign bit is not compared)
    wrong = np.flatnonzero(got[finite].view(np.uint64) != remainder[finite].view(np.uint64))
    assert wrong.size == 0, [(x[finite][i], got[finite][i], remainder[finite][i]) for i in wrong[:8]]
    half = rng.standard_normal(n) * 10.0 ** rng.integers(-9, 1, n)
    near_one = 1.0 - np.abs(rng.standard_normal(n)) * 10.0 ** rng.integers(-16, 0, n)
    probes = {2: ("atan2", np.sin(angle) * r, np.cos(angle) * r), 4: ("sin", half, half), 5: ("cos", half, half),
              6: ("acos", near_one, near_one)}
    for code, (name, x, y) in probes.items():
        got = device(code, x, y)
        host = orc.math(name, x, y, portable=True)
        exact = orc.quad_math(name, x, y)
        differ = int(np.count_nonzero(got.view(np.int64) != host.view(np.int64)))
        misrounded = int(np.count_nonzero(got.view(np.int64) != exact.view(np.int64)))
        print(f"device {name}: {differ}/{n} differ from the host portable function, {misrounded} misrounded",
              flush=True)
        assert differ == 0


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


@pytest.mark.parametrize("preset,steps", [("tiny", (1, 20)), ("small", (1, 120))])
def test_device_bit_equal_to_portable_oracle(preset, steps):
    """The device trajectory equals, bit for bit, the oracle's when the
    oracle uses the same portable transcendentals."""
    d, pressure, loads = flight_case(preset)
    structure = sw.Structure(d)
    structure.set_face_pressure(pressure)
    body = orc.OracleBody(d)
    body.use_portable_math(True)
    body.set_face_pressure(pressure)
    st = sw.flight_step_settings()
    for step in range(1, max(steps) + 1):
        structure.set_external_forces(loads)
        structure.step(st)
        body.set_external_forces(loads)
        body.step(st)
        if step in steps:
            expected = body.state()
            got = device_state(structure)
            report = compare(got, {k: expected[k] for k in STATE_KEYS}, f"{preset} step {step} vs portable oracle")
            assert all(eq for eq, _ in report.values()), f"{preset} step {step} is not bit-equal to the oracle"


def test_launch_shape_does_not_change_results():
    """Single-workgroup launches with LDS-resident nodes (256 and 128 lanes)
    and multi-workgroup grid launches, synchronised at agent scope or within
    one XCD, give bit-identical states."""
    d, pressure, loads = flight_case("small")
    st = sw.flight_step_settings()
    states = []
    for shape, local in ((None, True), ((1, 128), True), ((16, 256), False), ((64, 128), False), ((14, 256), True),
                         ((16, 256), True), ((32, 256), True), ((3, 256), True)):
        structure = sw.Structure(d)
        if shape:
            structure.set_launch(*shape)
        active = structure.set_xcd_local(local)
        # The small wing's default shape is 16 workgroups of 256 lanes.
        assert active == (local and (shape is None or shape[0] > 1)), (shape, local, active)
        structure.set_face_pressure(pressure)
        for _ in range(5):
            structure.set_external_forces(loads)
            structure.step(st)
        states.append(((shape, local), device_state(structure)))
    for shape, state in states[1:]:
        for key in state:
            assert bits_equal(state[key], states[0][1][key]), f"launch {shape}: {key} differs"


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
    # A compliant sphere (c > 0): the node moves out by (R - L) / (1 + c / dt^2),
    # and a second iteration leaves it at that fixed point.
    dt, c = 0.01, 2.0e-4
    expected = 0.02 + 0.03 / (1.0 + c / dt ** 2)
    for iterations in (1, 2):
        d = definition([[0, 0, 0], [0.02, 0, 0], [0, 0.3, 0]], [1.0, 1.0, 1.0], spheres=[(0, 0.05, c, [1, 2])])
        s = sw.Structure(d)
        s.step(settings(dt=dt, iterations=iterations))
        x, _, _ = s.node_states()
        assert math.isclose(x[1, 0], expected, rel_tol=1e-14) and x[1, 1] == 0.0, (iterations, x[1])
        s.close()
    # Release branch: a rigid sphere that already pushed the node out pulls it
    # back once a second, compliant sphere has moved it past the radius, so
    # after n iterations the node sits at 0.05 + 0.01 * 2^(1-n).
    for iterations in (1, 2, 3, 6):
        d = definition([[0, 0, 0], [0.02, 0, 0], [0, 0.5, 0], [-0.05, 0, 0]], [1.0, 1.0, 1.0, 1.0],
                       fixed=[1, 0, 1, 1], spheres=[(0, 0.05, 0.0, [1]), (3, 0.12, 1.0e-4, [1])])
        s = sw.Structure(d)
        s.step(settings(dt=0.01, iterations=iterations))
        x, _, _ = s.node_states()
        expected = 0.05 + 0.01 * 2.0 ** (1 - iterations)
        assert math.isclose(x[1, 0], expected, rel_tol=1e-14) and x[1, 1] == 0.0, (iterations, x[1])
        s.close()


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


def export_state(structure: sw.Structure, order: np.ndarray) -> dict:
    state = device_state(structure)
    state["dihedral_lambda_colour"] = state["dihedral_lambda"][order]
    return state


def test_long_horizon_matches_reference(golden):
    """The cfg0 wing with stiff hinges (preset small-stiff, which stays near
    rest under the flight loads): 240 flight steps on the device, every state
    array within 1e-10 of the reference's own CPU run at steps 1, 120 and 240
    (strided samples, relative to the array's max norm), and the line
    multipliers and payload in full."""
    reference = golden("small-stiff.digest.swfx")
    d, pressure, loads = flight_case("small-stiff")
    order, _ = sw.definition_colouring(d, 2)
    structure = sw.Structure(d)
    structure.set_face_pressure(pressure)
    st = sw.flight_step_settings()
    for step in range(1, 241):
        structure.set_external_forces(loads)
        structure.step(st)
        if step not in (1, 120, 240):
            continue
        got = export_state(structure, order.astype(np.int64))
        report = {}
        for key in ("position", "previous", "velocity", "membrane_residual", "membrane_lambda", "dihedral_lambda",
                    "constraint_lambda"):
            source = "dihedral_lambda_colour" if key == "dihedral_lambda" else key
            width = 3 if key in ("position", "previous", "velocity", "membrane_lambda") else 1
            rows = got[source].reshape(-1, width)[reference[f"step{step}/{key}/rows"].astype(np.int64)].reshape(-1)
            scale = float(reference[f"step{step}/{key}/maxabs"][0])
            report[key] = float(np.abs(rows - reference[f"step{step}/{key}/samples"]).max()) / max(scale, 1e-300)
        for key in ("segment_lambda", "payload"):
            report[key] = relative_error(got[key], reference[f"step{step}/{key}"])
        print(f"small-stiff step {step} vs reference:", {k: f"{e:.1e}" for k, e in report.items()}, flush=True)
        for key, err in report.items():
            assert err <= TOL, (step, key, err)
    structure.close()


def test_export_scale(golden):
    """The ~70k-triangle wing (43k nodes, 300 line segments, rigid pilot, 8
    contact spheres), 240 flight steps: bit-equal to the portable-math oracle
    at steps 1, 24 and 240; within 1e-10 of the reference's own CPU run after
    1 step (without spheres, which the reference lacks)."""
    if not (GOLDEN / "export.portable.swfx").exists():
        pytest.skip("export fixtures not generated")
    portable = golden("export.portable.swfx")
    d, pressure, loads = flight_case("export")
    structure = sw.Structure(d)
    structure.set_face_pressure(pressure)
    structure.set_launch(64, 256)
    st = sw.flight_step_settings()
    for step in range(1, 241):
        structure.set_external_forces(loads)
        structure.step(st)
        if step in (1, 24, 240):
            got = device_state(structure)
            for key in STATE_KEYS:
                digest = np.frombuffer(hashlib.sha256(np.ascontiguousarray(got[key]).tobytes()).digest(), np.uint64)
                same = np.array_equal(digest, portable[f"step{step}/{key}/digest"])
                if not same:
                    err = relative_error(got[key][::101], portable[f"step{step}/{key}/samples"])
                    print(f"export step {step} {key}: differs, sampled rel err {err:.2e}", flush=True)
                assert same, f"export step {step}: {key} is not bit-equal to the portable-math oracle"
            print(f"export step {step}: bit-equal to the portable-math oracle", flush=True)

    reference = golden("export.digest.swfx")
    order, _ = sw.definition_colouring(d, 2)
    d0, pressure0, loads0 = flight_case("export", strip_spheres=True)
    plain = sw.Structure(d0)
    plain.set_face_pressure(pressure0)
    plain.set_external_forces(loads0)
    plain.step(st)
    got = export_state(plain, order.astype(np.int64))
    for key in ("position", "previous", "velocity", "membrane_residual", "membrane_lambda", "dihedral_lambda",
                "constraint_lambda"):
        source = "dihedral_lambda_colour" if key == "dihedral_lambda" else key
        width = 3 if key in ("position", "previous", "velocity", "membrane_lambda") else 1
        rows = got[source].reshape(-1, width)[reference[f"step1/{key}/rows"].astype(np.int64)].reshape(-1)
        err = float(np.abs(rows - reference[f"step1/{key}/samples"]).max()) / float(reference[f"step1/{key}/maxabs"][0])
        print(f"export step 1 vs reference {key}: sampled rel err {err:.2e}", flush=True)
        assert err <= TOL
    for key in ("segment_lambda", "payload"):
        assert relative_error(got[key], reference[f"step1/{key}"]) <= TOL, key


def test_decoded_scene_stream_steps_like_the_reference(golden):
    """The reference's own scene-v2 stream of the tiny wing
    (tests/golden/tiny.scene.bin, written by writeScene), decoded by
    sw_scene_read, assembled and stepped on the device for 20 flight steps
    with the generated wing's face pressure: within 1e-10 of the reference's
    run of the same scene at steps 1, 5 and 20."""
    stream = (GOLDEN / "tiny.scene.bin").read_bytes()
    assembly = sw.read_scene(stream).assemble()
    d = assembly.definition
    generated = sw.generate_scene("tiny").assemble().definition
    for key, value in generated.arrays.items():
        assert np.array_equal(d.arrays[key], value), f"decoded scene assembles differently: {key}"
    structure = sw.Structure(d)
    structure.set_face_pressure(sw.generate_scene("tiny").face_pressure())
    loads = assembly.flight_loads()
    st = sw.flight_step_settings()
    for step in range(1, 21):
        structure.set_external_forces(loads)
        structure.step(st)
        if step in (1, 5, 20):
            compare(device_state(structure), reference_state(golden, "tiny", step), f"decoded tiny step {step}")
    structure.close()
