"""Closed-form pins of the CPU oracle (CPU only).

Each case has an answer fixed by the mathematics, mirroring the reference's
own known-answer tests (simwing_structure_test.cpp:104-177,
softwing_material_test.cpp:397-422) and the build-defined contact-sphere
family (no reference counterpart).
"""
import math

import numpy as np

import oracle as orc
from helpers import definition, settings

CABLE, DISTANCE, STITCH = 2, 1, 4


def test_free_fall_and_force_over_mass():
    # simwing_structure_test.cpp:104-127: with dt = 0.5 and F/m = 2, one step
    # gives v = 1.0 and x = 0.5 (semi-implicit Euler), exactly.
    d = definition([[0.0, 0.0, 0.0]], [2.0])
    body = orc.OracleBody(d)
    body.set_external_forces(np.array([4.0, 0.0, 0.0]))
    body.step(settings(dt=0.5, iterations=1))
    s = body.state()
    assert s["velocity"][0] == 1.0
    assert s["position"][0] == 0.5
    # Gravity only: x = g dt^2, v = g dt.
    body = orc.OracleBody(d)
    body.step(settings(dt=0.25, iterations=1, gravity=(0.0, 0.0, -8.0)))
    s = body.state()
    assert s["velocity"][2] == -2.0 and s["position"][2] == -0.5


def test_single_pair_distance_restores_rest_length():
    # Equal masses, rigid constraint: one projection moves both ends by half
    # the error along the pair (soft_body.cpp:1598-1640).
    d = definition([[0, 0, 0], [1.5, 0, 0]], [1.0, 1.0], constraints=[(DISTANCE, 0, 1, 1.0, 0.0)])
    body = orc.OracleBody(d)
    body.step(settings(dt=0.1, iterations=1))
    x = body.state()["position"].reshape(-1, 3)
    assert abs(np.linalg.norm(x[1] - x[0]) - 1.0) < 1e-15
    assert abs(x[0, 0] - 0.25) < 1e-15 and abs(x[1, 0] - 1.25) < 1e-15


def test_compliant_pair_matches_xpbd_closed_form():
    # One iteration: dl = -C / (w1 + w2 + alpha), alpha = compliance / dt^2.
    compliance, dt = 0.01, 0.1
    d = definition([[0, 0, 0], [2.0, 0, 0]], [1.0, 1.0], constraints=[(DISTANCE, 0, 1, 1.0, compliance)])
    body = orc.OracleBody(d)
    body.step(settings(dt=dt, iterations=1))
    alpha = compliance / dt ** 2
    dl = -1.0 / (2.0 + alpha)
    x = body.state()["position"].reshape(-1, 3)
    assert math.isclose(x[0, 0], -dl, rel_tol=1e-14)
    assert math.isclose(x[1, 0], 2.0 + dl, rel_tol=1e-14)


def test_slack_cable_is_a_no_op():
    d = definition([[0, 0, 0], [0.5, 0, 0]], [1.0, 1.0], constraints=[(CABLE, 0, 1, 1.0, 0.0)])
    body = orc.OracleBody(d)
    body.step(settings(dt=0.1, iterations=3))
    x = body.state()["position"].reshape(-1, 3)
    assert x[0, 0] == 0.0 and x[1, 0] == 0.5


def test_seam_stitch_closes():
    # softwing_material_test.cpp:397-422: a zero-rest rigid stitch closes to
    # round-off.
    d = definition([[0, 0, 0], [0.01, 0.02, -0.03]], [1.0, 3.0], constraints=[(STITCH, 0, 1, 0.0, 0.0)])
    body = orc.OracleBody(d)
    body.step(settings(dt=0.1, iterations=1))
    x = body.state()["position"].reshape(-1, 3)
    assert np.linalg.norm(x[1] - x[0]) <= 1e-15


def test_membrane_rest_chart_is_strain_free():
    # simwing_structure_test.cpp:256-295: geometry equal to the chart gives
    # zero strain, so the solve leaves positions unchanged (no gravity).
    pos = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.0]]
    chart = [0, 0, 1, 0, 0, 1]
    mat = [1e5, 1e5, 0.0, 5e4, 0.0, 0.0, 0.0, 1.0]
    d = definition(pos, [1.0, 1.0, 1.0, 1.0], fixed=[0, 0, 0, 1], triangles=[[0, 1, 2]],
                   membranes=[(0, chart, mat)])
    body = orc.OracleBody(d)
    body.step(settings(dt=0.01, iterations=4))
    s = body.state()
    assert np.array_equal(s["position"], np.asarray(pos, float).reshape(-1))
    assert np.all(s["membrane_lambda"] == 0.0)


def test_uniaxial_membrane_coupon_closed_form():
    # Right-angle element with two fixed corners and one free corner moved
    # along the warp axis: the warp strain e = 0.5(s^2 - 1) and its gradient
    # give a one-component solve; the free corner lands where the linearised
    # constraint is satisfied: x1 = s - sqrt(A)*e*s / (w*A*s^2 + alpha).
    s_stretch, w = 1.1, 1.0
    pos = [[0, 0, 0], [s_stretch, 0, 0], [0, 1, 0]]
    chart = [0, 0, 1, 0, 0, 1]
    k = 1.0e4
    mat = [k, k, 0.0, k / 2, 0.0, 0.0, 0.0, 1.0]
    dt = 0.01
    d = definition(pos, [1.0, 1.0 / w, 1.0], fixed=[1, 0, 1], triangles=[[0, 1, 2]], membranes=[(0, chart, mat)])
    body = orc.OracleBody(d)
    body.step(settings(dt=dt, iterations=1))
    x = body.state()["position"].reshape(-1, 3)
    area = 0.5
    strain = 0.5 * (s_stretch ** 2 - 1.0)
    alpha = (1.0 / k) / dt ** 2
    j = math.sqrt(area) * s_stretch  # d(sqrt(A) e_w)/dx1 along x
    dl = -(math.sqrt(area) * strain) / (w * j * j + alpha)
    assert math.isclose(x[1, 0], s_stretch + w * j * dl, rel_tol=1e-12)
    assert abs(x[1, 1]) < 1e-15 and abs(x[1, 2]) < 1e-15


def test_contact_sphere_pushes_node_to_radius():
    # Build-defined family: a rigid sphere (compliance 0) centred on node 0
    # moves an intruding node straight out to the radius; an outside node is
    # untouched and the centre never moves.
    d = definition([[0, 0, 0], [0.02, 0, 0], [0, 0.3, 0]], [1.0, 1.0, 1.0],
                   spheres=[(0, 0.05, 0.0, [1, 2])])
    body = orc.OracleBody(d)
    body.step(settings(dt=0.01, iterations=1))
    x = body.state()["position"].reshape(-1, 3)
    assert np.array_equal(x[0], [0, 0, 0])
    assert math.isclose(x[1, 0], 0.05, rel_tol=1e-15) and x[1, 1] == 0.0
    assert np.array_equal(x[2], [0, 0.3, 0])


def test_compliant_contact_sphere_closed_form():
    # Compliance c > 0: one projection moves an intruding node (w = 1) out
    # by dl = (R - L) / (1 + c / dt^2) along the centre line, dl becoming
    # the multiplier; a second iteration finds the compliant constraint at
    # its fixed point (-gap - alpha * lambda = 0 up to roundoff) and leaves
    # the node where it is.
    dt, c, radius, start = 0.01, 2.0e-4, 0.05, 0.02
    alpha = c / dt ** 2
    expected = start + (radius - start) / (1.0 + alpha)
    for iterations in (1, 2):
        d = definition([[0, 0, 0], [start, 0, 0], [0, 0.3, 0]], [1.0, 1.0, 1.0],
                       spheres=[(0, radius, c, [1, 2])])
        body = orc.OracleBody(d)
        body.step(settings(dt=dt, iterations=iterations))
        x = body.state()["position"].reshape(-1, 3)
        assert math.isclose(x[1, 0], expected, rel_tol=1e-14), (iterations, x[1, 0], expected)
        assert x[1, 1] == 0.0 and x[1, 2] == 0.0
        assert np.array_equal(x[0], [0, 0, 0]) and np.array_equal(x[2], [0, 0.3, 0])


def test_contact_sphere_release_branch_closed_form():
    # Two spheres on one node (w = 1, starting 0.02 along x): a rigid sphere
    # A (radius 0.05 about node 0 at the origin) and a compliant sphere B
    # (radius 0.12 about node 3 at x = -0.05, alpha = c / dt^2 = 1).
    # Iteration 1: A pushes the node out to 0.05 (lambda_A = 0.03), then B
    # moves it by (0.12 - 0.10) / (1 + 1) = 0.01 to 0.06 (lambda_B = 0.01).
    # From then on, A finds the node outside with lambda_A > 0 and releases
    # the excess, pulling it back to 0.05, and B's update (0.02 - lambda_B) / 2
    # halves each iteration, so after n iterations x = 0.05 + 0.01 * 2^(1-n).
    # Without the release (a sphere that ignores an outside node) the node
    # would stay at 0.06 from the second iteration on.
    for iterations in (1, 2, 3, 4, 6):
        d = definition([[0, 0, 0], [0.02, 0, 0], [0, 0.5, 0], [-0.05, 0, 0]], [1.0, 1.0, 1.0, 1.0],
                       fixed=[1, 0, 1, 1], spheres=[(0, 0.05, 0.0, [1]), (3, 0.12, 1.0e-4, [1])])
        body = orc.OracleBody(d)
        body.step(settings(dt=0.01, iterations=iterations))
        x = body.state()["position"].reshape(-1, 3)
        expected = 0.05 + 0.01 * 2.0 ** (1 - iterations)
        assert math.isclose(x[1, 0], expected, rel_tol=1e-14), (iterations, x[1, 0], expected)
        assert x[1, 1] == 0.0 and x[1, 2] == 0.0
