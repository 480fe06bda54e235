"""The long-double build of the oracle (the reference values of the adjoint
fixtures) computes the same function as the double build, and the
finite-difference fixtures are self-consistent."""
from pathlib import Path

import numpy as np
import pytest

import oracle as orc
import simwing_amd as sw
import swfx
from helpers import flight_case

GOLDEN = Path(__file__).resolve().parent / "golden"


def run(preset, steps, extended):
    d, p, l = flight_case(preset)
    b = orc.OracleBody(d, extended=extended)
    b.set_face_pressure(p)
    for _ in range(steps):
        b.set_external_forces(l)
        b.step(sw.flight_step_settings())
    return b.state()


def test_extended_build_is_the_same_function():
    a = run("tiny", 6, False)
    b = run("tiny", 6, True)
    for key in ("position", "velocity", "segment_lambda", "payload"):
        scale = max(1.0, np.abs(a[key]).max())
        assert np.abs(a[key] - b[key]).max() <= 1e-11 * scale, key
    # ... and is not the double build itself.
    assert not np.array_equal(a["position"], b["position"])


def test_tangent_build_values_are_the_double_build():
    """The tangent pairs' values follow the double build's operations one for
    one: with the portable transcendentals both return the same bits."""
    d, p, l = flight_case("tiny")
    bodies = [orc.OracleBody(d), orc.OracleBody(d, extended="tangent")]
    for b in bodies:
        b.use_portable_math(True)
        b.set_face_pressure(p)
    bodies[1].set_tangent(np.ones(d.arrays["membrane_triangle"].size), None, np.ones(d.triangle_count))
    for b in bodies:
        for _ in range(4):
            b.set_external_forces(l)
            b.step(sw.flight_step_settings())
    a, t = bodies[0].state(), bodies[1].state()
    for key in ("position", "velocity", "membrane_lambda", "segment_lambda", "payload"):
        assert np.array_equal(a[key], t[key]), key
    assert np.any(bodies[1].tangent_positions() != 0.0)
    with pytest.raises(orc.OracleError):
        bodies[0].tangent_positions()


@pytest.mark.parametrize("wing", ["tiny", "small"])
def test_difference_fixtures_are_consistent(wing):
    """The exact tangent derivatives, the long-double differences and the
    double differences tell one story: the long-double differences sit
    closer to the tangents than the double ones (less roundoff), and on the
    short tiny run both agree with the tangents to their noise."""
    f = swfx.read(str(GOLDEN / f"adjoint.{wing}.swfx"))
    n = f["family"].size
    assert n == f["index"].size == f["h"].size == f["fd_extended"].size == f["fd_double"].size == f["tangent"].size
    assert np.all(np.isfinite(f["tangent"])) and np.all(f["tangent"] != 0.0)
    assert f["tangent_direction"].size == 3 and np.all(np.isfinite(f["tangent_direction"]))
    ext = np.abs(f["fd_extended"] - f["tangent"]) / np.abs(f["tangent"])
    dbl = np.abs(f["fd_double"] - f["tangent"]) / np.abs(f["tangent"])
    assert np.median(ext) < np.median(dbl)
    pressure = f["family"] == 2
    assert np.median(ext[pressure]) < 1e-5
    if wing == "tiny":
        smooth = f["family"] != 0
        assert np.all(ext[smooth] < 1e-6) and np.all(dbl[smooth] < 1e-4)
    assert abs(f["loss_double"][0] - f["loss_extended"][0]) <= 1e-8 * f["loss_double"][0]
