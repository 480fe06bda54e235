"""Portable transcendentals (simwing_amd/csrc/portable_math.hpp) on the host.

The device kernels use these functions for every atan2, sin, cos and acos
of the substep (hinge angles, rotation increments, payload angular velocity),
and the oracle can use them too, so device and oracle trajectories can be
compared bit for bit. Pinned here by:
* the constant tables, recomputed with 80-digit decimal arithmetic;
* quadruple-precision libquadmath results rounded to double (the correctly
  rounded values) on seeded samples shaped like the substep's arguments and
  on wide ranges;
* the special values of C99 Annex F and odd/even symmetry.
"""
import math
import subprocess
import sys

import numpy as np
import pytest

import oracle as orc
from conftest import ROOT, bits_equal

N = 100_000


def samples(name: str, kind: str, rng):
    if name == "atan2":
        if kind == "hinge":
            angle = rng.uniform(-math.pi, math.pi, N)
            r = 10.0 ** rng.uniform(-3, 3, N)
            return np.sin(angle) * r, np.cos(angle) * r
        if kind == "flat":
            return rng.standard_normal(N) * 1e-9, np.abs(rng.standard_normal(N)) + 0.5
        return (rng.standard_normal(N) * 10.0 ** rng.integers(-30, 30, N),
                rng.standard_normal(N) * 10.0 ** rng.integers(-30, 30, N))
    if name in ("sin", "cos"):
        if kind == "small":
            return rng.standard_normal(N) * 10.0 ** rng.integers(-12, 0, N), None
        return rng.uniform(-1.0e5, 1.0e5, N), None
    if kind == "near-one":
        return 1.0 - np.abs(rng.standard_normal(N)) * 10.0 ** rng.integers(-16, 0, N), None
    return rng.uniform(-1.0, 1.0, N), None


CASES = [("atan2", "hinge"), ("atan2", "flat"), ("atan2", "wide"), ("sin", "small"), ("sin", "wide"),
         ("cos", "small"), ("cos", "wide"), ("acos", "near-one"), ("acos", "full")]


def test_tables_match_decimal_generation():
    script = ROOT / "simwing_amd" / "csrc" / "portable_math_tables.py"
    result = subprocess.run([sys.executable, str(script), "--check"], capture_output=True, text=True)
    assert result.returncode == 0, result.stdout + result.stderr


@pytest.mark.parametrize("name,kind", CASES)
def test_correctly_rounded_against_quad_precision(name, kind):
    rng = np.random.default_rng(100 + CASES.index((name, kind)))
    a, b = samples(name, kind, rng)
    got = orc.math(name, a, b, portable=True)
    exact = orc.quad_math(name, a, b)
    wrong = int(np.count_nonzero(got.view(np.int64) != exact.view(np.int64)))
    libc = orc.math(name, a, b, portable=False)
    libc_wrong = int(np.count_nonzero(libc.view(np.int64) != exact.view(np.int64)))
    print(f"{name}/{kind}: portable misrounded {wrong}/{N}, C library misrounded {libc_wrong}/{N}")
    assert wrong == 0


def test_atan2_special_values_follow_annex_f():
    inf, nan = math.inf, math.nan
    pairs = [(0.0, 1.0), (-0.0, 1.0), (0.0, -1.0), (-0.0, -1.0), (0.0, 0.0), (-0.0, 0.0), (0.0, -0.0),
             (-0.0, -0.0), (1.0, 0.0), (-1.0, 0.0), (1.0, -0.0), (inf, 1.0), (-inf, 1.0), (1.0, inf),
             (1.0, -inf), (-1.0, -inf), (inf, inf), (inf, -inf), (-inf, inf), (-inf, -inf), (5e-324, 1.0),
             (1.0, 5e-324), (1e308, 1e-308), (1e-308, 1e308)]
    y = np.array([p[0] for p in pairs])
    x = np.array([p[1] for p in pairs])
    assert bits_equal(orc.math("atan2", y, x, portable=True), np.array([math.atan2(a, b) for a, b in pairs]))
    out = orc.math("atan2", np.array([nan, 1.0]), np.array([1.0, nan]), portable=True)
    assert np.isnan(out).all()


def test_sin_cos_acos_special_values():
    z = np.array([0.0, -0.0])
    assert bits_equal(orc.math("sin", z, portable=True), z)
    assert bits_equal(orc.math("cos", z, portable=True), np.ones(2))
    assert np.isnan(orc.math("sin", np.array([math.inf, -math.inf, math.nan]), portable=True)).all()
    assert np.isnan(orc.math("cos", np.array([math.inf, math.nan]), portable=True)).all()
    got = orc.math("acos", np.array([1.0, -1.0, 0.0, -0.0]), portable=True)
    assert bits_equal(got, np.array([0.0, math.pi, math.pi / 2, math.pi / 2]))
    assert np.isnan(orc.math("acos", np.array([1.0 + 2**-52, -2.0, math.nan]), portable=True)).all()


def test_symmetries():
    rng = np.random.default_rng(11)
    x = rng.uniform(-10.0, 10.0, N)
    y = rng.uniform(-10.0, 10.0, N)
    assert bits_equal(orc.math("sin", -x, portable=True), -orc.math("sin", x, portable=True))
    assert bits_equal(orc.math("cos", -x, portable=True), orc.math("cos", x, portable=True))
    assert bits_equal(orc.math("atan2", -y, x, portable=True), -orc.math("atan2", y, x, portable=True))
