"""Writes the finite-difference fixtures that pin the device adjoint.

For each wing: a target shape from a hidden oracle run with every membrane's
compliance scaled by U(0.8, 1.2) and every line's rest length by
U(0.98, 1.02); the shape loss L = sum |x(end) - target|^2 at the definition's
parameters; and, for 32 (tiny: 12) random parameters drawn across the three
families (membrane log-compliance, line rest length, triangle pressure),
central differences (L(p + h) - L(p - h)) / 2h with h = 1e-6 (relative to
the parameter's magnitude for rest lengths and pressures) of

  * the oracle itself (double; its roundoff, amplified over the run, is the
    noise floor of these differences), and
  * the same oracle source computed in long double (x87, 64-bit mantissa),
    whose differences carry 2^11 times less of that noise;

and the exact derivatives of the oracle along the same parameters and along
one random direction per family, from the oracle's tangent build (the same
operation sequence carrying forward-mode tangent pairs, with the portable
transcendentals, so that its trajectory is the device's bit for bit): the
values the adjoint is checked against to roundoff (the tangent-adjoint
dot-product test).

Log-compliance theta is applied as the device library applies it: the
element's four stiffness entries times exp(-theta) before the cofactor
inverse. Loads: the flight loads at every step.

Output ``tests/golden/adjoint.<wing>.swfx``.

Usage: python tests/golden/make_adjoint_fixtures.py [tiny|small|variant10k ...]
"""
from __future__ import annotations

import math
import sys
import time
from multiprocessing import Pool
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
for path in (ROOT, HERE, ROOT / "tests", ROOT / "oracle"):
    sys.path.insert(0, str(path))

import oracle as orc  # noqa: E402
import simwing_amd as sw  # noqa: E402
import swfx  # noqa: E402
from helpers import flight_case  # noqa: E402

CASES = {
    # wing: steps, checkpoint interval (for the device run), parameter count, seed
    "tiny": (6, 4, 12, 0xAD0),
    "small": (120, 8, 32, 0xAD3),
    "variant10k": (120, 8, 32, 0xAD4),
}
H = 1.0e-6


def with_parameters(definition, theta, rest):
    d = definition.copy()
    mat = d.arrays["membrane_material"].reshape(-1, 8)
    for m in range(mat.shape[0]):
        if theta[m] != 0.0:
            scale = math.exp(-theta[m])
            for k in range(4):
                mat[m, k] = mat[m, k] * scale
    d.arrays["membrane_material"] = mat.reshape(-1)
    d.suspension["segment_rest"] = np.asarray(rest, dtype=np.float64).copy()
    return d


def final_positions(args):
    preset, theta, rest, pressure, steps, extended = args
    definition, _, loads = flight_case(preset)
    body = orc.OracleBody(with_parameters(definition, theta, rest), extended=extended)
    body.set_face_pressure(pressure)
    settings = sw.flight_step_settings()
    for _ in range(steps):
        body.set_external_forces(loads)
        body.step(settings)
    return body.state()["position"]


def tangent_loss(job):
    """d/ds L(p + s * direction) at s = 0 from the tangent build."""
    (preset, steps), seeds, target = job
    definition, pressure, loads = flight_case(preset)
    body = orc.OracleBody(definition, extended="tangent")
    body.use_portable_math(True)  # the device's transcendentals: the same trajectory bit for bit
    body.set_face_pressure(pressure)
    body.set_tangent(*seeds)
    settings = sw.flight_step_settings()
    for _ in range(steps):
        body.set_external_forces(loads)
        body.step(settings)
    x = body.state()["position"]
    return float(2.0 * (x - target) @ body.tangent_positions())


def loss(args, target):
    x = final_positions(args)
    d = x - target
    return float(np.dot(d, d))


def _loss_job(job):
    args, target = job
    return loss(args, target)


def make(preset: str, workers: int) -> None:
    steps, interval, count, seed = CASES[preset]
    definition, pressure, _ = flight_case(preset)
    rng = np.random.default_rng(seed)
    membranes = definition.arrays["membrane_triangle"].size
    rest = np.asarray(definition.suspension["segment_rest"], dtype=np.float64)
    pressure = np.asarray(pressure, dtype=np.float64)
    hidden_theta = np.log(rng.uniform(0.8, 1.2, membranes))
    hidden_rest = rest * rng.uniform(0.98, 1.02, rest.size)
    t0 = time.time()
    target = final_positions((preset, hidden_theta, hidden_rest, pressure, steps, False))
    print(f"{preset}: target in {time.time() - t0:.1f} s", flush=True)
    theta0 = np.zeros(membranes)
    families = rng.integers(0, 3, count)
    sizes = [membranes, rest.size, pressure.size]
    index = np.array([rng.integers(0, sizes[f]) for f in families], dtype=np.int64)
    jobs, steps_h = [], []
    for f, i in zip(families, index):
        base = [theta0, rest, pressure][f]
        h = H if f == 0 else H * max(1.0, abs(base[i]))
        steps_h.append(h)
        for sign in (1.0, -1.0):
            p = [theta0.copy(), rest.copy(), pressure.copy()]
            p[f][i] = base[i] + sign * h
            for extended in (False, True):
                jobs.append(((preset, p[0], p[1], p[2], steps, extended), target))
    jobs.append(((preset, theta0, rest, pressure, steps, False), target))
    jobs.append(((preset, theta0, rest, pressure, steps, True), target))
    tangent_jobs = []
    for f, i in zip(families, index):
        seeds = [np.zeros(membranes), np.zeros(rest.size), np.zeros(pressure.size)]
        seeds[f][i] = 1.0
        tangent_jobs.append(((preset, steps), seeds, target))
    directions = [rng.standard_normal(membranes), rng.standard_normal(rest.size) * rest,
                  rng.standard_normal(pressure.size) * 50.0]
    for f in range(3):
        seeds = [np.zeros(membranes), np.zeros(rest.size), np.zeros(pressure.size)]
        seeds[f] = directions[f]
        tangent_jobs.append(((preset, steps), seeds, target))
    t0 = time.time()
    with Pool(workers) as pool:
        tangents = pool.map(tangent_loss, tangent_jobs, chunksize=1)
        values = pool.map(_loss_job, jobs, chunksize=1)
    print(f"{preset}: {len(jobs) + len(tangent_jobs)} oracle runs in {time.time() - t0:.1f} s", flush=True)
    v = np.array(values[:-2]).reshape(count, 2, 2)  # [parameter][sign][double, extended]
    h = np.array(steps_h)
    fd_double = (v[:, 0, 0] - v[:, 1, 0]) / (2.0 * h)
    fd_extended = (v[:, 0, 1] - v[:, 1, 1]) / (2.0 * h)
    tangent = np.array(tangents[:count])
    for k in range(count):
        print(f"  {['theta', 'rest', 'pressure'][families[k]]:8s} {index[k]:6d} tangent {tangent[k]: .12e} "
              f"ext {(fd_extended[k] - tangent[k]) / abs(tangent[k]): .2e} "
              f"double {(fd_double[k] - tangent[k]) / abs(tangent[k]): .2e}", flush=True)
    swfx.write(str(HERE / f"adjoint.{preset}.swfx"), {
        "steps": np.array([steps], np.int64),
        "interval": np.array([interval], np.int64),
        "target": target,
        "family": families.astype(np.int64),
        "index": index,
        "h": h,
        "fd_extended": fd_extended,
        "fd_double": fd_double,
        "tangent": tangent,
        "direction_log_compliance": directions[0],
        "direction_rest_length": directions[1],
        "direction_pressure": directions[2],
        "tangent_direction": np.array(tangents[count:]),
        "loss_double": np.array([values[-2]]),
        "loss_extended": np.array([values[-1]]),
    })


if __name__ == "__main__":
    names = sys.argv[1:] or ["tiny", "small"]
    for name in names:
        make(name, workers=8)
