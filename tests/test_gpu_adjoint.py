"""Device adjoint (needs an MI355X): gradients of the shape loss
sum |x(end) - target|^2 with respect to membrane log-compliance, line rest
length and triangle pressure.

Pinned by the oracle (tests/golden/adjoint.*.swfx, written by
make_adjoint_fixtures.py): per parameter and along one random direction per
family, the exact derivative of the oracle from its tangent build (the
tangent-adjoint dot-product test, to roundoff), and central differences
(h = 1e-6) of its long-double build, whose roundoff noise is far below
that of the double build's differences (both are recorded). Also: the run leaves the structure
untouched and repeats bit for bit, central differences of the device
forward pass, and the torch.autograd.Function harness."""
from pathlib import Path

import numpy as np
import pytest
import torch

import simwing_amd as sw
import swfx
from helpers import flight_case
from simwing_amd.torch_adjoint import shape_loss

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
FAMILIES = ("grad_log_compliance", "grad_rest_length", "grad_pressure")
# Unit roundoff of double over that of the x87 extended format: the expected
# noise of the extended build's differences relative to the double build's
# (reported only).
EXTENDED_NOISE = 2.0 ** -11


def setup(preset, steps, interval):
    d, pressure, loads = flight_case(preset)
    s = sw.Structure(d)
    s.set_face_pressure(pressure)
    s.set_external_forces(loads)
    adj = sw.Adjoint(s, steps, interval)
    rest = np.asarray(d.suspension["segment_rest"], dtype=np.float64)
    return d, s, adj, np.asarray(pressure, dtype=np.float64), rest


@pytest.mark.parametrize("wing", ["tiny", "small"])
def test_adjoint_matches_oracle_derivatives(wing):
    f = swfx.read(str(GOLDEN / f"adjoint.{wing}.swfx"))
    steps, interval = int(f["steps"][0]), int(f["interval"][0])
    print(f"adjoint {wing}: {steps} steps, checkpoints every {interval}", flush=True)
    d, s, adj, pressure, rest = setup(wing, steps, interval)
    r = adj.run(f["target"])
    # The device trajectory equals the oracle's to ~1e-10; the loss is a sum
    # of squares of differences of that size from the target.
    assert abs(r["loss"] - f["loss_double"][0]) <= 1e-7 * f["loss_double"][0]
    worst_tangent = worst_fd = 0.0
    kinks = []
    for k in range(f["family"].size):
        family, index = int(f["family"][k]), int(f["index"][k])
        ad = r[FAMILIES[family]][index]
        exact = f["tangent"][k]
        fd = f["fd_extended"][k]
        noise = EXTENDED_NOISE * abs(f["fd_double"][k] - fd)
        rel_t = abs(ad - exact) / abs(exact)
        rel_fd = abs(ad - fd) / abs(fd)
        worst_tangent = max(worst_tangent, rel_t)
        worst_fd = max(worst_fd, rel_fd)
        print(f"  {FAMILIES[family]:20s} {index:6d} adjoint {ad: .12e} tangent rel {rel_t:.1e} "
              f"differences rel {rel_fd:.1e} (their noise {noise / abs(fd):.1e})", flush=True)
        assert rel_t <= 1e-8, (family, index, ad, exact)
        # A kink inside [p - h, p + h] (a line going slack or taut in some
        # iteration: the max(0, .) of the unilateral projection): the two
        # differences agree with each other, being no noise, but not with the
        # one-sided derivative. Reported, not compared.
        spread = abs(f["fd_double"][k] - fd)
        if abs(fd - exact) > 1e-5 * abs(exact) and spread < 0.5 * abs(fd - exact):
            kinks.append((FAMILIES[family], index))
            print(f"    near a kink: differences {fd: .6e} / {f['fd_double'][k]: .6e}", flush=True)
            continue
        # Elsewhere: within 1e-6 plus the differences' own noise, of which
        # the double build's spread is an upper bound.
        assert abs(ad - fd) <= 1e-6 * abs(fd) + spread, (family, index, ad, fd)
    for family in range(3):
        direction = f[("direction_log_compliance", "direction_rest_length", "direction_pressure")[family]]
        ad = float(r[FAMILIES[family]] @ direction)
        exact = f["tangent_direction"][family]
        print(f"  direction {FAMILIES[family]:20s} adjoint {ad: .12e} tangent {exact: .12e}", flush=True)
        assert abs(ad - exact) <= 1e-8 * abs(exact)
    print(f"  worst relative error: tangent {worst_tangent:.1e}, differences {worst_fd:.1e}; "
          f"{len(kinks)} of {f['family'].size} parameters near a kink: {kinks}", flush=True)
    assert len(kinks) <= f["family"].size // 4
    fwd, rev, nbytes = adj.timing()
    print(f"  device ms: forward {fwd:.2f}, replay + reverse {rev:.2f}; {nbytes / 2**20:.1f} MiB", flush=True)
    adj.close()
    s.close()


def test_adjoint_matches_oracle_tangent_variant10k():
    """The cfg3 10k-triangle variant (6,642 nodes, 120 steps x 48 iterations,
    checkpoints every 8): every one of the 32 random parameters' adjoint
    components and the three directional derivatives within 1e-8 of the
    oracle's tangent build (tests/golden/adjoint.variant10k.swfx). This wing
    folds under the flight loads, so its exact derivatives are ~1e58-1e68
    and central differences at h = 1e-6 measure nothing (they read ~1e-6 of
    the derivative's size); the tangent, which follows the device's
    operations bit for bit, is the check."""
    f = swfx.read(str(GOLDEN / "adjoint.variant10k.swfx"))
    steps, interval = int(f["steps"][0]), int(f["interval"][0])
    d, s, adj, pressure, rest = setup("variant10k", steps, interval)
    r = adj.run(f["target"])
    # loss_double comes from the oracle with the C library's transcendentals;
    # on a folding wing the two math modes part by ~1e-3 over 120 steps (the
    # tangent build uses the device's portable ones, so it is the tight check).
    assert abs(r["loss"] - f["loss_double"][0]) <= 1e-2 * f["loss_double"][0], (r["loss"], f["loss_double"][0])
    worst = 0.0
    for k in range(f["family"].size):
        family, index = int(f["family"][k]), int(f["index"][k])
        ad, exact = r[FAMILIES[family]][index], f["tangent"][k]
        rel = abs(ad - exact) / abs(exact)
        worst = max(worst, rel)
        assert rel <= 1e-8, (FAMILIES[family], index, ad, exact)
    for family in range(3):
        direction = f[("direction_log_compliance", "direction_rest_length", "direction_pressure")[family]]
        ad = float(r[FAMILIES[family]] @ direction)
        exact = f["tangent_direction"][family]
        print(f"  direction {FAMILIES[family]:20s} adjoint {ad: .12e} tangent {exact: .12e}", flush=True)
        assert abs(ad - exact) <= 1e-8 * abs(exact)
    print(f"variant10k: loss {r['loss']:.9e}, worst relative error against the tangent {worst:.1e}", flush=True)
    adj.close()
    s.close()


def test_adjoint_leaves_structure_and_repeats():
    d, s, adj, pressure, rest = setup("tiny", 5, 2)
    before = s.node_states()
    target = np.asarray(d.arrays["node_position"]).reshape(-1, 3) + 0.01
    first = adj.run(target)
    second = adj.run(target)
    for key in ("loss",) + FAMILIES:
        assert np.array_equal(first[key], second[key]), key
    after = s.node_states()
    for a, b in zip(before, after):
        assert np.array_equal(a, b), "an adjoint run changed the structure's state"
    assert s.accepted_step_count() == 0
    # The structure still steps as if no adjoint had run.
    d2, p2, l2 = flight_case("tiny")
    fresh = sw.Structure(d2)
    fresh.set_face_pressure(p2)
    for t in (s, fresh):
        t.set_external_forces(l2)
        t.step(sw.flight_step_settings())
    assert np.array_equal(s.node_states()[0], fresh.node_states()[0])
    adj.close()
    s.close()
    fresh.close()


def test_adjoint_starts_from_the_current_state_and_restores_face_pressure():
    """A run starts from the state the structure holds when it starts (here
    after a step taken since configure) and puts that state back, face
    pressure included, after running at other pressures."""
    d, s, adj, pressure, rest = setup("tiny", 3, 2)
    settings = sw.flight_step_settings()
    _, p2, l2 = flight_case("tiny")
    target = np.asarray(d.arrays["node_position"]).reshape(-1, 3) + 0.01
    s.set_external_forces(l2)
    s.step(settings)
    out = adj.run(target, pressure=1.5 * pressure, gradients=False)

    expected = sw.Structure(flight_case("tiny")[0])
    expected.set_face_pressure(p2)
    expected.set_external_forces(l2)
    expected.step(settings)
    expected.set_face_pressure(1.5 * p2)
    for _ in range(3):
        expected.set_external_forces(l2)
        expected.step(settings)
    assert np.array_equal(out["final_position"].reshape(-1), expected.node_states()[0].reshape(-1))

    # Afterwards the structure steps on at its own face pressure.
    after = sw.Structure(flight_case("tiny")[0])
    after.set_face_pressure(p2)
    for t in (s, after):
        t.set_external_forces(l2)
        t.step(settings)
    after.set_external_forces(l2)
    after.step(settings)
    assert np.array_equal(s.node_states()[0], after.node_states()[0])
    assert s.accepted_step_count() == 2
    for t in (adj, s, expected, after):
        t.close()


def test_adjoint_sequential_and_pipelined_reverse_passes_agree():
    """Three checkpoint intervals, so the pipelined reverse pass reuses its
    first trace buffer. 200 sweep workgroups fit beside the forward grid
    (pipelined), 250 do not (the intervals are replayed and reversed in
    sequence over one buffer); both agree with the default launch to
    rounding (sphere-centre adjoints are accumulated atomically, so other
    partitions may differ in the last bits) and each repeats bit for bit."""
    d, s, adj, pressure, rest = setup("tiny", 6, 2)
    target = np.asarray(d.arrays["node_position"]).reshape(-1, 3) + 0.01
    runs = {}
    for workgroups in (None, 200, 250):
        if workgroups is not None:
            adj.set_workgroups(workgroups)
        first, second = adj.run(target), adj.run(target)
        for key in FAMILIES:
            assert np.array_equal(first[key], second[key]), (workgroups, key)
        runs[workgroups] = first
    for workgroups in (200, 250):
        for key in FAMILIES:
            a, b = runs[None][key], runs[workgroups][key]
            assert np.all(np.isfinite(b)), (workgroups, key)
            assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a)), (workgroups, key)
    adj.close()
    s.close()


def test_adjoint_matches_device_differences_tiny():
    """Directional derivatives against central differences of the device
    forward pass (h = 1e-5): the pressure and rest-length families are far
    above the differencing noise, log-compliance is not (its tolerance is
    that noise, estimated from two step sizes)."""
    d, s, adj, pressure, rest = setup("tiny", 5, 2)
    rng = np.random.default_rng(7)
    x0 = np.asarray(d.arrays["node_position"]).reshape(-1, 3)
    hidden = (np.log(rng.uniform(0.8, 1.2, adj.membrane_count)), rest * rng.uniform(0.98, 1.02, rest.size))
    target = adj.run(x0, *hidden, gradients=False)["final_position"]
    base = [np.zeros(adj.membrane_count), rest, pressure]
    r = adj.run(target, *base)
    for which, scale in ((0, 1.0), (1, rest), (2, 50.0)):
        direction = np.random.default_rng(11 + which).standard_normal(base[which].size) * scale
        fd = []
        for h in (1e-5, 2e-5):
            params = [b.copy() for b in base]
            params[which] = base[which] + h * direction
            plus = adj.run(target, *params, gradients=False)["loss"]
            params[which] = base[which] - h * direction
            minus = adj.run(target, *params, gradients=False)["loss"]
            fd.append((plus - minus) / (2 * h))
        ad = float(r[FAMILIES[which]] @ direction)
        tol = 1e-6 * abs(fd[0]) + 4.0 * abs(fd[0] - fd[1])
        print(f"{FAMILIES[which]}: adjoint {ad:.12e} differences {fd[0]:.12e} / {fd[1]:.12e}", flush=True)
        assert abs(ad - fd[0]) <= tol, (which, ad, fd)
    adj.close()
    s.close()


def test_autograd_harness_tiny():
    d, s, adj, pressure, rest = setup("tiny", 3, 2)
    target = np.asarray(d.arrays["node_position"]).reshape(-1, 3) + 0.02
    lc = torch.zeros(adj.membrane_count, dtype=torch.float64, requires_grad=True)
    rl = torch.tensor(rest, dtype=torch.float64, requires_grad=True)
    pr = torch.tensor(pressure, dtype=torch.float64, requires_grad=True)
    loss = shape_loss(lc, rl, pr, adj, target)
    (2.0 * loss).backward()
    r = adj.run(target, np.zeros(adj.membrane_count), rest, pressure)
    assert loss.item() == r["loss"]
    assert np.array_equal(lc.grad.numpy(), 2.0 * r["grad_log_compliance"])
    assert np.array_equal(rl.grad.numpy(), 2.0 * r["grad_rest_length"])
    assert np.array_equal(pr.grad.numpy(), 2.0 * r["grad_pressure"])
    adj.close()
    s.close()


def test_autograd_harness_with_device_tensors():
    """Parameter tensors on the GPU get their gradients on the GPU."""
    d, s, adj, pressure, rest = setup("tiny", 3, 2)
    target = np.asarray(d.arrays["node_position"]).reshape(-1, 3) + 0.02
    dev = torch.device("cuda:0")
    lc = torch.zeros(adj.membrane_count, dtype=torch.float64, device=dev, requires_grad=True)
    rl = torch.tensor(rest, dtype=torch.float64, device=dev, requires_grad=True)
    pr = torch.tensor(pressure, dtype=torch.float64, requires_grad=True)  # a CPU leaf beside them
    loss = shape_loss(lc, rl, pr, adj, target)
    loss.backward()
    r = adj.run(target, np.zeros(adj.membrane_count), rest, pressure)
    assert loss.device == dev and lc.grad.device == dev and rl.grad.device == dev and pr.grad.device.type == "cpu"
    assert np.array_equal(lc.grad.cpu().numpy(), r["grad_log_compliance"])
    assert np.array_equal(rl.grad.cpu().numpy(), r["grad_rest_length"])
    assert np.array_equal(pr.grad.numpy(), r["grad_pressure"])
    adj.close()
    s.close()
