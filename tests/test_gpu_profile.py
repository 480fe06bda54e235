"""Step profile (needs an MI355X): the soft-body fields of
softwing::StepPerformanceProfile (soft_body.h:136-170) from the device.

Visit counts follow the reference's rules (soft_body.cpp:1078, 1222-1245,
1389-1404) and are pinned by the reference's own load-path cascade test
(tests/softwing_material_test.cpp:332-395: 42 cable visits for a seven-level
cascade swept in three pairs over two iterations), which this file rebuilds
node for node; the cascade's positions are also checked bit for bit against
the oracle and the reference's conditioning claim (three sweep pairs leave
less than 0.85 of the plain run's worst cable extension). The device times
are checked for consistency: every part non-negative, the parts together
about the substep kernel's event time."""
import numpy as np
import pytest

import oracle as orc
import simwing_amd as sw
from helpers import definition, flight_case, settings

pytestmark = pytest.mark.gpu

CABLE, TIE = 2, 3  # SW_CONSTRAINT_CABLE, SW_CONSTRAINT_SUSPENSION_TIE


def cascade():
    """softwing_material_test.cpp:332-350: a fixed node and seven levels one
    metre apart, 0.05 kg each but the 90 kg last; cables between levels 1-6
    first, then the ties 0-1 and 6-7 (the Playground's build order)."""
    positions = [[0.0, 0.0, 0.0]] + [[0.0, 0.0, -float(level)] for level in range(1, 8)]
    masses = [1.0] + [90.0 if level == 7 else 0.05 for level in range(1, 8)]
    constraints = [(CABLE, level, level + 1, 1.0, 0.0) for level in range(1, 6)]
    constraints += [(TIE, 0, 1, 1.0, 0.0), (TIE, 6, 7, 1.0, 0.0)]
    return definition(positions, masses, fixed=[1] + [0] * 7, constraints=constraints)


def worst_extension(x):
    pairs = [(level, level + 1) for level in range(1, 6)] + [(0, 1), (6, 7)]
    return max(np.linalg.norm(x[b] - x[a]) - 1.0 for a, b in pairs)


def test_cable_cascade_visits_and_conditioning():
    forces = np.zeros(24)
    forces[3 * 7 + 2] = -9000.0
    extensions = {}
    for pairs in (0, 3):
        st = settings(dt=1.0 / 60.0, iterations=2)
        st.cable_sweep_pairs = pairs
        d = cascade()
        s = sw.Structure(d)
        s.set_profiling(True)
        s.set_external_forces(forces)
        s.step(st)
        x = s.node_states()[0].reshape(-1, 3)
        body = orc.OracleBody(d)
        body.set_external_forces(forces)
        body.step(st)
        assert np.array_equal(x.reshape(-1), body.state()["position"]), f"{pairs} pairs: device differs from oracle"
        extensions[pairs] = worst_extension(x)
        p = s.profile()
        assert p["cable_constraint_visits"] == 2 * pairs * 7
        assert p["distance_constraint_visits"] == 2 * 7 and p["constraint_iterations"] == 2
        assert p["structural_substeps"] == 1
        s.close()
    # softwing_material_test.cpp:385-388
    assert extensions[3] < 0.85 * extensions[0], extensions
    assert 2 * 3 * 7 == 42


def test_flight_profile_counts_and_times():
    d, pressure, loads = flight_case("small")
    s = sw.Structure(d)
    s.set_face_pressure(pressure)
    st = sw.flight_step_settings()
    s.set_external_forces(loads)
    s.step(st)  # code objects loaded before the profiled steps
    s.set_profiling(True)
    s.set_kernel_timing(True)
    steps = 4
    for _ in range(steps):
        s.set_external_forces(loads)
        s.step(st)
    p = s.profile()
    kernel_ms, launches = s.kernel_time()
    a = d.arrays
    it = st.constraint_iterations
    kinds = np.asarray(a["constraint_kind"])
    stitches = int((kinds == 4).sum())
    load_path = int(((kinds == CABLE) | (kinds == TIE)).sum())
    assert p["structural_substeps"] == steps and p["constraint_iterations"] == steps * it
    assert p["distance_constraint_visits"] == steps * (it * kinds.size + stitches)
    assert p["cable_constraint_visits"] == steps * 2 * st.cable_sweep_pairs * load_path
    assert p["membrane_constraint_visits"] == steps * it * a["membrane_triangle"].size
    assert p["bending_constraint_visits"] == steps * it * a["dihedral_rest_angle"].size
    assert p["suspension_constraint_visits"] == steps * it * len(d.suspension["segment_rest"])
    assert p["contact_constraint_visits"] == (steps * it if a["sphere_nodes"].size else 0)
    parts = [p[k] for k in ("prediction_ns", "distance_constraint_ns", "cable_constraint_ns", "membrane_constraint_ns",
                            "bending_constraint_ns", "suspension_constraint_ns", "contact_constraint_ns",
                            "membrane_diagnostics_ns", "finalization_ns")]
    print({k: v for k, v in p.items() if k.endswith("_ns")}, f"kernel events {kernel_ms:.3f} ms / {launches}",
          flush=True)
    assert launches == steps
    assert p["bending_constraint_ns"] > 0 and p["membrane_constraint_ns"] > 0 and p["prediction_ns"] > 0
    total_parts_ms = sum(parts) * 1e-6
    assert 0.8 * kernel_ms <= total_parts_ms <= 1.05 * kernel_ms, (total_parts_ms, kernel_ms)
    assert p["soft_body_total_ns"] * 1e-6 >= total_parts_ms
    s.close()
