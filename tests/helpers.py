"""Small hand-built structural definitions for closed-form tests."""
from __future__ import annotations

import numpy as np

import simwing_amd as sw


def definition(positions, masses, fixed=None, constraints=(), triangles=(), membranes=(), hinges=(),
               spheres=()):
    """constraints: (kind, a, b, rest, compliance); membranes: (triangle, chart6, material8);
    hinges: (a, b, c, d, rest, compliance); spheres: (centre, radius, compliance, [nodes])."""
    n = len(masses)
    arrays = {
        "node_position": np.asarray(positions, np.float64).reshape(-1),
        "node_mass": np.asarray(masses, np.float64),
        "node_fixed": np.asarray(fixed if fixed is not None else [0] * n, np.uint8),
        "triangle_nodes": np.asarray(triangles, np.uint64).reshape(-1),
        "constraint_kind": np.asarray([c[0] for c in constraints], np.uint8),
        "constraint_nodes": np.asarray([[c[1], c[2]] for c in constraints], np.uint64).reshape(-1),
        "constraint_rest": np.asarray([c[3] for c in constraints], np.float64),
        "constraint_compliance": np.asarray([c[4] for c in constraints], np.float64),
        "membrane_triangle": np.asarray([m[0] for m in membranes], np.uint64),
        "membrane_chart": np.asarray([m[1] for m in membranes], np.float64).reshape(-1),
        "membrane_material": np.asarray([m[2] for m in membranes], np.float64).reshape(-1),
        "membrane_role": np.ones(len(membranes), np.uint8),
        "dihedral_nodes": np.asarray([h[:4] for h in hinges], np.uint64).reshape(-1),
        "dihedral_rest_angle": np.asarray([h[4] for h in hinges], np.float64),
        "dihedral_compliance": np.asarray([h[5] for h in hinges], np.float64),
        "sphere_centre": np.asarray([s[0] for s in spheres], np.uint64),
        "sphere_radius": np.asarray([s[1] for s in spheres], np.float64),
        "sphere_compliance": np.asarray([s[2] for s in spheres], np.float64),
        "sphere_offsets": np.cumsum([0] + [len(s[3]) for s in spheres]).astype(np.uint64),
        "sphere_nodes": np.asarray([v for s in spheres for v in s[3]], np.uint64),
    }
    return sw.DefinitionArrays(arrays)


def settings(dt=0.1, iterations=1, gravity=(0.0, 0.0, 0.0), damping=0.0):
    s = sw.flight_step_settings(iterations)
    s.time_step = dt
    s.gravity[:] = gravity
    s.velocity_damping = damping
    s.cable_sweep_pairs = 0
    return s


def flight_case(preset: str, strip_spheres: bool = False):
    """The benchmark workload: definition, face pressure and flight loads."""
    scene = sw.generate_scene(preset)
    assembly = scene.assemble()
    definition = assembly.definition
    if strip_spheres:
        definition = definition.copy()
        for key in ("sphere_centre", "sphere_radius", "sphere_compliance", "sphere_nodes"):
            definition.arrays[key] = definition.arrays[key][:0]
        definition.arrays["sphere_offsets"] = np.zeros(1, np.uint64)
    return definition, scene.face_pressure(), assembly.flight_loads()
