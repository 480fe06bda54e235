"""Product host logic against the reference (CPU only): scene assembly and
sweep schedules are integer-exact (and float-exact) with the reference's
assembleSceneStructure and colourings."""
import hashlib
import re
from pathlib import Path

import numpy as np
import pytest

import simwing_amd as sw
from simwing_amd import native
from conftest import GOLDEN, ROOT

ARRAYS = ("node_position", "node_mass", "node_fixed", "triangle_nodes", "constraint_kind", "constraint_nodes",
          "constraint_rest", "constraint_compliance", "membrane_triangle", "membrane_chart", "membrane_material",
          "dihedral_nodes", "dihedral_rest_angle", "dihedral_compliance")


def suspension_tables(d: sw.DefinitionArrays) -> dict:
    s = d.suspension
    return {
        "segments": np.stack([s["segment_ids"], s["segment_from_kind"].astype(np.uint64), s["segment_from_id"],
                              s["segment_to_kind"].astype(np.uint64), s["segment_to_id"]], 1).reshape(-1),
        "segment_values": np.stack([s["segment_rest"], s["segment_stiffness"], s["segment_damping"]], 1).reshape(-1),
        "suspension_attachments": np.stack([s["attachment_ids"], s["attachment_nodes"]], 1).reshape(-1),
        "suspension_junctions": np.stack([s["junction_ids"], s["junction_nodes"]], 1).reshape(-1),
        "harness_ids": s["harness_ids"],
        "harness_local": s["harness_local"],
    }


def same_bytes(a, b) -> bool:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b).astype(a.dtype)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("preset", ["tiny", "small"])
def test_assembly_bit_equal_to_reference(golden, preset):
    fixture = golden(f"{preset}.assembly.swfx")
    assembly = sw.generate_scene(preset).assemble()
    d = assembly.definition
    for name in ARRAYS:
        assert same_bytes(d.arrays[name], fixture[name]), f"{preset}: {name} differs from assembleSceneStructure"
    for name, value in suspension_tables(d).items():
        assert same_bytes(value, fixture[name]), f"{preset}: suspension {name} differs"
    assert same_bytes(np.array([assembly.fabric_mass, assembly.seam_mass]), fixture["total_masses"])
    assert same_bytes(sw.generate_scene(preset).face_pressure(), fixture["face_pressure"])


@pytest.mark.parametrize("preset", ["tiny", "small"])
def test_schedules_match_reference(golden, preset):
    d = sw.generate_scene(preset).assemble().definition
    colouring = golden(f"{preset}.colouring.swfx")
    order, offsets = sw.definition_colouring(d, 0)
    assert np.array_equal(order, colouring["distance_order"])
    assert np.array_equal(offsets, colouring["distance_offsets"])
    hinge, _ = sw.definition_colouring(d, 2)
    assert np.array_equal(hinge, golden(f"{preset}.assembly.swfx")["dihedral_colour_order"])


def test_export_scale_assembly_digests(golden):
    path = GOLDEN / "export.digest.swfx"
    if not path.exists():
        pytest.skip("export fixture not generated")
    fixture = golden("export.digest.swfx")
    assembly = sw.generate_scene("export").assemble()
    d = assembly.definition
    tables = {**{k: d.arrays[k] for k in ARRAYS}, **suspension_tables(d)}
    for name, value in tables.items():
        expected = fixture[f"digest/{name}"]
        got = np.frombuffer(hashlib.sha256(np.ascontiguousarray(value).tobytes()).digest(), np.uint64)
        assert np.array_equal(got, expected), f"export: {name} differs from assembleSceneStructure"
    order, offsets = sw.definition_colouring(d, 0)
    for name, value in (("distance_order", order), ("distance_offsets", offsets)):
        got = np.frombuffer(hashlib.sha256(value.tobytes()).digest(), np.uint64)
        assert np.array_equal(got, fixture[f"digest/{name}"])
    hinge, _ = sw.definition_colouring(d, 2)
    got = np.frombuffer(hashlib.sha256(hinge.tobytes()).digest(), np.uint64)
    assert np.array_equal(got, fixture["dihedral_colour_order_digest"])


def brute_force_greedy(items: np.ndarray, nodes: int) -> np.ndarray:
    """First-fit colour of each item by its definition: the smallest colour
    not used by an earlier item sharing a node."""
    colours = np.zeros(len(items), np.int64)
    for i, item in enumerate(items):
        used = {colours[j] for j in range(i) if set(items[j]) & set(item)}
        c = 0
        while c in used:
            c += 1
        colours[i] = c
    return colours


@pytest.mark.parametrize("family,arity", [(0, 2), (1, 3), (2, 4)])
def test_colourings_are_greedy_and_node_disjoint(family, arity):
    d = sw.generate_scene("tiny").assemble().definition
    a = d.arrays
    if family == 0:
        items = a["constraint_nodes"].reshape(-1, 2)
    elif family == 1:
        tri = a["triangle_nodes"].reshape(-1, 3)
        items = tri[a["membrane_triangle"].astype(np.int64)]
    else:
        items = a["dihedral_nodes"].reshape(-1, 4)
    order, offsets = sw.definition_colouring(d, family)
    colours = brute_force_greedy(items.tolist(), d.node_count)
    sizes = np.bincount(colours)
    ranked = sorted(range(sizes.size), key=lambda c: (-sizes[c], c))
    expected = np.concatenate([np.flatnonzero(colours == c) for c in ranked])
    assert np.array_equal(order.astype(np.int64), expected)
    for k in range(offsets.size - 1):
        block = items[order[offsets[k]:offsets[k + 1]].astype(np.int64)].reshape(-1)
        assert np.unique(block).size == block.size, "a colour shares a node"


def test_suspension_levels_reproduce_string_order():
    d = sw.generate_scene("small").assemble().definition
    order, offsets = sw.definition_colouring(d, 3)
    s = d.suspension
    names = [f"segment:{i}" for i in s["segment_ids"]]
    assert sorted(names) == [names[i] for i in sorted(range(len(names)), key=lambda i: names[i])]
    # Every pair of segments sharing a node keeps its string order across levels.
    level = np.zeros(order.size, np.int64)
    for k in range(offsets.size - 1):
        level[order[offsets[k]:offsets[k + 1]].astype(np.int64)] = k
    rank = {i: r for r, i in enumerate(sorted(range(len(names)), key=lambda i: names[i]))}
    ends = lambda i: {(int(s["segment_from_kind"][i]), int(s["segment_from_id"][i])),
                      (int(s["segment_to_kind"][i]), int(s["segment_to_id"][i]))}
    for i in range(len(names)):
        for j in range(len(names)):
            if rank[i] < rank[j]:
                shared = ends(i) & ends(j)
                payload = any(k == 3 for k, _ in ends(i)) and any(k == 3 for k, _ in ends(j))
                if shared or payload:
                    assert level[i] < level[j]


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "simwing_xpbd.h").read_text()
    declared = set(re.findall(r"\b(sw_[a-z_]+)\s*\(", header))
    assert declared == set(native.declared_symbols())
    lib = native.library()
    for name in declared:
        assert hasattr(lib, name), name


def test_definition_validation_errors():
    d = sw.generate_scene("tiny").assemble().definition.copy()
    d.arrays["constraint_rest"][0] = -1.0
    with pytest.raises(sw.NativeError) as error:
        sw.definition_colouring(d, 0)
    assert error.value.status == native.SW_INVALID_ARGUMENT
