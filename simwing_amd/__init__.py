"""MI355X XPBD structural substep: host-side mirror of SimWing's structural
adapter (reference src/fsi/structure.h:273-309) over the C-ABI in
include/simwing_xpbd.h.

``Structure`` keeps the reference adapter's method names and error
behaviour (exceptions carry the C status that mirrors the reference's C++
exception class); ``generate_scene`` / ``assemble`` produce the seeded
synthetic wings and the integer-exact structural definition
(scene_structure.cpp:894-1282). All compute runs in the HIP library; there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import native
from .native import NativeError, check, f64p, u8p, u32p, u64p

__all__ = [
    "NativeError",
    "DefinitionArrays",
    "Scene",
    "Assembly",
    "Structure",
    "Batch",
    "Adjoint",
    "BatchAdjoint",
    "generate_scene",
    "flight_step_settings",
    "flight_assembly_settings",
]

flight_step_settings = native.flight_step_settings
flight_assembly_settings = native.flight_assembly_settings

_DEF_KEYS = {
    "node_position": np.float64,
    "node_mass": np.float64,
    "node_fixed": np.uint8,
    "triangle_nodes": np.uint64,
    "constraint_kind": np.uint8,
    "constraint_nodes": np.uint64,
    "constraint_rest": np.float64,
    "constraint_compliance": np.float64,
    "membrane_triangle": np.uint64,
    "membrane_chart": np.float64,
    "membrane_material": np.float64,
    "membrane_role": np.uint8,
    "dihedral_nodes": np.uint64,
    "dihedral_rest_angle": np.float64,
    "dihedral_compliance": np.float64,
    "sphere_centre": np.uint64,
    "sphere_radius": np.float64,
    "sphere_compliance": np.float64,
    "sphere_offsets": np.uint64,
    "sphere_nodes": np.uint64,
}
_SUSP_KEYS = {
    "attachment_ids": np.uint64,
    "attachment_nodes": np.uint64,
    "junction_ids": np.uint64,
    "junction_nodes": np.uint64,
    "harness_ids": np.uint64,
    "harness_local": np.float64,
    "segment_ids": np.uint64,
    "segment_from_kind": np.uint8,
    "segment_from_id": np.uint64,
    "segment_to_kind": np.uint8,
    "segment_to_id": np.uint64,
    "segment_rest": np.float64,
    "segment_stiffness": np.float64,
    "segment_damping": np.float64,
    "segment_role": np.uint32,
}


@dataclass
class DefinitionArrays:
    """Owning SoA structural definition (StructureDefinition, structure.h:168-176)."""

    arrays: dict
    suspension: dict | None = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def node_count(self) -> int:
        return int(self.arrays["node_mass"].size)

    @property
    def triangle_count(self) -> int:
        return int(self.arrays["triangle_nodes"].size // 3)

    def copy(self) -> "DefinitionArrays":
        susp = None if self.suspension is None else {k: (v.copy() if isinstance(v, np.ndarray) else v)
                                                      for k, v in self.suspension.items()}
        return DefinitionArrays({k: v.copy() for k, v in self.arrays.items()}, susp)

    def to_ctypes(self) -> native.Definition:
        a = {k: np.ascontiguousarray(self.arrays.get(k, np.zeros(0, dt)), dtype=dt) for k, dt in _DEF_KEYS.items()}
        if a["sphere_offsets"].size == 0:
            a["sphere_offsets"] = np.zeros(1, np.uint64)
        if a["membrane_role"].size == 0:
            a["membrane_role"] = np.ones(a["membrane_triangle"].size, np.uint8)
        keep = list(a.values())
        d = native.Definition()
        d.node_count = a["node_mass"].size
        d.node_position = f64p(a["node_position"])
        d.node_mass = f64p(a["node_mass"])
        d.node_fixed = u8p(a["node_fixed"])
        d.triangle_count = a["triangle_nodes"].size // 3
        d.triangle_nodes = u64p(a["triangle_nodes"])
        d.constraint_count = a["constraint_kind"].size
        d.constraint_kind = u8p(a["constraint_kind"])
        d.constraint_nodes = u64p(a["constraint_nodes"])
        d.constraint_rest = f64p(a["constraint_rest"])
        d.constraint_compliance = f64p(a["constraint_compliance"])
        d.membrane_count = a["membrane_triangle"].size
        d.membrane_triangle = u64p(a["membrane_triangle"])
        d.membrane_chart = f64p(a["membrane_chart"])
        d.membrane_material = f64p(a["membrane_material"])
        d.membrane_role = u8p(a["membrane_role"])
        d.dihedral_count = a["dihedral_rest_angle"].size
        d.dihedral_nodes = u64p(a["dihedral_nodes"])
        d.dihedral_rest_angle = f64p(a["dihedral_rest_angle"])
        d.dihedral_compliance = f64p(a["dihedral_compliance"])
        d.sphere_count = a["sphere_centre"].size
        d.sphere_centre = u64p(a["sphere_centre"])
        d.sphere_radius = f64p(a["sphere_radius"])
        d.sphere_compliance = f64p(a["sphere_compliance"])
        d.sphere_offsets = u64p(a["sphere_offsets"])
        d.sphere_nodes = u64p(a["sphere_nodes"])
        if self.suspension is not None:
            s = self.suspension
            sa = {k: np.ascontiguousarray(s[k], dtype=dt) for k, dt in _SUSP_KEYS.items()}
            keep.extend(sa.values())
            sv = native.SuspensionDefinition()
            sv.pilot_stable_id = int(s["pilot_stable_id"])
            sv.pilot_mass = float(s["pilot_mass"])
            for name in ("pilot_com_local", "pilot_com_world", "pilot_velocity", "pilot_angular_velocity",
                         "pilot_inertia"):
                getattr(sv, name)[:] = [float(x) for x in s[name]]
            sv.pilot_orientation[:] = [float(x) for x in s["pilot_orientation"]]
            sv.attachment_count = sa["attachment_ids"].size
            sv.attachment_ids = u64p(sa["attachment_ids"])
            sv.attachment_nodes = u64p(sa["attachment_nodes"])
            sv.junction_count = sa["junction_ids"].size
            sv.junction_ids = u64p(sa["junction_ids"])
            sv.junction_nodes = u64p(sa["junction_nodes"])
            sv.harness_count = sa["harness_ids"].size
            sv.harness_ids = u64p(sa["harness_ids"])
            sv.harness_local = f64p(sa["harness_local"])
            sv.segment_count = sa["segment_ids"].size
            sv.segment_ids = u64p(sa["segment_ids"])
            sv.segment_from_kind = u8p(sa["segment_from_kind"])
            sv.segment_from_id = u64p(sa["segment_from_id"])
            sv.segment_to_kind = u8p(sa["segment_to_kind"])
            sv.segment_to_id = u64p(sa["segment_to_id"])
            sv.segment_rest = f64p(sa["segment_rest"])
            sv.segment_stiffness = f64p(sa["segment_stiffness"])
            sv.segment_damping = f64p(sa["segment_damping"])
            sv.segment_role = u32p(sa["segment_role"])
            sv.solver_iterations = int(s["solver_iterations"])
            sv.attachment_tolerance = float(s["attachment_tolerance"])
            sv.minimum_line_length = float(s["minimum_line_length"])
            sv.maximum_line_residual = float(s["maximum_line_residual"])
            sv.maximum_control_work = float(s["maximum_control_work"])
            keep.append(sv)
            d.suspension = C.pointer(sv)
        self._keep = keep
        return d

    @staticmethod
    def from_ctypes(d: native.Definition) -> "DefinitionArrays":
        def arr(ptr, count, dtype):
            if count == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

        n, t, c, m, h, s = (d.node_count, d.triangle_count, d.constraint_count, d.membrane_count,
                            d.dihedral_count, d.sphere_count)
        arrays = {
            "node_position": arr(d.node_position, 3 * n, np.float64),
            "node_mass": arr(d.node_mass, n, np.float64),
            "node_fixed": arr(d.node_fixed, n, np.uint8),
            "triangle_nodes": arr(d.triangle_nodes, 3 * t, np.uint64),
            "constraint_kind": arr(d.constraint_kind, c, np.uint8),
            "constraint_nodes": arr(d.constraint_nodes, 2 * c, np.uint64),
            "constraint_rest": arr(d.constraint_rest, c, np.float64),
            "constraint_compliance": arr(d.constraint_compliance, c, np.float64),
            "membrane_triangle": arr(d.membrane_triangle, m, np.uint64),
            "membrane_chart": arr(d.membrane_chart, 6 * m, np.float64),
            "membrane_material": arr(d.membrane_material, 8 * m, np.float64),
            "membrane_role": arr(d.membrane_role, m, np.uint8),
            "dihedral_nodes": arr(d.dihedral_nodes, 4 * h, np.uint64),
            "dihedral_rest_angle": arr(d.dihedral_rest_angle, h, np.float64),
            "dihedral_compliance": arr(d.dihedral_compliance, h, np.float64),
            "sphere_centre": arr(d.sphere_centre, s, np.uint64),
            "sphere_radius": arr(d.sphere_radius, s, np.float64),
            "sphere_compliance": arr(d.sphere_compliance, s, np.float64),
            "sphere_offsets": arr(d.sphere_offsets, s + 1, np.uint64),
        }
        arrays["sphere_nodes"] = arr(d.sphere_nodes, int(arrays["sphere_offsets"][-1]) if s else 0, np.uint64)
        susp = None
        if d.suspension:
            sv = d.suspension.contents
            susp = {
                "pilot_stable_id": sv.pilot_stable_id,
                "pilot_mass": sv.pilot_mass,
                "pilot_com_local": list(sv.pilot_com_local),
                "pilot_com_world": list(sv.pilot_com_world),
                "pilot_velocity": list(sv.pilot_velocity),
                "pilot_angular_velocity": list(sv.pilot_angular_velocity),
                "pilot_orientation": list(sv.pilot_orientation),
                "pilot_inertia": list(sv.pilot_inertia),
                "attachment_ids": arr(sv.attachment_ids, sv.attachment_count, np.uint64),
                "attachment_nodes": arr(sv.attachment_nodes, sv.attachment_count, np.uint64),
                "junction_ids": arr(sv.junction_ids, sv.junction_count, np.uint64),
                "junction_nodes": arr(sv.junction_nodes, sv.junction_count, np.uint64),
                "harness_ids": arr(sv.harness_ids, sv.harness_count, np.uint64),
                "harness_local": arr(sv.harness_local, 3 * sv.harness_count, np.float64),
                "segment_ids": arr(sv.segment_ids, sv.segment_count, np.uint64),
                "segment_from_kind": arr(sv.segment_from_kind, sv.segment_count, np.uint8),
                "segment_from_id": arr(sv.segment_from_id, sv.segment_count, np.uint64),
                "segment_to_kind": arr(sv.segment_to_kind, sv.segment_count, np.uint8),
                "segment_to_id": arr(sv.segment_to_id, sv.segment_count, np.uint64),
                "segment_rest": arr(sv.segment_rest, sv.segment_count, np.float64),
                "segment_stiffness": arr(sv.segment_stiffness, sv.segment_count, np.float64),
                "segment_damping": arr(sv.segment_damping, sv.segment_count, np.float64),
                "segment_role": arr(sv.segment_role, sv.segment_count, np.uint32),
                "solver_iterations": sv.solver_iterations,
                "attachment_tolerance": sv.attachment_tolerance,
                "minimum_line_length": sv.minimum_line_length,
                "maximum_line_residual": sv.maximum_line_residual,
                "maximum_control_work": sv.maximum_control_work,
            }
        return DefinitionArrays(arrays, susp)


def definition_colouring(definition: DefinitionArrays, family: int):
    """Host-side colour schedule of a family (0 distance, 1 membrane, 2 hinge,
    3 suspension levels) without a device."""
    lib = native.library()
    view = definition.to_ctypes()
    nc, no = C.c_size_t(), C.c_size_t()
    check(lib.sw_definition_colouring(C.byref(view), family, None, C.byref(nc), None, C.byref(no)))
    order = np.zeros(nc.value, np.uint64)
    offsets = np.zeros(no.value, np.uint64)
    check(lib.sw_definition_colouring(C.byref(view), family, u64p(order), None, u64p(offsets), None))
    return order, offsets


class Scene:
    """Scene-v2 entities held by the library: a seeded synthetic wing
    (``Scene(preset)``) or a decoded binary stream (``read_scene``)."""

    def __init__(self, preset: str | None, _handle=None):
        if _handle is None:
            lib = native.library()
            handle = C.c_void_p()
            check(lib.sw_scene_generate(preset.encode(), C.byref(handle)))
            _handle = handle
        self._handle = _handle
        self.preset = preset

    def to_bytes(self) -> bytes:
        """writeScene's compact binary stream (sw_scene_write)."""
        lib = native.library()
        size = C.c_size_t()
        check(lib.sw_scene_write(self._handle, None, 0, C.byref(size)))
        out = np.zeros(size.value, np.uint8)
        check(lib.sw_scene_write(self._handle, u8p(out), out.size, C.byref(size)))
        return out.tobytes()

    def __del__(self):
        if getattr(self, "_handle", None):
            native.library().sw_scene_destroy(self._handle)
            self._handle = None

    def face_pressure(self) -> np.ndarray:
        ptr = C.POINTER(C.c_double)()
        count = C.c_size_t()
        check(native.library().sw_scene_face_pressure(self._handle, C.byref(ptr), C.byref(count)))
        return np.ctypeslib.as_array(ptr, shape=(count.value,)).copy() if count.value else np.zeros(0)

    def assemble(self, settings: native.AssemblySettings | None = None) -> "Assembly":
        return Assembly(self, settings or native.flight_assembly_settings())


def definition_fingerprint(definition: "DefinitionArrays") -> int:
    """definitionFingerprint() of a definition, on the host (sw_definition_fingerprint)."""
    view = definition.to_ctypes()
    out = C.c_uint64()
    check(native.library().sw_definition_fingerprint(C.byref(view), C.byref(out)))
    return out.value


def generate_scene(preset: str) -> Scene:
    return Scene(preset)


def read_scene(data: bytes) -> Scene:
    """readScene: a scene from its compact binary stream (sw_scene_read)."""
    buf = np.frombuffer(bytes(data), np.uint8).copy()
    handle = C.c_void_p()
    check(native.library().sw_scene_read(u8p(buf), buf.size, C.byref(handle)))
    return Scene(None, _handle=handle)


class Assembly:
    """assembleSceneStructure output: the definition and its mappings."""

    def __init__(self, scene: Scene, settings: native.AssemblySettings):
        lib = native.library()
        handle = C.c_void_p()
        check(lib.sw_scene_assemble(scene._handle, C.byref(settings), C.byref(handle)))
        self._handle = handle
        view = native.Definition()
        check(lib.sw_assembly_definition(handle, C.byref(view)))
        self.definition = DefinitionArrays.from_ctypes(view)
        vn, jn = C.c_size_t(), C.c_size_t()
        fm, sm = C.c_double(), C.c_double()
        check(lib.sw_assembly_info(handle, C.byref(vn), C.byref(jn), C.byref(fm), C.byref(sm)))
        self.vertex_nodes, self.junction_nodes = vn.value, jn.value
        self.fabric_mass, self.seam_mass = fm.value, sm.value

    def __del__(self):
        if getattr(self, "_handle", None):
            native.library().sw_assembly_destroy(self._handle)
            self._handle = None

    def flight_loads(self) -> np.ndarray:
        out = np.zeros(3 * self.definition.node_count)
        check(native.library().sw_assembly_flight_loads(self._handle, f64p(out)))
        return out


def _diagnostics_dict(d: native.Diagnostics) -> dict:
    out = {}
    for name, _ in native.Diagnostics._fields_:
        value = getattr(d, name)
        out[name] = list(value) if hasattr(value, "__len__") else value
    return out


def _diagnostics_struct(values: dict) -> native.Diagnostics:
    d = native.Diagnostics()
    for name, _ in native.Diagnostics._fields_:
        if name in values:
            value = values[name]
            if hasattr(getattr(d, name), "__len__"):
                getattr(d, name)[:] = [float(x) for x in value]
            else:
                setattr(d, name, value)
    return d


class FrameMapping:
    """Stable identity of a structure's diagnostic frames
    (StructureFrameMappingDefinition, structure_frame.h:35-39): one id per
    frame vertex (the N nodes, then the H harness points), per triangle (with
    its two side regions) and per line (the C constraints, then the S
    segments, a segment line carrying its segment's stable id) plus a role
    per line. Line endpoints are derived from the definition."""

    def __init__(self, vertex_ids, triangle_ids, triangle_negative_region, triangle_positive_region, line_ids,
                 line_roles):
        self.vertex_ids = np.ascontiguousarray(vertex_ids, np.uint64)
        self.triangle_ids = np.ascontiguousarray(triangle_ids, np.uint64)
        self.triangle_negative_region = np.ascontiguousarray(triangle_negative_region, np.uint64)
        self.triangle_positive_region = np.ascontiguousarray(triangle_positive_region, np.uint64)
        self.line_ids = np.ascontiguousarray(line_ids, np.uint64)
        self.line_roles = np.ascontiguousarray(line_roles, np.uint32)

    @classmethod
    def sequential(cls, definition: "DefinitionArrays") -> "FrameMapping":
        """Vertices 1..N+H, triangles 1..T between side regions 1 and 2,
        segment lines under their stable ids with their roles, constraint
        lines numbered on from the largest segment id with their kind as
        role."""
        a, sus = definition.arrays, definition.suspension
        h = 0 if sus is None else sus["harness_ids"].size
        t = a["triangle_nodes"].size // 3
        c = a["constraint_kind"].size
        seg_ids = np.zeros(0, np.uint64) if sus is None else sus["segment_ids"].astype(np.uint64)
        seg_roles = np.zeros(0, np.uint32) if sus is None else sus["segment_role"].astype(np.uint32)
        base = int(seg_ids.max()) + 1 if seg_ids.size else 1
        return cls(np.arange(1, definition.node_count + h + 1, dtype=np.uint64),
                   np.arange(1, t + 1, dtype=np.uint64), np.ones(t, np.uint64), np.full(t, 2, np.uint64),
                   np.concatenate([np.arange(base, base + c, dtype=np.uint64), seg_ids]),
                   np.concatenate([a["constraint_kind"].astype(np.uint32), seg_roles]))

    def _ctypes(self) -> native.FrameMapping:
        m = native.FrameMapping()
        m.vertex_count = self.vertex_ids.size
        m.vertex_ids = u64p(self.vertex_ids)
        m.triangle_count = self.triangle_ids.size
        m.triangle_ids = u64p(self.triangle_ids)
        m.triangle_negative_region = u64p(self.triangle_negative_region)
        m.triangle_positive_region = u64p(self.triangle_positive_region)
        m.line_count = self.line_ids.size
        m.line_ids = u64p(self.line_ids)
        m.line_roles = u32p(self.line_roles)
        return m


def frame_context(scene_checksum: str, solver_commit: str, time_step: float, coupling_iteration: int = 0,
                  coupling_residuals=(0.0,) * 5, fluid_mass: float = 0.0, total_momentum=(0.0,) * 3,
                  total_energy: float = 0.0, interface_force_residual=(0.0,) * 3,
                  interface_moment_residual=(0.0,) * 3, interface_power_residual: float = 0.0) -> native.FrameContext:
    """StructureFrameContext (structure_frame.h:86-93)."""
    c = native.FrameContext()
    c.scene_checksum = scene_checksum.encode()
    c.solver_commit = solver_commit.encode()
    c.time_step = time_step
    c.coupling_iteration = coupling_iteration
    c.coupling_residuals[:] = [float(x) for x in coupling_residuals]
    c.fluid_mass = fluid_mass
    c.total_momentum[:] = [float(x) for x in total_momentum]
    c.total_energy = total_energy
    c.interface_force_residual[:] = [float(x) for x in interface_force_residual]
    c.interface_moment_residual[:] = [float(x) for x in interface_moment_residual]
    c.interface_power_residual = interface_power_residual
    return c


def frame_encode(definition: "DefinitionArrays", mapping: FrameMapping, context: native.FrameContext, step: int,
                 simulation_time: float, fields: dict, diagnostics: dict) -> bytes:
    """SWVF bytes of the frame of caller-held fields (sw_frame_encode)."""
    lib = native.library()
    view = definition.to_ctypes()
    m = mapping._ctypes()
    arrays = [np.ascontiguousarray(fields[k], np.float64).reshape(-1) for k in (
        "vertex_position", "vertex_velocity", "vertex_pending_force", "line_length", "line_violation")]
    d = _diagnostics_struct(diagnostics)
    size = C.c_size_t()
    check(lib.sw_frame_encode(C.byref(view), C.byref(m), C.byref(context), step, simulation_time,
                              *[f64p(x) for x in arrays], C.byref(d), None, 0, C.byref(size)))
    out = np.zeros(size.value, np.uint8)
    check(lib.sw_frame_encode(C.byref(view), C.byref(m), C.byref(context), step, simulation_time,
                              *[f64p(x) for x in arrays], C.byref(d), u8p(out), out.size, C.byref(size)))
    return out.tobytes()


class Structure:
    """simwing::fsi::Structure over the device (structure.h:273-309)."""

    def __init__(self, definition: DefinitionArrays, device: int = 0):
        lib = native.library()
        self._definition = definition
        view = definition.to_ctypes()
        handle = C.c_void_p()
        check(lib.sw_structure_create(C.byref(view), device, C.byref(handle)))
        self._handle = handle
        self._owned = True
        self._dependents = []  # adjoints borrowing this structure: closed first
        self.node_count = definition.node_count

    @classmethod
    def _borrow(cls, handle, definition: DefinitionArrays, owner) -> "Structure":
        """A wing owned by a batch (kept alive by `owner`)."""
        s = cls.__new__(cls)
        s._definition = definition
        s._handle = handle
        s._owned = False
        s._owner = owner
        s._dependents = []
        s.node_count = definition.node_count
        return s

    def close(self) -> None:
        for dependent in getattr(self, "_dependents", []):
            dependent.close()
        if getattr(self, "_handle", None):
            if getattr(self, "_owned", True):
                native.library().sw_structure_destroy(self._handle)
            self._handle = None

    def __del__(self):
        self.close()

    @property
    def definition(self) -> DefinitionArrays:
        return self._definition

    def frame_fields(self, noalloc: bool = False) -> dict:
        """Diagnostic-frame fields (sw_structure_frame_fields, or with
        noalloc=True sw_structure_frame_fields_noalloc, which stages them in
        the library's static device array): vertices are the nodes then the
        harness points, lines the constraints then the suspension segments
        (definition order)."""
        d = self._definition
        h = 0 if d.suspension is None else int(np.asarray(d.suspension["harness_ids"]).size)
        lines = int(d.arrays["constraint_kind"].size) + (
            0 if d.suspension is None else int(np.asarray(d.suspension["segment_rest"]).size))
        v = self.node_count + h
        out = {"vertex_position": np.zeros(3 * v), "vertex_velocity": np.zeros(3 * v),
               "vertex_pending_force": np.zeros(3 * v), "line_length": np.zeros(lines),
               "line_violation": np.zeros(lines)}
        entry = "sw_structure_frame_fields_noalloc" if noalloc else "sw_structure_frame_fields"
        check(getattr(native.library(), entry)(self._handle, *[f64p(out[k]) for k in (
            "vertex_position", "vertex_velocity", "vertex_pending_force", "line_length", "line_violation")]))
        for k in ("vertex_position", "vertex_velocity", "vertex_pending_force"):
            out[k] = out[k].reshape(-1, 3)
        return out

    def frame(self, mapping: FrameMapping | None = None, context: native.FrameContext | None = None) -> bytes:
        """The diagnostic frame of the committed state as SWVF bytes
        (sw_structure_frame: buildStructureFrame, then serializeFrame)."""
        lib = native.library()
        mapping = mapping or FrameMapping.sequential(self._definition)  # owns the arrays m points into
        m = mapping._ctypes()
        c = context or frame_context("scene", "simwing-amd", 0.0)
        size = C.c_size_t()
        check(lib.sw_structure_frame(self._handle, C.byref(m), C.byref(c), None, 0, C.byref(size)))
        out = np.zeros(size.value, np.uint8)
        check(lib.sw_structure_frame(self._handle, C.byref(m), C.byref(c), u8p(out), out.size, C.byref(size)))
        return out.tobytes()

    def definition_fingerprint(self) -> int:
        value = C.c_uint64()
        check(native.library().sw_structure_fingerprint(self._handle, C.byref(value)))
        return value.value

    def accepted_step_count(self) -> int:
        steps = C.c_uint64()
        check(native.library().sw_structure_progress(self._handle, C.byref(steps), None))
        return steps.value

    def simulation_time_seconds(self) -> float:
        time = C.c_double()
        check(native.library().sw_structure_progress(self._handle, None, C.byref(time)))
        return time.value

    def clear_external_forces(self) -> None:
        check(native.library().sw_structure_clear_external_forces(self._handle))

    def add_external_force(self, node: int, force) -> None:
        f = np.ascontiguousarray(force, dtype=np.float64)
        check(native.library().sw_structure_add_external_force(self._handle, node, f64p(f)))

    def set_external_forces(self, forces: np.ndarray) -> None:
        f = np.ascontiguousarray(forces, dtype=np.float64).reshape(-1)
        check(native.library().sw_structure_set_external_forces(self._handle, f64p(f), f.size // 3))

    def set_face_pressure(self, pressure: np.ndarray) -> None:
        p = np.ascontiguousarray(pressure, dtype=np.float64)
        check(native.library().sw_structure_set_face_pressure(self._handle, f64p(p), p.size))

    def step(self, settings: native.StepSettings | None = None) -> dict:
        d = native.Diagnostics()
        s = settings or native.flight_step_settings()
        check(native.library().sw_structure_step(self._handle, C.byref(s), C.byref(d)))
        return _diagnostics_dict(d)

    def step_many(self, count: int, settings: native.StepSettings | None = None) -> None:
        s = settings or native.flight_step_settings()
        check(native.library().sw_structure_step_many(self._handle, C.byref(s), int(count)))

    def diagnostics(self) -> dict:
        d = native.Diagnostics()
        check(native.library().sw_structure_diagnostics(self._handle, C.byref(d)))
        return _diagnostics_dict(d)

    def node_states(self):
        n = self.node_count
        x, p, v = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        check(native.library().sw_structure_node_states(self._handle, f64p(x), f64p(p), f64p(v)))
        return x.reshape(n, 3), p.reshape(n, 3), v.reshape(n, 3)

    def payload_state(self) -> np.ndarray:
        out = np.zeros(13)
        check(native.library().sw_structure_payload_state(self._handle, f64p(out)))
        return out

    def solver_state(self) -> dict:
        a = self._definition.arrays
        c = a["constraint_kind"].size
        m = a["membrane_triangle"].size
        h = a["dihedral_rest_angle"].size
        s = 0 if self._definition.suspension is None else self._definition.suspension["segment_ids"].size
        out = {"constraint_lambda": np.zeros(c), "membrane_lambda": np.zeros(3 * m),
               "membrane_resultant": np.zeros(3 * m), "membrane_residual": np.zeros(m),
               "dihedral_lambda": np.zeros(h), "segment_lambda": np.zeros(s)}
        check(native.library().sw_structure_solver_state(
            self._handle, f64p(out["constraint_lambda"]), f64p(out["membrane_lambda"]),
            f64p(out["membrane_resultant"]), f64p(out["membrane_residual"]), f64p(out["dihedral_lambda"]),
            f64p(out["segment_lambda"])))
        return out

    def suspension_state(self) -> dict | None:
        """suspensionState() (sw_structure_suspension_state): None without a
        suspension."""
        d = self._definition
        present = C.c_int()
        if d.suspension is None:
            check(native.library().sw_structure_suspension_state(self._handle, C.byref(present), None, None, None,
                                                                  None, None, None, None))
            return None
        h = int(np.asarray(d.suspension["harness_ids"]).size)
        sn = int(np.asarray(d.suspension["segment_ids"]).size)
        out = {"payload": np.zeros(13), "harness_position": np.zeros(3 * h), "harness_velocity": np.zeros(3 * h),
               "segment_ids": np.zeros(sn, np.uint64), "segment_length": np.zeros(sn),
               "segment_commanded_rest": np.zeros(sn), "segment_tension": np.zeros(sn)}
        check(native.library().sw_structure_suspension_state(
            self._handle, C.byref(present), f64p(out["payload"]), f64p(out["harness_position"]),
            f64p(out["harness_velocity"]), u64p(out["segment_ids"]), f64p(out["segment_length"]),
            f64p(out["segment_commanded_rest"]), f64p(out["segment_tension"])))
        return out if present.value else None

    def checkpoint(self) -> "Checkpoint":
        handle = C.c_void_p()
        check(native.library().sw_structure_checkpoint(self._handle, C.byref(handle)))
        return Checkpoint(handle)

    def restore(self, checkpoint: "Checkpoint") -> None:
        check(native.library().sw_structure_restore(self._handle, checkpoint._handle))

    def serialize_checkpoint(self, checkpoint: "Checkpoint") -> bytes:
        """The reference's SWST bytes of a checkpoint of this structure
        (sw_checkpoint_serialize, serializeStructureCheckpoint)."""
        lib = native.library()
        size = C.c_size_t()
        check(lib.sw_checkpoint_serialize(self._handle, checkpoint._handle, None, 0, C.byref(size)))
        out = (C.c_uint8 * size.value)()
        check(lib.sw_checkpoint_serialize(self._handle, checkpoint._handle, out, size.value, C.byref(size)))
        return bytes(out)

    def deserialize_checkpoint(self, data: bytes) -> "Checkpoint":
        """A checkpoint of this structure read from SWST bytes
        (sw_checkpoint_deserialize, deserializeStructureCheckpoint)."""
        buffer = (C.c_uint8 * len(data)).from_buffer_copy(data)
        handle = C.c_void_p()
        check(native.library().sw_checkpoint_deserialize(self._handle, buffer, len(data), C.byref(handle)))
        return Checkpoint(handle)

    def colouring(self, family: int):
        lib = native.library()
        nc, no = C.c_size_t(), C.c_size_t()
        check(lib.sw_structure_colouring(self._handle, family, None, C.byref(nc), None, C.byref(no)))
        order = np.zeros(nc.value, np.uint64)
        offsets = np.zeros(no.value, np.uint64)
        check(lib.sw_structure_colouring(self._handle, family, u64p(order), None, u64p(offsets), None))
        return order, offsets

    def set_launch(self, workgroups: int, threads: int) -> None:
        check(native.library().sw_structure_set_launch(self._handle, workgroups, threads))

    def set_cooperative(self, cooperative: bool) -> None:
        check(native.library().sw_structure_set_cooperative(self._handle, int(cooperative)))

    def set_xcd_local(self, enabled: bool) -> bool:
        """One-XCD synchronisation of multi-workgroup launches
        (sw_structure_set_xcd_local); returns whether it is active."""
        active = C.c_int()
        check(native.library().sw_structure_set_xcd_local(self._handle, int(enabled), C.byref(active)))
        return bool(active.value)

    def set_kernel_timing(self, enabled: bool) -> None:
        check(native.library().sw_structure_set_kernel_timing(self._handle, int(enabled)))

    def set_profiling(self, enabled: bool) -> None:
        """Step profile on (totals from zero) or off (sw_structure_set_profiling)."""
        check(native.library().sw_structure_set_profiling(self._handle, int(enabled)))

    def profile(self) -> dict:
        """StepPerformanceProfile's soft-body fields over the profiled steps
        (sw_structure_profile): nanoseconds and visit counts."""
        out = native.StepProfile()
        check(native.library().sw_structure_profile(self._handle, C.byref(out)))
        return {name: int(getattr(out, name)) for name, _ in native.StepProfile._fields_}

    def reset_profile(self) -> None:
        check(native.library().sw_structure_reset_profile(self._handle))

    def kernel_time(self) -> tuple[float, int]:
        """Summed substep-kernel time (ms, HIP events on the structure's
        stream) and launch count since timing was enabled."""
        ms = C.c_double()
        launches = C.c_longlong()
        check(native.library().sw_structure_kernel_time(self._handle, C.byref(ms), C.byref(launches)))
        return ms.value, launches.value

    def launch(self) -> tuple[int, int]:
        w, t = C.c_int(), C.c_int()
        check(native.library().sw_structure_get_launch(self._handle, C.byref(w), C.byref(t)))
        return w.value, t.value


class Checkpoint:
    """Immutable structure checkpoint (StructureCheckpoint, structure.h:253-271)."""

    def __init__(self, handle):
        self._handle = handle

    def info(self) -> dict:
        """The public fields: version, definition fingerprint, accepted step
        count, simulation time and node count (sw_checkpoint_info)."""
        version, fingerprint, steps = C.c_uint32(), C.c_uint64(), C.c_uint64()
        time, nodes = C.c_double(), C.c_size_t()
        check(native.library().sw_checkpoint_info(self._handle, C.byref(version), C.byref(fingerprint),
                                                  C.byref(steps), C.byref(time), C.byref(nodes)))
        return {"version": version.value, "fingerprint": fingerprint.value, "accepted_steps": steps.value,
                "time": time.value, "node_count": nodes.value}

    def nodes(self) -> dict:
        """Node states, pending loads and the last applied load (sw_checkpoint_nodes)."""
        n = self.info()["node_count"]
        out = {k: np.zeros(3 * n) for k in ("position", "previous", "velocity", "pending_force")}
        out["last_applied"] = np.zeros(3)
        check(native.library().sw_checkpoint_nodes(self._handle, *[f64p(out[k]) for k in (
            "position", "previous", "velocity", "pending_force", "last_applied")]))
        return out

    def __del__(self):
        if getattr(self, "_handle", None):
            native.library().sw_checkpoint_release(self._handle)
            self._handle = None


class Batch:
    """Many uneven wings stepped together in one persistent launch."""

    def __init__(self, definitions: list[DefinitionArrays], device: int = 0):
        lib = native.library()
        views = [d.to_ctypes() for d in definitions]
        self._definitions = definitions
        self._views = views
        ptrs = (C.POINTER(native.Definition) * len(views))(*[C.pointer(v) for v in views])
        handle = C.c_void_p()
        check(lib.sw_batch_create(ptrs, len(views), device, C.byref(handle)))
        self._handle = handle
        self._dependents = []  # adjoints borrowing this batch: closed first

    def close(self) -> None:
        for dependent in getattr(self, "_dependents", []):
            dependent.close()
        if getattr(self, "_handle", None):
            native.library().sw_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        self.close()

    def set_face_pressure(self, wing: int, pressure: np.ndarray) -> None:
        p = np.ascontiguousarray(pressure, dtype=np.float64)
        check(native.library().sw_batch_set_face_pressure(self._handle, wing, f64p(p), p.size))

    def set_external_forces(self, wing: int, forces: np.ndarray) -> None:
        f = np.ascontiguousarray(forces, dtype=np.float64).reshape(-1)
        check(native.library().sw_batch_set_external_forces(self._handle, wing, f64p(f), f.size // 3))

    def step(self, count: int, settings: native.StepSettings | None = None) -> None:
        s = settings or native.flight_step_settings()
        check(native.library().sw_batch_step(self._handle, C.byref(s), int(count)))

    def node_states(self, wing: int):
        n = self._definitions[wing].node_count
        x, v = np.zeros(3 * n), np.zeros(3 * n)
        check(native.library().sw_batch_node_states(self._handle, wing, f64p(x), f64p(v)))
        return x.reshape(n, 3), v.reshape(n, 3)

    def set_launch(self, workgroups_cap: int, threads: int) -> None:
        check(native.library().sw_batch_set_launch(self._handle, workgroups_cap, threads))

    def workgroups(self) -> np.ndarray:
        """Workgroups each wing runs on in the current launch plan."""
        out = (C.c_int * len(self._definitions))()
        check(native.library().sw_batch_workgroups(self._handle, out, len(self._definitions)))
        return np.array(out[:], dtype=np.int64)

    def set_cooperative(self, cooperative: bool) -> None:
        check(native.library().sw_batch_set_cooperative(self._handle, int(cooperative)))

    def structure(self, wing: int) -> Structure:
        """The wing's own structure (owned by the batch)."""
        handle = C.c_void_p()
        check(native.library().sw_batch_structure(self._handle, wing, C.byref(handle)))
        return Structure._borrow(handle, self._definitions[wing], self)


class Adjoint:
    """Reverse-mode gradients of the shape loss sum |x(end) - target|^2 over
    `steps` steps from the structure's current state (loads pending now are
    applied at every step), with respect to per-membrane log-compliance,
    per-segment rest length and per-triangle pressure (sw_adjoint_*)."""

    def __init__(self, structure: Structure, steps: int, checkpoint_interval: int = 8,
                 settings: native.StepSettings | None = None):
        self._structure = structure  # kept alive: the adjoint borrows it
        d = structure.definition
        self.membrane_count = int(d.arrays["membrane_triangle"].size)
        self.segment_count = 0 if d.suspension is None else int(np.asarray(d.suspension["segment_rest"]).size)
        self.triangle_count = d.triangle_count
        self.node_count = d.node_count
        s = settings or native.flight_step_settings()
        handle = C.c_void_p()
        check(native.library().sw_adjoint_configure(structure._handle, C.byref(s), int(steps),
                                                    int(checkpoint_interval), C.byref(handle)))
        self._handle = handle
        structure._dependents.append(self)

    def close(self) -> None:
        if getattr(self, "_handle", None):
            native.library().sw_adjoint_destroy(self._handle)
            self._handle = None

    def __del__(self):
        self.close()

    @staticmethod
    def _param(values, count: int):
        if values is None:
            return None, None
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if a.size != count:
            raise ValueError(f"expected {count} values, got {a.size}")
        return a, f64p(a)

    def run(self, target, log_compliance=None, rest_length=None, pressure=None, gradients: bool = True) -> dict:
        """One forward + reverse run. Returns loss, final positions (N x 3)
        and, when `gradients`, the three gradients (definition order)."""
        t = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
        if t.size != 3 * self.node_count:
            raise ValueError("target needs 3 values per node")
        lc, lcp = self._param(log_compliance, self.membrane_count)
        rl, rlp = self._param(rest_length, self.segment_count)
        pr, prp = self._param(pressure, self.triangle_count)
        loss = C.c_double()
        final = np.zeros(3 * self.node_count)
        out = {}
        if gradients:
            out["grad_log_compliance"] = np.zeros(self.membrane_count)
            out["grad_rest_length"] = np.zeros(self.segment_count)
            out["grad_pressure"] = np.zeros(self.triangle_count)
        g = [f64p(out[k]) if gradients else None
             for k in ("grad_log_compliance", "grad_rest_length", "grad_pressure")]
        check(native.library().sw_adjoint_run(self._handle, lcp, rlp, prp, f64p(t), C.byref(loss), f64p(final),
                                              *g))
        out["loss"] = loss.value
        out["final_position"] = final.reshape(-1, 3)
        return out

    def set_workgroups(self, workgroups: int) -> None:
        check(native.library().sw_adjoint_set_workgroups(self._handle, int(workgroups)))

    def timing(self) -> tuple[float, float, int]:
        """Device ms of the last run's forward pass and of its replay +
        reverse kernels, and the device bytes the adjoint holds."""
        f, r, b = C.c_double(), C.c_double(), C.c_size_t()
        check(native.library().sw_adjoint_timing(self._handle, C.byref(f), C.byref(r), C.byref(b)))
        return f.value, r.value, int(b.value)

    def timing_detail(self) -> tuple[float, float, float]:
        """The last run's reverse pass split (ms): recorded replay, Jacobian
        passes, sweeps."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(native.library().sw_adjoint_timing_detail(self._handle, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value


class BatchAdjoint:
    """Rest-length gradients of every wing's shape loss, the wings stepped
    together (sw_batch_adjoint_*)."""

    def __init__(self, batch: Batch, steps: int, checkpoint_interval: int = 16, jacobian_chunk: int = 0,
                 settings: native.StepSettings | None = None):
        self._batch = batch
        self.wings = len(batch._definitions)
        self.segment_counts = [0 if d.suspension is None else int(np.asarray(d.suspension["segment_rest"]).size)
                               for d in batch._definitions]
        self.node_counts = [d.node_count for d in batch._definitions]
        s = settings or native.flight_step_settings()
        handle = C.c_void_p()
        check(native.library().sw_batch_adjoint_configure(batch._handle, C.byref(s), int(steps),
                                                          int(checkpoint_interval), int(jacobian_chunk),
                                                          C.byref(handle)))
        self._handle = handle
        batch._dependents.append(self)

    def close(self) -> None:
        if getattr(self, "_handle", None):
            native.library().sw_batch_adjoint_destroy(self._handle)
            self._handle = None

    def __del__(self):
        self.close()

    def run(self, targets, rest_lengths=None, gradients: bool = True) -> dict:
        n = self.wings
        keep = []

        def table(arrays, sizes):
            ptrs = (native._F64P * n)()
            for w in range(n):
                if arrays is not None and arrays[w] is not None:
                    a = np.ascontiguousarray(arrays[w], dtype=np.float64).reshape(-1)
                    if a.size != sizes[w]:
                        raise ValueError(f"wing {w}: expected {sizes[w]} values, got {a.size}")
                    keep.append(a)
                    ptrs[w] = f64p(a)
            return ptrs

        t = table(targets, [3 * c for c in self.node_counts])
        r = table(rest_lengths, self.segment_counts) if rest_lengths is not None else None
        losses = np.zeros(n)
        grads = [np.zeros(c) for c in self.segment_counts]
        finals = [np.zeros(3 * c) for c in self.node_counts]
        g = table(grads, self.segment_counts) if gradients else None
        fp = table(finals, [3 * c for c in self.node_counts])
        # `keep` holds every array the pointer tables point into until the call returns.
        check(native.library().sw_batch_adjoint_run(self._handle, r, t, f64p(losses), g, fp))
        out = {"loss": losses, "final_position": [f.reshape(-1, 3) for f in finals]}
        if gradients:
            out["grad_rest_length"] = grads
        return out

    def timing(self) -> tuple[float, float, int]:
        f, r, b = C.c_double(), C.c_double(), C.c_size_t()
        check(native.library().sw_batch_adjoint_timing(self._handle, C.byref(f), C.byref(r), C.byref(b)))
        return f.value, r.value, int(b.value)

    def workgroups(self) -> np.ndarray:
        out = (C.c_int * self.wings)()
        check(native.library().sw_batch_adjoint_workgroups(self._handle, out, self.wings))
        return np.array(out[:], dtype=np.int64)

    def timing_detail(self) -> tuple[float, float, float]:
        """The last run's reverse pass split (ms): replays, Jacobian passes, sweeps."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(native.library().sw_batch_adjoint_timing_detail(self._handle, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value
