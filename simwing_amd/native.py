"""ctypes declarations of the C-ABI in include/simwing_xpbd.h.

The library is loaded from the package directory; a missing library is an
error (there is no CPU fallback for the product path).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PACKAGE = Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE / "libsimwing_xpbd.so"

SW_OK = 0
SW_INVALID_ARGUMENT = 1
SW_OUT_OF_RANGE = 2
SW_SUSPENSION_ERROR = 3
SW_LOGIC_ERROR = 4
SW_DEVICE_ERROR = 5
SW_BAD_ALLOC = 6

# Constraint kinds (sw_definition.constraint_kind).
SW_CONSTRAINT_DISTANCE = 1
SW_CONSTRAINT_CABLE = 2
SW_CONSTRAINT_SUSPENSION_TIE = 3
SW_CONSTRAINT_SEAM_STITCH = 4

_U64P = C.POINTER(C.c_uint64)
_F64P = C.POINTER(C.c_double)
_U8P = C.POINTER(C.c_uint8)
_U32P = C.POINTER(C.c_uint32)


class SuspensionDefinition(C.Structure):
    _fields_ = [
        ("pilot_stable_id", C.c_uint64),
        ("pilot_mass", C.c_double),
        ("pilot_com_local", C.c_double * 3),
        ("pilot_com_world", C.c_double * 3),
        ("pilot_velocity", C.c_double * 3),
        ("pilot_angular_velocity", C.c_double * 3),
        ("pilot_orientation", C.c_double * 4),
        ("pilot_inertia", C.c_double * 3),
        ("attachment_count", C.c_size_t),
        ("attachment_ids", _U64P),
        ("attachment_nodes", _U64P),
        ("junction_count", C.c_size_t),
        ("junction_ids", _U64P),
        ("junction_nodes", _U64P),
        ("harness_count", C.c_size_t),
        ("harness_ids", _U64P),
        ("harness_local", _F64P),
        ("segment_count", C.c_size_t),
        ("segment_ids", _U64P),
        ("segment_from_kind", _U8P),
        ("segment_from_id", _U64P),
        ("segment_to_kind", _U8P),
        ("segment_to_id", _U64P),
        ("segment_rest", _F64P),
        ("segment_stiffness", _F64P),
        ("segment_damping", _F64P),
        ("segment_role", _U32P),
        ("solver_iterations", C.c_int),
        ("attachment_tolerance", C.c_double),
        ("minimum_line_length", C.c_double),
        ("maximum_line_residual", C.c_double),
        ("maximum_control_work", C.c_double),
    ]


class Definition(C.Structure):
    _fields_ = [
        ("node_count", C.c_size_t),
        ("node_position", _F64P),
        ("node_mass", _F64P),
        ("node_fixed", _U8P),
        ("triangle_count", C.c_size_t),
        ("triangle_nodes", _U64P),
        ("constraint_count", C.c_size_t),
        ("constraint_kind", _U8P),
        ("constraint_nodes", _U64P),
        ("constraint_rest", _F64P),
        ("constraint_compliance", _F64P),
        ("membrane_count", C.c_size_t),
        ("membrane_triangle", _U64P),
        ("membrane_chart", _F64P),
        ("membrane_material", _F64P),
        ("membrane_role", _U8P),
        ("dihedral_count", C.c_size_t),
        ("dihedral_nodes", _U64P),
        ("dihedral_rest_angle", _F64P),
        ("dihedral_compliance", _F64P),
        ("suspension", C.POINTER(SuspensionDefinition)),
        ("sphere_count", C.c_size_t),
        ("sphere_centre", _U64P),
        ("sphere_radius", _F64P),
        ("sphere_compliance", _F64P),
        ("sphere_offsets", _U64P),
        ("sphere_nodes", _U64P),
    ]


class StepSettings(C.Structure):
    _fields_ = [
        ("time_step", C.c_double),
        ("substeps", C.c_int),
        ("constraint_iterations", C.c_int),
        ("cable_sweep_pairs", C.c_int),
        ("gravity", C.c_double * 3),
        ("velocity_damping", C.c_double),
        ("damping_reference_velocity", C.c_double * 3),
        ("worker_threads", C.c_uint),
    ]


class StepProfile(C.Structure):
    """sw_step_profile: StepPerformanceProfile's soft-body fields."""
    _fields_ = [(name, C.c_uint64) for name in (
        "soft_body_total_ns", "transaction_snapshot_ns", "prediction_ns", "distance_constraint_ns",
        "cable_constraint_ns", "membrane_constraint_ns", "bending_constraint_ns", "suspension_constraint_ns",
        "contact_constraint_ns", "membrane_diagnostics_ns", "finalization_ns", "structural_substeps",
        "constraint_iterations", "distance_constraint_visits", "cable_constraint_visits",
        "membrane_constraint_visits", "bending_constraint_visits", "suspension_constraint_visits",
        "contact_constraint_visits")]


class Diagnostics(C.Structure):
    _fields_ = [
        ("node_count", C.c_size_t),
        ("dynamic_node_count", C.c_size_t),
        ("triangle_count", C.c_size_t),
        ("constraint_count", C.c_size_t),
        ("membrane_count", C.c_size_t),
        ("dihedral_count", C.c_size_t),
        ("contact_pair_count", C.c_size_t),
        ("active_contact_count", C.c_size_t),
        ("suspension_segment_count", C.c_size_t),
        ("total_dynamic_mass", C.c_double),
        ("center_of_mass", C.c_double * 3),
        ("linear_momentum", C.c_double * 3),
        ("kinetic_energy", C.c_double),
        ("maximum_distance_error", C.c_double),
        ("maximum_cable_extension", C.c_double),
        ("maximum_absolute_membrane_strain", C.c_double),
        ("maximum_membrane_residual", C.c_double),
        ("maximum_contact_penetration", C.c_double),
        ("maximum_suspension_residual", C.c_double),
        ("pending_external_force", C.c_double * 3),
        ("last_applied_external_force", C.c_double * 3),
        ("finite", C.c_int),
    ]


class FrameMapping(C.Structure):
    _fields_ = [
        ("vertex_count", C.c_size_t),
        ("vertex_ids", _U64P),
        ("triangle_count", C.c_size_t),
        ("triangle_ids", _U64P),
        ("triangle_negative_region", _U64P),
        ("triangle_positive_region", _U64P),
        ("line_count", C.c_size_t),
        ("line_ids", _U64P),
        ("line_roles", _U32P),
    ]


class FrameContext(C.Structure):
    _fields_ = [
        ("scene_checksum", C.c_char_p),
        ("solver_commit", C.c_char_p),
        ("time_step", C.c_double),
        ("coupling_iteration", C.c_uint32),
        ("coupling_residuals", C.c_double * 5),
        ("fluid_mass", C.c_double),
        ("total_momentum", C.c_double * 3),
        ("total_energy", C.c_double),
        ("interface_force_residual", C.c_double * 3),
        ("interface_moment_residual", C.c_double * 3),
        ("interface_power_residual", C.c_double),
    ]


class AssemblySettings(C.Structure):
    _fields_ = [
        ("seam_coincidence_tolerance", C.c_double),
        ("suspension_iterations", C.c_int),
        ("suspension_attachment_tolerance", C.c_double),
        ("suspension_minimum_line_length", C.c_double),
        ("suspension_maximum_line_residual", C.c_double),
        ("suspension_maximum_control_work", C.c_double),
        ("contact_sphere_compliance", C.c_double),
    ]


def flight_step_settings(iterations: int = 48) -> StepSettings:
    """The surface-flight solver settings (surface_flight_case.cpp:94-114)."""
    s = StepSettings()
    s.time_step = 1.0 / 120.0
    s.substeps = 1
    s.constraint_iterations = iterations
    s.cable_sweep_pairs = 3
    s.gravity[:] = (0.0, 0.0, -9.80665)
    s.velocity_damping = 8.0
    s.damping_reference_velocity[:] = (0.0, 0.0, 0.0)
    s.worker_threads = 12
    return s


def flight_assembly_settings(iterations: int = 48) -> AssemblySettings:
    a = AssemblySettings()
    a.seam_coincidence_tolerance = 1.0e-9
    a.suspension_iterations = iterations
    a.suspension_attachment_tolerance = 1.0e-10
    a.suspension_minimum_line_length = 1.0e-10
    a.suspension_maximum_line_residual = 0.10
    a.suspension_maximum_control_work = 1.0e6
    a.contact_sphere_compliance = 0.0
    return a


# Every function the header declares, with its ctypes signature.
_SIGNATURES = {
    "sw_last_error": (C.c_char_p, []),
    "sw_last_suspension_phase": (C.c_int, []),
    "sw_last_suspension_entity": (C.c_char_p, []),
    "sw_structure_create": (C.c_int, [C.POINTER(Definition), C.c_int, C.POINTER(C.c_void_p)]),
    "sw_structure_destroy": (None, [C.c_void_p]),
    "sw_structure_fingerprint": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "sw_structure_progress": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "sw_structure_clear_external_forces": (C.c_int, [C.c_void_p]),
    "sw_structure_add_external_force": (C.c_int, [C.c_void_p, C.c_size_t, _F64P]),
    "sw_structure_set_external_forces": (C.c_int, [C.c_void_p, _F64P, C.c_size_t]),
    "sw_structure_set_face_pressure": (C.c_int, [C.c_void_p, _F64P, C.c_size_t]),
    "sw_structure_step": (C.c_int, [C.c_void_p, C.POINTER(StepSettings), C.POINTER(Diagnostics)]),
    "sw_structure_diagnostics": (C.c_int, [C.c_void_p, C.POINTER(Diagnostics)]),
    "sw_structure_step_many": (C.c_int, [C.c_void_p, C.POINTER(StepSettings), C.c_int]),
    "sw_structure_node_states": (C.c_int, [C.c_void_p, _F64P, _F64P, _F64P]),
    "sw_structure_payload_state": (C.c_int, [C.c_void_p, _F64P]),
    "sw_structure_solver_state": (C.c_int, [C.c_void_p, _F64P, _F64P, _F64P, _F64P, _F64P, _F64P]),
    "sw_structure_checkpoint": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "sw_structure_restore": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sw_checkpoint_release": (None, [C.c_void_p]),
    "sw_structure_colouring": (C.c_int, [C.c_void_p, C.c_int, _U64P, C.POINTER(C.c_size_t), _U64P,
                                         C.POINTER(C.c_size_t)]),
    "sw_definition_colouring": (C.c_int, [C.POINTER(Definition), C.c_int, _U64P, C.POINTER(C.c_size_t), _U64P,
                                          C.POINTER(C.c_size_t)]),
    "sw_device_math": (C.c_int, [C.c_int, C.c_int, _F64P, _F64P, _F64P, C.c_size_t]),
    "sw_structure_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "sw_structure_get_launch": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "sw_structure_set_kernel_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "sw_structure_set_cooperative": (C.c_int, [C.c_void_p, C.c_int]),
    "sw_structure_kernel_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "sw_structure_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "sw_structure_profile": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sw_structure_reset_profile": (C.c_int, [C.c_void_p]),
    "sw_scene_generate": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "sw_scene_destroy": (None, [C.c_void_p]),
    "sw_scene_face_pressure": (C.c_int, [C.c_void_p, C.POINTER(_F64P), C.POINTER(C.c_size_t)]),
    "sw_scene_assemble": (C.c_int, [C.c_void_p, C.POINTER(AssemblySettings), C.POINTER(C.c_void_p)]),
    "sw_assembly_destroy": (None, [C.c_void_p]),
    "sw_assembly_definition": (C.c_int, [C.c_void_p, C.POINTER(Definition)]),
    "sw_assembly_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "sw_assembly_flight_loads": (C.c_int, [C.c_void_p, _F64P]),
    "sw_structure_frame_fields": (C.c_int, [C.c_void_p, _F64P, _F64P, _F64P, _F64P, _F64P]),
    "sw_structure_frame_fields_noalloc": (C.c_int, [C.c_void_p, _F64P, _F64P, _F64P, _F64P, _F64P]),
    "sw_structure_set_xcd_local": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "sw_definition_fingerprint": (C.c_int, [C.POINTER(Definition), C.POINTER(C.c_uint64)]),
    "sw_checkpoint_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_double), C.POINTER(C.c_size_t)]),
    "sw_checkpoint_nodes": (C.c_int, [C.c_void_p, _F64P, _F64P, _F64P, _F64P, _F64P]),
    "sw_checkpoint_serialize": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "sw_checkpoint_deserialize": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "sw_structure_suspension_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), _F64P, _F64P, _F64P,
                                                C.POINTER(C.c_uint64), _F64P, _F64P, _F64P]),
    "sw_scene_read": (C.c_int, [_U8P, C.c_size_t, C.POINTER(C.c_void_p)]),
    "sw_scene_write": (C.c_int, [C.c_void_p, _U8P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "sw_structure_frame": (C.c_int, [C.c_void_p, C.POINTER(FrameMapping), C.POINTER(FrameContext), _U8P,
                                     C.c_size_t, C.POINTER(C.c_size_t)]),
    "sw_frame_encode": (C.c_int, [C.c_void_p, C.POINTER(FrameMapping), C.POINTER(FrameContext), C.c_uint64,
                                  C.c_double, _F64P, _F64P, _F64P, _F64P, _F64P, C.POINTER(Diagnostics), _U8P,
                                  C.c_size_t, C.POINTER(C.c_size_t)]),
    "sw_batch_create": (C.c_int, [C.POINTER(C.POINTER(Definition)), C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]),
    "sw_batch_destroy": (None, [C.c_void_p]),
    "sw_batch_set_face_pressure": (C.c_int, [C.c_void_p, C.c_size_t, _F64P, C.c_size_t]),
    "sw_batch_set_external_forces": (C.c_int, [C.c_void_p, C.c_size_t, _F64P, C.c_size_t]),
    "sw_batch_step": (C.c_int, [C.c_void_p, C.POINTER(StepSettings), C.c_int]),
    "sw_batch_node_states": (C.c_int, [C.c_void_p, C.c_size_t, _F64P, _F64P]),
    "sw_batch_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "sw_batch_workgroups": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_size_t]),
    "sw_batch_set_cooperative": (C.c_int, [C.c_void_p, C.c_int]),
    "sw_batch_structure": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "sw_adjoint_configure": (C.c_int, [C.c_void_p, C.POINTER(StepSettings), C.c_uint64, C.c_uint32,
                                       C.POINTER(C.c_void_p)]),
    "sw_adjoint_destroy": (None, [C.c_void_p]),
    "sw_batch_adjoint_configure": (C.c_int, [C.c_void_p, C.POINTER(StepSettings), C.c_uint64, C.c_uint32, C.c_uint32,
                                             C.POINTER(C.c_void_p)]),
    "sw_batch_adjoint_destroy": (None, [C.c_void_p]),
    "sw_batch_adjoint_run": (C.c_int, [C.c_void_p, C.POINTER(_F64P), C.POINTER(_F64P), _F64P, C.POINTER(_F64P),
                                       C.POINTER(_F64P)]),
    "sw_batch_adjoint_timing": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_size_t)]),
    "sw_batch_adjoint_workgroups": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_size_t]),
    "sw_batch_adjoint_timing_detail": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                 C.POINTER(C.c_double)]),
    "sw_adjoint_run": (C.c_int, [C.c_void_p, _F64P, _F64P, _F64P, _F64P, C.POINTER(C.c_double), _F64P, _F64P,
                                 _F64P, _F64P]),
    "sw_adjoint_set_workgroups": (C.c_int, [C.c_void_p, C.c_int]),
    "sw_adjoint_timing": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_size_t)]),
    "sw_adjoint_timing_detail": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                           C.POINTER(C.c_double)]),
}

_lib = None


def library() -> C.CDLL:
    """Loads libsimwing_xpbd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            raise RuntimeError(f"{LIBRARY_PATH} is missing: run __graft_entry__.build() first")
        lib = C.CDLL(str(LIBRARY_PATH))
        for name, (restype, argtypes) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = lib
    return _lib


def declared_symbols() -> list[str]:
    return list(_SIGNATURES)


class NativeError(RuntimeError):
    def __init__(self, status: int, message: str, phase: int = -1, entity: str = ""):
        super().__init__(f"[{status}] {message}")
        self.status = status
        self.phase = phase
        self.entity = entity


def check(status: int) -> None:
    if status == SW_OK:
        return
    lib = library()
    message = lib.sw_last_error().decode(errors="replace")
    phase = lib.sw_last_suspension_phase() if status == SW_SUSPENSION_ERROR else -1
    entity = lib.sw_last_suspension_entity().decode(errors="replace") if status == SW_SUSPENSION_ERROR else ""
    raise NativeError(status, message, phase, entity)


def f64p(array: np.ndarray):
    return array.ctypes.data_as(_F64P)


def u64p(array: np.ndarray):
    return array.ctypes.data_as(_U64P)


def u8p(array: np.ndarray):
    return array.ctypes.data_as(_U8P)


def u32p(array: np.ndarray):
    return array.ctypes.data_as(_U32P)
