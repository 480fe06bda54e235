"""ORACLE (test infrastructure only): ctypes wrapper of liboracle_xpbd.so,
the CPU restatement of the reference's XPBD substep. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIBRARY_PATH = Path(__file__).resolve().parent / "liboracle_xpbd.so"
# The same restatement computed in long double (x87 extended precision):
# its central differences are the low-noise derivative check of the adjoint.
EXTENDED_LIBRARY_PATH = Path(__file__).resolve().parent / "liboracle_xpbd_ext.so"
# ... and with forward-mode tangent pairs (exact directional derivatives).
TANGENT_LIBRARY_PATH = Path(__file__).resolve().parent / "liboracle_xpbd_tan.so"
_libs = {}
_PATHS = {False: LIBRARY_PATH, True: EXTENDED_LIBRARY_PATH, "tangent": TANGENT_LIBRARY_PATH}


def library(extended=False):
    """The double build, the long-double build (extended=True) or the
    tangent build (extended="tangent")."""
    if extended not in _libs:
        path = _PATHS[extended]
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run make -C oracle")
        lib = C.CDLL(str(path))
        lib.oracle_last_error.restype = C.c_char_p
        lib.oracle_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        lib.oracle_destroy.argtypes = [C.c_void_p]
        lib.oracle_destroy.restype = None
        lib.oracle_set_face_pressure.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        lib.oracle_set_external_forces.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        lib.oracle_step.argtypes = [C.c_void_p, C.c_void_p]
        lib.oracle_state.argtypes = [C.c_void_p] + [C.c_void_p] * 11
        lib.oracle_frame_fields.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        lib.oracle_colouring.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_size_t), C.c_void_p,
                                         C.POINTER(C.c_size_t)]
        lib.oracle_set_angle_perturbation.argtypes = [C.c_void_p, C.c_int]
        lib.oracle_set_angle_perturbation.restype = None
        lib.oracle_set_portable_math.argtypes = [C.c_void_p, C.c_int]
        lib.oracle_set_portable_math.restype = None
        for name in ("oracle_math",):
            getattr(lib, name).argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
            getattr(lib, name).restype = None
        lib.oracle_quad_math.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
        lib.oracle_quad_math.restype = None
        lib.oracle_set_tangent.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.oracle_tangent_state.argtypes = [C.c_void_p, C.c_void_p]
        _libs[extended] = lib
    return _libs[extended]


MATH_FUNCTIONS = {"atan2": 2, "sin": 4, "cos": 5, "acos": 6}


def math(name: str, a: np.ndarray, b: np.ndarray | None = None, portable: bool = True) -> np.ndarray:
    """The oracle's transcendental `name` on the host: the portable functions
    (portable=True) or the C library's."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(a if b is None else b, np.float64)
    out = np.empty_like(a)
    library().oracle_math(int(portable), MATH_FUNCTIONS[name], _p(a), _p(b), _p(out), a.size)
    return out


def quad_math(name: str, a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    """`name` in quadruple precision (libquadmath), rounded to double."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(a if b is None else b, np.float64)
    out = np.empty_like(a)
    library().oracle_quad_math(MATH_FUNCTIONS[name], _p(a), _p(b), _p(out), a.size)
    return out


class OracleError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"[{status}] {message}")
        self.status = status


def _check(status, lib=None):
    if status != 0:
        raise OracleError(status, (lib or library()).oracle_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleBody:
    """CPU XPBD body in the reference's coloured sweep order (`extended`:
    True computes in long double, "tangent" with tangent pairs)."""

    def __init__(self, definition, extended=False):
        lib = library(extended)
        self._lib = lib
        view = definition.to_ctypes()
        self._definition = definition
        handle = C.c_void_p()
        _check(lib.oracle_create(C.byref(view), C.byref(handle)), lib)
        self._handle = handle
        a = definition.arrays
        self.n = a["node_mass"].size
        self.c = a["constraint_kind"].size
        self.m = a["membrane_triangle"].size
        self.h = a["dihedral_rest_angle"].size
        self.s = 0 if definition.suspension is None else definition.suspension["segment_ids"].size

    def __del__(self):
        if getattr(self, "_handle", None):
            self._lib.oracle_destroy(self._handle)
            self._handle = None

    def set_face_pressure(self, pressure):
        p = np.ascontiguousarray(pressure, dtype=np.float64)
        _check(self._lib.oracle_set_face_pressure(self._handle, _p(p), p.size), self._lib)

    def set_external_forces(self, forces):
        f = np.ascontiguousarray(forces, dtype=np.float64).reshape(-1)
        _check(self._lib.oracle_set_external_forces(self._handle, _p(f), f.size // 3), self._lib)

    def step(self, settings):
        _check(self._lib.oracle_step(self._handle, C.byref(settings)), self._lib)

    def use_portable_math(self, enabled: bool = True):
        """Steps use the portable transcendentals of the device kernels."""
        self._lib.oracle_set_portable_math(self._handle, int(enabled))

    def set_tangent(self, d_log_compliance=None, d_rest=None, d_pressure=None):
        """Tangent build: seeds one parameter direction (after the face
        pressure is set; definition order)."""
        arrays = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                  for a in (d_log_compliance, d_rest, d_pressure)]
        self._seeds = arrays
        _check(self._lib.oracle_set_tangent(self._handle, *[_p(a) for a in arrays]), self._lib)

    def tangent_positions(self) -> np.ndarray:
        out = np.zeros(3 * self.n)
        _check(self._lib.oracle_tangent_state(self._handle, _p(out)), self._lib)
        return out

    def perturb_angles(self, ulps: int):
        self._lib.oracle_set_angle_perturbation(self._handle, ulps)

    def state(self) -> dict:
        out = {
            "position": np.zeros(3 * self.n), "previous": np.zeros(3 * self.n), "velocity": np.zeros(3 * self.n),
            "constraint_lambda": np.zeros(self.c), "membrane_lambda": np.zeros(3 * self.m),
            "membrane_resultant": np.zeros(3 * self.m), "membrane_residual": np.zeros(self.m),
            "dihedral_lambda": np.zeros(self.h), "segment_lambda": np.zeros(self.s), "payload": np.zeros(13),
            "suspension_residual": np.zeros(1),
        }
        _check(self._lib.oracle_state(self._handle, *[_p(out[k]) for k in (
            "position", "previous", "velocity", "constraint_lambda", "membrane_lambda", "membrane_resultant",
            "membrane_residual", "dihedral_lambda", "segment_lambda", "payload", "suspension_residual")]), self._lib)
        return out

    def frame_fields(self) -> dict:
        """Diagnostic-frame fields (oracle_frame_fields): vertices are the
        nodes then the harness points, lines the constraints then the
        segments, all in definition order."""
        hcount = 0 if self._definition.suspension is None else self._definition.suspension["harness_ids"].size
        v = self.n + hcount
        out = {"vertex_position": np.zeros(3 * v), "vertex_velocity": np.zeros(3 * v),
               "vertex_pending_force": np.zeros(3 * v), "line_length": np.zeros(self.c + self.s),
               "line_violation": np.zeros(self.c + self.s)}
        _check(self._lib.oracle_frame_fields(self._handle, *[_p(out[k]) for k in (
            "vertex_position", "vertex_velocity", "vertex_pending_force", "line_length", "line_violation")]),
            self._lib)
        return out

    def colouring(self, family):
        lib = self._lib
        nc, no = C.c_size_t(), C.c_size_t()
        _check(lib.oracle_colouring(self._handle, family, None, C.byref(nc), None, C.byref(no)))
        order = np.zeros(nc.value, np.uint64)
        offsets = np.zeros(no.value, np.uint64)
        _check(lib.oracle_colouring(self._handle, family, _p(order), None, _p(offsets), None))
        return order, offsets
