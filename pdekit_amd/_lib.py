"""ctypes binding of libpdekit_rd.so (the C ABI declared in include/pdekit_rd.h).

The library is built in-tree by ``__graft_entry__.build()``; importing this
module without it raises immediately -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpdekit_rd.so")

PDRD_OK = 0
STATUS_NAMES = {
    -1: "PDRD_EINVAL",
    -2: "PDRD_ENOMEM",
    -3: "PDRD_EHIP",
    -4: "PDRD_ENOTSUP",
    -5: "PDRD_ENEGPRESSURE",
    -6: "PDRD_ENONFINITE",
}

# enumerations (mirror include/pdekit_rd.h)
TRIAG, QUAD, TETRA = 0, 1, 2
EQUIDIST, WARPBLEND = 0, 1
ADVECTION2D, BURGERS2D, EULER2D, EULER3D, ROTATION2D = 0, 1, 2, 3, 4
LDA, N, B, PSI, BPSI = 0, 1, 2, 3, 4
SOLUTION, RESIDUAL, TIME_STEP, WAVE_SPEED, DUAL_VOLUME, BLEND_COEFF, ART_VISC = 0, 1, 2, 3, 4, 5, 6
BLEND_LF = 0
OPT_SCATTER, OPT_PRECISION, OPT_BLOCK, OPT_TIMING, OPT_PATCH_CELLS, OPT_REFRESH, OPT_CHAIN = 0, 1, 2, 3, 4, 5, 6
WEAK_WALL, WEAK_SUBINLET, WEAK_SUBOUTLET, WEAK_FARFIELD, WEAK_SYMMETRY, WEAK_SUPERINLET, WEAK_DIRICHLET = range(7)
NOALLOC_SCRATCH_BYTES = 32 << 20
NOALLOC_MAX_ACC = NOALLOC_SCRATCH_BYTES // 8 - 8192 - 32  # doubles: nb_nodes * (m + 1) at most

# every symbol include/pdekit_rd.h declares
EXPORTED = (
    "pdrd_fe_describe", "pdrd_fe_fill", "pdrd_reference_nodes", "pdrd_quadrature",
    "pdrd_mesh_generate", "pdrd_mesh_free", "pdrd_mesh_jitter", "pdrd_mesh_bump", "pdrd_mesh_clone",
    "pdrd_mesh_info", "pdrd_mesh_coords", "pdrd_mesh_node_coords", "pdrd_mesh_cell_shapes",
    "pdrd_mesh_geo_conn", "pdrd_mesh_sol_conn", "pdrd_mesh_boundary_name", "pdrd_mesh_boundary_nodes",
    "pdrd_create", "pdrd_add_cells", "pdrd_initialize", "pdrd_destroy", "pdrd_set_option",
    "pdrd_set_stream", "pdrd_synchronize", "pdrd_set_field", "pdrd_get_field", "pdrd_field_device_ptr",
    "pdrd_add_strong_dirichlet", "pdrd_assemble_rhs", "pdrd_assemble_lhs_and_rhs", "pdrd_solve",
    "pdrd_compute_residual_norm", "pdrd_residual_norm_L2", "pdrd_iterate", "pdrd_check_status",
    "pdrd_method_info", "pdrd_last_timing", "pdrd_last_error", "pdrd_version", "pdrd_get_option",
    "pdrd_batch_create", "pdrd_batch_destroy", "pdrd_batch_set_graphs", "pdrd_batch_iterate", "pdrd_add_weak_bc", "pdrd_weak_bc_info",
    "pdrd_weak_bc_points", "pdrd_set_weak_bc_state",
    "pdrd_mesh_read_gmsh", "pdrd_mesh_write_gmsh", "pdrd_gmsh_append_node_data", "pdrd_gmsh_read_node_data",
    "pdrd_set_blending_coeff", "pdrd_set_artificial_viscosity", "pdrd_set_sources", "pdrd_compute_blending_coeff",
    "pdrd_compute_artificial_viscosity", "pdrd_element_jacobians", "pdrd_assemble_rhs_noalloc",
)


class FeDesc(C.Structure):
    _fields_ = [
        ("shape", C.c_int), ("geo_order", C.c_int), ("sol_order", C.c_int), ("quad_order", C.c_int),
        ("point_set", C.c_int), ("dim", C.c_int), ("ng", C.c_int), ("nd", C.c_int), ("nf", C.c_int),
        ("nq", C.c_int), ("table_size", C.c_size_t),
    ]


class PdrdError(RuntimeError):
    """A negative status returned by the C ABI."""

    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: {STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    P, I, U32, D, SZ, I64 = C.c_void_p, C.c_int, C.c_uint32, C.c_double, C.c_size_t, C.c_int64
    PD, PU32, PI64, PI = C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.POINTER(C.c_int64), C.POINTER(C.c_int)
    sig = {
        "pdrd_fe_describe": [I, I, I, I, I, C.POINTER(FeDesc)],
        "pdrd_fe_fill": [C.POINTER(FeDesc), PD],
        "pdrd_reference_nodes": [I, I, I, PD, PI],
        "pdrd_quadrature": [I, I, PD, PD, PI],
        "pdrd_mesh_generate": [I, PI64, PD, C.POINTER(P)],
        "pdrd_mesh_free": [P],
        "pdrd_mesh_jitter": [P, PD, D, C.c_uint64],
        "pdrd_mesh_bump": [P, I, PD],
        "pdrd_mesh_clone": [P, I, I],
        "pdrd_mesh_info": [P, PI64],
        "pdrd_mesh_coords": [P, PD],
        "pdrd_mesh_node_coords": [P, PD],
        "pdrd_mesh_cell_shapes": [P, C.POINTER(C.c_int32)],
        "pdrd_mesh_geo_conn": [P, PU32, PI64],
        "pdrd_mesh_sol_conn": [P, PU32, PI64],
        "pdrd_mesh_boundary_name": [P, I, C.c_char_p, I],
        "pdrd_mesh_boundary_nodes": [P, C.c_char_p, PU32, PI64],
        "pdrd_mesh_read_gmsh": [C.c_char_p, C.POINTER(P)],
        "pdrd_mesh_write_gmsh": [P, C.c_char_p],
        "pdrd_gmsh_append_node_data": [C.c_char_p, C.c_char_p, I, U32, PD, D, U32, I],
        "pdrd_gmsh_read_node_data": [C.c_char_p, I, C.POINTER(C.c_char_p), U32, PD, PI],
        "pdrd_create": [I, U32, PD, U32, I, I, PD, I, C.POINTER(P)],
        "pdrd_add_cells": [P, C.POINTER(FeDesc), PD, U32, PU32, PU32],
        "pdrd_initialize": [P],
        "pdrd_destroy": [P],
        "pdrd_set_option": [P, I, I],
        "pdrd_set_stream": [P, P],
        "pdrd_synchronize": [P],
        "pdrd_set_field": [P, I, PD],
        "pdrd_get_field": [P, I, PD],
        "pdrd_field_device_ptr": [P, I, C.POINTER(PD)],
        "pdrd_add_strong_dirichlet": [P, U32, PU32, PD],
        "pdrd_assemble_rhs": [P],
        "pdrd_assemble_lhs_and_rhs": [P, D],
        "pdrd_solve": [P],
        "pdrd_compute_residual_norm": [P, PD],
        "pdrd_residual_norm_L2": [P, I, PD],
        "pdrd_iterate": [P, D, I, I, PD],
        "pdrd_get_option": [P, I, PI],
        "pdrd_add_weak_bc": [P, I, U32, PU32, PD, I, I],
        "pdrd_weak_bc_info": [P, I, PI64],
        "pdrd_weak_bc_points": [P, I, PD],
        "pdrd_set_weak_bc_state": [P, I, PD, SZ],
        "pdrd_batch_create": [C.POINTER(P)],
        "pdrd_batch_destroy": [P],
        "pdrd_batch_set_graphs": [P, C.c_int],
        "pdrd_batch_iterate": [P, I, C.POINTER(P), PD, I, I, PD],
        "pdrd_set_blending_coeff": [P, PD],
        "pdrd_set_artificial_viscosity": [P, PD],
        "pdrd_set_sources": [P, PD],
        "pdrd_compute_blending_coeff": [P, I],
        "pdrd_compute_artificial_viscosity": [P, D, D],
        "pdrd_element_jacobians": [P, I, PD, PD, PD],
        # device pointers for tables, connectivity, coords, u, V, outputs, status; stream
        "pdrd_assemble_rhs_noalloc": [I, I, PD, I, C.POINTER(FeDesc), P, U32, P, P, U32, P, U32, P, P, P, P, P,
                                      P],
        "pdrd_check_status": [P],
        "pdrd_method_info": [P, PI64],
        "pdrd_last_timing": [P, PD, PD],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.pdrd_last_error.restype = C.c_char_p
    L.pdrd_last_error.argtypes = []
    L.pdrd_version.restype = C.c_char_p
    L.pdrd_version.argtypes = []
    _lib = L
    return L


def check(rc: int, where: str) -> None:
    if rc != PDRD_OK:
        msg = lib().pdrd_last_error().decode(errors="replace")
        raise PdrdError(rc, where, msg)


def ptr(a, ctype=C.c_double):
    """Pointer to the data of a contiguous numpy array."""
    return a.ctypes.data_as(C.POINTER(ctype))
