"""FE-value and quadrature inputs of the cell loop (host side).

Mirrors what pdekit's FunctionSpaceCells::set_reference_fe_values and
FluxSpaceMetric::allocate_buffer prepare once per element type
(src/interpolation/FunctionSpaceCells.hpp:135-190,
src/interpolation/FluxSpaceMetric.hpp:237-456, FEValues.cpp:157-279).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

SHAPE_DIM = {L.TRIAG: 2, L.QUAD: 2, L.TETRA: 3}


@dataclass
class ElementTables:
    desc: L.FeDesc
    packed: np.ndarray
    qp: np.ndarray
    w: np.ndarray
    Vg: np.ndarray
    dVg: np.ndarray
    V: np.ndarray
    dV: np.ndarray
    Vf: np.ndarray
    dVf: np.ndarray
    Vgf: np.ndarray
    Vsf: np.ndarray
    sol_nodes: np.ndarray
    flux_nodes: np.ndarray

    @property
    def nd(self) -> int:
        return self.desc.nd

    @property
    def ng(self) -> int:
        return self.desc.ng

    @property
    def nf(self) -> int:
        return self.desc.nf

    @property
    def nq(self) -> int:
        return self.desc.nq

    @property
    def dim(self) -> int:
        return self.desc.dim


def element_tables(shape: int, sol_order: int, quad_order: int, geo_order: int = 1,
                   point_set: int = L.WARPBLEND) -> ElementTables:
    lib = L.lib()
    d = L.FeDesc()
    L.check(lib.pdrd_fe_describe(shape, geo_order, sol_order, quad_order, point_set, C.byref(d)),
            "pdrd_fe_describe")
    buf = np.zeros(d.table_size, dtype=np.float64)
    L.check(lib.pdrd_fe_fill(C.byref(d), L.ptr(buf)), "pdrd_fe_fill")
    D, ng, nd, nf, nq = d.dim, d.ng, d.nd, d.nf, d.nq
    sizes = [("qp", (nq, D)), ("w", (nq,)), ("Vg", (nq, ng)), ("dVg", (D, nq, ng)), ("V", (nq, nd)),
             ("dV", (D, nq, nd)), ("Vf", (nq, nf)), ("dVf", (D, nq, nf)), ("Vgf", (nf, ng)),
             ("Vsf", (nf, nd)), ("sol_nodes", (nd, D)), ("flux_nodes", (nf, D))]
    out = {}
    off = 0
    for name, shp in sizes:
        n = int(np.prod(shp))
        out[name] = buf[off:off + n].reshape(shp)
        off += n
    assert off == buf.size
    return ElementTables(desc=d, packed=buf, **out)


def reference_nodes(shape: int, order: int, point_set: int = L.WARPBLEND) -> np.ndarray:
    lib = L.lib()
    n = C.c_int(0)
    L.check(lib.pdrd_reference_nodes(shape, order, point_set, None, C.byref(n)), "pdrd_reference_nodes")
    xi = np.zeros((n.value, SHAPE_DIM[shape]))
    L.check(lib.pdrd_reference_nodes(shape, order, point_set, L.ptr(xi), C.byref(n)), "pdrd_reference_nodes")
    return xi


def quadrature(shape: int, order: int):
    lib = L.lib()
    nq = C.c_int(0)
    L.check(lib.pdrd_quadrature(shape, order, None, None, C.byref(nq)), "pdrd_quadrature")
    xi = np.zeros((nq.value, SHAPE_DIM[shape]))
    w = np.zeros(nq.value)
    L.check(lib.pdrd_quadrature(shape, order, L.ptr(xi), L.ptr(w), C.byref(nq)), "pdrd_quadrature")
    return xi, w
