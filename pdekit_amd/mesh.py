"""Mesh inputs of the cell loop (host side): structured generators, jitter,
bump mappings and the continuous P1 -> Pk DOF numbering.

Mirrors src/mesh/io/MeshCreator.hpp:82-99, 436-600 (make_unit_quad) and
src/mesh/containers/DofMap.hpp:2225-2290, 2556-3073 (clone_continuous).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

NB_VERT = {L.TRIAG: 3, L.QUAD: 4, L.TETRA: 4}


def _nb_nodes(shape: int, p: int) -> int:
    if shape == L.TRIAG:
        return (p + 1) * (p + 2) // 2
    if shape == L.QUAD:
        return (p + 1) ** 2
    return (p + 1) * (p + 2) * (p + 3) // 6


@dataclass
class CellBlock:
    """Cells of one shape, in mesh order."""
    shape: int
    cells: np.ndarray      # global cell ids
    geo_conn: np.ndarray   # [ne][ng] uint32
    sol_conn: np.ndarray   # [ne][nd] uint32


@dataclass
class Mesh:
    dim: int
    coords: np.ndarray            # [nv][dim]
    cell_shape: np.ndarray        # [ncells]
    order: int
    nb_nodes: int
    node_coords: np.ndarray       # [N][dim]
    blocks: list = field(default_factory=list)
    boundaries: dict = field(default_factory=dict)  # name -> solution node ids
    generator: object = field(default=None, repr=False, compare=False)  # the MeshGen it was built from

    @property
    def nb_cells(self) -> int:
        return int(self.cell_shape.size)


class MeshGen:
    """Owner of a pdrd_meshgen handle."""

    def __init__(self, kind: int, iparams, rparams=None):
        lib = L.lib()
        ip = np.zeros(8, dtype=np.int64)
        ip[:len(iparams)] = iparams
        rp = np.zeros(8, dtype=np.float64)
        if rparams is not None:
            rp[:len(rparams)] = rparams
        h = C.c_void_p()
        L.check(lib.pdrd_mesh_generate(kind, L.ptr(ip, C.c_int64), L.ptr(rp), C.byref(h)), "pdrd_mesh_generate")
        self._h = h
        self._cloned = False

    @classmethod
    def from_gmsh(cls, path: str) -> "MeshGen":
        """GmshReader::read_mesh_from_file (Gmsh 2.2 ASCII, straight-sided cells)."""
        self = cls.__new__(cls)
        h = C.c_void_p()
        L.check(L.lib().pdrd_mesh_read_gmsh(str(path).encode(), C.byref(h)), "pdrd_mesh_read_gmsh")
        self._h = h
        self._cloned = False
        return self

    def write_gmsh(self, path: str):
        L.check(L.lib().pdrd_mesh_write_gmsh(self._h, str(path).encode()), "pdrd_mesh_write_gmsh")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib().pdrd_mesh_free(h)
            self._h = None

    def info(self):
        out = np.zeros(5, dtype=np.int64)
        L.check(L.lib().pdrd_mesh_info(self._h, L.ptr(out, C.c_int64)), "pdrd_mesh_info")
        return out

    def jitter(self, h, f: float, seed: int):
        hh = np.asarray(h, dtype=np.float64)
        L.check(L.lib().pdrd_mesh_jitter(self._h, L.ptr(hh), float(f), C.c_uint64(seed)), "pdrd_mesh_jitter")
        return self

    def bump(self, mode: int, params):
        p = np.zeros(5)
        p[:len(params)] = params
        L.check(L.lib().pdrd_mesh_bump(self._h, mode, L.ptr(p)), "pdrd_mesh_bump")
        return self

    def clone(self, order: int, point_set: int = L.WARPBLEND):
        L.check(L.lib().pdrd_mesh_clone(self._h, order, point_set), "pdrd_mesh_clone")
        self._cloned = True
        self._order = order
        return self

    def build(self) -> Mesh:
        if not self._cloned:
            raise RuntimeError("clone() the mesh to a solution order first")
        lib = L.lib()
        dim, nv, nc, nn, nb = (int(x) for x in self.info())
        coords = np.zeros((nv, dim))
        L.check(lib.pdrd_mesh_coords(self._h, L.ptr(coords)), "pdrd_mesh_coords")
        ncoords = np.zeros((nn, dim))
        L.check(lib.pdrd_mesh_node_coords(self._h, L.ptr(ncoords)), "pdrd_mesh_node_coords")
        shapes = np.zeros(nc, dtype=np.int32)
        L.check(lib.pdrd_mesh_cell_shapes(self._h, L.ptr(shapes, C.c_int32)), "pdrd_mesh_cell_shapes")
        size = C.c_int64(0)
        L.check(lib.pdrd_mesh_geo_conn(self._h, None, C.byref(size)), "pdrd_mesh_geo_conn")
        gconn = np.zeros(size.value, dtype=np.uint32)
        L.check(lib.pdrd_mesh_geo_conn(self._h, L.ptr(gconn, C.c_uint32), C.byref(size)), "pdrd_mesh_geo_conn")
        L.check(lib.pdrd_mesh_sol_conn(self._h, None, C.byref(size)), "pdrd_mesh_sol_conn")
        sconn = np.zeros(size.value, dtype=np.uint32)
        L.check(lib.pdrd_mesh_sol_conn(self._h, L.ptr(sconn, C.c_uint32), C.byref(size)), "pdrd_mesh_sol_conn")
        p = self._order
        ng_of = np.array([NB_VERT[int(s)] for s in range(3)])
        nd_of = np.array([_nb_nodes(int(s), p) for s in range(3)])
        g_off = np.concatenate([[0], np.cumsum(ng_of[shapes])])
        s_off = np.concatenate([[0], np.cumsum(nd_of[shapes])])
        blocks = []
        for s in sorted(set(int(x) for x in np.unique(shapes))):
            cells = np.nonzero(shapes == s)[0]
            ng, nd = int(ng_of[s]), int(nd_of[s])
            gi = g_off[cells][:, None] + np.arange(ng)[None, :]
            si = s_off[cells][:, None] + np.arange(nd)[None, :]
            blocks.append(CellBlock(shape=s, cells=cells, geo_conn=np.ascontiguousarray(gconn[gi]),
                                    sol_conn=np.ascontiguousarray(sconn[si])))
        bnd = {}
        name = C.create_string_buffer(64)
        for b in range(nb):
            L.check(lib.pdrd_mesh_boundary_name(self._h, b, name, 64), "pdrd_mesh_boundary_name")
            n = C.c_int64(0)
            L.check(lib.pdrd_mesh_boundary_nodes(self._h, name.value, None, C.byref(n)), "pdrd_mesh_boundary_nodes")
            ids = np.zeros(n.value, dtype=np.uint32)
            L.check(lib.pdrd_mesh_boundary_nodes(self._h, name.value, L.ptr(ids, C.c_uint32), C.byref(n)),
                    "pdrd_mesh_boundary_nodes")
            bnd[name.value.decode()] = ids
        return Mesh(dim=dim, coords=coords, cell_shape=shapes, order=p, nb_nodes=nn, node_coords=ncoords,
                    blocks=blocks, boundaries=bnd, generator=self)


def unit_quad_triag(N: int, order: int = 1, to_unit_square: bool = True) -> Mesh:
    """MeshCreator::make_unit_quad(N) with triangles; optionally mapped from
    [-1,1]^2 to [0,1]^2 by translating (1,1) and scaling by 0.5 as the
    reference's tests do (test/utest-solver/utest-Euler-continuous-rdm.cpp:462-471)."""
    m = MeshGen(0, [N]).clone(order).build()
    if to_unit_square:
        m.coords = (m.coords + 1.0) * 0.5
        m.node_coords = (m.node_coords + 1.0) * 0.5
    return m


def rectangle(nx: int, ny: int, extent=(0.0, 1.0, 0.0, 1.0), split: int = 2, quad_rows: int = 0,
              order: int = 1, jitter: float = 0.0, seed: int = 0, bump=None) -> Mesh:
    g = MeshGen(1, [nx, ny, split, quad_rows], list(extent))
    if jitter:
        hx = (extent[1] - extent[0]) / nx
        hy = (extent[3] - extent[2]) / ny
        g.jitter([hx, hy], jitter, seed)
    if bump is not None:
        g.bump(0, bump)
    return g.clone(order).build()


def box_tetra(nx: int, ny: int, nz: int, extent=(1.0, 1.0, 1.0), order: int = 1, jitter: float = 0.0,
              seed: int = 0, bump=None) -> Mesh:
    g = MeshGen(2, [nx, ny, nz], list(extent))
    if jitter:
        g.jitter([extent[0] / nx, extent[1] / ny, extent[2] / nz], jitter, seed)
    if bump is not None:
        g.bump(1, bump)
    return g.clone(order).build()
