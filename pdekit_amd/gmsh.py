"""Gmsh 2.2 ASCII meshes and nodal data: the GmshReader / GmshWriter pair of
src/mesh/io/gmsh/GmshReader.hpp (read_mesh_from_file :118-176,
read_nodal_function_from_file :487-) and GmshWriter.hpp
(write_mesh_to_file :145-200, append_nodal_function_to_file :202-264), over
the library's native reader and writer.

Cells are the elements of the highest dimension; high-order cells must be
straight-sided and are read by their corners, the solution nodes coming from
the continuous P1 -> Pk numbering (whose local node order is Gmsh's).
Boundaries are the physical groups one dimension down, by name.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .mesh import Mesh, MeshGen


class GmshReader:
    def read_mesh_from_file(self, filename: str, order: int = 1, point_set: int = L.WARPBLEND) -> Mesh:
        """The mesh of `filename` with continuous solution nodes of `order`."""
        return MeshGen.from_gmsh(filename).clone(order, point_set).build()

    def read_nodal_function_from_file(self, mesh: Mesh, filename: str, field_names) -> np.ndarray:
        """[N][len(field_names)]: column f from the $NodeData block named
        field_names[f] (names without their quotes); raises if one is missing."""
        names = [str(n) for n in field_names]
        out = np.zeros((mesh.nb_nodes, len(names)))
        arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
        found = C.c_int(0)
        L.check(L.lib().pdrd_gmsh_read_node_data(str(filename).encode(), len(names), arr, mesh.nb_nodes,
                                                 L.ptr(out), C.byref(found)), "pdrd_gmsh_read_node_data")
        if found.value != len(names):
            raise KeyError(f"{filename}: {len(names) - found.value} of the fields {names} are missing")
        return out


class GmshWriter:
    def write_mesh_to_file(self, mesh: Mesh, filename: str):
        """Solution nodes, boundary facets and cells of `mesh` (built by the
        library's mesh generators or GmshReader)."""
        if mesh.generator is None:
            raise ValueError("the mesh carries no generator handle")
        mesh.generator.write_gmsh(filename)

    def append_nodal_function_to_file(self, mesh: Mesh, filename: str, values, field_name: str,
                                      time: float = 0.0, timestep: int = 0):
        """values [N][m] as m $NodeData blocks "<field>_<k>"; a 1-D array is a
        scalar field written as "<field>"."""
        v = np.asarray(values, dtype=np.float64)
        scalar = v.ndim == 1
        v = np.ascontiguousarray(v.reshape(mesh.nb_nodes, -1))
        L.check(L.lib().pdrd_gmsh_append_node_data(str(filename).encode(), field_name.encode(), v.shape[1],
                                                   mesh.nb_nodes, L.ptr(v), float(time), int(timestep),
                                                   int(scalar)), "pdrd_gmsh_append_node_data")
