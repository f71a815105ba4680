// Gmsh 2.2 ASCII meshes and nodal data (src/mesh/io/gmsh/GmshReader.hpp,
// GmshReader.cpp:60-243, GmshWriter.hpp:145-330, PDEKitToGmsh.cpp).
//
// The cell loop runs on straight-sided P1/Q1 geometry, so a file's
// high-order cells (tri6/10, quad9/16, tet10/20) are read by their corner
// vertices after checking that every other node lies where the straight
// cell puts it (curved cells are refused); the solution nodes then come from
// clone_continuous, whose local node order is Gmsh's.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mesh.hpp"

namespace pdrd {

// Cells: the elements of the highest topological dimension, in file order;
// vertices: the file's nodes (ids 1..n -> 0..n-1), coordinates cut to the
// mesh dimension. Boundaries: one per physical group of dimension dim - 1,
// named by $PhysicalNames (or by the decimal tag when unnamed), holding the
// corner vertices of its elements. Throws std::invalid_argument on binary
// files, unknown element types, node ids outside 1..n or curved cells.
GeoMesh read_gmsh(const std::string &path);

// write_mesh_to_file(mesh, dofs, file): the solution nodes of `s` as $Nodes,
// every boundary's facets on the domain boundary (physical tag k + 1, named
// after the boundary) and the cells (physical group "interior") as
// $Elements, in the Gmsh element type of the solution order.
void write_gmsh(const std::string &path, const GeoMesh &g, const SolMesh &s);

// append_nodal_function_to_file: one $NodeData block per component, named
// "<field>_<k>" (or "<field>" for m = 1 when scalar is set), real tag time,
// integer tags (step, 1, n), values written with 17 significant digits.
void append_gmsh_node_data(const std::string &path, const std::string &field, int m, uint32_t n,
                           const double *values, double time, uint32_t step, bool scalar);

// read_nodal_function_from_file: every $NodeData block whose name (without
// its quotes) is names[f] fills column f of values[n][names.size()]. Returns
// the number of names found.
int read_gmsh_node_data(const std::string &path, const std::vector<std::string> &names, uint32_t n,
                        double *values);

}  // namespace pdrd
