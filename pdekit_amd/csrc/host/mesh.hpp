// Host-side mesh inputs of the cell loop: structured generators (the
// reference's MeshCreator::make_unit_quad and the configurations built on
// it), seeded vertex jitter, bump mappings, P1 -> Pk continuous DOF
// numbering and element colouring.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "fe_tables.hpp"

namespace pdrd {

// A mesh of P1/Q1 geometry cells, possibly of mixed shape. Cells keep one
// global order (the order the reference's Tria / DofMap iterates them in).
struct GeoMesh
{
  int dim = 2;
  std::vector<double> coords;         // [nv][dim]
  std::vector<int> cell_shape;        // [ncells] Shape
  std::vector<uint32_t> cell_offset;  // [ncells+1] into cell_vertices
  std::vector<uint32_t> cell_vertices;
  // boundary vertex sets by name (e.g. "Edge0".."Edge3", "xmin", ...)
  std::vector<std::string> bnd_names;
  std::vector<std::vector<uint32_t>> bnd_vertices;

  uint32_t nb_vertices() const { return static_cast<uint32_t>(coords.size() / dim); }
  uint32_t nb_cells() const { return static_cast<uint32_t>(cell_shape.size()); }
};

// MeshCreator::make_unit_quad (src/mesh/io/MeshCreator.hpp:436-600,
// MeshCreator.cpp:52-75): N x N points on [-1,1]^2, row-major ids, each quad
// (k, k+1, k+N+1, k+N) split into triangles (k, k+1, k+N+1) and
// (k+N+1, k+N, k). Boundaries Edge0 (y=-1), Edge1 (x=1), Edge2 (y=1),
// Edge3 (x=-1).
GeoMesh make_unit_quad_triag(uint32_t N);

// Rectangle [x0,x1] x [y0,y1] with nx x ny cells. split = 2: the
// make_unit_quad diagonal split; split = 4: criss-cross split with one
// centre vertex per cell (ids after the grid vertices); split = 1: Q1 quads.
// quad_rows: for split 2/4, the lowest quad_rows cell rows stay Q1 quads
// (mixed meshes). Boundaries "ymin", "xmax", "ymax", "xmin".
GeoMesh make_rectangle(uint32_t nx, uint32_t ny, double x0, double x1, double y0, double y1,
                       int split, uint32_t quad_rows);

// Box [0,lx] x [0,ly] x [0,lz] with nx x ny x nz hexahedra, each split into
// 6 tetrahedra around the main diagonal. Boundaries xmin..zmax.
GeoMesh make_box_tetra(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz);

// Move every vertex that lies on no boundary by U(-f*h_d, f*h_d) per axis d,
// drawn from std::mt19937_64(seed) in vertex order, axis by axis.
void jitter_interior(GeoMesh &m, const std::vector<double> &h, double f, uint64_t seed);

// 2D channel: y -> yb(x) + y * (H - yb(x)) / H with yb = amp*exp(-k (x-xc)^2).
void bump_map_2d(GeoMesh &m, double H, double amp, double k, double xc);
// 3D box: z -> zb + z (H - zb)/H with zb = amp*exp(-k((x-xc)^2+(y-yc)^2)).
void bump_map_3d(GeoMesh &m, double H, double amp, double k, double xc, double yc);

// Continuous Pk DOF numbering (DofMap::clone_continuous,
// src/mesh/containers/DofMap.hpp:2556-3073): P1 vertices keep their
// ascending order; every higher-order node shared by several cells is
// numbered once, in order of first appearance in a cell-order scan.
struct SolMesh
{
  int order = 1;
  uint32_t nb_nodes = 0;
  std::vector<uint32_t> cell_offset;
  std::vector<uint32_t> cell_nodes;
  std::vector<double> node_coords;  // [nb_nodes][dim]
  // For every node, the P1 vertices of the mesh entity that carries it
  // (1 for vertex nodes, 2 for edge nodes, 0 for cell-interior nodes).
  std::vector<uint32_t> host_count;
  std::vector<uint32_t> host_vertices;  // [nb_nodes][2]
};

SolMesh clone_continuous(const GeoMesh &m, int order, PointSet ps);

// Solution nodes lying on a named boundary: nodes whose carrying vertex or
// edge has all its P1 vertices in the boundary's vertex set.
std::vector<uint32_t> boundary_nodes(const GeoMesh &m, const SolMesh &s, const std::string &name);

// Greedy colouring of a subset of cells so that no two cells of one colour
// share a node. Returns the colour of each listed cell.
std::vector<uint32_t> greedy_colour(const std::vector<uint32_t> &cells, const SolMesh &s,
                                    uint32_t *nb_colours);

}  // namespace pdrd
