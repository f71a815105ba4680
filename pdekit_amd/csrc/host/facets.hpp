// Boundary facets for the weak boundary conditions of the explicit RD method
// (pdekit's WeakBC over mesh::BoundaryFacets, src/solver/rdm/bc/WeakBC.hpp:200-352):
// the facets of the cells that lie on the domain boundary and whose solution
// nodes all belong to a given node set, with their outward unit normals,
// measures and the facet quadrature tables.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace pdrd {

// One block of cells as the facet search needs it.
struct FacetBlockIn
{
  int shape = 0;  // 0 triangle, 1 quadrilateral, 2 tetrahedron
  int order = 1;  // solution order
  int ng = 0, nd = 0;
  uint32_t ne = 0;
  const uint32_t *geo = nullptr;  // [ne][ng]
  const uint32_t *sol = nullptr;  // [ne][nd]
};

struct FacetSet
{
  int nfn = 0;                   // solution nodes per facet
  int nfv = 0;                   // vertices per facet
  std::vector<uint32_t> nodes;   // [nf][nfn]: vertices first, then the P2 mid-edge nodes
  std::vector<uint32_t> verts;   // [nf][nfv]
  std::vector<double> normal;    // [nf][dim] outward unit normal
  std::vector<double> jac;       // [nf] |dx/dxi| of the facet map (straight facets)
  size_t size() const { return nfn ? nodes.size() / nfn : 0; }
};

// Facets on the domain boundary (held by one cell only) whose solution nodes
// all lie in `in_set` (indexed by solution node). Supported cells: triangles,
// quadrilaterals and tetrahedra of order 1 or 2; throws std::invalid_argument
// for others or for mixed facet kinds.
FacetSet boundary_facets(int dim, const double *coords, const std::vector<FacetBlockIn> &blocks,
                         const std::vector<char> &in_set);

// Facet rule of polynomial order `order`: Gauss-Legendre on [-1, 1] for edges
// (max(2, order / 2 + 1) points, QdPointSetLineGauss.cpp) or the triangle
// Gauss rule (reference triangle, weights sum to 2) for faces; V[q][n] the
// Lagrange values of the facet's nodes (line P1: -1, 1; line P2: -1, 1, 0;
// triangle P1: the reference vertices; triangle P2: vertices, then the mid
// nodes of edges (0,1), (1,2), (2,0)).
void facet_tables(int dim, int nfn, int order, std::vector<double> &w, std::vector<double> &V);

}  // namespace pdrd
