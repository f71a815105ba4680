// Host-side finite-element inputs of the residual-distribution cell loop:
// quadrature rules, reference-element nodes (equidistant / warp-blend) and
// Lagrange shape-function tables at arbitrary reference points.
//
// These are the tables that pdekit's FEValues / FunctionSpaceCells /
// FluxSpaceMetric build once at set-up time
// (src/interpolation/FEValues.cpp:157-279,
//  src/interpolation/FluxSpaceMetric.hpp:237-456,
//  src/mesh/shape_function/ShapeFunction.cpp:24-240).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pdrd {

enum class Shape : int { Triag = 0, Quad = 1, Tetra = 2 };
enum class PointSet : int { Equidist = 0, Warpblend = 1 };

int shape_dim(Shape s);
// Number of Lagrange nodes of a reference element of order p.
int shape_nb_nodes(Shape s, int p);
// Number of P1 (vertex) nodes of the shape.
int shape_nb_vertices(Shape s);

// Reference-element node coordinates, row-major [n][dim].
// Reference triangle (-1,-1),(1,-1),(-1,1); quad [-1,1]^2 counter-clockwise;
// tetrahedron (-1,-1,-1),(1,-1,-1),(-1,1,-1),(-1,-1,1).
// Node order: vertices, then edge nodes edge by edge, then interior nodes
// (src/mesh/std_region/WarpblendStdRegionTriag.cpp:87-395,
//  EquidistStdRegionQuad.cpp:264-, EquidistStdRegionTetra.cpp:522-).
std::vector<double> reference_nodes(Shape s, int p, PointSet ps);

// Gauss quadrature of the given polynomial order on the reference element.
// The point/weight values are the literal tables of
// src/mesh/point_set/QdPointSet{Triag,Quad,Tetra}Gauss.cpp.
void quadrature(Shape s, int order, std::vector<double> &xi, std::vector<double> &w);

// Lagrange basis of order p on the nodes of a reference element.
class LagrangeBasis
{
  public:
  LagrangeBasis(Shape s, int p, PointSet ps);
  LagrangeBasis(Shape s, int p, const std::vector<double> &nodes);

  int nb_nodes() const { return m_n; }
  int dim() const { return m_dim; }
  const std::vector<double> &nodes() const { return m_nodes; }

  // values [npts][n]; near-zero (<1e-14) entries are clipped and every row
  // is rescaled to sum to one, as ShapeFunction.cpp:149-210 does.
  void values(const std::vector<double> &pts, std::vector<double> &V) const;
  // derivatives [dim][npts][n] with respect to the reference coordinates
  // (no clipping, ShapeFunction.cpp:212-240).
  void derivatives(const std::vector<double> &pts, std::vector<double> &dV) const;

  private:
  void build();
  void monomials(const double *x, long double *val, long double *grad) const;

  Shape m_shape;
  int m_p;
  int m_dim;
  int m_n;
  std::vector<double> m_nodes;
  std::vector<int> m_exp;           // [n][dim] monomial exponents
  std::vector<long double> m_invV;  // [n modes][n nodes]
};

// One element type of the cell loop: solution order p on a geometry of order
// geo_p, P+1 flux element, quadrature of order qorder. All tables row-major.
struct ElementTables
{
  Shape shape;
  int geo_order, sol_order, quad_order;
  PointSet point_set;
  int dim, ng, nd, nf, nq;
  std::vector<double> qp;    // [nq][dim]
  std::vector<double> w;     // [nq]
  std::vector<double> Vg;    // [nq][ng]
  std::vector<double> dVg;   // [dim][nq][ng]
  std::vector<double> V;     // [nq][nd]
  std::vector<double> dV;    // [dim][nq][nd]
  std::vector<double> Vf;    // [nq][nf]
  std::vector<double> dVf;   // [dim][nq][nf]
  std::vector<double> Vgf;   // [nf][ng]
  std::vector<double> Vsf;   // [nf][nd]
  std::vector<double> sol_nodes;   // [nd][dim]
  std::vector<double> flux_nodes;  // [nf][dim]
};

ElementTables make_element_tables(Shape s, int geo_order, int sol_order, int quad_order,
                                  PointSet ps);

// An orthonormal basis, under the rule (qp[nq][dim], w[nq]), of the
// polynomials of total degree < p on the reference simplex: psi[nq][npm]
// (modified Gram-Schmidt, twice, of the monomials in long double). The L2
// projection onto them is what the smoothness indicator of
// ArtificialViscosity.hpp:371-540 takes from the modal expansion by leaving
// out its leading terms.
std::vector<double> lower_degree_basis(Shape s, int p, const std::vector<double> &qp, const std::vector<double> &w,
                                       int *npm);

// The smoothness indicator on quadrilaterals (ArtificialViscosity.hpp:
// 371-540 with ModalExpansionQuad): hm[nq][nd] maps nodal values to
// u - u_{p-1} at the points qp[nq][2], u_{p-1} being the modal expansion
// (normalised tensor Legendre modes P_m(x) P_n(y) / sqrt(g_m g_n), g_k =
// 2 / (2k + 1), mode id m (p + 1) + n) without the modes the reference flags
// as leading. ModalExpansionQuad::is_leading_expansion_term (ModalExpansionQuad.
// cpp:240-257) advances its flag index only for the 2p + 1 modes with m == p
// or n == p, so its flags fall on mode ids 0 .. 2p; those are the modes left
// out here, as there. nodes: the reference nodes of the solution [nd][2].
std::vector<double> quad_zeroed_modes(int p, const std::vector<double> &qp, const std::vector<double> &nodes);

// P1 / Q1 shape functions of the reference vertices at reference points
// pts[n][dim]: [n][nvert].
std::vector<double> vertex_functions(Shape s, const std::vector<double> &pts);

// Packed layout used by the C ABI: the arrays above concatenated in this
// order: qp, w, Vg, dVg, V, dV, Vf, dVf, Vgf, Vsf, sol_nodes, flux_nodes.
size_t packed_size(const ElementTables &t);
void pack(const ElementTables &t, double *out);

}  // namespace pdrd
