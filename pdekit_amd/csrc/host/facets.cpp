#include "facets.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <stdexcept>

#include "fe_tables.hpp"

namespace pdrd {

namespace {

struct LocalFacet
{
  int nv;        // vertices
  int v[3];      // local vertex indices
  int nm;        // P2 mid-edge nodes on the facet
  int mid[3];    // their local indices: edge (v0,v1), then (v1,v2), (v2,v0)
};

// Local facets of the reference cells in the node order of reference_nodes:
// triangle edges (0,1), (1,2), (2,0) with P2 mid nodes 3, 4, 5;
// quadrilateral edges (0,1), (1,2), (2,3), (3,0) with Q2 mid nodes 4..7;
// tetrahedron faces opposite each vertex, whose P2 mid nodes are those of the
// edges (0,1), (1,2), (2,0), (3,0), (3,2), (3,1) (local 4..9).
std::vector<LocalFacet> local_facets(int shape, int order)
{
  if (order > 2) throw std::invalid_argument("weak boundary conditions: cells of order 1 or 2");
  if (shape == 0) {
    std::vector<LocalFacet> f = {{2, {0, 1, 0}, 0, {0}}, {2, {1, 2, 0}, 0, {0}}, {2, {2, 0, 0}, 0, {0}}};
    if (order == 2)
      for (int e = 0; e < 3; ++e) { f[e].nm = 1; f[e].mid[0] = 3 + e; }
    return f;
  }
  if (shape == 1) {
    std::vector<LocalFacet> f = {{2, {0, 1, 0}, 0, {0}}, {2, {1, 2, 0}, 0, {0}}, {2, {2, 3, 0}, 0, {0}},
                                 {2, {3, 0, 0}, 0, {0}}};
    if (order == 2)
      for (int e = 0; e < 4; ++e) { f[e].nm = 1; f[e].mid[0] = 4 + e; }
    return f;
  }
  std::vector<LocalFacet> f = {{3, {1, 2, 3}, 0, {0}}, {3, {0, 2, 3}, 0, {0}}, {3, {0, 1, 3}, 0, {0}},
                               {3, {0, 1, 2}, 0, {0}}};
  if (order == 2) {
    const int e[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 2}, {3, 1}};
    for (LocalFacet &lf : f) {
      lf.nm = 3;
      for (int k = 0; k < 3; ++k) {
        const int a = lf.v[k], b = lf.v[(k + 1) % 3];
        for (int j = 0; j < 6; ++j)
          if ((e[j][0] == a && e[j][1] == b) || (e[j][0] == b && e[j][1] == a)) lf.mid[k] = 4 + j;
      }
    }
  }
  return f;
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n in long double).
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < n; ++i) {
    long double t = std::cos(pi * (i + 0.75L) / (n + 0.5L));
    long double dp = 0.0L;
    for (int it = 0; it < 100; ++it) {
      long double p0 = 1.0L, p1 = t;
      for (int k = 2; k <= n; ++k) {
        const long double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      const long double pn = (n == 1) ? t : p1;
      const long double pm = (n == 1) ? 1.0L : p0;
      dp = n * (t * pn - pm) / (t * t - 1.0L);
      const long double dt = pn / dp;
      t -= dt;
      if (std::fabs(static_cast<double>(dt)) < 1e-30) break;
    }
    // ascending order
    x[n - 1 - i] = static_cast<double>(t);
    w[n - 1 - i] = static_cast<double>(2.0L / ((1.0L - t * t) * dp * dp));
  }
}

}  // namespace

FacetSet boundary_facets(int dim, const double *coords, const std::vector<FacetBlockIn> &blocks,
                         const std::vector<char> &in_set)
{
  // occurrences of every facet (sorted vertex tuple) over all cells
  std::map<std::array<uint32_t, 3>, int> count;
  auto key_of = [](const uint32_t *cv, const LocalFacet &lf) {
    std::array<uint32_t, 3> k = {UINT32_MAX, UINT32_MAX, UINT32_MAX};
    for (int i = 0; i < lf.nv; ++i) k[i] = cv[lf.v[i]];
    std::sort(k.begin(), k.begin() + lf.nv);
    return k;
  };
  for (const FacetBlockIn &b : blocks) {
    const std::vector<LocalFacet> lfs = local_facets(b.shape, b.order);
    for (uint32_t c = 0; c < b.ne; ++c)
      for (const LocalFacet &lf : lfs) count[key_of(b.geo + static_cast<size_t>(c) * b.ng, lf)]++;
  }
  FacetSet fs;
  for (const FacetBlockIn &b : blocks) {
    const std::vector<LocalFacet> lfs = local_facets(b.shape, b.order);
    for (uint32_t c = 0; c < b.ne; ++c) {
      const uint32_t *cv = b.geo + static_cast<size_t>(c) * b.ng;
      const uint32_t *cs = b.sol + static_cast<size_t>(c) * b.nd;
      for (const LocalFacet &lf : lfs) {
        if (count[key_of(cv, lf)] != 1) continue;
        const int nfn = lf.nv + lf.nm;
        uint32_t fn[6];
        for (int i = 0; i < lf.nv; ++i) fn[i] = cs[lf.v[i]];
        for (int i = 0; i < lf.nm; ++i) fn[lf.nv + i] = cs[lf.mid[i]];
        bool all = true;
        for (int i = 0; i < nfn; ++i) all = all && in_set[fn[i]];
        if (!all) continue;
        if (fs.nfn == 0) {
          fs.nfn = nfn;
          fs.nfv = lf.nv;
        } else if (fs.nfn != nfn || fs.nfv != lf.nv) {
          throw std::invalid_argument("weak boundary condition over facets of different kinds");
        }
        for (int i = 0; i < nfn; ++i) fs.nodes.push_back(fn[i]);
        double x[3][3] = {{0}};
        for (int i = 0; i < lf.nv; ++i) {
          fs.verts.push_back(cv[lf.v[i]]);
          for (int d = 0; d < dim; ++d) x[i][d] = coords[static_cast<size_t>(cv[lf.v[i]]) * dim + d];
        }
        // centroid of the cell, to orient the normal outwards
        double cen[3] = {0, 0, 0};
        for (int g = 0; g < b.ng; ++g)
          for (int d = 0; d < dim; ++d) cen[d] += coords[static_cast<size_t>(cv[g]) * dim + d] / b.ng;
        double nrm[3] = {0, 0, 0}, jac = 0.0;
        if (dim == 2) {
          const double tx = x[1][0] - x[0][0], ty = x[1][1] - x[0][1];
          const double len = std::sqrt(tx * tx + ty * ty);
          nrm[0] = ty / len;
          nrm[1] = -tx / len;
          jac = 0.5 * len;  // reference edge [-1, 1]
        } else {
          const double a[3] = {x[1][0] - x[0][0], x[1][1] - x[0][1], x[1][2] - x[0][2]};
          const double e[3] = {x[2][0] - x[0][0], x[2][1] - x[0][1], x[2][2] - x[0][2]};
          const double cr[3] = {a[1] * e[2] - a[2] * e[1], a[2] * e[0] - a[0] * e[2], a[0] * e[1] - a[1] * e[0]};
          const double len = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
          for (int d = 0; d < 3; ++d) nrm[d] = cr[d] / len;
          jac = 0.25 * len;  // reference triangle of area 2: |J| = area / 2
        }
        double dot = 0.0;
        for (int d = 0; d < dim; ++d) dot += nrm[d] * (x[0][d] - cen[d]);
        if (dot < 0)
          for (int d = 0; d < dim; ++d) nrm[d] = -nrm[d];
        for (int d = 0; d < dim; ++d) fs.normal.push_back(nrm[d]);
        fs.jac.push_back(jac);
      }
    }
  }
  return fs;
}

void facet_tables(int dim, int nfn, int order, std::vector<double> &w, std::vector<double> &V)
{
  std::vector<double> xi;
  if (dim == 2) {
    const int n = std::max(2, order / 2 + 1);
    gauss_legendre(n, xi, w);
    V.clear();
    for (int q = 0; q < n; ++q) {
      const double s = xi[q];
      if (nfn == 2) {
        V.push_back(0.5 * (1.0 - s));
        V.push_back(0.5 * (1.0 + s));
      } else {
        V.push_back(0.5 * s * (s - 1.0));
        V.push_back(0.5 * s * (s + 1.0));
        V.push_back(1.0 - s * s);
      }
    }
    return;
  }
  quadrature(Shape::Triag, order, xi, w);
  V.clear();
  for (size_t q = 0; q < w.size(); ++q) {
    const double a = xi[2 * q], b = xi[2 * q + 1];
    const double L0 = -0.5 * (a + b), L1 = 0.5 * (1.0 + a), L2 = 0.5 * (1.0 + b);
    if (nfn == 3) {
      V.push_back(L0);
      V.push_back(L1);
      V.push_back(L2);
    } else {
      // P2 face: vertices, then the mid nodes of (0,1), (1,2), (2,0)
      V.push_back(L0 * (2.0 * L0 - 1.0));
      V.push_back(L1 * (2.0 * L1 - 1.0));
      V.push_back(L2 * (2.0 * L2 - 1.0));
      V.push_back(4.0 * L0 * L1);
      V.push_back(4.0 * L1 * L2);
      V.push_back(4.0 * L2 * L0);
    }
  }
}

}  // namespace pdrd
