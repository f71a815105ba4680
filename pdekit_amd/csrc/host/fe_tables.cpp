#include "fe_tables.hpp"

#include <array>
#include <cmath>
#include <stdexcept>

namespace pdrd {

namespace {

void fail(const std::string &msg) { throw std::invalid_argument(msg); }

// Gauss-Jordan inverse with partial pivoting, extended precision.
void invert_ld(std::vector<long double> &a, int n, std::vector<long double> &inv)
{
  inv.assign(static_cast<size_t>(n) * n, 0.0L);
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0L;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    long double best = std::fabs(a[c * n + c]);
    for (int r = c + 1; r < n; ++r) {
      const long double v = std::fabs(a[r * n + c]);
      if (v > best) { best = v; piv = r; }
    }
    if (best == 0.0L) fail("singular Vandermonde matrix");
    if (piv != c) {
      for (int k = 0; k < n; ++k) {
        std::swap(a[c * n + k], a[piv * n + k]);
        std::swap(inv[c * n + k], inv[piv * n + k]);
      }
    }
    const long double d = 1.0L / a[c * n + c];
    for (int k = 0; k < n; ++k) { a[c * n + k] *= d; inv[c * n + k] *= d; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const long double f = a[r * n + c];
      if (f == 0.0L) continue;
      for (int k = 0; k < n; ++k) {
        a[r * n + k] -= f * a[c * n + k];
        inv[r * n + k] -= f * inv[c * n + k];
      }
    }
  }
}

// Interior Gauss-Lobatto-Legendre points of order p (roots of P'_p), ascending.
std::vector<double> gll_interior(int p)
{
  std::vector<double> x;
  const int m = p - 1;
  for (int i = 0; i < m; ++i) {
    // Chebyshev-Gauss-Lobatto guess, then Newton on P'_p in long double.
    long double r = -std::cos(static_cast<long double>(M_PI) * (i + 1) / p);
    for (int it = 0; it < 100; ++it) {
      // Legendre P_k and derivatives by recurrence
      long double p0 = 1.0L, p1 = r, d0 = 0.0L, d1 = 1.0L, dd0 = 0.0L, dd1 = 0.0L;
      for (int k = 1; k < p; ++k) {
        const long double p2 = ((2 * k + 1) * r * p1 - k * p0) / (k + 1);
        const long double d2 = ((2 * k + 1) * (p1 + r * d1) - k * d0) / (k + 1);
        const long double dd2 = ((2 * k + 1) * (2 * d1 + r * dd1) - k * dd0) / (k + 1);
        p0 = p1; p1 = p2; d0 = d1; d1 = d2; dd0 = dd1; dd1 = dd2;
      }
      const long double step = d1 / dd1;
      r -= step;
      if (std::fabs(step) < 1e-19L) break;
    }
    x.push_back(static_cast<double>(r));
  }
  return x;
}

// Warp function of the Hesthaven-Warburton construction at points r
// (src/mesh/std_region/WarpblendStdRegionTriag.cpp:402-520): interpolate the
// displacement LGL - equidistant with the equidistant Lagrange basis, clip
// tiny basis values, and divide by (1 - r^2) away from the end points.
std::vector<double> warpfactor(int p, const std::vector<double> &r)
{
  const int np = p + 1;
  std::vector<double> lgl(np), xeq(np);
  lgl[0] = -1.0;
  lgl[1] = 1.0;
  const std::vector<double> inner = gll_interior(p);
  for (int i = 0; i < np - 2; ++i) lgl[i + 2] = inner[i];
  const double dx = 2. / (np - 1);
  xeq[0] = -1.0;
  xeq[1] = 1.0;
  for (int i = 1; i < np - 1; ++i) xeq[i + 1] = -1. + i * dx;

  std::vector<double> warp(r.size(), 0.0);
  for (size_t k = 0; k < r.size(); ++k) {
    double acc = 0.0;
    for (int j = 0; j < np; ++j) {
      long double l = 1.0L;
      for (int i = 0; i < np; ++i)
        if (i != j) l *= (static_cast<long double>(r[k]) - xeq[i]) / (xeq[j] - xeq[i]);
      double lv = static_cast<double>(l);
      if (std::fabs(lv) < 1e-14) lv = 0.0;
      acc += lv * (lgl[j] - xeq[j]);
    }
    if (std::fabs(r[k]) < (1. - 1.e-10)) acc /= (1.0 - r[k] * r[k]);
    warp[k] = acc;
  }
  return warp;
}

// Barycentric coordinates of the equidistant triangle lattice in the
// vertex/edge/peeled-interior order of the reference element.
void triag_lattice(int p, std::vector<double> &L0, std::vector<double> &L1, std::vector<double> &L2)
{
  const int n = (p + 1) * (p + 2) / 2;
  L0.assign(n, 0.0);
  L1.assign(n, 0.0);
  L2.assign(n, 0.0);
  std::vector<double> X(n), Y(n);
  X[0] = -1; Y[0] = -1;
  X[1] = 1;  Y[1] = -1;
  X[2] = -1; Y[2] = 1;
  const double dx = 2. / p, dy = 2. / p;
  int in = 3;
  for (int i = 1; i < p; ++i) { X[in] = -1. + i * dx; Y[in] = -1.0; ++in; }
  for (int i = 1; i < p; ++i) { X[in] = 1. - i * dx; Y[in] = -1. + i * dy; ++in; }
  for (int i = 1; i < p; ++i) { X[in] = -1.; Y[in] = 1. - i * dy; ++in; }
  int offset = 1;
  int cur = p - 3;
  while (cur >= 0) {
    X[in] = -1. + offset * dx; Y[in] = -1. + offset * dy; ++in;
    if (cur == 0) break;
    X[in] = 1. - (2 * offset) * dx; Y[in] = -1. + offset * dy; ++in;
    X[in] = -1. + offset * dx; Y[in] = 1. - (2 * offset) * dy; ++in;
    if (cur == 1) break;
    for (int i = 1; i < cur; ++i) { X[in] = -1. + (offset + i) * dx; Y[in] = -1. + offset * dy; ++in; }
    for (int i = 1; i < cur; ++i) { X[in] = 1. - (2 * offset + i) * dx; Y[in] = -1. + (offset + i) * dy; ++in; }
    for (int i = 1; i < cur; ++i) { X[in] = -1. + offset * dx; Y[in] = 1. - (2 * offset + i) * dy; ++in; }
    cur -= 3;
    ++offset;
  }
  if (in != n) fail("triangle lattice size mismatch");
  for (int i = 0; i < n; ++i) {
    if (i < 3) {
      L0[i] = (i == 0); L1[i] = (i == 1); L2[i] = (i == 2);
      continue;
    }
    L0[i] = 0.25 * ((-2.) * (X[i] + 1.) - 2. * (Y[i] - 1.));
    L1[i] = 0.25 * (2. * (X[i] + 1.));
    L2[i] = 1. - L0[i] - L1[i];
  }
}

std::vector<double> triag_nodes(int p, PointSet ps)
{
  const int n = (p + 1) * (p + 2) / 2;
  std::vector<double> c(2 * n);
  if (p == 0) { c[0] = -1. / 3.; c[1] = -1. / 3.; return c; }
  std::vector<double> L0, L1, L2;
  triag_lattice(p, L0, L1, L2);
  if (ps == PointSet::Equidist || p <= 2) {
    for (int i = 0; i < n; ++i) {
      c[2 * i] = -L0[i] + L1[i] - L2[i];
      c[2 * i + 1] = -L0[i] - L1[i] + L2[i];
    }
    // exact literals for vertices/mid-edges as the reference writes them
    return c;
  }
  // Warp & blend (Hesthaven-Warburton Nodes2D), alpha from the optimised table.
  static const double alpopt[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                                    1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
  const double alpha = (p < 16) ? alpopt[p - 1] : 5. / 3.;
  std::vector<double> x0(n), x1(n), b0(n), b1(n), b2(n), d(n);
  for (int i = 0; i < n; ++i) {
    x0[i] = -L0[i] + L1[i];
    x1[i] = (-L0[i] - L1[i] + 2. * L2[i]) / std::sqrt(3.0);
    b0[i] = 4. * L0[i] * L1[i];
    b1[i] = 4. * L2[i] * L1[i];
    b2[i] = 4. * L2[i] * L0[i];
  }
  for (int i = 0; i < n; ++i) d[i] = L1[i] - L0[i];
  std::vector<double> w0 = warpfactor(p, d);
  for (int i = 0; i < n; ++i) d[i] = L2[i] - L1[i];
  std::vector<double> w1 = warpfactor(p, d);
  for (int i = 0; i < n; ++i) d[i] = L0[i] - L2[i];
  std::vector<double> w2 = warpfactor(p, d);
  const double PI = 3.14159265358979323846;
  for (int i = 0; i < n; ++i) {
    const double ww0 = b0[i] * w0[i] * (1. + (alpha * L2[i]) * (alpha * L2[i]));
    const double ww1 = b1[i] * w1[i] * (1. + (alpha * L0[i]) * (alpha * L0[i]));
    const double ww2 = b2[i] * w2[i] * (1. + (alpha * L1[i]) * (alpha * L1[i]));
    x0[i] = x0[i] + 1. * ww0 + std::cos(2. * PI / 3.) * ww1 + std::cos(4. * PI / 3.) * ww2;
    x1[i] = x1[i] + 0. * ww0 + std::sin(2. * PI / 3.) * ww1 + std::sin(4. * PI / 3.) * ww2;
  }
  for (int i = 0; i < n; ++i) {
    const double l2 = (std::sqrt(3.0) * x1[i] + 1.0) / 3.0;
    const double l0 = (-3.0 * x0[i] - std::sqrt(3.0) * x1[i] + 2.0) / 6.0;
    const double l1 = (3.0 * x0[i] - std::sqrt(3.0) * x1[i] + 2.0) / 6.0;
    c[2 * i] = -l0 + l1 - l2;
    c[2 * i + 1] = -l0 - l1 + l2;
  }
  // vertices stay exact
  c[0] = -1; c[1] = -1; c[2] = 1; c[3] = -1; c[4] = -1; c[5] = 1;
  return c;
}

std::vector<double> quad_nodes(int p, PointSet ps)
{
  const int n = (p + 1) * (p + 1);
  std::vector<double> c(2 * n);
  if (p == 0) { c[0] = 0; c[1] = 0; return c; }
  // 1D positions: equidistant, or for warp-blend from Q3 on the
  // Gauss-Lobatto-Legendre points (WarpblendStdRegionQuad.cpp:79-110)
  std::vector<double> pt(p + 1);
  for (int i = 0; i <= p; ++i) pt[i] = -1. + i * (2. / p);
  if (ps == PointSet::Warpblend && p >= 3) {
    const std::vector<double> g = gll_interior(p);
    for (int i = 1; i < p; ++i) pt[i] = (std::fabs(g[i - 1]) < 1e-16) ? 0.0 : g[i - 1];
  }
  int in = 0;
  auto put = [&](int i, int j) { c[2 * in] = pt[i]; c[2 * in + 1] = pt[j]; ++in; };
  // peeled squares from outside in: corners, then the nodes of each edge
  int lo = 0, hi = p, cur = p;
  while (cur >= 0) {
    if (cur == 0) { put(lo, lo); break; }
    put(lo, lo); put(hi, lo); put(hi, hi); put(lo, hi);
    for (int i = 1; i < cur; ++i) put(lo + i, lo);
    for (int i = 1; i < cur; ++i) put(hi, lo + i);
    for (int i = 1; i < cur; ++i) put(hi - i, hi);
    for (int i = 1; i < cur; ++i) put(lo, hi - i);
    ++lo; --hi; cur -= 2;
  }
  if (in != n) fail("quad lattice size mismatch");
  return c;
}

std::vector<double> tetra_nodes(int p, PointSet ps)
{
  if (p > 3) fail("tetrahedral nodes above P3 are not built");
  const int n = (p + 1) * (p + 2) * (p + 3) / 6;
  std::vector<double> c(3 * n);
  if (p == 0) { c[0] = c[1] = c[2] = -0.5; return c; }
  const double v[4][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
  int in = 0;
  for (int k = 0; k < 4; ++k) {
    for (int d = 0; d < 3; ++d) c[3 * in + d] = v[k][d];
    ++in;
  }
  // edge nodes from each edge's first vertex to its second
  // (EquidistStdRegionTetra.cpp:589-646): fractions k / p, or for warp-blend
  // the Gauss-Lobatto-Legendre points, where the warp & blend construction
  // (WarpblendStdRegionTetra.cpp:88-386) puts the nodes of every edge
  std::vector<double> frac(p > 1 ? p - 1 : 0);
  for (int k = 1; k < p; ++k) frac[k - 1] = static_cast<double>(k) / p;
  if (ps == PointSet::Warpblend && p >= 3) {
    const std::vector<double> g = gll_interior(p);
    for (int k = 1; k < p; ++k) frac[k - 1] = 0.5 * (1.0 + g[k - 1]);
  }
  const int e[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 2}, {3, 1}};
  for (int k = 0; k < 6; ++k)
    for (int i = 0; i < p - 1; ++i) {
      for (int d = 0; d < 3; ++d) {
        const double a = v[e[k][0]][d], b = v[e[k][1]][d];
        c[3 * in + d] = (p == 2) ? 0.5 * (a + b) : a + frac[i] * (b - a);
      }
      ++in;
    }
  if (p == 3) {
    // one node per face at its centroid, faces (0,2,1), (0,1,3), (0,3,2),
    // (3,1,2) (EquidistStdRegionTetra.cpp:648-1020); the warp & blend
    // construction leaves face centroids in place at P3 (its face warp is
    // symmetric under the face's rotations and the other faces' blends
    // vanish there)
    const int f[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {3, 1, 2}};
    for (int k = 0; k < 4; ++k) {
      for (int d = 0; d < 3; ++d) c[3 * in + d] = (v[f[k][0]][d] + v[f[k][1]][d] + v[f[k][2]][d]) / 3.0;
      ++in;
    }
  }
  if (in != n) fail("tetra lattice size mismatch");
  return c;
}

}  // namespace

int shape_dim(Shape s) { return s == Shape::Tetra ? 3 : 2; }

int shape_nb_nodes(Shape s, int p)
{
  switch (s) {
    case Shape::Triag: return (p + 1) * (p + 2) / 2;
    case Shape::Quad: return (p + 1) * (p + 1);
    case Shape::Tetra: return (p + 1) * (p + 2) * (p + 3) / 6;
  }
  return 0;
}

int shape_nb_vertices(Shape s)
{
  switch (s) {
    case Shape::Triag: return 3;
    case Shape::Quad: return 4;
    case Shape::Tetra: return 4;
  }
  return 0;
}

std::vector<double> reference_nodes(Shape s, int p, PointSet ps)
{
  switch (s) {
    case Shape::Triag: return triag_nodes(p, ps);
    case Shape::Quad: return quad_nodes(p, ps);
    case Shape::Tetra: return tetra_nodes(p, ps);
  }
  fail("unknown shape");
  return {};
}

// ---------------------------------------------------------------------------
// Quadrature tables (values as listed in QdPointSet*Gauss.cpp)
// ---------------------------------------------------------------------------

void quadrature(Shape s, int order, std::vector<double> &xi, std::vector<double> &w)
{
  xi.clear();
  w.clear();
  auto add2 = [&](double a, double b, double wt) { xi.push_back(a); xi.push_back(b); w.push_back(wt); };
  auto add3 = [&](double a, double b, double c, double wt) {
    xi.push_back(a); xi.push_back(b); xi.push_back(c); w.push_back(wt);
  };
  if (s == Shape::Triag) {
    switch (order) {
      case 1:
        add2(-1. / 3., -1. / 3., 2.0);
        return;
      case 2:
        add2(-2. / 3., -2. / 3., 2. / 3.);
        add2(-2. / 3., 1. / 3., 2. / 3.);
        add2(1. / 3., -2. / 3., 2. / 3.);
        return;
      case 3:
        add2(-1. / 3., -1. / 3., -27. / 24.);
        add2(-0.6, -0.6, 25. / 24.);
        add2(0.2, -0.6, 25. / 24.);
        add2(-0.6, 0.2, 25. / 24.);
        return;
      case 4: {
        const double a = -0.108103018168070, b = -0.783793963663860;
        const double c = -0.816847572980458, d = 0.633695145960918;
        const double wa = 0.446763179356022, wc = 0.219903487310644;
        add2(a, a, wa); add2(a, b, wa); add2(b, a, wa);
        add2(c, c, wc); add2(c, d, wc); add2(d, c, wc);
        return;
      }
      case 5: {
        const double a = -0.059715871789770, b = -0.880568256420460;
        const double c = -0.797426985353088, d = 0.594853970706174;
        const double wa = 0.264788305577012, wc = 0.251878361089654;
        add2(-0.333333333333333, -0.333333333333333, 0.450000000000000);
        add2(a, a, wa); add2(a, b, wa); add2(b, a, wa);
        add2(c, c, wc); add2(c, d, wc); add2(d, c, wc);
        return;
      }
      case 6: {
        const double a = -0.501426509658180, b = 0.002853019316358;
        const double c = -0.873821971016996, d = 0.747643942033992;
        const double e = -0.379295097932432, f = 0.273004998242798, g = -0.893709900310366;
        const double wa = 0.233572551452758, wc = 0.101689812740414, we = 0.165702151236748;
        add2(a, a, wa); add2(a, b, wa); add2(b, a, wa);
        add2(c, c, wc); add2(c, d, wc); add2(d, c, wc);
        add2(e, f, we); add2(f, g, we); add2(g, e, we);
        add2(e, g, we); add2(f, e, we); add2(g, f, we);
        return;
      }
      default: fail("triangle quadrature order not tabulated");
    }
  }
  if (s == Shape::Quad) {
    switch (order) {
      case 1:
        add2(0.0, 0.0, 4.0);
        return;
      case 2:
      case 3: {
        const double a = 0.577350269189626;
        add2(a, a, 1.0); add2(a, -a, 1.0); add2(-a, a, 1.0); add2(-a, -a, 1.0);
        return;
      }
      case 4: {
        const double a = 0.683130051063973, b = 0.881917103688197;
        const double wa = 0.816326530612245, wb = 0.183673469387755;
        add2(a, 0.0, wa); add2(-a, 0.0, wa); add2(0.0, a, wa); add2(0.0, -a, wa);
        add2(b, b, wb); add2(b, -b, wb); add2(-b, b, wb); add2(-b, -b, wb);
        return;
      }
      default: fail("quadrilateral quadrature order not tabulated");
    }
  }
  if (s == Shape::Tetra) {
    switch (order) {
      case 1:
        add3(-0.5, -0.5, -0.5, 4. / 3.);
        return;
      case 2: {
        const double a = -0.7236067977, b = 0.1708203932;
        add3(a, a, a, 1. / 3.); add3(b, a, a, 1. / 3.); add3(a, b, a, 1. / 3.); add3(a, a, b, 1. / 3.);
        return;
      }
      case 3:
        add3(-0.5, -0.5, -0.5, -16. / 15.);
        add3(-2. / 3., -2. / 3., -2. / 3., 0.6);
        add3(-2. / 3., -2. / 3., 0.0, 0.6);
        add3(-2. / 3., 0.0, -2. / 3., 0.6);
        add3(0.0, -2. / 3., -2. / 3., 0.6);
        return;
      case 4: {
        // the 11-point rule with a negative centre weight, at the
        // reference's 10 digits (QdPointSetTetraGauss.cpp:268-331)
        const double a = -0.8571428571, b = 0.5714285714, c = -0.2011928476, d = -0.7988071523;
        const double w1 = 0.0609777777, w2 = 0.1991111111;
        add3(-0.5, -0.5, -0.5, -0.1052444444);
        add3(a, a, a, w1); add3(a, a, b, w1); add3(a, b, a, w1); add3(b, a, a, w1);
        add3(c, c, d, w2); add3(c, d, c, w2); add3(d, c, c, w2);
        add3(c, d, d, w2); add3(d, c, d, w2); add3(d, d, c, w2);
        return;
      }
      default: fail("tetrahedral quadrature order not tabulated");
    }
  }
  fail("unknown shape");
}

// ---------------------------------------------------------------------------
// Lagrange basis
// ---------------------------------------------------------------------------

LagrangeBasis::LagrangeBasis(Shape s, int p, PointSet ps)
    : m_shape(s), m_p(p), m_dim(shape_dim(s)), m_n(shape_nb_nodes(s, p)), m_nodes(reference_nodes(s, p, ps))
{
  build();
}

LagrangeBasis::LagrangeBasis(Shape s, int p, const std::vector<double> &nodes)
    : m_shape(s), m_p(p), m_dim(shape_dim(s)), m_n(shape_nb_nodes(s, p)), m_nodes(nodes)
{
  if (static_cast<int>(nodes.size()) != m_n * m_dim) fail("node array size mismatch");
  build();
}

void LagrangeBasis::build()
{
  m_exp.clear();
  if (m_shape == Shape::Triag) {
    for (int t = 0; t <= m_p; ++t)
      for (int j = 0; j <= t; ++j) { m_exp.push_back(t - j); m_exp.push_back(j); }
  } else if (m_shape == Shape::Quad) {
    for (int j = 0; j <= m_p; ++j)
      for (int i = 0; i <= m_p; ++i) { m_exp.push_back(i); m_exp.push_back(j); }
  } else {
    for (int t = 0; t <= m_p; ++t)
      for (int k = 0; k <= t; ++k)
        for (int j = 0; j <= t - k; ++j) {
          m_exp.push_back(t - k - j); m_exp.push_back(j); m_exp.push_back(k);
        }
  }
  if (static_cast<int>(m_exp.size()) != m_n * m_dim) fail("monomial count mismatch");
  std::vector<long double> V(static_cast<size_t>(m_n) * m_n), grad(static_cast<size_t>(m_dim) * m_n);
  for (int r = 0; r < m_n; ++r) monomials(&m_nodes[r * m_dim], &V[r * m_n], grad.data());
  invert_ld(V, m_n, m_invV);
}

// Monomials in shifted coordinates x' = (xi + 1) / 2 and their gradients with
// respect to xi ([dim][n]).
void LagrangeBasis::monomials(const double *x, long double *val, long double *grad) const
{
  long double xs[3] = {0, 0, 0};
  for (int d = 0; d < m_dim; ++d) xs[d] = (static_cast<long double>(x[d]) + 1.0L) * 0.5L;
  auto ipow = [](long double b, int e) {
    long double r = 1.0L;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
  };
  for (int m = 0; m < m_n; ++m) {
    long double v = 1.0L;
    for (int d = 0; d < m_dim; ++d) v *= ipow(xs[d], m_exp[m * m_dim + d]);
    val[m] = v;
    for (int d = 0; d < m_dim; ++d) {
      const int e = m_exp[m * m_dim + d];
      long double g = (e == 0) ? 0.0L : e * ipow(xs[d], e - 1) * 0.5L;
      for (int o = 0; o < m_dim; ++o)
        if (o != d) g *= ipow(xs[o], m_exp[m * m_dim + o]);
      grad[d * m_n + m] = g;
    }
  }
}

void LagrangeBasis::values(const std::vector<double> &pts, std::vector<double> &Vout) const
{
  const int np = static_cast<int>(pts.size()) / m_dim;
  Vout.assign(static_cast<size_t>(np) * m_n, 0.0);
  std::vector<long double> mono(m_n), grad(static_cast<size_t>(m_dim) * m_n);
  for (int q = 0; q < np; ++q) {
    monomials(&pts[q * m_dim], mono.data(), grad.data());
    double row_sum = 0.0;
    for (int j = 0; j < m_n; ++j) {
      long double s = 0.0L;
      for (int m = 0; m < m_n; ++m) s += mono[m] * m_invV[m * m_n + j];
      double v = static_cast<double>(s);
      if (std::fabs(v) < 1.e-14) v = 0.0;
      Vout[q * m_n + j] = v;
      row_sum += v;
    }
    if (row_sum != 1.0) {
      const double factor = 1.0 / row_sum;
      for (int j = 0; j < m_n; ++j) Vout[q * m_n + j] *= factor;
    }
  }
}

void LagrangeBasis::derivatives(const std::vector<double> &pts, std::vector<double> &dV) const
{
  const int np = static_cast<int>(pts.size()) / m_dim;
  dV.assign(static_cast<size_t>(m_dim) * np * m_n, 0.0);
  std::vector<long double> mono(m_n), grad(static_cast<size_t>(m_dim) * m_n);
  for (int q = 0; q < np; ++q) {
    monomials(&pts[q * m_dim], mono.data(), grad.data());
    for (int d = 0; d < m_dim; ++d)
      for (int j = 0; j < m_n; ++j) {
        long double s = 0.0L;
        for (int m = 0; m < m_n; ++m) s += grad[d * m_n + m] * m_invV[m * m_n + j];
        dV[(static_cast<size_t>(d) * np + q) * m_n + j] = static_cast<double>(s);
      }
  }
}

// ---------------------------------------------------------------------------

ElementTables make_element_tables(Shape s, int geo_order, int sol_order, int quad_order, PointSet ps)
{
  ElementTables t;
  t.shape = s;
  t.geo_order = geo_order;
  t.sol_order = sol_order;
  t.quad_order = quad_order;
  t.point_set = ps;
  t.dim = shape_dim(s);
  quadrature(s, quad_order, t.qp, t.w);
  t.nq = static_cast<int>(t.w.size());

  // Geometry: Lagrange of the geometry order (P1/Q1 vertices).
  LagrangeBasis geo(s, geo_order, PointSet::Equidist);
  LagrangeBasis sol(s, sol_order, ps);
  LagrangeBasis flux(s, sol_order + 1, ps);
  t.ng = geo.nb_nodes();
  t.nd = sol.nb_nodes();
  t.nf = flux.nb_nodes();
  geo.values(t.qp, t.Vg);
  geo.derivatives(t.qp, t.dVg);
  sol.values(t.qp, t.V);
  sol.derivatives(t.qp, t.dV);
  flux.values(t.qp, t.Vf);
  flux.derivatives(t.qp, t.dVf);
  geo.values(flux.nodes(), t.Vgf);
  sol.values(flux.nodes(), t.Vsf);
  t.sol_nodes = sol.nodes();
  t.flux_nodes = flux.nodes();
  return t;
}

size_t packed_size(const ElementTables &t)
{
  return t.qp.size() + t.w.size() + t.Vg.size() + t.dVg.size() + t.V.size() + t.dV.size() +
         t.Vf.size() + t.dVf.size() + t.Vgf.size() + t.Vsf.size() + t.sol_nodes.size() +
         t.flux_nodes.size();
}

void pack(const ElementTables &t, double *out)
{
  const std::vector<double> *parts[] = {&t.qp, &t.w,   &t.Vg,  &t.dVg,       &t.V,
                                        &t.dV, &t.Vf,  &t.dVf, &t.Vgf,       &t.Vsf,
                                        &t.sol_nodes, &t.flux_nodes};
  size_t off = 0;
  for (const auto *p : parts) {
    for (double v : *p) out[off++] = v;
  }
}

std::vector<double> quad_zeroed_modes(int p, const std::vector<double> &qp, const std::vector<double> &nodes)
{
  // normalised tensor Legendre modes, mode m (p + 1) + n = P_m(x) P_n(y) / sqrt(g_m g_n)
  const int nm = (p + 1) * (p + 1);
  auto modes = [&](double x, double y, long double *out) {
    long double Px[8], Py[8];
    Px[0] = Py[0] = 1.0L;
    Px[1] = x;
    Py[1] = y;
    for (int k = 1; k < p; ++k) {
      Px[k + 1] = ((2 * k + 1) * x * Px[k] - k * Px[k - 1]) / (k + 1);
      Py[k + 1] = ((2 * k + 1) * y * Py[k] - k * Py[k - 1]) / (k + 1);
    }
    for (int m = 0; m <= p; ++m)
      for (int n = 0; n <= p; ++n)
        out[m * (p + 1) + n] = Px[m] * Py[n] * std::sqrt((2.0L * m + 1) * (2.0L * n + 1)) / 2.0L;
  };
  if (p < 1 || p > 6) fail("quadrilateral modal basis of order 1-6");
  const int nd = static_cast<int>(nodes.size() / 2);
  if (nd != nm) fail("quadrilateral nodes do not match the modal basis");
  // modal coefficients of nodal values: the inverse of the modes at the nodes
  std::vector<long double> Vn(static_cast<size_t>(nd) * nm), inv;
  for (int i = 0; i < nd; ++i) modes(nodes[2 * i], nodes[2 * i + 1], &Vn[static_cast<size_t>(i) * nm]);
  invert_ld(Vn, nd, inv);
  const int nq = static_cast<int>(qp.size() / 2);
  std::vector<double> hm(static_cast<size_t>(nq) * nd);
  std::vector<long double> vq(nm);
  for (int q = 0; q < nq; ++q) {
    modes(qp[2 * q], qp[2 * q + 1], vq.data());
    for (int n = 0; n < nd; ++n) {
      long double t = 0.0L;
      for (int k = 0; k < 2 * p + 1; ++k) t += vq[k] * inv[static_cast<size_t>(k) * nd + n];
      hm[static_cast<size_t>(q) * nd + n] = static_cast<double>(t);
    }
  }
  return hm;
}

std::vector<double> lower_degree_basis(Shape s, int p, const std::vector<double> &qp, const std::vector<double> &w,
                                       int *npm)
{
  if (s == Shape::Quad) fail("the smoothness indicator is built for simplices");
  const int dim = shape_dim(s);
  const int nq = static_cast<int>(w.size());
  std::vector<std::array<int, 3>> exps;
  for (int i = 0; i < p; ++i)
    for (int j = 0; j + i < p; ++j)
      if (dim == 2)
        exps.push_back({i, j, 0});
      else
        for (int k = 0; i + j + k < p; ++k) exps.push_back({i, j, k});
  const int n = static_cast<int>(exps.size());
  std::vector<long double> B(static_cast<size_t>(nq) * n);
  for (int q = 0; q < nq; ++q)
    for (int k = 0; k < n; ++k) {
      long double v = 1.0L;
      for (int d = 0; d < dim; ++d)
        for (int e = 0; e < exps[k][d]; ++e) v *= qp[static_cast<size_t>(q) * dim + d];
      B[static_cast<size_t>(q) * n + k] = v;
    }
  auto dot = [&](int a, int b) {
    long double t = 0.0L;
    for (int q = 0; q < nq; ++q) t += w[q] * B[static_cast<size_t>(q) * n + a] * B[static_cast<size_t>(q) * n + b];
    return t;
  };
  for (int k = 0; k < n; ++k) {
    for (int pass = 0; pass < 2; ++pass)
      for (int j = 0; j < k; ++j) {
        const long double c = dot(k, j);
        for (int q = 0; q < nq; ++q) B[static_cast<size_t>(q) * n + k] -= c * B[static_cast<size_t>(q) * n + j];
      }
    const long double nrm = std::sqrt(dot(k, k));
    if (!(nrm > 0.0L)) fail("degenerate lower-degree basis");
    for (int q = 0; q < nq; ++q) B[static_cast<size_t>(q) * n + k] /= nrm;
  }
  *npm = n;
  return std::vector<double>(B.begin(), B.end());
}

std::vector<double> vertex_functions(Shape s, const std::vector<double> &pts)
{
  const int dim = shape_dim(s);
  const size_t np = pts.size() / dim;
  std::vector<double> out;
  for (size_t i = 0; i < np; ++i) {
    const double x = pts[i * dim], y = pts[i * dim + 1];
    if (s == Shape::Triag) {
      out.insert(out.end(), {-(x + y) / 2, (1 + x) / 2, (1 + y) / 2});
    } else if (s == Shape::Quad) {
      out.insert(out.end(), {(1 - x) * (1 - y) / 4, (1 + x) * (1 - y) / 4, (1 + x) * (1 + y) / 4, (1 - x) * (1 + y) / 4});
    } else {
      const double z = pts[i * dim + 2];
      out.insert(out.end(), {-(1 + x + y + z) / 2, (1 + x) / 2, (1 + y) / 2, (1 + z) / 2});
    }
  }
  return out;
}

}  // namespace pdrd
