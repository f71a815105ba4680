#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <random>
#include <stdexcept>
#include <unordered_map>

namespace pdrd {

namespace {

void fail(const std::string &msg) { throw std::invalid_argument(msg); }

void add_cell(GeoMesh &m, Shape s, std::initializer_list<uint32_t> v)
{
  m.cell_shape.push_back(static_cast<int>(s));
  for (uint32_t x : v) m.cell_vertices.push_back(x);
  m.cell_offset.push_back(static_cast<uint32_t>(m.cell_vertices.size()));
}

void add_boundary(GeoMesh &m, const std::string &name, std::vector<uint32_t> verts)
{
  std::sort(verts.begin(), verts.end());
  verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
  m.bnd_names.push_back(name);
  m.bnd_vertices.push_back(std::move(verts));
}

// Local topology of a reference element of order p: the vertex pair of every
// edge and the local indices of the edge's interior nodes, ordered from the
// edge's first vertex to its second.
struct EdgeTable
{
  std::vector<std::pair<int, int>> edges;
  int first_edge_node;  // local index of the first edge node
};

EdgeTable edge_table(Shape s, int p)
{
  EdgeTable t;
  switch (s) {
    case Shape::Triag: t.edges = {{0, 1}, {1, 2}, {2, 0}}; t.first_edge_node = 3; break;
    case Shape::Quad: t.edges = {{0, 1}, {1, 2}, {2, 3}, {3, 0}}; t.first_edge_node = 4; break;
    case Shape::Tetra:
      t.edges = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 2}, {3, 1}};
      t.first_edge_node = 4;
      if (p > 2) fail("tetrahedra above P2 are not numbered");
      break;
  }
  return t;
}

}  // namespace

GeoMesh make_unit_quad_triag(uint32_t N)
{
  if (N < 2) fail("make_unit_quad_triag needs N >= 2");
  GeoMesh m;
  m.dim = 2;
  m.coords.resize(2 * static_cast<size_t>(N) * N);
  const double delta = 2.0 / (N - 1);
  size_t pt = 0;
  for (uint32_t j = 0; j < N; ++j)
    for (uint32_t i = 0; i < N; ++i) {
      m.coords[2 * pt] = -1. + i * delta;
      m.coords[2 * pt + 1] = -1. + j * delta;
      ++pt;
    }
  m.cell_offset.push_back(0);
  uint32_t k = 0;
  for (uint32_t j = 0; j < N - 1; ++j) {
    for (uint32_t i = 0; i < N - 1; ++i) {
      add_cell(m, Shape::Triag, {k, k + 1, k + N + 1});
      add_cell(m, Shape::Triag, {k + N + 1, k + N, k});
      ++k;
      if (i == N - 2) ++k;
    }
  }
  std::vector<uint32_t> e0, e1, e2, e3;
  for (uint32_t i = 0; i < N; ++i) {
    e0.push_back(i);
    e1.push_back(i * N + N - 1);
    e2.push_back((N - 1) * N + i);
    e3.push_back(i * N);
  }
  add_boundary(m, "Edge0", e0);
  add_boundary(m, "Edge1", e1);
  add_boundary(m, "Edge2", e2);
  add_boundary(m, "Edge3", e3);
  return m;
}

GeoMesh make_rectangle(uint32_t nx, uint32_t ny, double x0, double x1, double y0, double y1,
                       int split, uint32_t quad_rows)
{
  if (nx < 1 || ny < 1) fail("make_rectangle needs nx, ny >= 1");
  if (split != 1 && split != 2 && split != 4) fail("split must be 1, 2 or 4");
  GeoMesh m;
  m.dim = 2;
  const uint32_t Nx = nx + 1, Ny = ny + 1;
  const double hx = (x1 - x0) / nx, hy = (y1 - y0) / ny;
  for (uint32_t j = 0; j < Ny; ++j)
    for (uint32_t i = 0; i < Nx; ++i) {
      m.coords.push_back(x0 + i * hx);
      m.coords.push_back(y0 + j * hy);
    }
  const uint32_t nb_grid = Nx * Ny;
  if (split == 4) {
    for (uint32_t j = 0; j < ny; ++j)
      for (uint32_t i = 0; i < nx; ++i) {
        m.coords.push_back(x0 + (i + 0.5) * hx);
        m.coords.push_back(y0 + (j + 0.5) * hy);
      }
  }
  m.cell_offset.push_back(0);
  for (uint32_t j = 0; j < ny; ++j) {
    for (uint32_t i = 0; i < nx; ++i) {
      const uint32_t k = j * Nx + i;
      if (split == 1 || j < quad_rows) {
        add_cell(m, Shape::Quad, {k, k + 1, k + Nx + 1, k + Nx});
      } else if (split == 2) {
        add_cell(m, Shape::Triag, {k, k + 1, k + Nx + 1});
        add_cell(m, Shape::Triag, {k + Nx + 1, k + Nx, k});
      } else {
        const uint32_t c = nb_grid + j * nx + i;
        add_cell(m, Shape::Triag, {k, k + 1, c});
        add_cell(m, Shape::Triag, {k + 1, k + Nx + 1, c});
        add_cell(m, Shape::Triag, {k + Nx + 1, k + Nx, c});
        add_cell(m, Shape::Triag, {k + Nx, k, c});
      }
    }
  }
  std::vector<uint32_t> b0, b1, b2, b3;
  for (uint32_t i = 0; i < Nx; ++i) { b0.push_back(i); b2.push_back((Ny - 1) * Nx + i); }
  for (uint32_t j = 0; j < Ny; ++j) { b1.push_back(j * Nx + Nx - 1); b3.push_back(j * Nx); }
  add_boundary(m, "ymin", b0);
  add_boundary(m, "xmax", b1);
  add_boundary(m, "ymax", b2);
  add_boundary(m, "xmin", b3);
  return m;
}

GeoMesh make_box_tetra(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz)
{
  if (nx < 1 || ny < 1 || nz < 1) fail("make_box_tetra needs nx, ny, nz >= 1");
  GeoMesh m;
  m.dim = 3;
  const uint32_t Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  for (uint32_t k = 0; k < Nz; ++k)
    for (uint32_t j = 0; j < Ny; ++j)
      for (uint32_t i = 0; i < Nx; ++i) {
        m.coords.push_back(lx * i / nx);
        m.coords.push_back(ly * j / ny);
        m.coords.push_back(lz * k / nz);
      }
  auto vid = [&](uint32_t i, uint32_t j, uint32_t k) { return (k * Ny + j) * Nx + i; };
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  m.cell_offset.push_back(0);
  for (uint32_t k = 0; k < nz; ++k)
    for (uint32_t j = 0; j < ny; ++j)
      for (uint32_t i = 0; i < nx; ++i) {
        for (const auto &p : perms) {
          uint32_t ijk[3] = {i, j, k};
          uint32_t v[4];
          v[0] = vid(ijk[0], ijk[1], ijk[2]);
          ijk[p[0]]++;
          v[1] = vid(ijk[0], ijk[1], ijk[2]);
          ijk[p[1]]++;
          v[2] = vid(ijk[0], ijk[1], ijk[2]);
          v[3] = vid(i + 1, j + 1, k + 1);
          // orientation: positive volume like the reference tetrahedron
          const double *a = &m.coords[3 * v[0]], *b = &m.coords[3 * v[1]];
          const double *c = &m.coords[3 * v[2]], *d = &m.coords[3 * v[3]];
          double e1[3], e2[3], e3[3];
          for (int q = 0; q < 3; ++q) { e1[q] = b[q] - a[q]; e2[q] = c[q] - a[q]; e3[q] = d[q] - a[q]; }
          const double det = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) -
                             e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                             e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
          if (det < 0) std::swap(v[2], v[3]);
          add_cell(m, Shape::Tetra, {v[0], v[1], v[2], v[3]});
        }
      }
  std::vector<uint32_t> s[6];
  for (uint32_t k = 0; k < Nz; ++k)
    for (uint32_t j = 0; j < Ny; ++j)
      for (uint32_t i = 0; i < Nx; ++i) {
        const uint32_t v = vid(i, j, k);
        if (i == 0) s[0].push_back(v);
        if (i == nx) s[1].push_back(v);
        if (j == 0) s[2].push_back(v);
        if (j == ny) s[3].push_back(v);
        if (k == 0) s[4].push_back(v);
        if (k == nz) s[5].push_back(v);
      }
  const char *names[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
  for (int q = 0; q < 6; ++q) add_boundary(m, names[q], s[q]);
  return m;
}

void jitter_interior(GeoMesh &m, const std::vector<double> &h, double f, uint64_t seed)
{
  if (static_cast<int>(h.size()) != m.dim) fail("jitter needs one spacing per axis");
  std::vector<char> on_bnd(m.nb_vertices(), 0);
  for (const auto &b : m.bnd_vertices)
    for (uint32_t v : b) on_bnd[v] = 1;
  std::mt19937_64 rng(seed);
  for (uint32_t v = 0; v < m.nb_vertices(); ++v) {
    if (on_bnd[v]) continue;
    for (int d = 0; d < m.dim; ++d) {
      const double u = static_cast<double>(rng() >> 11) * (1.0 / 9007199254740992.0);
      m.coords[v * m.dim + d] += (2.0 * u - 1.0) * f * h[d];
    }
  }
}

void bump_map_2d(GeoMesh &m, double H, double amp, double k, double xc)
{
  for (uint32_t v = 0; v < m.nb_vertices(); ++v) {
    const double x = m.coords[2 * v], y = m.coords[2 * v + 1];
    const double yb = amp * std::exp(-k * (x - xc) * (x - xc));
    m.coords[2 * v + 1] = yb + y * (H - yb) / H;
  }
}

void bump_map_3d(GeoMesh &m, double H, double amp, double k, double xc, double yc)
{
  for (uint32_t v = 0; v < m.nb_vertices(); ++v) {
    const double x = m.coords[3 * v], y = m.coords[3 * v + 1], z = m.coords[3 * v + 2];
    const double zb = amp * std::exp(-k * ((x - xc) * (x - xc) + (y - yc) * (y - yc)));
    m.coords[3 * v + 2] = zb + z * (H - zb) / H;
  }
}

SolMesh clone_continuous(const GeoMesh &m, int order, PointSet ps)
{
  if (order < 1) fail("order must be >= 1");
  SolMesh s;
  s.order = order;
  const uint32_t nc = m.nb_cells();
  const uint32_t nv = m.nb_vertices();
  // P1 vertices used by cells, compacted in ascending id
  std::vector<uint32_t> vnew(nv, UINT32_MAX);
  for (uint32_t v : m.cell_vertices) vnew[v] = 0;
  uint32_t nb_p1 = 0;
  std::vector<uint32_t> vold;
  for (uint32_t v = 0; v < nv; ++v)
    if (vnew[v] == 0) { vnew[v] = nb_p1++; vold.push_back(v); }

  // per shape: reference nodes, geometry basis at the solution nodes
  struct ShapeInfo
  {
    bool ready = false;
    int nd = 0, nvert = 0;
    EdgeTable et;
    std::vector<double> Vmap;  // [nd][nvert]
  };
  ShapeInfo info[3];
  auto get_info = [&](Shape sh) -> ShapeInfo & {
    ShapeInfo &si = info[static_cast<int>(sh)];
    if (!si.ready) {
      si.nd = shape_nb_nodes(sh, order);
      si.nvert = shape_nb_vertices(sh);
      si.et = edge_table(sh, order);
      LagrangeBasis geo(sh, 1, PointSet::Equidist);
      LagrangeBasis sol(sh, order, ps);
      geo.values(sol.nodes(), si.Vmap);
      si.ready = true;
    }
    return si;
  };

  // edge (lo << 32 | hi) -> id of the node nearest to the edge's first
  // vertex as the first cell met it, and whether that cell ran lo -> hi
  struct EdgeRun { uint32_t base; bool lo_to_hi; };
  std::unordered_map<uint64_t, EdgeRun> edge_runs;
  uint32_t next_id = nb_p1;
  s.cell_offset.assign(nc + 1, 0);
  s.node_coords.assign(static_cast<size_t>(nb_p1) * m.dim, 0.0);
  s.host_count.assign(nb_p1, 1);
  s.host_vertices.assign(2 * static_cast<size_t>(nb_p1), 0);
  for (uint32_t i = 0; i < nb_p1; ++i) {
    for (int d = 0; d < m.dim; ++d) s.node_coords[i * m.dim + d] = m.coords[vold[i] * m.dim + d];
    s.host_vertices[2 * i] = vold[i];
    s.host_vertices[2 * i + 1] = vold[i];
  }

  auto new_node = [&](const ShapeInfo &si, const uint32_t *cv, int loc, uint32_t hc, uint32_t ha,
                      uint32_t hb) {
    for (int d = 0; d < m.dim; ++d) {
      double x = 0.0;
      for (int a = 0; a < si.nvert; ++a) x += si.Vmap[loc * si.nvert + a] * m.coords[cv[a] * m.dim + d];
      s.node_coords.push_back(x);
    }
    s.host_count.push_back(hc);
    s.host_vertices.push_back(ha);
    s.host_vertices.push_back(hb);
    return next_id++;
  };

  const int ne = order - 1;
  for (uint32_t c = 0; c < nc; ++c) {
    const Shape sh = static_cast<Shape>(m.cell_shape[c]);
    ShapeInfo &si = get_info(sh);
    const uint32_t *cv = &m.cell_vertices[m.cell_offset[c]];
    if (static_cast<int>(m.cell_offset[c + 1] - m.cell_offset[c]) != si.nvert)
      fail("cell vertex count does not match its shape");
    std::vector<uint32_t> nodes(si.nd, UINT32_MAX);
    for (int a = 0; a < si.nvert; ++a) nodes[a] = vnew[cv[a]];
    // Scan the higher-order nodes in local order; edges are met in local
    // order too since their nodes are stored edge by edge.
    for (size_t e = 0; e < si.et.edges.size(); ++e) {
      const uint32_t va = cv[si.et.edges[e].first], vb = cv[si.et.edges[e].second];
      const uint32_t lo = std::min(va, vb), hi = std::max(va, vb);
      const uint64_t ek = (static_cast<uint64_t>(lo) << 32) | hi;
      auto it = edge_runs.find(ek);
      for (int i = 1; i <= ne; ++i) {
        const int loc = si.et.first_edge_node + static_cast<int>(e) * ne + (i - 1);
        if (it == edge_runs.end()) {
          nodes[loc] = new_node(si, cv, loc, 2, lo, hi);
        } else {
          const int pos = (va < vb) ? i : (order - i);  // 1-based from lo
          const int k = it->second.lo_to_hi ? (pos - 1) : (ne - pos);
          nodes[loc] = it->second.base + static_cast<uint32_t>(k);
        }
      }
      if (it == edge_runs.end() && ne > 0) {
        const int loc0 = si.et.first_edge_node + static_cast<int>(e) * ne;
        edge_runs.emplace(ek, EdgeRun{nodes[loc0], va < vb});
      }
    }
    const int first_interior = si.et.first_edge_node + static_cast<int>(si.et.edges.size()) * ne;
    for (int loc = first_interior; loc < si.nd; ++loc) nodes[loc] = new_node(si, cv, loc, 0, 0, 0);
    for (int a = 0; a < si.nd; ++a) s.cell_nodes.push_back(nodes[a]);
    s.cell_offset[c + 1] = static_cast<uint32_t>(s.cell_nodes.size());
  }
  s.nb_nodes = next_id;
  return s;
}

std::vector<uint32_t> boundary_nodes(const GeoMesh &m, const SolMesh &s, const std::string &name)
{
  size_t b = 0;
  for (; b < m.bnd_names.size(); ++b)
    if (m.bnd_names[b] == name) break;
  if (b == m.bnd_names.size()) fail("unknown boundary " + name);
  const std::vector<uint32_t> &set = m.bnd_vertices[b];
  auto in = [&](uint32_t v) { return std::binary_search(set.begin(), set.end(), v); };
  std::vector<uint32_t> out;
  for (uint32_t n = 0; n < s.nb_nodes; ++n) {
    const uint32_t hc = s.host_count[n];
    if (hc == 0) continue;
    if (in(s.host_vertices[2 * n]) && in(s.host_vertices[2 * n + 1])) out.push_back(n);
  }
  return out;
}

std::vector<uint32_t> greedy_colour(const std::vector<uint32_t> &cells, const SolMesh &s,
                                    uint32_t *nb_colours)
{
  std::vector<uint64_t> used(s.nb_nodes, 0);
  std::vector<uint32_t> colour(cells.size());
  uint32_t ncol = 0;
  for (size_t i = 0; i < cells.size(); ++i) {
    const uint32_t c = cells[i];
    uint64_t mask = 0;
    for (uint32_t k = s.cell_offset[c]; k < s.cell_offset[c + 1]; ++k) mask |= used[s.cell_nodes[k]];
    if (mask == ~0ULL) fail("more than 64 colours needed");
    uint32_t col = 0;
    while (mask & (1ULL << col)) ++col;
    colour[i] = col;
    ncol = std::max(ncol, col + 1);
    for (uint32_t k = s.cell_offset[c]; k < s.cell_offset[c + 1]; ++k) used[s.cell_nodes[k]] |= (1ULL << col);
  }
  if (nb_colours) *nb_colours = ncol;
  return colour;
}

}  // namespace pdrd
