#include "gmsh.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace pdrd {

namespace {

void fail(const std::string &msg) { throw std::invalid_argument(msg); }

// Gmsh element types read here (PDEKitToGmsh.cpp): topological dimension,
// nodes, shape of the cell (-1 for lines and points) and order.
struct GmshType
{
  int dim, nn, shape, order;
};

bool gmsh_type(int t, GmshType &g)
{
  switch (t) {
    case 15: g = {0, 1, -1, 0}; return true;                                // point
    case 1: g = {1, 2, -1, 1}; return true;                                 // line
    case 8: g = {1, 3, -1, 2}; return true;
    case 26: g = {1, 4, -1, 3}; return true;
    case 2: g = {2, 3, static_cast<int>(Shape::Triag), 1}; return true;     // triangle
    case 9: g = {2, 6, static_cast<int>(Shape::Triag), 2}; return true;
    case 21: g = {2, 10, static_cast<int>(Shape::Triag), 3}; return true;
    case 3: g = {2, 4, static_cast<int>(Shape::Quad), 1}; return true;      // quadrilateral
    case 10: g = {2, 9, static_cast<int>(Shape::Quad), 2}; return true;
    case 36: g = {2, 16, static_cast<int>(Shape::Quad), 3}; return true;
    case 4: g = {3, 4, static_cast<int>(Shape::Tetra), 1}; return true;     // tetrahedron
    case 11: g = {3, 10, static_cast<int>(Shape::Tetra), 2}; return true;
    case 29: g = {3, 20, static_cast<int>(Shape::Tetra), 3}; return true;
    default: return false;
  }
}

int nb_corners(int shape) { return shape == static_cast<int>(Shape::Triag) ? 3 : 4; }

const std::vector<std::pair<int, int>> &cell_edges(int shape)
{
  static const std::vector<std::pair<int, int>> tri = {{0, 1}, {1, 2}, {2, 0}};
  static const std::vector<std::pair<int, int>> quad = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
  static const std::vector<std::pair<int, int>> tet = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 2}, {3, 1}};
  if (shape == static_cast<int>(Shape::Triag)) return tri;
  if (shape == static_cast<int>(Shape::Quad)) return quad;
  return tet;
}

// Weights over the cell's corners of every node of a Gmsh cell of order p
// (<= 3), in Gmsh's node order: corners; the nodes of each edge from its
// first vertex to its second; then the face / interior nodes (triangle P3:
// centroid; quadrilateral: the inner lattice, peeled like the outer one;
// tetrahedron P3: the centroids of faces (0,2,1), (0,1,3), (0,3,2), (3,1,2)).
std::vector<std::vector<double>> gmsh_node_weights(int shape, int p)
{
  const int nc = nb_corners(shape);
  std::vector<std::vector<double>> w;
  for (int a = 0; a < nc; ++a) {
    std::vector<double> x(nc, 0.0);
    x[a] = 1.0;
    w.push_back(x);
  }
  for (const auto &e : cell_edges(shape))
    for (int k = 1; k < p; ++k) {
      std::vector<double> x(nc, 0.0);
      x[e.first] = 1.0 - static_cast<double>(k) / p;
      x[e.second] = static_cast<double>(k) / p;
      w.push_back(x);
    }
  auto bilinear = [&](double s, double t) {
    w.push_back({(1 - s) * (1 - t), s * (1 - t), s * t, (1 - s) * t});
  };
  if (shape == static_cast<int>(Shape::Triag) && p == 3) w.push_back({1. / 3, 1. / 3, 1. / 3});
  if (shape == static_cast<int>(Shape::Quad) && p == 2) bilinear(0.5, 0.5);
  if (shape == static_cast<int>(Shape::Quad) && p == 3) {
    bilinear(1. / 3, 1. / 3);
    bilinear(2. / 3, 1. / 3);
    bilinear(2. / 3, 2. / 3);
    bilinear(1. / 3, 2. / 3);
  }
  if (shape == static_cast<int>(Shape::Tetra) && p == 3) {
    const int faces[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {3, 1, 2}};
    for (const auto &f : faces) {
      std::vector<double> x(4, 0.0);
      for (int a : f) x[a] = 1. / 3;
      w.push_back(x);
    }
  }
  return w;
}

std::string next_line(std::istream &in, const char *what)
{
  std::string s;
  if (!std::getline(in, s)) fail(std::string("Gmsh file ends inside ") + what);
  if (!s.empty() && s.back() == '\r') s.pop_back();
  return s;
}

std::string unquote(const std::string &s)
{
  if (s.size() >= 2 && s.front() == '"' && s.back() == '"') return s.substr(1, s.size() - 2);
  return s;
}

}  // namespace

GeoMesh read_gmsh(const std::string &path)
{
  std::ifstream in(path);
  if (!in) fail("cannot open " + path);
  std::map<std::pair<int, int>, std::string> names;  // (dim, tag) -> name
  std::vector<double> xyz;                           // [n][3]
  struct Elem { int type; int tag; std::vector<uint32_t> nodes; };
  std::vector<Elem> elems;
  bool have_nodes = false, have_elems = false;
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line == "$MeshFormat") {
      std::istringstream h(next_line(in, "$MeshFormat"));
      double version = 0;
      int file_type = -1;
      h >> version >> file_type;
      if (version < 2.0 || version >= 3.0) fail("Gmsh format 2.x expected");
      if (file_type != 0) fail("binary Gmsh files are not read");
    } else if (line == "$PhysicalNames") {
      const long n = std::stol(next_line(in, "$PhysicalNames"));
      for (long i = 0; i < n; ++i) {
        std::istringstream r(next_line(in, "$PhysicalNames"));
        int dim, tag;
        r >> dim >> tag;
        std::string rest;
        std::getline(r, rest);
        const size_t a = rest.find('"'), b = rest.rfind('"');
        if (a == std::string::npos || b == a) fail("unquoted physical name");
        names[{dim, tag}] = rest.substr(a + 1, b - a - 1);
      }
    } else if (line == "$Nodes") {
      const long n = std::stol(next_line(in, "$Nodes"));
      xyz.assign(3 * static_cast<size_t>(n), 0.0);
      std::vector<char> seen(n, 0);
      for (long i = 0; i < n; ++i) {
        std::istringstream r(next_line(in, "$Nodes"));
        long id;
        double x, y, z;
        if (!(r >> id >> x >> y >> z)) fail("bad $Nodes line");
        if (id < 1 || id > n || seen[id - 1]) fail("Gmsh node ids must be 1..n, each once");
        seen[id - 1] = 1;
        xyz[3 * (id - 1)] = x;
        xyz[3 * (id - 1) + 1] = y;
        xyz[3 * (id - 1) + 2] = z;
      }
      have_nodes = true;
    } else if (line == "$Elements") {
      const long n = std::stol(next_line(in, "$Elements"));
      elems.reserve(n);
      for (long i = 0; i < n; ++i) {
        std::istringstream r(next_line(in, "$Elements"));
        long id;
        int type, ntags;
        if (!(r >> id >> type >> ntags)) fail("bad $Elements line");
        GmshType g;
        if (!gmsh_type(type, g)) fail("Gmsh element type " + std::to_string(type) + " is not read");
        Elem e;
        e.type = type;
        e.tag = 0;
        for (int t = 0; t < ntags; ++t) {
          int v;
          r >> v;
          if (t == 0) e.tag = v;
        }
        e.nodes.resize(g.nn);
        for (int k = 0; k < g.nn; ++k) {
          long v;
          if (!(r >> v)) fail("short element line");
          e.nodes[k] = static_cast<uint32_t>(v - 1);
        }
        elems.push_back(std::move(e));
      }
      have_elems = true;
    }
  }
  if (!have_nodes || !have_elems) fail("Gmsh file without $Nodes or $Elements");
  const uint32_t nnodes = static_cast<uint32_t>(xyz.size() / 3);
  int tdim = 0;
  for (const Elem &e : elems) {
    GmshType g;
    gmsh_type(e.type, g);
    tdim = std::max(tdim, g.dim);
    for (uint32_t v : e.nodes)
      if (v >= nnodes) fail("element node id beyond $Nodes");
  }
  if (tdim < 2) fail("Gmsh mesh of dimension 2 or 3 expected");
  GeoMesh m;
  m.dim = tdim;
  m.coords.resize(static_cast<size_t>(nnodes) * tdim);
  for (uint32_t v = 0; v < nnodes; ++v)
    for (int d = 0; d < tdim; ++d) m.coords[static_cast<size_t>(v) * tdim + d] = xyz[3 * v + d];
  m.cell_offset.push_back(0);
  std::map<int, std::vector<uint32_t>> bverts;  // physical tag -> corner vertices
  std::map<std::pair<int, int>, std::vector<std::vector<double>>> wcache;
  for (const Elem &e : elems) {
    GmshType g;
    gmsh_type(e.type, g);
    if (g.dim == tdim) {
      const int nc = nb_corners(g.shape);
      if (g.order > 1) {
        // straight-sided check: every node where the P1 / Q1 cell puts it
        auto key = std::make_pair(g.shape, g.order);
        if (!wcache.count(key)) wcache[key] = gmsh_node_weights(g.shape, g.order);
        const auto &w = wcache[key];
        double h = 0.0;
        for (int a = 1; a < nc; ++a)
          for (int d = 0; d < 3; ++d) h = std::max(h, std::fabs(xyz[3 * e.nodes[a] + d] - xyz[3 * e.nodes[0] + d]));
        for (int k = nc; k < g.nn; ++k)
          for (int d = 0; d < 3; ++d) {
            double x = 0.0;
            for (int a = 0; a < nc; ++a) x += w[k][a] * xyz[3 * e.nodes[a] + d];
            if (std::fabs(x - xyz[3 * e.nodes[k] + d]) > 1e-8 * h)
              fail("curved high-order Gmsh cells are not read (straight-sided cells only)");
          }
      }
      m.cell_shape.push_back(g.shape);
      for (int a = 0; a < nc; ++a) m.cell_vertices.push_back(e.nodes[a]);
      m.cell_offset.push_back(static_cast<uint32_t>(m.cell_vertices.size()));
    } else if (g.dim == tdim - 1) {
      const int nc = (g.dim == 1) ? 2 : nb_corners(g.shape);
      std::vector<uint32_t> &bv = bverts[e.tag];
      for (int a = 0; a < nc; ++a) bv.push_back(e.nodes[a]);
    }
  }
  if (m.cell_shape.empty()) fail("Gmsh file without cells");
  for (auto &kv : bverts) {
    std::vector<uint32_t> v = kv.second;
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    auto it = names.find({tdim - 1, kv.first});
    m.bnd_names.push_back(it != names.end() ? it->second : std::to_string(kv.first));
    m.bnd_vertices.push_back(std::move(v));
  }
  return m;
}

void write_gmsh(const std::string &path, const GeoMesh &g, const SolMesh &s)
{
  const int p = s.order;
  const int dim = g.dim;
  if (p > 3 || (dim == 3 && p > 2)) fail("Gmsh output of this order is not written");
  auto cell_type = [&](int shape) {
    if (shape == static_cast<int>(Shape::Triag)) return p == 1 ? 2 : (p == 2 ? 9 : 21);
    if (shape == static_cast<int>(Shape::Quad)) return p == 1 ? 3 : (p == 2 ? 10 : 36);
    return p == 1 ? 4 : 11;
  };
  const int facet_type = (dim == 2) ? (p == 1 ? 1 : (p == 2 ? 8 : 26)) : (p == 1 ? 2 : 9);
  const uint32_t nc = g.nb_cells();

  // facets on the domain boundary: (sorted corner key) -> occurrences
  auto facets_of = [&](uint32_t c, std::vector<std::vector<uint32_t>> &out) {
    out.clear();
    const int shape = g.cell_shape[c];
    const uint32_t *cn = &s.cell_nodes[s.cell_offset[c]];
    const int nv = nb_corners(shape);
    const auto &edges = cell_edges(shape);
    if (dim == 2) {
      for (size_t e = 0; e < edges.size(); ++e) {
        std::vector<uint32_t> f = {cn[edges[e].first], cn[edges[e].second]};
        for (int k = 0; k < p - 1; ++k) f.push_back(cn[nv + e * (p - 1) + k]);
        out.push_back(f);
      }
      return;
    }
    const int faces[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (const auto &fc : faces) {
      std::vector<uint32_t> f = {cn[fc[0]], cn[fc[1]], cn[fc[2]]};
      if (p == 2)
        for (int k = 0; k < 3; ++k) {
          const int a = fc[k], b = fc[(k + 1) % 3];
          for (size_t e = 0; e < edges.size(); ++e)
            if ((edges[e].first == a && edges[e].second == b) || (edges[e].first == b && edges[e].second == a))
              f.push_back(cn[nv + e]);
        }
      out.push_back(f);
    }
  };
  const int fc_corners = (dim == 2) ? 2 : 3;
  auto key_of = [&](const std::vector<uint32_t> &f) {
    std::array<uint32_t, 3> k = {UINT32_MAX, UINT32_MAX, UINT32_MAX};
    for (int i = 0; i < fc_corners; ++i) k[i] = f[i];
    std::sort(k.begin(), k.begin() + fc_corners);
    return k;
  };
  std::map<std::array<uint32_t, 3>, int> count;
  std::vector<std::vector<uint32_t>> fl;
  for (uint32_t c = 0; c < nc; ++c) {
    facets_of(c, fl);
    for (const auto &f : fl) count[key_of(f)]++;
  }
  // solution-node membership of every boundary (corner nodes are vertex nodes)
  std::vector<uint32_t> vert_of_node(s.nb_nodes, UINT32_MAX);
  for (uint32_t i = 0; i < s.nb_nodes; ++i)
    if (s.host_count[i] == 1) vert_of_node[i] = s.host_vertices[2 * i];
  std::vector<std::vector<std::vector<uint32_t>>> bfacets(g.bnd_names.size());
  for (size_t b = 0; b < g.bnd_names.size(); ++b) {
    std::vector<char> in_set(g.nb_vertices(), 0);
    for (uint32_t v : g.bnd_vertices[b]) in_set[v] = 1;
    for (uint32_t c = 0; c < nc; ++c) {
      facets_of(c, fl);
      for (const auto &f : fl) {
        if (count[key_of(f)] != 1) continue;
        bool all = true;
        for (int i = 0; i < fc_corners; ++i) all = all && vert_of_node[f[i]] != UINT32_MAX && in_set[vert_of_node[f[i]]];
        if (all) bfacets[b].push_back(f);
      }
    }
  }

  FILE *fp = std::fopen(path.c_str(), "w");
  if (!fp) fail("cannot write " + path);
  std::fprintf(fp, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$PhysicalNames\n%zu\n", g.bnd_names.size() + 1);
  for (size_t b = 0; b < g.bnd_names.size(); ++b) std::fprintf(fp, "%d %zu \"%s\"\n", dim - 1, b + 1, g.bnd_names[b].c_str());
  const size_t interior = g.bnd_names.size() + 1;
  std::fprintf(fp, "%d %zu \"interior\"\n$EndPhysicalNames\n$Nodes\n%u\n", dim, interior, s.nb_nodes);
  for (uint32_t i = 0; i < s.nb_nodes; ++i) {
    std::fprintf(fp, "%u", i + 1);
    for (int d = 0; d < 3; ++d) std::fprintf(fp, " %.17g", d < dim ? s.node_coords[static_cast<size_t>(i) * dim + d] : 0.0);
    std::fprintf(fp, "\n");
  }
  size_t nel = nc;
  for (const auto &bf : bfacets) nel += bf.size();
  std::fprintf(fp, "$EndNodes\n$Elements\n%zu\n", nel);
  size_t id = 1;
  for (size_t b = 0; b < bfacets.size(); ++b)
    for (const auto &f : bfacets[b]) {
      std::fprintf(fp, "%zu %d 2 %zu %zu", id++, facet_type, b + 1, b + 1);
      for (uint32_t v : f) std::fprintf(fp, " %u", v + 1);
      std::fprintf(fp, "\n");
    }
  for (uint32_t c = 0; c < nc; ++c) {
    std::fprintf(fp, "%zu %d 2 %zu %zu", id++, cell_type(g.cell_shape[c]), interior, interior);
    for (uint32_t k = s.cell_offset[c]; k < s.cell_offset[c + 1]; ++k) std::fprintf(fp, " %u", s.cell_nodes[k] + 1);
    std::fprintf(fp, "\n");
  }
  std::fprintf(fp, "$EndElements\n");
  if (std::fclose(fp) != 0) fail("error writing " + path);
}

void append_gmsh_node_data(const std::string &path, const std::string &field, int m, uint32_t n,
                           const double *values, double time, uint32_t step, bool scalar)
{
  FILE *fp = std::fopen(path.c_str(), "a");
  if (!fp) fail("cannot append to " + path);
  for (int k = 0; k < m; ++k) {
    const std::string name = (scalar && m == 1) ? field : field + "_" + std::to_string(k);
    std::fprintf(fp, "$NodeData\n1\n\"%s\"\n1\n%.17g\n3\n%u\n1\n%u\n", name.c_str(), time, step, n);
    for (uint32_t i = 0; i < n; ++i) std::fprintf(fp, "%u %.17g\n", i + 1, values[static_cast<size_t>(i) * m + k]);
    std::fprintf(fp, "$EndNodeData\n");
  }
  if (std::fclose(fp) != 0) fail("error writing " + path);
}

int read_gmsh_node_data(const std::string &path, const std::vector<std::string> &names, uint32_t n, double *values)
{
  std::ifstream in(path);
  if (!in) fail("cannot open " + path);
  const size_t nf = names.size();
  std::vector<char> found(nf, 0);
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line != "$NodeData") continue;
    const int nstr = std::stoi(next_line(in, "$NodeData"));
    std::string name;
    for (int i = 0; i < nstr; ++i) {
      const std::string t = unquote(next_line(in, "$NodeData"));
      if (i == 0) name = t;
    }
    const int nreal = std::stoi(next_line(in, "$NodeData"));
    for (int i = 0; i < nreal; ++i) next_line(in, "$NodeData");
    const int nint = std::stoi(next_line(in, "$NodeData"));
    long entries = 0;
    for (int i = 0; i < nint; ++i) entries = std::stol(next_line(in, "$NodeData"));
    size_t f = 0;
    while (f < nf && names[f] != name) ++f;
    for (long i = 0; i < entries; ++i) {
      std::istringstream r(next_line(in, "$NodeData"));
      long idx;
      double v;
      if (!(r >> idx >> v)) fail("bad $NodeData line");
      if (f == nf) continue;
      if (idx < 1 || idx > static_cast<long>(n)) fail("$NodeData node id beyond the mesh");
      values[static_cast<size_t>(idx - 1) * nf + f] = v;
    }
    if (f < nf) found[f] = 1;
  }
  int c = 0;
  for (char x : found) c += x;
  return c;
}

}  // namespace pdrd
