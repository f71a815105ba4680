// C ABI of the host-side inputs: FE tables and mesh generation / numbering.
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>

#include "../../include/pdekit_rd.h"
#include "host/capi_common.hpp"
#include "host/fe_tables.hpp"
#include "host/gmsh.hpp"
#include "host/mesh.hpp"

struct pdrd_meshgen
{
  pdrd::GeoMesh geo;
  pdrd::SolMesh sol;
  bool cloned = false;
};

namespace pdrd {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg)
{
  g_last_error = msg;
  return code;
}

}  // namespace pdrd

using pdrd::set_error;

#define PDRD_TRY try {
#define PDRD_CATCH                                                 \
  }                                                                \
  catch (const std::invalid_argument &e)                           \
  {                                                                \
    return set_error(PDRD_EINVAL, e.what());                       \
  }                                                                \
  catch (const std::exception &e)                                  \
  {                                                                \
    return set_error(PDRD_EINVAL, e.what());                       \
  }

extern "C" {

const char *pdrd_last_error(void) { return pdrd::g_last_error.c_str(); }

const char *pdrd_version(void) { return "pdekit_rd 0.1 (gfx950)"; }

int pdrd_fe_describe(int shape, int geo_order, int sol_order, int quad_order, int point_set,
                     pdrd_fe_desc *out)
{
  PDRD_TRY
  if (!out) return set_error(PDRD_EINVAL, "null output");
  if (shape < 0 || shape > 2) return set_error(PDRD_EINVAL, "unknown shape");
  const pdrd::ElementTables t = pdrd::make_element_tables(
      static_cast<pdrd::Shape>(shape), geo_order, sol_order, quad_order, static_cast<pdrd::PointSet>(point_set));
  out->shape = shape;
  out->geo_order = geo_order;
  out->sol_order = sol_order;
  out->quad_order = quad_order;
  out->point_set = point_set;
  out->dim = t.dim;
  out->ng = t.ng;
  out->nd = t.nd;
  out->nf = t.nf;
  out->nq = t.nq;
  out->table_size = pdrd::packed_size(t);
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_fe_fill(const pdrd_fe_desc *d, double *tables)
{
  PDRD_TRY
  if (!d || !tables) return set_error(PDRD_EINVAL, "null argument");
  const pdrd::ElementTables t = pdrd::make_element_tables(static_cast<pdrd::Shape>(d->shape), d->geo_order,
                                                          d->sol_order, d->quad_order,
                                                          static_cast<pdrd::PointSet>(d->point_set));
  if (pdrd::packed_size(t) != d->table_size) return set_error(PDRD_EINVAL, "descriptor does not match tables");
  pdrd::pack(t, tables);
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_reference_nodes(int shape, int order, int point_set, double *xi, int *n)
{
  PDRD_TRY
  const auto nodes = pdrd::reference_nodes(static_cast<pdrd::Shape>(shape), order,
                                           static_cast<pdrd::PointSet>(point_set));
  if (n) *n = static_cast<int>(nodes.size()) / pdrd::shape_dim(static_cast<pdrd::Shape>(shape));
  if (xi) std::memcpy(xi, nodes.data(), nodes.size() * sizeof(double));
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_quadrature(int shape, int order, double *xi, double *w, int *nq)
{
  PDRD_TRY
  std::vector<double> x, ww;
  pdrd::quadrature(static_cast<pdrd::Shape>(shape), order, x, ww);
  if (nq) *nq = static_cast<int>(ww.size());
  if (xi) std::memcpy(xi, x.data(), x.size() * sizeof(double));
  if (w) std::memcpy(w, ww.data(), ww.size() * sizeof(double));
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_mesh_generate(int kind, const int64_t *ip, const double *rp, pdrd_meshgen_t *out)
{
  PDRD_TRY
  if (!out || !ip) return set_error(PDRD_EINVAL, "null argument");
  auto *m = new pdrd_meshgen();
  try {
    if (kind == 0) {
      m->geo = pdrd::make_unit_quad_triag(static_cast<uint32_t>(ip[0]));
    } else if (kind == 1) {
      if (!rp) throw std::invalid_argument("rectangle needs extents");
      m->geo = pdrd::make_rectangle(static_cast<uint32_t>(ip[0]), static_cast<uint32_t>(ip[1]), rp[0], rp[1], rp[2],
                                    rp[3], static_cast<int>(ip[2]), static_cast<uint32_t>(ip[3]));
    } else if (kind == 2) {
      if (!rp) throw std::invalid_argument("box needs extents");
      m->geo = pdrd::make_box_tetra(static_cast<uint32_t>(ip[0]), static_cast<uint32_t>(ip[1]),
                                    static_cast<uint32_t>(ip[2]), rp[0], rp[1], rp[2]);
    } else {
      throw std::invalid_argument("unknown mesh kind");
    }
  } catch (...) {
    delete m;
    throw;
  }
  *out = m;
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_mesh_free(pdrd_meshgen_t m)
{
  delete m;
  return PDRD_OK;
}

int pdrd_mesh_jitter(pdrd_meshgen_t m, const double *h, double f, uint64_t seed)
{
  PDRD_TRY
  if (!m || !h) return set_error(PDRD_EINVAL, "null argument");
  if (m->cloned) return set_error(PDRD_EINVAL, "jitter before cloning");
  pdrd::jitter_interior(m->geo, std::vector<double>(h, h + m->geo.dim), f, seed);
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_mesh_bump(pdrd_meshgen_t m, int mode, const double *p)
{
  PDRD_TRY
  if (!m || !p) return set_error(PDRD_EINVAL, "null argument");
  if (m->cloned) return set_error(PDRD_EINVAL, "bump before cloning");
  if (mode == 0)
    pdrd::bump_map_2d(m->geo, p[0], p[1], p[2], p[3]);
  else
    pdrd::bump_map_3d(m->geo, p[0], p[1], p[2], p[3], p[4]);
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_mesh_clone(pdrd_meshgen_t m, int order, int point_set)
{
  PDRD_TRY
  if (!m) return set_error(PDRD_EINVAL, "null mesh");
  m->sol = pdrd::clone_continuous(m->geo, order, static_cast<pdrd::PointSet>(point_set));
  m->cloned = true;
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_mesh_info(pdrd_meshgen_t m, int64_t *info)
{
  if (!m || !info) return set_error(PDRD_EINVAL, "null argument");
  info[0] = m->geo.dim;
  info[1] = m->geo.nb_vertices();
  info[2] = m->geo.nb_cells();
  info[3] = m->cloned ? m->sol.nb_nodes : 0;
  info[4] = static_cast<int64_t>(m->geo.bnd_names.size());
  return PDRD_OK;
}

int pdrd_mesh_coords(pdrd_meshgen_t m, double *c)
{
  if (!m || !c) return set_error(PDRD_EINVAL, "null argument");
  std::memcpy(c, m->geo.coords.data(), m->geo.coords.size() * sizeof(double));
  return PDRD_OK;
}

int pdrd_mesh_node_coords(pdrd_meshgen_t m, double *c)
{
  if (!m || !c) return set_error(PDRD_EINVAL, "null argument");
  if (!m->cloned) return set_error(PDRD_EINVAL, "mesh not cloned");
  std::memcpy(c, m->sol.node_coords.data(), m->sol.node_coords.size() * sizeof(double));
  return PDRD_OK;
}

int pdrd_mesh_cell_shapes(pdrd_meshgen_t m, int32_t *s)
{
  if (!m || !s) return set_error(PDRD_EINVAL, "null argument");
  for (size_t i = 0; i < m->geo.cell_shape.size(); ++i) s[i] = m->geo.cell_shape[i];
  return PDRD_OK;
}

int pdrd_mesh_geo_conn(pdrd_meshgen_t m, uint32_t *conn, int64_t *size)
{
  if (!m) return set_error(PDRD_EINVAL, "null mesh");
  if (size) *size = static_cast<int64_t>(m->geo.cell_vertices.size());
  if (conn) std::memcpy(conn, m->geo.cell_vertices.data(), m->geo.cell_vertices.size() * sizeof(uint32_t));
  return PDRD_OK;
}

int pdrd_mesh_sol_conn(pdrd_meshgen_t m, uint32_t *conn, int64_t *size)
{
  if (!m) return set_error(PDRD_EINVAL, "null mesh");
  if (!m->cloned) return set_error(PDRD_EINVAL, "mesh not cloned");
  if (size) *size = static_cast<int64_t>(m->sol.cell_nodes.size());
  if (conn) std::memcpy(conn, m->sol.cell_nodes.data(), m->sol.cell_nodes.size() * sizeof(uint32_t));
  return PDRD_OK;
}

int pdrd_mesh_boundary_name(pdrd_meshgen_t m, int idx, char *name, int cap)
{
  if (!m || !name || cap <= 0) return set_error(PDRD_EINVAL, "null argument");
  if (idx < 0 || idx >= static_cast<int>(m->geo.bnd_names.size())) return set_error(PDRD_EINVAL, "bad index");
  std::strncpy(name, m->geo.bnd_names[idx].c_str(), cap - 1);
  name[cap - 1] = 0;
  return PDRD_OK;
}

int pdrd_mesh_boundary_nodes(pdrd_meshgen_t m, const char *name, uint32_t *nodes, int64_t *n)
{
  PDRD_TRY
  if (!m || !name) return set_error(PDRD_EINVAL, "null argument");
  if (!m->cloned) return set_error(PDRD_EINVAL, "mesh not cloned");
  const auto v = pdrd::boundary_nodes(m->geo, m->sol, name);
  if (n) *n = static_cast<int64_t>(v.size());
  if (nodes) std::memcpy(nodes, v.data(), v.size() * sizeof(uint32_t));
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_mesh_read_gmsh(const char *path, pdrd_meshgen_t *out)
{
  PDRD_TRY
  if (!path || !out) return set_error(PDRD_EINVAL, "null argument");
  auto *m = new pdrd_meshgen();
  try {
    m->geo = pdrd::read_gmsh(path);
  } catch (...) {
    delete m;
    throw;
  }
  *out = m;
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_mesh_write_gmsh(pdrd_meshgen_t m, const char *path)
{
  PDRD_TRY
  if (!m || !path) return set_error(PDRD_EINVAL, "null argument");
  if (!m->cloned) return set_error(PDRD_EINVAL, "mesh not cloned");
  pdrd::write_gmsh(path, m->geo, m->sol);
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_gmsh_append_node_data(const char *path, const char *field, int m, uint32_t n, const double *values,
                               double time, uint32_t step, int scalar)
{
  PDRD_TRY
  if (!path || !field || (n && !values) || m < 1) return set_error(PDRD_EINVAL, "bad argument");
  pdrd::append_gmsh_node_data(path, field, m, n, values, time, step, scalar != 0);
  return PDRD_OK;
  PDRD_CATCH
}

int pdrd_gmsh_read_node_data(const char *path, int nfields, const char *const *names, uint32_t n, double *values,
                             int *found)
{
  PDRD_TRY
  if (!path || nfields < 1 || !names || (n && !values)) return set_error(PDRD_EINVAL, "bad argument");
  std::vector<std::string> nm;
  for (int f = 0; f < nfields; ++f) {
    if (!names[f]) return set_error(PDRD_EINVAL, "null field name");
    nm.emplace_back(names[f]);
  }
  const int c = pdrd::read_gmsh_node_data(path, nm, n, values);
  if (found) *found = c;
  return PDRD_OK;
  PDRD_CATCH
}

}  // extern "C"
