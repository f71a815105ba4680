// pdekit_rd.hpp -- C++ host mirror of pdekit's explicit continuous RD method
// over the C ABI of pdekit_rd.h.
//
// PGRDMethodExplicit<Physics, Scheme> keeps the member names and call order of
// pdekit's PGRDMethodExplicit<MeshConfig, Physics, SchemeTag>
// (src/solver/rdm/PGRDMethodExplicit.hpp:19-193, RDMethod.hpp:33-139):
//   configure_mesh_data -> initialize_work_data -> set_vec_function ->
//   add_boundary_condition -> { assemble_lhs_and_rhs(CFL); solve(); } ->
//   compute_residual_norm.
// Nodal fields are std::vector<double> in pdekit's AoS [N][m] layout. Errors
// throw pdrd::Error carrying the C ABI status and message.
#ifndef PDEKIT_RD_HPP
#define PDEKIT_RD_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pdekit_rd.h"

namespace pdrd {

class Error : public std::runtime_error
{
  public:
  Error(int code, const std::string &where)
      : std::runtime_error(where + ": status " + std::to_string(code) + ": " + pdrd_last_error()), m_code(code)
  {
  }
  int code() const { return m_code; }

  private:
  int m_code;
};

inline void check(int rc, const char *where)
{
  if (rc != PDRD_OK) throw Error(rc, where);
}

// Physics tags (template arguments of the reference class).
struct Advection2D { static constexpr int id = PDRD_ADVECTION2D, neq = 1, dim = 2; };
struct Burgers2D { static constexpr int id = PDRD_BURGERS2D, neq = 1, dim = 2; };
struct Euler2DCons { static constexpr int id = PDRD_EULER2D, neq = 4, dim = 2; };
struct Euler3DCons { static constexpr int id = PDRD_EULER3D, neq = 5, dim = 3; };
// Scheme tags.
struct PGLDA { static constexpr int id = PDRD_LDA; };
struct PGN { static constexpr int id = PDRD_N; };
struct PGB { static constexpr int id = PDRD_B; };
struct PSI { static constexpr int id = PDRD_PSI; };

enum class SolverVecFn { solution = PDRD_SOLUTION, residuals = PDRD_RESIDUAL };
enum class SolverScalFn { time_step = PDRD_TIME_STEP, wave_speed = PDRD_WAVE_SPEED, dual_volume = PDRD_DUAL_VOLUME };

// One block of cells of one element type: geometry and solution connectivity
// (row-major [ne][ng] / [ne][nd]).
struct CellBlock
{
  int shape = PDRD_TRIAG;
  int order = 1;
  std::vector<uint32_t> geo_conn, sol_conn;
};

// A mesh as the method needs it: vertex coordinates, number of solution
// nodes, cell blocks.
struct Mesh
{
  int dim = 2;
  std::vector<double> coords;
  uint32_t nb_nodes = 0;
  std::vector<CellBlock> blocks;
};

template <typename Physics, typename Scheme>
class PGRDMethodExplicit
{
  public:
  static constexpr int NEQ = Physics::neq;

  explicit PGRDMethodExplicit(std::vector<double> params = {}) : m_params(std::move(params)) {}
  PGRDMethodExplicit(const PGRDMethodExplicit &) = delete;
  PGRDMethodExplicit &operator=(const PGRDMethodExplicit &) = delete;
  ~PGRDMethodExplicit()
  {
    if (m_h) pdrd_destroy(m_h);
  }

  void configure_mesh_data(const Mesh &mesh) { m_mesh = &mesh; }

  // quad_order: Gauss order of the cell quadrature (the reference's
  // initialize_work_data(SFunc::Lagrange, PointSetID::Gauss, quad_order, ...)).
  void initialize_work_data(int quad_order, int point_set = PDRD_WARPBLEND)
  {
    if (!m_mesh) throw std::logic_error("configure_mesh_data first");
    const Mesh &m = *m_mesh;
    check(pdrd_create(m.dim, static_cast<uint32_t>(m.coords.size() / m.dim), m.coords.data(), m.nb_nodes, Physics::id,
                      Scheme::id, m_params.empty() ? nullptr : m_params.data(), static_cast<int>(m_params.size()),
                      &m_h),
          "pdrd_create");
    for (const CellBlock &b : m.blocks) {
      pdrd_fe_desc d;
      check(pdrd_fe_describe(b.shape, 1, b.order, quad_order, point_set, &d), "pdrd_fe_describe");
      std::vector<double> tab(d.table_size);
      check(pdrd_fe_fill(&d, tab.data()), "pdrd_fe_fill");
      const uint32_t ne = static_cast<uint32_t>(b.geo_conn.size() / d.ng);
      check(pdrd_add_cells(m_h, &d, tab.data(), ne, b.geo_conn.data(), b.sol_conn.data()), "pdrd_add_cells");
    }
    check(pdrd_initialize(m_h), "pdrd_initialize");
  }

  void set_vec_function(SolverVecFn which, const std::vector<double> &values)
  {
    check(pdrd_set_field(m_h, static_cast<int>(which), values.data()), "pdrd_set_field");
  }

  std::vector<double> vec_function(SolverVecFn which) const
  {
    std::vector<double> out(static_cast<size_t>(m_mesh->nb_nodes) * NEQ);
    check(pdrd_get_field(m_h, static_cast<int>(which), out.data()), "pdrd_get_field");
    return out;
  }

  std::vector<double> scal_function(SolverScalFn which) const
  {
    std::vector<double> out(m_mesh->nb_nodes);
    check(pdrd_get_field(m_h, static_cast<int>(which), out.data()), "pdrd_get_field");
    return out;
  }

  // add_boundary_condition("StrongDirichlet", name, boundary) with the values
  // of the BC expression at the listed nodes ([n][NEQ]).
  void add_boundary_condition(const std::string &type, const std::vector<uint32_t> &nodes,
                              const std::vector<double> &values)
  {
    if (type != "StrongDirichlet") throw std::invalid_argument("only StrongDirichlet is built on the GPU path");
    check(pdrd_add_strong_dirichlet(m_h, static_cast<uint32_t>(nodes.size()), nodes.data(), values.data()),
          "pdrd_add_strong_dirichlet");
  }

  void assemble_rhs() { check(pdrd_assemble_rhs(m_h), "pdrd_assemble_rhs"); }
  void assemble_lhs_and_rhs(double cfl) { check(pdrd_assemble_lhs_and_rhs(m_h, cfl), "pdrd_assemble_lhs_and_rhs"); }
  void solve() { check(pdrd_solve(m_h), "pdrd_solve"); }

  std::vector<double> compute_residual_norm() const
  {
    std::vector<double> n(NEQ);
    check(pdrd_compute_residual_norm(m_h, n.data()), "pdrd_compute_residual_norm");
    return n;
  }

  std::vector<double> norm_L2_residual(bool exclude_dirichlet = false) const
  {
    std::vector<double> n(NEQ);
    check(pdrd_residual_norm_L2(m_h, exclude_dirichlet ? 1 : 0, n.data()), "pdrd_residual_norm_L2");
    return n;
  }

  // n_iter x { assemble_lhs_and_rhs(cfl); solve(); } on the device.
  std::vector<double> iterate(double cfl, int n_iter, bool exclude_dirichlet = false)
  {
    std::vector<double> n(NEQ);
    check(pdrd_iterate(m_h, cfl, n_iter, exclude_dirichlet ? 1 : 0, n.data()), "pdrd_iterate");
    return n;
  }

  void check_status() const { check(pdrd_check_status(m_h), "pdrd_check_status"); }
  pdrd_method_t handle() const { return m_h; }

  private:
  std::vector<double> m_params;
  const Mesh *m_mesh = nullptr;
  pdrd_method_t m_h = nullptr;
};

// make_unit_quad(N) of the reference (MeshCreator.hpp:436-600) cloned to
// order P, mapped to [0,1]^2, with the node ids of its boundary "Edge0".."Edge3".
inline Mesh make_unit_quad(int N, int order, std::vector<std::vector<uint32_t>> *edges = nullptr)
{
  pdrd_meshgen_t g = nullptr;
  const int64_t ip[1] = {N};
  check(pdrd_mesh_generate(0, ip, nullptr, &g), "pdrd_mesh_generate");
  struct Guard { pdrd_meshgen_t g; ~Guard() { pdrd_mesh_free(g); } } guard{g};
  check(pdrd_mesh_clone(g, order, PDRD_WARPBLEND), "pdrd_mesh_clone");
  int64_t info[5];
  check(pdrd_mesh_info(g, info), "pdrd_mesh_info");
  Mesh m;
  m.dim = 2;
  m.coords.resize(static_cast<size_t>(info[1]) * 2);
  check(pdrd_mesh_coords(g, m.coords.data()), "pdrd_mesh_coords");
  for (double &x : m.coords) x = (x + 1.0) * 0.5;
  m.nb_nodes = static_cast<uint32_t>(info[3]);
  CellBlock b;
  b.shape = PDRD_TRIAG;
  b.order = order;
  int64_t n = 0;
  check(pdrd_mesh_geo_conn(g, nullptr, &n), "pdrd_mesh_geo_conn");
  b.geo_conn.resize(n);
  check(pdrd_mesh_geo_conn(g, b.geo_conn.data(), &n), "pdrd_mesh_geo_conn");
  check(pdrd_mesh_sol_conn(g, nullptr, &n), "pdrd_mesh_sol_conn");
  b.sol_conn.resize(n);
  check(pdrd_mesh_sol_conn(g, b.sol_conn.data(), &n), "pdrd_mesh_sol_conn");
  m.blocks.push_back(std::move(b));
  if (edges) {
    edges->clear();
    const char *names[4] = {"Edge0", "Edge1", "Edge2", "Edge3"};
    for (const char *nm : names) {
      int64_t k = 0;
      check(pdrd_mesh_boundary_nodes(g, nm, nullptr, &k), "pdrd_mesh_boundary_nodes");
      std::vector<uint32_t> ids(k);
      check(pdrd_mesh_boundary_nodes(g, nm, ids.data(), &k), "pdrd_mesh_boundary_nodes");
      edges->push_back(std::move(ids));
    }
  }
  return m;
}

}  // namespace pdrd

#endif  // PDEKIT_RD_HPP
