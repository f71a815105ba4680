// C++ driver of the GPU cell loop with pdekit's call pattern
// (examples/testcases-rdm/Advection-Continuous-RDS-2D-Explicit.cpp:52-181):
// make_unit_quad(N) on [0,1]^2, constant advection a = (0.8, 0.6), strong
// Dirichlet inflow u = 1 for y < 0.5 else 0 on Edge0 (y = 0) and Edge3 (x = 0),
// u0 = 0, explicit pseudo-time iterations at CFL 0.9 until the L2 norm of the
// residual on the non-Dirichlet nodes drops below the tolerance.
//
// usage: advection_rds_2d [N=129] [scheme: lda|n|b|psi = n] [tol=1e-12] [max_iter=20000]
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>

#include "pdekit_rd.hpp"

template <typename Scheme>
int run(int N, double tol, int max_iter)
{
  using namespace pdrd;
  std::vector<std::vector<uint32_t>> edges;
  const Mesh mesh = make_unit_quad(N, 1, &edges);
  PGRDMethodExplicit<Advection2D, Scheme> scheme({0.8, 0.6});
  scheme.configure_mesh_data(mesh);
  scheme.initialize_work_data(4);
  scheme.set_vec_function(SolverVecFn::solution, std::vector<double>(mesh.nb_nodes, 0.0));
  std::set<uint32_t> inflow(edges[0].begin(), edges[0].end());
  inflow.insert(edges[3].begin(), edges[3].end());
  std::vector<uint32_t> nodes(inflow.begin(), inflow.end());
  std::vector<double> vals(nodes.size());
  for (size_t i = 0; i < nodes.size(); ++i) vals[i] = mesh.coords[2 * nodes[i] + 1] < 0.5 ? 1.0 : 0.0;
  scheme.add_boundary_condition("StrongDirichlet", nodes, vals);

  const auto t0 = std::chrono::steady_clock::now();
  int it = 0;
  double norm = 1.0;
  const int chunk = 100;
  while (it < max_iter) {
    norm = scheme.iterate(0.9, chunk, true)[0];
    it += chunk;
    if (norm <= tol) break;
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  scheme.check_status();
  const std::vector<double> u = scheme.vec_function(SolverVecFn::solution);
  const auto mm = std::minmax_element(u.begin(), u.end());
  std::printf("cells=%zu nodes=%u iterations=%d residual=%.3e converged=%d umin=%.6e umax=%.6e seconds=%.3f\n",
              mesh.blocks[0].geo_conn.size() / 3, mesh.nb_nodes, it, norm, norm <= tol ? 1 : 0, *mm.first,
              *mm.second, secs);
  return norm <= tol ? 0 : 1;
}

int main(int argc, char **argv)
{
  const int N = argc > 1 ? std::atoi(argv[1]) : 129;
  const char *s = argc > 2 ? argv[2] : "n";
  const double tol = argc > 3 ? std::atof(argv[3]) : 1e-12;
  const int max_iter = argc > 4 ? std::atoi(argv[4]) : 20000;
  try {
    if (!std::strcmp(s, "lda")) return run<pdrd::PGLDA>(N, tol, max_iter);
    if (!std::strcmp(s, "b")) return run<pdrd::PGB>(N, tol, max_iter);
    if (!std::strcmp(s, "psi")) return run<pdrd::PSI>(N, tol, max_iter);
    return run<pdrd::PGN>(N, tol, max_iter);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
}
