/*
 * CPU oracle for the residual-distribution cell loop -- TEST INFRASTRUCTURE.
 *
 * A plain FP64 restatement of pdekit's explicit continuous RD residual
 * assembly (PGRDMethodExplicit + PGRDMExplicitCellWorker + the PGLDA / PGN /
 * PGB element schemes + RDTimeUpdate + StrongDirichletBC), used only as the
 * checker of the GPU path (tests, smoke check, CPU baseline of the bench).
 * It shares no code with the GPU library; it receives the same mesh and FE
 * tables as inputs.
 */
#ifndef PDRD_ORACLE_H
#define PDRD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ORC_ROTATION2D: RotationAdvection2D<AroundOrigin2D>, the advection velocity
 * (y, -x) at every point (RotationAdvection2D.hpp:33-45, 201-224). */
enum { ORC_ADVECTION2D = 0, ORC_BURGERS2D = 1, ORC_EULER2D = 2, ORC_EULER3D = 3, ORC_ROTATION2D = 4 };
enum { ORC_LDA = 0, ORC_N = 1, ORC_B = 2, ORC_PSI = 3, ORC_BPSI = 4 };

/* One block of cells of a single element type. Tables row-major:
 * w[nq], dVg[dim][nq][ng], V[nq][nd], dV[dim][nq][nd], dVf[dim][nq][nf],
 * Vsf[nf][nd], Vgf[nf][ng]; connectivity geo[ne][ng], sol[ne][nd]. */
typedef struct orc_block
{
  int ng, nd, nf, nq;
  int vol_nq; /* dual-volume quadrature (order = cell order) */
  uint32_t ne;
  const double *w, *dVg, *V, *dV, *dVf, *Vsf, *Vgf;
  const double *vol_w, *vol_dVg; /* [vol_nq], [dim][vol_nq][ng] */
  const uint32_t *geo_conn, *sol_conn;
  const double *Vg; /* [nq][ng] geometry functions at the quadrature points */
} orc_block;

/* Weak boundary conditions: WeakWall, WeakSubInlet, WeakSubOutlet,
 * WeakFarfield, WeakSymmetry, WeakSuperInlet (Euler) and WeakDirichlet (any
 * physics, a constant boundary state). */
enum {
  ORC_WEAK_WALL = 0,
  ORC_WEAK_SUBINLET = 1,
  ORC_WEAK_SUBOUTLET = 2,
  ORC_WEAK_FARFIELD = 3,
  ORC_WEAK_SYMMETRY = 4,
  ORC_WEAK_SUPERINLET = 5,
  ORC_WEAK_DIRICHLET = 6
};

/* One weak BC over its boundary facets. Facet rule on the reference facet:
 * w[nq] (edge [-1,1] / triangle of area 2), V[nq][nfn] the values of the
 * facet's Lagrange nodes. The facet's normal and measure are computed here
 * from its vertices, oriented away from `inside` (a point of its cell). */
typedef struct orc_weak
{
  int type;
  double params[5]; /* SubInlet: total temperature, total pressure; SubOutlet: pressure;
                       Farfield / SuperInlet / Dirichlet: the boundary state */
  int order;        /* solution order (WeakWall's relaxation coefficient) */
  uint32_t nfacet;
  int nfn, nfv, nq;
  const uint32_t *nodes; /* [nfacet][nfn] solution nodes */
  const uint32_t *verts; /* [nfacet][nfv] vertices */
  const double *inside;  /* [nfacet][dim] */
  const double *w, *V;
  /* WeakDirichlet: the boundary state at every facet point [nfacet][nq][m]
   * (an expression of the point, WeakDirichletBC.hpp:216-229), or NULL for
   * the constant params */
  const double *qstate;
  /* the coordinates of every facet point [nfacet][nq][dim], or NULL; needed
   * by the rotating advection, whose velocity is the point's (y, -x) */
  const double *qpoints;
} orc_weak;

typedef struct orc_problem
{
  int dim, physics, scheme, nblocks;
  uint32_t nv, nn;
  const double *coords; /* [nv][dim] */
  double params[4];     /* advection: ax, ay; Euler: gamma, R */
  const orc_block *blocks;
  int nweak; /* weak BCs applied after the cell loop, in this order */
  const orc_weak *weak;
  /* set_blending_coeff: nodal blending coefficient [N] that PGB uses instead
   * of its cell-wise theta, or NULL; set_artificial_viscosity: nodal
   * artificial viscosity [N] of PGLDA's shock-capturing term, or NULL */
  const double *blend;
  const double *art_visc;
  /* set_vec_function(SolverVecFn::sources, f): nodal source field [N][m] or
   * NULL. With it the cell worker calls compute_adv_reaction_residuals
   * (PGRDMExplicitCellWorker.hpp:535-544), whose residual at a quadrature
   * point is div F - s_q with s_q = V . source (PGLDA.hpp:553-741, PGN.hpp:
   * 470-660, PGB.hpp:825-1050) and which applies no artificial viscosity.
   * The implicit forms (orc_cell_jacobians) take no source. */
  const double *sources;
} orc_problem;

int orc_neq(int physics);

/* RDTimeUpdate::setup: dual volume V_i = sum of the volumes of all cells that
 * hold node i. */
int orc_dual_volume(const orc_problem *p, double *vdual);

/* One assemble_rhs sweep: res[N][m] = sum_cells Phi/V_dual, ws[N] = sum of
 * element wave speeds. Returns 0, or -5 if a non-positive pressure was met. */
int orc_assemble_rhs(const orc_problem *p, const double *u, const double *vdual, double *res,
                     double *ws);

/* Per-cell residuals and wave speed of one block (before the scatter):
 * phi[ne][nd][m], wave[ne]. */
int orc_cell_residuals(const orc_problem *p, int block, const double *u, double *phi,
                       double *wave);

/* compute_adv_res_and_jacobian of every cell of one block (the implicit
 * forms of PGLDA / PGN / PGB): phi[ne][nd][m], jac[ne][nd*m][nd*m]
 * (row-major, row = node * m + equation), wave[ne]. Returns -1 for PSI. */
int orc_cell_jacobians(const orc_problem *p, int block, const double *u, double *phi, double *jac,
                       double *wave);

/* n explicit iterations of assemble_lhs_and_rhs(CFL) + solve() with strong
 * Dirichlet values g[nb][m] written into u after every cell loop
 * (StrongDirichletBC::apply). norms[n][m] receives norm_L2(residual) over all
 * nodes per iteration when non-null. On return u holds the new state and
 * res / dt the last residual and time step. */
int orc_iterate(const orc_problem *p, double cfl, int n_iter, uint32_t nb_dirichlet,
                const uint32_t *dirichlet_nodes, const double *dirichlet_values, double *u,
                double *res, double *dt, double *norms);

/* Providers of the two stabilisation inputs. Per block: p1w[nd][nvert], the
 * P1 / Q1 shape functions of the cell's vertices at its solution nodes
 * (vertices are the first nvert nodes). */
/* RDBlendingCoeff "LF_Blend" (LFBlendingCoeffImpl.hpp:79-136) followed by
 * MeshFunctionTools::interp_from_cells_to_nodes_p1_max: theta[N]. */
int orc_lf_blend(const orc_problem *p, const double *u, const double *const *p1w, const int *nvert,
                 double *theta);
/* The smoothness-indicator quadrature of one block: rule of order
 * max(2p, 1) with weights w[nq], nodal basis V[nq][nd], geometry derivatives
 * dVg[dim][nq][ng], and psi[nq][npm] an orthonormal basis (under the rule)
 * of the polynomials of degree < p on the reference simplex. */
typedef struct orc_av_rule
{
  int p, nq, npm;
  const double *w, *V, *dVg, *psi;
  /* quadrilaterals: hm[nq][nd], the modal terms the reference zeroes in
   * V_modal_p_m_1 evaluated at the points, applied to the nodal values
   * (u - u_{p-1} at q = sum_n hm[q][n] u_n); null on simplices */
  const double *hm;
} orc_av_rule;
/* ArtificialViscosity::compute_artificial_viscosity_nodal
 * (ArtificialViscosity.hpp:371-560): nu[N]. */
int orc_art_visc(const orc_problem *p, const double *u, const orc_av_rule *rules, const double *const *p1w,
                 const int *nvert, double kappa, double factor, double *nu);

/* orc_iterate with the cell loop split over `threads` threads, each summing
 * into its own nodal arrays (added in thread order afterwards): the
 * multi-core timing baseline. No norms. sweep_s (n_iter doubles, or null)
 * receives the steady_clock duration of every iteration (cell loop,
 * boundary conditions, dt and update). */
int orc_iterate_threads(const orc_problem *p, double cfl, int n_iter, uint32_t nb_dirichlet,
                        const uint32_t *dirichlet_nodes, const double *dirichlet_values, double *u, double *res,
                        double *dt, int threads, double *sweep_s);

/* norm_L2 per field: sqrt(sum_i r_i^2 / N) (MeshFunctionNorm.hpp). */
void orc_norm_L2(const double *f, uint32_t n, int m, double *out);

/* Small dense helpers exposed for the invariant tests. */
int orc_invert(int n, const double *a, double *inv);
/* Euler eigen-structure at state u, direction g: R, L, D (diagonal) */
int orc_euler_eigen(int dim, double gamma, const double *u, const double *g, double *R, double *L,
                    double *D);
/* Closed-form K+ (PGLDA's build_K_mat with max(0, .)) */
int orc_euler_kplus(int dim, double gamma, const double *u, const double *g, double *K, double *eig);

#ifdef __cplusplus
}
#endif

#endif
