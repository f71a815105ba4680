/*
 * pdekit_rd.h -- C ABI of the MI355X residual-distribution cell loop.
 *
 * The reference's plug-in surface for this path is the C++ class template
 *   PGRDMethodExplicit<MeshConfig, Physics, SchemeTag>
 *     (src/solver/rdm/PGRDMethodExplicit.hpp:19-193, virtual API in
 *      src/solver/rdm/RDMethod.hpp:33-139)
 * with the element scheme selected by CellSchemeSelector<Physics, Tag>
 *     (src/solver/rdm/cellsplitters/CellSchemeSelector.hpp:17; PGLDA.hpp:746,
 *      PGN.hpp:673-676, PGB.hpp:1071).
 * Every entry point below names the member it replaces. Plain pointers and
 * sizes only; all arrays are row-major, nodal fields are AoS [N][m] doubles
 * as in VectorMeshFunction (src/interpolation/mesh_function/
 * VectorMeshFunction.hpp:145-249). Every call returns PDRD_OK (0) or a
 * negative status; pdrd_last_error() describes the last failure of the
 * calling thread. One handle serves one in-flight call at a time
 * (stream-ordered); distinct handles may be used concurrently.
 */
#ifndef PDEKIT_RD_H
#define PDEKIT_RD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
enum pdrd_status {
  PDRD_OK = 0,
  PDRD_EINVAL = -1,        /* bad argument / size / state */
  PDRD_ENOMEM = -2,        /* device allocation failed */
  PDRD_EHIP = -3,          /* HIP runtime error */
  PDRD_ENOTSUP = -4,       /* combination not built */
  PDRD_ENEGPRESSURE = -5,  /* non-positive pressure met in a cell (the reference
                              prints and continues, Euler2DCons.hpp:188-201) */
  PDRD_ENONFINITE = -6     /* NaN/Inf in the residual */
};

/* ---- enumerations ------------------------------------------------------- */
enum pdrd_shape { PDRD_TRIAG = 0, PDRD_QUAD = 1, PDRD_TETRA = 2 };
enum pdrd_point_set { PDRD_EQUIDIST = 0, PDRD_WARPBLEND = 1 };
/* Physics template argument of PGRDMethodExplicit:
 * ADVECTION2D = RotationAdvection2D with a constant velocity (params[0..1]),
 * BURGERS2D = Burgers2D, EULER2D = Euler2DCons, EULER3D = Euler3DCons
 * (params[0] = gamma, params[1] = R). */
/* PDRD_ADVECTION2D: RotationAdvection2D with a constant velocity (params ax,
 * ay; ConstAdvection2D is (0, 1)); PDRD_ROTATION2D: RotationAdvection2D with
 * AroundOrigin2D, the velocity (y, -x) at every point (RotationAdvection2D.hpp:
 * 33-45, 201-224; no parameters; not with mixed precision or weak BCs). */
enum pdrd_physics {
  PDRD_ADVECTION2D = 0,
  PDRD_BURGERS2D = 1,
  PDRD_EULER2D = 2,
  PDRD_EULER3D = 3,
  PDRD_ROTATION2D = 4
};
/* SchemeTag: PGLDA, PGN, PGB; PSI = N residuals limited with the x+ limiter
 * of LFResidualLimiter.hpp:84-110; BPSI = the PSI / LDA blend
 * (1 - theta) LDA + theta PSI with PGB's cell theta (equation 0 of the N
 * residuals, PGB.hpp:437-462). PSI and BPSI have no implicit form. */
enum pdrd_scheme { PDRD_LDA = 0, PDRD_N = 1, PDRD_B = 2, PDRD_PSI = 3, PDRD_BPSI = 4 };
/* Nodal fields held by a method (RDMethod::vec_function / scal_function). */
enum pdrd_field {
  PDRD_SOLUTION = 0,   /* [N][m] */
  PDRD_RESIDUAL = 1,   /* [N][m] */
  PDRD_TIME_STEP = 2,  /* [N]    */
  PDRD_WAVE_SPEED = 3, /* [N]    sum of element wave speeds (RDTimeUpdate) */
  PDRD_DUAL_VOLUME = 4, /* [N]   */
  PDRD_BLEND_COEFF = 5, /* [N]   nodal blending coefficient of the B scheme (once set or computed) */
  PDRD_ART_VISC = 6     /* [N]   nodal artificial viscosity of the LDA scheme (once set or computed) */
};
/* Options of a method. */
enum pdrd_option {
  PDRD_OPT_SCATTER = 0,   /* 0 = FP64 atomics, 1 = colour-by-colour launches,
                             2 = patch assembly (default): Morton-ordered patches of
                             cells summed in LDS, shared nodes summed by a fix-up pass;
                             1 and 2 are deterministic */
  PDRD_OPT_PRECISION = 1, /* 0 = plain FP64, 1 = mixed (patch assembly only; not PSI, not
                             in batches): a linearised defect
                             correction. An FP64 sweep sets the linearisation point u_ref
                             with its FP64 residual and time step; the sweeps in between are
                             linear, res = R(u_ref) + sum_e J_e (u - u_ref)_e / V with the
                             element Jacobians J_e of the implicit form in FP32 and J delta
                             in FP32, dt = dt(u_ref); nodal sums, update and norms in FP64.
                             Its fixed point is the FP64 steady state. The last iteration of
                             every pdrd_iterate call is an FP64 sweep (FP64 outputs and norm);
                             pdrd_assemble_rhs / _lhs_and_rhs / solve stay FP64 */
  PDRD_OPT_BLOCK = 2,     /* threads per block of the cell kernels (64, 128, 192, 256) */
  PDRD_OPT_TIMING = 3,    /* 1: pdrd_iterate runs eagerly and times the cell loop and the
                             nodal pass of every iteration with HIP events on the method's
                             stream; pdrd_last_timing then returns per-iteration means;
                             2: the same over the linear sweeps of the mixed form only */
  PDRD_OPT_PATCH_CELLS = 4, /* cells per patch workgroup, set before pdrd_initialize: 0 (each
                             kernel's default: 256 scalar P1/P2, 128 otherwise), 64, 128 or
                             256 (capped at the default); read back after pdrd_initialize
                             it gives the largest in use */
  PDRD_OPT_REFRESH = 5,    /* mixed precision: every this many iterations (default 50) an FP64
                             sweep with fresh element Jacobians; 0 or 1: every sweep FP64 (the
                             plain iteration, bit for bit) */
  PDRD_OPT_CHAIN = 6       /* patch assembly, pdrd_iterate: 1 = the shared nodes of a sweep
                             are summed and stepped in the next sweep's gather (one launch
                             per element type and sweep; bitwise the same results), 0 = a
                             fix-up pass after every sweep; -1 (default) = 1 for the scalar
                             physics, 0 for the Euler systems (as measured faster) */
};

/* ---- FE inputs (src/interpolation, src/mesh/point_set, shape_function) -- */
/* Sizes of one element type: solution order sol_order on a geometry of
 * order geo_order, P+1 flux element, Gauss quadrature of order quad_order. */
typedef struct pdrd_fe_desc {
  int shape, geo_order, sol_order, quad_order, point_set;
  int dim, ng, nd, nf, nq;
  size_t table_size; /* doubles in the packed table */
} pdrd_fe_desc;

/* FunctionSpaceCells::set_reference_fe_values + FluxSpaceMetric::
 * allocate_buffer (FunctionSpaceCells.hpp:135-190, FluxSpaceMetric.hpp:237-456,
 * FEValues.cpp:157-279). */
int pdrd_fe_describe(int shape, int geo_order, int sol_order, int quad_order, int point_set,
                     pdrd_fe_desc *out);
/* Packed tables, concatenated in this order (row-major):
 * qp[nq][dim], w[nq], Vg[nq][ng], dVg[dim][nq][ng], V[nq][nd], dV[dim][nq][nd],
 * Vf[nq][nf], dVf[dim][nq][nf], Vgf[nf][ng], Vsf[nf][nd], sol_nodes[nd][dim],
 * flux_nodes[nf][dim]. */
int pdrd_fe_fill(const pdrd_fe_desc *desc, double *tables);
/* Reference nodes and Gauss rules on their own. */
int pdrd_reference_nodes(int shape, int order, int point_set, double *xi, int *n);
int pdrd_quadrature(int shape, int order, double *xi, double *w, int *nq);

/* ---- mesh inputs (src/mesh/io/MeshCreator, src/mesh/containers/DofMap) -- */
typedef struct pdrd_meshgen *pdrd_meshgen_t;
/* kind: 0 = MeshCreator::make_unit_quad(N = i0) on [-1,1]^2,
 *       1 = rectangle nx=i0, ny=i1 on [r0,r1]x[r2,r3], split i2 (1,2,4),
 *           quad_rows i3,
 *       2 = box tetrahedra nx=i0, ny=i1, nz=i2 on [0,r0]x[0,r1]x[0,r2]. */
int pdrd_mesh_generate(int kind, const int64_t *iparams, const double *rparams,
                       pdrd_meshgen_t *out);
int pdrd_mesh_free(pdrd_meshgen_t m);
/* seeded jitter of non-boundary vertices: U(-f h_d, f h_d) per axis */
int pdrd_mesh_jitter(pdrd_meshgen_t m, const double *h, double f, uint64_t seed);
/* mode 0: 2D bump y -> yb + y (H - yb)/H, yb = amp exp(-k (x-xc)^2);
 * mode 1: 3D bump in z with centre (xc, yc). p = {H, amp, k, xc, yc}. */
int pdrd_mesh_bump(pdrd_meshgen_t m, int mode, const double *p);
/* DofMap::clone_continuous to order P (DofMap.hpp:2225-2290, 2556-3073). */
int pdrd_mesh_clone(pdrd_meshgen_t m, int order, int point_set);
/* counts: dim, nb_vertices, nb_cells, nb_nodes (after clone), nb_boundaries */
int pdrd_mesh_info(pdrd_meshgen_t m, int64_t *info5);
int pdrd_mesh_coords(pdrd_meshgen_t m, double *coords);          /* [nv][dim] */
int pdrd_mesh_node_coords(pdrd_meshgen_t m, double *coords);     /* [N][dim]  */
int pdrd_mesh_cell_shapes(pdrd_meshgen_t m, int32_t *shapes);    /* [ncells]  */
/* geometry / solution connectivity, packed cell after cell */
int pdrd_mesh_geo_conn(pdrd_meshgen_t m, uint32_t *conn, int64_t *size);
int pdrd_mesh_sol_conn(pdrd_meshgen_t m, uint32_t *conn, int64_t *size);
int pdrd_mesh_boundary_name(pdrd_meshgen_t m, int idx, char *name, int cap);
int pdrd_mesh_boundary_nodes(pdrd_meshgen_t m, const char *name, uint32_t *nodes, int64_t *n);

/* ---- Gmsh 2.2 ASCII (src/mesh/io/gmsh/GmshReader.hpp, GmshReader.cpp:60-243,
 * GmshWriter.hpp:145-330) -------------------------------------------------- */
/* GmshReader::read_mesh_from_file: cells are the elements of the highest
 * dimension (triangles, quadrilaterals, tetrahedra of order 1-3, file order);
 * vertices are the file's nodes (ids 1..n); high-order cells must be
 * straight-sided and are kept by their corners (pdrd_mesh_clone then gives
 * the solution nodes in Gmsh's local order); boundaries are the physical
 * groups of dimension dim - 1, named by $PhysicalNames. */
int pdrd_mesh_read_gmsh(const char *path, pdrd_meshgen_t *out);
/* GmshWriter::write_mesh_to_file(mesh, dofs, file) of the cloned solution
 * nodes: $Nodes, boundary facets and cells in the element type of the
 * solution order (P1-P3 in 2D, P1-P2 tetrahedra). */
int pdrd_mesh_write_gmsh(pdrd_meshgen_t m, const char *path);
/* GmshWriter::append_nodal_function_to_file: values[n][m] as m $NodeData
 * blocks "<field>_<k>" (scalar != 0 and m == 1: "<field>"). */
int pdrd_gmsh_append_node_data(const char *path, const char *field, int m, uint32_t n, const double *values,
                               double time, uint32_t step, int scalar);
/* GmshReader::read_nodal_function_from_file: the $NodeData blocks named
 * names[f] (without quotes) fill column f of values[n][nfields]; *found =
 * how many of the names were present. */
int pdrd_gmsh_read_node_data(const char *path, int nfields, const char *const *names, uint32_t n, double *values,
                             int *found);

/* ---- the method --------------------------------------------------------- */
typedef struct pdrd_method *pdrd_method_t;

/* PGRDMethodExplicit() + RDMethod::configure_mesh_data(mesh, geo, sol)
 * (RDMethod.hpp:158-173): vertex coordinates [nv][dim] and the number of
 * solution nodes. The method lives on the current HIP device. */
int pdrd_create(int dim, uint32_t nb_vertices, const double *vertex_coords, uint32_t nb_nodes,
                int physics, int scheme, const double *params, int nb_params, pdrd_method_t *out);
/* One block of cells of one element type (the per-type DataMap entries of
 * PGRDMExplicitCellWorker::configure_cell_spaces, PGRDMExplicitCellWorker.hpp:178-284).
 * geo_conn [ne][ng], sol_conn [ne][nd]; tables as pdrd_fe_fill writes them. */
int pdrd_add_cells(pdrd_method_t h, const pdrd_fe_desc *fe, const double *tables, uint32_t ne,
                   const uint32_t *geo_conn, const uint32_t *sol_conn);
/* initialize_work_data (PGRDMethodExplicit.hpp:231-299) + RDTimeUpdate::setup
 * (RDTimeUpdate.hpp:71-194): uploads the mesh, builds the colouring and the
 * dual volumes. */
int pdrd_initialize(pdrd_method_t h);
int pdrd_destroy(pdrd_method_t h);
int pdrd_set_option(pdrd_method_t h, int option, int value);
/* The value of an option; PDRD_OPT_SCATTER gives the assembly in use (the
 * patch assembly falls back to 1 at pdrd_initialize when a patch does not fit
 * its limits). */
int pdrd_get_option(pdrd_method_t h, int option, int *value);
/* HIP stream (hipStream_t) the method enqueues on; NULL = the method's own. */
int pdrd_set_stream(pdrd_method_t h, void *stream);
int pdrd_synchronize(pdrd_method_t h);

/* set_vec_function / vec_function: host <-> device copies of a field. */
int pdrd_set_field(pdrd_method_t h, int field, const double *host);
int pdrd_get_field(pdrd_method_t h, int field, double *host);
/* Device address of a field for callers that keep data resident. */
int pdrd_field_device_ptr(pdrd_method_t h, int field, double **dptr);

/* add_boundary_condition("WeakWall" | "WeakSubInlet" | "WeakSubOutlet" |
 * "WeakFarfield" | "WeakSymmetry" | "WeakSuperInlet" | "WeakDirichlet", name,
 * boundary) (src/solver/rdm/bc/WeakWall.hpp, WeakSubInlet.hpp,
 * WeakSubOutlet.hpp, WeakFarfield.hpp, WeakSymmetry.hpp, WeakSuperInlet.hpp,
 * WeakDirichletBC.hpp over WeakBC::apply, WeakBC.hpp:225-352): the boundary is
 * given by its solution nodes; its facets are the cell facets on the domain
 * boundary whose nodes all belong to it (triangles P1/P2, quadrilaterals Q1,
 * tetrahedra P1). WeakDirichlet applies to every physics, the others to the
 * Euler equations. params: WeakSubInlet total temperature, total pressure;
 * WeakSubOutlet pressure; WeakFarfield and WeakDirichlet the boundary state
 * (m values); none for the others. quad_order: the method's quadrature order
 * (the facet rule has order quad_order + 1). Added before pdrd_initialize;
 * applied in every sweep after the cells, with the patch assembly. */
enum pdrd_weak_bc {
  PDRD_WEAK_WALL = 0,
  PDRD_WEAK_SUBINLET = 1,
  PDRD_WEAK_SUBOUTLET = 2,
  PDRD_WEAK_FARFIELD = 3,
  PDRD_WEAK_SYMMETRY = 4,
  PDRD_WEAK_SUPERINLET = 5,
  PDRD_WEAK_DIRICHLET = 6
};
int pdrd_add_weak_bc(pdrd_method_t h, int type, uint32_t n, const uint32_t *nodes, const double *params,
                     int nparams, int quad_order);
/* After pdrd_initialize: facets, nodes per facet, facet quadrature points of
 * weak BC k (in the order added). */
int pdrd_weak_bc_info(pdrd_method_t h, int k, int64_t *info3);
/* WeakDirichlet with an expression of the point (set_expression,
 * WeakDirichletBC.hpp:149-154, 216-229): the coordinates of weak BC k's
 * facet points xq[nf][nq][dim] (after pdrd_initialize; nf, nq from
 * pdrd_weak_bc_info), and the boundary state at each of them,
 * state[nf][nq][m] (count = nf * nq * m doubles, checked), which then
 * replaces the constant one (NULL restores it). The caller evaluates its
 * expression at the points; the reference's expressions may also read the
 * state, which a precomputed table cannot. */
int pdrd_weak_bc_points(pdrd_method_t h, int k, double *xq);
int pdrd_set_weak_bc_state(pdrd_method_t h, int k, const double *state, size_t count);

/* Nodal inputs of the cell loop (PGRDMExplicitCellWorker.hpp:307-330, 501-533).
 * set_blending_coeff(f): PGB takes theta_i = f[node i] instead of its
 * cell-wise theta, Phi_i = lda_i + theta_i n_i, and keeps the LDA residuals
 * of a cell whose nodes are all below 1e-6 (PGB.hpp:141-148, 467-477).
 * set_artificial_viscosity(nu): PGLDA adds w |J| nu_q grad(phi_i) . grad(u)
 * at every quadrature point, nu_q interpolated from the nodes like the state
 * (PGLDA.hpp:245-256). theta / nu: host [N], copied to the device; NULL
 * switches the input off. Other schemes ignore them; both run outside
 * batches. */
int pdrd_set_blending_coeff(pdrd_method_t h, const double *theta);
int pdrd_set_artificial_viscosity(pdrd_method_t h, const double *nu);
/* set_vec_function(SolverVecFn::sources, f) (PGRDMethodExplicit.hpp:350-356):
 * a nodal source field [N][m] (host, copied; NULL switches it off). With it
 * the cells compute the reaction residuals (compute_adv_reaction_residuals,
 * PGRDMExplicitCellWorker.hpp:535-544): the residual at a quadrature point
 * is div F - s_q with s_q the source interpolated like the solution
 * (PGLDA.hpp:553-741, PGN.hpp:470-660, PGB.hpp:825-1050). Runs with the
 * patch assembly, in batches too; with the B scheme an external blending
 * coefficient may accompany it (PGB.hpp:825-1050 takes both), an artificial
 * viscosity set beside it is not applied (the reference's reaction LDA has
 * no shock-capturing term). pdrd_element_jacobians ignores it, as
 * compute_adv_res_and_jacobian does. */
int pdrd_set_sources(pdrd_method_t h, const double *sources);
/* Providers, on the device from the current solution, each storing its
 * result as the method's input (and switching it on):
 * RDBlendingCoeff::calculate with "LF_Blend" (LFBlendingCoeffImpl.hpp:79-136
 * + MeshFunctionTools::interp_from_cells_to_nodes_p1_max, RDBlendingCoeff.hpp:
 * 150-165); ArtificialViscosity::compute_artificial_viscosity_nodal(kappa,
 * factor) (ArtificialViscosity.hpp:371-560; triangles, tetrahedra and, with the
 * modes ModalExpansionQuad flags as leading, quadrilaterals). */
enum pdrd_blend_kind { PDRD_BLEND_LF = 0 };
int pdrd_compute_blending_coeff(pdrd_method_t h, int kind);
int pdrd_compute_artificial_viscosity(pdrd_method_t h, double kappa, double factor);

/* compute_adv_res_and_jacobian (PGLDA.hpp:315-505, PGN.hpp:294-400,
 * PGB.hpp:503-800): for every cell of block `block` (the order of its
 * pdrd_add_cells call) the implicit form's element residual phi[ne][nd][m],
 * the element Jacobian jac[ne][nd*m][nd*m] (row-major, row = node * m +
 * equation) and the element wave speed wave[ne], from the current solution.
 * Host buffers; a NULL one is not copied. PSI has no implicit form
 * (PDRD_ENOTSUP). */
int pdrd_element_jacobians(pdrd_method_t h, int block, double *phi, double *jac, double *wave);

/* add_boundary_condition("StrongDirichlet", ...) with precomputed values:
 * u[nodes[k]] = values[k][0..m) after every cell loop
 * (StrongDirichletBC::apply, StrongDirichletBC.hpp:202-233). */
int pdrd_add_strong_dirichlet(pdrd_method_t h, uint32_t n, const uint32_t *nodes, const double *values);

/* assemble_rhs (PGRDMethodExplicit.hpp:427-463): residual = sum Phi / V_dual,
 * wave speeds, then the boundary conditions. */
int pdrd_assemble_rhs(pdrd_method_t h);
/* assemble_lhs_and_rhs(CFL) (PGRDMethodExplicit.hpp:416-423): assemble_rhs +
 * RDTimeUpdate::compute_local_time_step (dt = CFL V / sum lambda). */
int pdrd_assemble_lhs_and_rhs(pdrd_method_t h, double cfl);
/* solve (PGRDMethodExplicit.hpp:467-477): u -= dt * res. */
int pdrd_solve(pdrd_method_t h);
/* compute_residual_norm (PGRDMethodExplicit.hpp:481-514): sqrt(sum (r dt)^2 / N). */
int pdrd_compute_residual_norm(pdrd_method_t h, double *norm);
/* norm_L2(residual) (MeshFunctionNorm.hpp:92-): sqrt(sum r^2 / n) per field;
 * exclude_dirichlet != 0 leaves out the strong-Dirichlet rows (n counts the
 * remaining nodes). */
int pdrd_residual_norm_L2(pdrd_method_t h, int exclude_dirichlet, double *norm);
/* n_iter times { assemble_lhs_and_rhs(cfl); solve() } on the device (the
 * driver loop of examples/testcases-rdm/Advection-Continuous-RDS-2D-Explicit.cpp:160-163);
 * norm receives norm_L2 of the last residual (exclude_dirichlet as above).
 * The patch assembly launches its sweeps directly (chained per
 * PDRD_OPT_CHAIN); the atomic / colour assemblies replay one captured HIP
 * graph per iteration. */
int pdrd_iterate(pdrd_method_t h, double cfl, int n_iter, int exclude_dirichlet, double *norm);

/* Status of the last cell loop: PDRD_OK or PDRD_ENEGPRESSURE. */
int pdrd_check_status(pdrd_method_t h);

/* Counts of the method: nb_nodes, neq, nb_blocks, total cells, nb_colours
 * (max over blocks), device bytes held. */
int pdrd_method_info(pdrd_method_t h, int64_t *info6);

/* Timing of the last pdrd_assemble_rhs (cell loop / nodal pass) or, with
 * PDRD_OPT_TIMING, the per-iteration means of the last pdrd_iterate, measured
 * with HIP events on the method's stream (ms). Without PDRD_OPT_TIMING,
 * pdrd_iterate reports its whole duration as cell_ms and 0 as node_ms. After
 * pdrd_element_jacobians, cell_ms is the Jacobian kernel's duration. */
int pdrd_last_timing(pdrd_method_t h, double *cell_ms, double *node_ms);

/* Batches (no counterpart in the reference, whose driver loops run one
 * problem at a time): n_iter iterations of each of `count` independent
 * methods, with the patches of every method of one element type / physics /
 * scheme in one launch per iteration and the shared-node sums of every method
 * with the same number of equations in another, launched directly on the
 * batch's stream (or, after pdrd_batch_set_graphs(b, 1), captured once per
 * call into a HIP graph replayed for all but the last iteration: measured
 * slower on MI355X, whose direct launches cost less than the replays; p99 of
 * the serving stream 2.67 s direct against 2.87 s from graphs). Each method advances
 * exactly as pdrd_iterate(methods[i], cfl[i], n_iter, exclude_dirichlet)
 * would (same kernels, same order of sums, so the same bits); methods with a
 * nodal input (blending coefficient, artificial viscosity) or a source field
 * join the launches of their own kernel. Methods must be initialised on the
 * batch's device with the patch assembly, in FP64 (PDRD_OPT_PRECISION 0);
 * norms receives the norms of the methods one after the other (neq doubles
 * each). */
typedef struct pdrd_batch *pdrd_batch_t;
int pdrd_batch_create(pdrd_batch_t *out);
int pdrd_batch_destroy(pdrd_batch_t b);
int pdrd_batch_set_graphs(pdrd_batch_t b, int on);
int pdrd_batch_iterate(pdrd_batch_t b, int count, const pdrd_method_t *methods, const double *cfl, int n_iter,
                       int exclude_dirichlet, double *norms);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The cell loop of pdrd_assemble_rhs with FP64 atomics
 * (PDRD_OPT_SCATTER = 0), without boundary conditions, over one block of
 * cells held in caller-owned device memory:
 *   tables        the packed FE tables of pdrd_fe_fill (fe->table_size doubles),
 *   geo_conn      [ng][ne] and sol_conn [nd][ne] (structure of arrays),
 *   vertex_coords [nb_vertices][dim], u [nb_nodes][m], dual_volume [nb_nodes]
 *                 (the PDRD_DUAL_VOLUME field of a method on the same mesh);
 * outputs residual [nb_nodes][m] = sum Phi / V and wave_speed [nb_nodes], and,
 * when status is not null, the status word at *status (device int: bit 0
 * non-positive pressure, bit 1 connectivity out of range -- such cells are
 * skipped). params / nb_params as pdrd_create. Stream-ordered on `stream`
 * (NULL: the null stream); returns once enqueued. Its scratch (FE tables in
 * the device layout, nodal accumulators, status) is one static device array
 * of PDRD_NOALLOC_SCRATCH_BYTES per device, set up at the first call on each
 * device: nb_nodes * (m + 1) must not exceed PDRD_NOALLOC_MAX_ACC doubles
 * (PDRD_EINVAL otherwise). Every call on a device shares that array, so calls
 * that overlap on one device race; a call clears what it needs first. */
#define PDRD_NOALLOC_SCRATCH_BYTES (32u << 20)
#define PDRD_NOALLOC_MAX_ACC ((PDRD_NOALLOC_SCRATCH_BYTES / 8u) - 8192u - 32u)
int pdrd_assemble_rhs_noalloc(int physics, int scheme, const double *params, int nb_params, const pdrd_fe_desc *fe,
                              const double *tables, uint32_t ne, const uint32_t *geo_conn, const uint32_t *sol_conn,
                              uint32_t nb_vertices, const double *vertex_coords, uint32_t nb_nodes, const double *u,
                              const double *dual_volume, double *residual, double *wave_speed, int *status,
                              void *stream);

const char *pdrd_last_error(void);
const char *pdrd_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PDEKIT_RD_H */
