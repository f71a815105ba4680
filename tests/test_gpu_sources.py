"""Source fields on the GPU (set_vec_function(sources): the reaction
residuals, div F - s_q at every quadrature point) against the oracle,
<= 1e-12 normwise relative: one sweep of every parity mesh and scheme with a
seeded random source field, 10 iterations with strong Dirichlet inflow
values (chained and unchained sweeps bit for bit alike), a linear
manufactured steady state of advection whose residual vanishes, the source
switched off again, the source beside the B scheme's external blending
coefficient and beside the LDA scheme's artificial viscosity (which its
reaction residuals do not apply), and the refused combinations."""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.mesh import unit_quad_triag
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import CASES, oracle_problem, quad_orders, rel_err

pytestmark = pytest.mark.gpu

NAMES = ["adv_p1", "adv_p2", "adv_p3", "adv_q2", "burgers_p2", "burgers_q1", "burgers_q2_mixed", "euler2d_mixed",
         "euler2d_p2", "euler2d_q2_mixed", "euler3d_p1", "euler3d_p2"]


def _setup(name, scheme):
    build, physics, params, state = CASES[name]
    mesh = build()
    u = state(mesh)
    m = PGRDMethodExplicit(physics, scheme, params=params if params else None)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    m.set_vec_function("solution", u)
    P = oracle_problem(mesh, physics, scheme, params, quad_orders(mesh))
    src = np.random.default_rng(17).uniform(-0.5, 0.5, u.shape) * np.abs(u).max(0)
    return mesh, u, m, P, src


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("scheme", [L.LDA, L.N, L.B, L.PSI])
def test_source_sweep_matches_oracle(name, scheme):
    print(f"sources {name} scheme {scheme}", flush=True)
    mesh, u, m, P, src = _setup(name, scheme)
    m.set_vec_function("sources", src)
    m.assemble_rhs()
    m.check_status()
    P.set_sources(src)
    ref, ref_ws, rc = P.assemble_rhs(u)
    assert rc == 0
    assert rel_err(m.vec_function("residuals"), ref) <= 1e-12
    assert rel_err(m.vec_function("wave_speed"), ref_ws) <= 1e-12
    # switched off: the plain residual, bit for bit
    m.set_sources(None)
    m.assemble_rhs()
    plain = m.vec_function("residuals")
    m.destroy()
    _, _, m2, _, _ = _setup(name, scheme)
    m2.assemble_rhs()
    assert np.array_equal(plain, m2.vec_function("residuals"))
    m2.destroy()


@pytest.mark.parametrize("chain", [0, 1])
def test_source_iterations_match_oracle(chain):
    mesh = unit_quad_triag(33)
    x, y = mesh.node_coords[:, 0], mesh.node_coords[:, 1]
    src = (0.5 + np.sin(3 * x) * np.cos(2 * y))[:, None]
    bn = np.union1d(mesh.boundaries["Edge0"], mesh.boundaries["Edge3"]).astype(np.uint32)
    g = np.where(mesh.node_coords[bn, 1] < 0.5, 1.0, 0.0)[:, None]
    m = PGRDMethodExplicit(L.ADVECTION2D, L.N, params=(0.8, 0.6))
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=4)
    m.set_option(L.OPT_CHAIN, chain)
    m.set_vec_function("solution", np.zeros((mesh.nb_nodes, 1)))
    m.add_boundary_condition("StrongDirichlet", "inflow", bn, g)
    m.set_vec_function("sources", src)
    norm = m.iterate(0.9, 10)
    P = oracle_problem(mesh, L.ADVECTION2D, L.N, (0.8, 0.6))
    P.set_sources(src)
    uo, res, dt, norms, rc = P.iterate(np.zeros(mesh.nb_nodes), 0.9, 10, bn, g)
    assert rc == 0
    assert rel_err(m.vec_function("solution"), uo) <= 1e-12
    assert abs(norm[0] - norms[-1, 0]) <= 1e-12 * norms[-1, 0]
    m.destroy()


def test_manufactured_steady_state_on_the_gpu():
    mesh = unit_quad_triag(17)
    x, y = mesh.node_coords[:, 0], mesh.node_coords[:, 1]
    u = (0.3 + 1.1 * x - 0.7 * y)[:, None]
    m = PGRDMethodExplicit(L.ADVECTION2D, L.LDA, params=(0.8, 0.6))
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=4)
    m.set_vec_function("solution", u)
    m.assemble_rhs()
    r0 = np.abs(m.vec_function("residuals")).max()
    m.set_vec_function("sources", np.full((mesh.nb_nodes, 1), 0.8 * 1.1 - 0.6 * 0.7))
    m.assemble_rhs()
    assert r0 > 1e-2 and np.abs(m.vec_function("residuals")).max() <= 1e-12 * r0
    m.destroy()


@pytest.mark.parametrize("name", ["burgers_p2", "euler2d_mixed", "adv_p2", "euler3d_p1", "euler3d_p2"])
def test_source_with_blending_coeff_matches_oracle(name):
    """The B scheme's reaction residuals with an external blending coefficient
    (PGB.hpp:825-1050, skip below 1e-6 at :878): one sweep against the oracle,
    with a coefficient that is partly below the skip level."""
    print(f"sources + theta {name}", flush=True)
    mesh, u, m, P, src = _setup(name, L.B)
    theta = np.random.default_rng(5).uniform(0.0, 1.0, mesh.nb_nodes)
    theta[::3] = 0.0
    m.set_vec_function("sources", src)
    m.set_blending_coeff(theta)
    m.assemble_rhs()
    m.check_status()
    P.set_sources(src)
    P.set_blending_coeff(theta)
    ref, ref_ws, rc = P.assemble_rhs(u)
    assert rc == 0
    assert rel_err(m.vec_function("residuals"), ref) <= 1e-12
    assert rel_err(m.vec_function("wave_speed"), ref_ws) <= 1e-12
    m.destroy()


def test_source_with_blending_coeff_iterations_match_oracle():
    mesh = unit_quad_triag(33)
    x, y = mesh.node_coords[:, 0], mesh.node_coords[:, 1]
    src = (0.5 + np.sin(3 * x) * np.cos(2 * y))[:, None]
    theta = 0.5 + 0.5 * np.sin(5 * x + y)
    bn = np.union1d(mesh.boundaries["Edge0"], mesh.boundaries["Edge3"]).astype(np.uint32)
    g = np.where(mesh.node_coords[bn, 1] < 0.5, 1.0, 0.0)[:, None]
    m = PGRDMethodExplicit(L.ADVECTION2D, L.B, params=(0.8, 0.6))
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=4)
    m.set_vec_function("solution", np.zeros((mesh.nb_nodes, 1)))
    m.add_boundary_condition("StrongDirichlet", "inflow", bn, g)
    m.set_blending_coeff(theta)
    m.set_vec_function("sources", src)
    m.iterate(0.9, 10)
    P = oracle_problem(mesh, L.ADVECTION2D, L.B, (0.8, 0.6))
    P.set_sources(src)
    P.set_blending_coeff(theta)
    uo, res, dt, norms, rc = P.iterate(np.zeros(mesh.nb_nodes), 0.9, 10, bn, g)
    assert rc == 0
    assert rel_err(m.vec_function("solution"), uo) <= 1e-12
    m.destroy()


@pytest.mark.parametrize("name", ["adv_p2", "euler2d_mixed"])
def test_source_ignores_artificial_viscosity(name):
    """The LDA scheme's reaction residuals apply no artificial viscosity
    (PGLDA.hpp:553-741): with a source field the viscosity changes nothing,
    bit for bit, and the result is the oracle's."""
    mesh, u, m, P, src = _setup(name, L.LDA)
    nu = np.random.default_rng(6).uniform(0.0, 0.05, mesh.nb_nodes)
    m.set_vec_function("sources", src)
    m.assemble_rhs()
    r_src = m.vec_function("residuals")
    m.set_artificial_viscosity(nu)
    m.assemble_rhs()
    assert np.array_equal(m.vec_function("residuals"), r_src)
    P.set_sources(src)
    P.set_artificial_viscosity(nu)
    ref, _, rc = P.assemble_rhs(u)
    assert rc == 0
    assert rel_err(r_src, ref) <= 1e-12
    m.destroy()


def test_refused_combinations():
    mesh, u, m, P, src = _setup("burgers_p2", L.B)
    m.set_vec_function("sources", src)
    with pytest.raises(L.PdrdError):
        m.set_option(L.OPT_SCATTER, 1)
    m.set_sources(None)
    m.set_option(L.OPT_SCATTER, 1)
    with pytest.raises(L.PdrdError):
        m.set_vec_function("sources", src)
    m.destroy()
