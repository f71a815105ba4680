"""The cell loop's nodal inputs and their providers on the GPU against the
oracle, ≤ 1e-12 normwise relative:

* the external blending coefficient of the B scheme and the artificial
  viscosity of the LDA scheme (random nodal fields, one cell in eight with
  theta below 1e-6 everywhere), one sweep in every assembly (patch, colour,
  atomic) on scalar P1 / P2 / P3, Euler 2D P1 + Q1 and P2 + Q2, Euler 3D P1,
  and 10 iterations of a method that iterated before the input was set;
* the LF_Blend and smoothness-indicator providers computed on the device
  from the method's state (triangles P1-P3, quadrilaterals, tetrahedra).
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from pdekit_amd import _lib as L
from pdekit_amd.mesh import box_tetra, rectangle
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import euler_state, oracle_problem, quad_orders, rel_err

pytestmark = pytest.mark.gpu

SCATTER = {"patch": 2, "colour": 1, "atomic": 0}


def _scalar_state(mesh):
    x = mesh.node_coords
    return (np.tanh(8.0 * (x[:, 0] - 0.45 - 0.2 * x[:, 1])) + 1.3)[:, None].copy()


CASES = {
    "burgers_p1": (lambda: rectangle(10, 9, split=2, jitter=0.15, seed=2), L.BURGERS2D, (), _scalar_state),
    "burgers_p2": (lambda: rectangle(8, 7, split=2, order=2, jitter=0.15, seed=2), L.BURGERS2D, (), _scalar_state),
    "adv_p3": (lambda: rectangle(6, 6, split=2, order=3, jitter=0.1, seed=4), L.ADVECTION2D, (0.8, 0.6),
               _scalar_state),
    "euler2d_mixed": (lambda: rectangle(12, 6, extent=(-0.5, 1.5, 0.0, 1.0), split=2, quad_rows=3,
                                        bump=(1.0, 0.1, 25.0, 0.5)), L.EULER2D, (1.4, 287.05),
                      lambda m: euler_state(m, pert=0.05)),
    "euler2d_q2": (lambda: rectangle(8, 4, extent=(-0.5, 1.5, 0.0, 1.0), split=2, quad_rows=2, order=2,
                                     bump=(1.0, 0.1, 25.0, 0.5)), L.EULER2D, (1.4, 287.05),
                   lambda m: euler_state(m, pert=0.05)),
    "euler3d_p1": (lambda: box_tetra(4, 4, 3, extent=(1.0, 1.0, 0.5), jitter=0.1, seed=4), L.EULER3D,
                   (1.4, 287.05), lambda m: euler_state(m, dim=3, pert=0.05)),
    # the B scheme's two-pass quadrature of the P2 tetrahedra with its nodal input
    "euler3d_p2": (lambda: box_tetra(3, 3, 2, extent=(1.0, 1.0, 0.5), order=2, jitter=0.1, seed=4), L.EULER3D,
                   (1.4, 287.05), lambda m: euler_state(m, dim=3, pert=0.05)),
}


def _nodal(mesh, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(0.0, 1.0, mesh.nb_nodes)
    # every node of every eighth cell of the first block below 1e-6
    conn = mesh.blocks[0].sol_conn
    small = np.unique(conn[::8])
    v[small] = rng.uniform(0.0, 9e-7, small.size)
    return v


@pytest.mark.parametrize("scatter", sorted(SCATTER))
@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("scheme", [L.B, L.LDA])
def test_nodal_input_sweep(name, scheme, scatter):
    print(f"nodal input {name} scheme {scheme} scatter {scatter}", flush=True)
    build, physics, params, state = CASES[name]
    mesh = build()
    u = state(mesh)
    m = PGRDMethodExplicit(physics, scheme, params=params if params else None)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh), scatter=SCATTER[scatter])
    assert m.get_option(L.OPT_SCATTER) == SCATTER[scatter]
    P = oracle_problem(mesh, physics, scheme, params)
    v = _nodal(mesh, 3) * (1.0 if scheme == L.B else 0.05)
    if scheme == L.B:
        m.set_blending_coeff(v)
        P.set_blending_coeff(v)
    else:
        m.set_artificial_viscosity(v)
        P.set_artificial_viscosity(v)
    m.set_vec_function("solution", u)
    m.assemble_lhs_and_rhs(0.5)
    m.check_status()
    ref, ref_ws, rc = P.assemble_rhs(u)
    assert rc == 0
    assert rel_err(m.vec_function("residuals"), ref) <= 1e-12
    assert rel_err(m.vec_function("wave_speed"), ref_ws) <= 1e-12
    m.destroy()


@pytest.mark.parametrize("name", ["burgers_p2", "euler2d_mixed", "euler3d_p1"])
@pytest.mark.parametrize("scheme", [L.B, L.LDA])
def test_nodal_input_iterations_after_plain_ones(name, scheme):
    build, physics, params, state = CASES[name]
    mesh = build()
    u = state(mesh)
    m = PGRDMethodExplicit(physics, scheme, params=params if params else None)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    m.set_vec_function("solution", u)
    m.iterate(0.3, 3)  # plain iterations first: the input must reach the later ones
    m.set_vec_function("solution", u)
    P = oracle_problem(mesh, physics, scheme, params)
    v = _nodal(mesh, 5) * (1.0 if scheme == L.B else 0.05)
    (m.set_blending_coeff if scheme == L.B else m.set_artificial_viscosity)(v)
    (P.set_blending_coeff if scheme == L.B else P.set_artificial_viscosity)(v)
    m.iterate(0.3, 10)
    uo, _, _, _, rc = P.iterate(u, 0.3, 10)
    assert rc == 0
    assert rel_err(m.vec_function("solution"), uo) <= 1e-12
    # switched off again: the plain scheme
    (m.set_blending_coeff if scheme == L.B else m.set_artificial_viscosity)(None)
    m.set_vec_function("solution", u)
    m.iterate(0.3, 10)
    uo0, _, _, _, _ = oracle_problem(mesh, physics, scheme, params).iterate(u, 0.3, 10)
    assert rel_err(m.vec_function("solution"), uo0) <= 1e-12


PROVIDER_MESHES = {
    "tri_p1": lambda: rectangle(9, 8, split=2, jitter=0.15, seed=1),
    "tri_p2": lambda: rectangle(7, 6, split=2, order=2, jitter=0.1, seed=1),
    "tri_p3": lambda: rectangle(5, 5, split=2, order=3, jitter=0.1, seed=1),
    "mixed_q1": lambda: rectangle(9, 8, split=2, quad_rows=4, jitter=0.1, seed=1),
    "mixed_q2": lambda: rectangle(7, 6, split=2, quad_rows=3, order=2, jitter=0.1, seed=1),
    "tet_p1": lambda: box_tetra(4, 4, 3, jitter=0.1, seed=4),
    "tet_p2": lambda: box_tetra(3, 3, 2, order=2, jitter=0.1, seed=4),
}


@pytest.mark.parametrize("name", sorted(PROVIDER_MESHES))
def test_providers_match_oracle(name):
    print(f"providers {name}", flush=True)
    mesh = PROVIDER_MESHES[name]()
    physics, params = (L.BURGERS2D, ()) if mesh.dim == 2 else (L.EULER3D, (1.4, 287.05))
    x = mesh.node_coords
    if mesh.dim == 2:
        u = (1.0 + 0.4 * np.tanh(6 * (x[:, 0] - 0.5)) + 0.05 * np.sin(9 * x[:, 1]))[:, None]
    else:
        u = euler_state(mesh, dim=3, pert=0.0)
        u[:, 0] = 1.0 + 0.4 * np.tanh(6 * (x[:, 0] - 0.5)) + 0.05 * np.sin(9 * x[:, 1])
    m = PGRDMethodExplicit(physics, L.B, params=params if params else None)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    m.set_vec_function("solution", u)
    P = oracle_problem(mesh, physics, L.B, params)
    th = m.compute_blending_coeff("LF_Blend")
    ref = P.lf_blend(u)
    assert np.abs(th - ref).max() <= 1e-14
    nu = m.compute_artificial_viscosity(kappa=5.0, factor=1.5)
    nref = P.art_visc(u, kappa=5.0, factor=1.5)
    assert nref.max() > 0
    assert rel_err(nu, nref) <= 1e-12
