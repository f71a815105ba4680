"""Pins of the oracle's rotating advection (RotationAdvection2D with
AroundOrigin2D: velocity (y, -x) at every point, evaluated at the P+1 flux
nodes for the fluxes and at the quadrature points for the upwind
parameters):

* every function of the radius is steady (div (a u) = u div a + a . grad u
  = 0): u = x^2 + y^2 on P2 triangles (exactly represented, its flux
  x^2 + y^2 times a linear velocity is cubic, exact in the P3 flux element)
  has zero LDA and B residuals (the N dissipation of a non-constant state
  does not vanish), a constant on P1 zero residuals in every scheme;
* u = x: div(a u) = y, so the residuals of all nodes (times their dual
  volumes) add up to the integral of y over the domain, +1/2 on the unit
  square for every scheme -- a velocity of the wrong sense gives -1/2, a
  constant or mis-oriented one neither;
* the problem is linear: the implicit forms' Jacobian times u reproduces the
  LDA and N residuals.
"""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.mesh import rectangle, unit_quad_triag
from tests.helpers import oracle_problem, quad_orders

ROT = L.ROTATION2D


def _mesh(order):
    if order == 1:
        return unit_quad_triag(11)
    return rectangle(7, 6, split=2, order=order, jitter=0.1, seed=5)


@pytest.mark.parametrize("order,kind,scheme", [(2, "radial", L.LDA), (2, "radial", L.B), (3, "radial", L.LDA),
                                               (1, "constant", L.LDA), (1, "constant", L.N), (1, "constant", L.B),
                                               (1, "constant", L.PSI)])
def test_radial_states_are_steady(order, kind, scheme):
    mesh = _mesh(order)
    x, y = mesh.node_coords[:, 0], mesh.node_coords[:, 1]
    u = (x * x + y * y) if kind == "radial" else np.full_like(x, 0.7)
    P = oracle_problem(mesh, ROT, scheme, (), quad_orders(mesh))
    r, ws, rc = P.assemble_rhs(u[:, None])
    assert rc == 0
    ref, _, _ = P.assemble_rhs(x[:, None])  # a state that is not steady
    assert np.abs(ref).max() > 1e-2
    assert np.abs(r).max() <= 1e-12 * np.abs(ref).max()
    assert ws.min() > 0


@pytest.mark.parametrize("scheme", [L.LDA, L.N, L.B, L.PSI])
@pytest.mark.parametrize("order", [1, 2])
def test_total_residual_of_u_equal_x(scheme, order):
    mesh = unit_quad_triag(9, order=order)
    x = mesh.node_coords[:, 0]
    P = oracle_problem(mesh, ROT, scheme, (), quad_orders(mesh))
    r, _, rc = P.assemble_rhs(x[:, None])
    assert rc == 0
    total = float((r[:, 0] * P.dual_volume()).sum())
    assert abs(total - 0.5) <= 1e-12


@pytest.mark.parametrize("scheme", [L.LDA, L.N])
def test_residual_is_jacobian_times_state(scheme):
    mesh = _mesh(2)
    u = np.random.default_rng(3).uniform(-1, 1, (mesh.nb_nodes, 1))
    P = oracle_problem(mesh, ROT, scheme, (), quad_orders(mesh))
    phi, jac, wave, rc = P.cell_jacobians(0, u)
    assert rc == 0
    ue = u[mesh.blocks[0].sol_conn, 0]
    pr, _, _ = P.cell_residuals(0, u)
    assert np.abs(np.einsum("eij,ej->ei", jac, ue) - phi.reshape(ue.shape)).max() <= 1e-12 * np.abs(phi).max()
    if scheme == L.LDA:
        assert np.abs(phi - pr).max() <= 1e-12 * np.abs(pr).max()
