"""Manufactured smooth solution of the 2D Euler equations of the reference's
end-to-end test (utest-Euler-continuous-rdm.cpp:94-300, the table of
fields at 156-172), its source term (div F of the exact state, here by
complex-step differentiation of the fluxes), the constant initial state, the
density L2 error, and the oracle problem with the source field and
WeakDirichlet states at the facet points of the four edges."""
import numpy as np

from oracle import pyoracle as O
from pdekit_amd import _lib as L
from pdekit_amd.fe import element_tables
from tests.helpers import boundary_facets, facet_rule, oracle_problem

GAMMA = 1.4
QORDER = 4


def _primitive(x, y):
    pi = np.pi
    rho = 0.15 * np.sin(1.0 * pi * x) - 0.1 * np.cos(0.5 * pi * y) + 1.0
    u = 5.0 * np.sin(1.5 * pi * x) - 7.0 * np.cos(0.6 * pi * y) + 70.0
    v = -15.0 * np.cos(0.5 * pi * x) + 8.5 * np.sin(2.0 / 3.0 * pi * y) + 90.0
    p = 0.2e5 * np.cos(2.0 * pi * x) + 0.5e5 * np.sin(1.0 * pi * y) + 1.0e5
    return rho, u, v, p


def exact_state(x, y):
    rho, u, v, p = _primitive(x, y)
    return np.stack([rho, rho * u, rho * v, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)], -1)


def _fluxes(x, y):
    rho, u, v, p = _primitive(x, y)
    E = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)
    fx = np.stack([rho * u, rho * u * u + p, rho * u * v, (E + p) * u], -1)
    fy = np.stack([rho * v, rho * u * v, rho * v * v + p, (E + p) * v], -1)
    return fx, fy


def source(x, y, h=1e-30):
    """div F(U(x, y)) by complex steps: d f / dx = Im f(x + ih) / h."""
    fx, _ = _fluxes(x + 1j * h, y + 0j)
    _, fy = _fluxes(x + 0j, y + 1j * h)
    return (fx.imag + fy.imag) / h


def initial(n):
    rho, u, v, p = 1.0, 77.2, 82.6, 2.94e5
    return np.tile([rho, rho * u, rho * v, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)], (n, 1))


def density_l2_error(mesh, u):
    """L2 norm over the cells (rule of order 4) of the nodal density error."""
    x, y = mesh.node_coords[:, 0], mesh.node_coords[:, 1]
    err = u[:, 0] - exact_state(x, y)[:, 0]
    t = element_tables(L.TRIAG, 1, QORDER)
    blk = mesh.blocks[0]
    X = mesh.coords[blk.geo_conn]
    # |J| of the map from the reference triangle (area 2, weights summing to 2): area / 2
    det = np.abs((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1]) -
                 (X[:, 2, 0] - X[:, 0, 0]) * (X[:, 1, 1] - X[:, 0, 1])) / 4
    eq = err[blk.sol_conn] @ t.V.T  # [ne][nq]
    return float(np.sqrt((det[:, None] * t.w[None, :] * eq ** 2).sum()))


def oracle(mesh):
    P = oracle_problem(mesh, L.EULER2D, L.LDA, (GAMMA, 287.05), {L.TRIAG: QORDER})
    for e in ["Edge0", "Edge1", "Edge2", "Edge3"]:
        fn, fv, ins = boundary_facets(mesh, mesh.boundaries[e])
        w, V = facet_rule(2, fn.shape[1], QORDER + 1)
        xq = np.einsum("qv,fvd->fqd", V, mesh.coords[fv])  # P1 facets: the nodes are the vertices
        P.add_weak(O.WEAK_DIRICHLET, initial(1)[0], 1, fn, fv, ins, w, V,
                   qstate=exact_state(xq[..., 0], xq[..., 1]))
    x, y = mesh.node_coords[:, 0], mesh.node_coords[:, 1]
    P.set_sources(source(x, y))
    return P
