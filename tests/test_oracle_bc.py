"""Pins of the oracle's weak boundary conditions (WeakWall, WeakSubInlet,
WeakSubOutlet of the Euler equations) by closed forms on straight
boundaries with a uniform state, where the facet integral is exact:

* the boundary part of V_dual * residual summed over a boundary's nodes is
  |boundary| * dF, with dF restated here in numpy: the wall's
  F(u_g).n - F(u).n with u_g = u without its normal velocity (same pressure),
  and K-(n) (u_g - u) for the subsonic inlet / outlet with K- assembled from
  numpy's eigendecomposition of the complex-step flux Jacobian;
* a wall parallel to the flow and an outlet at the state's own pressure add
  no residual; every boundary adds V_dual * max |eigenvalue| to the wave
  speed of its nodes.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from pdekit_amd.mesh import box_tetra, rectangle
from tests.helpers import add_oracle_weak, boundary_facets, oracle_problem

GAMMA = 1.4
R_GAS = 287.046


def _state(rho, vel, p):
    vel = np.asarray(vel, dtype=float)
    E = p / (GAMMA - 1) + 0.5 * rho * vel @ vel
    return np.concatenate([[rho], rho * vel, [E]])


def _flux_n(u, n):
    dim = len(n)
    rho, m, E = u[0], u[1:1 + dim], u[1 + dim]
    v = m / rho
    p = (GAMMA - 1) * (E - 0.5 * rho * v @ v)
    un = v @ n
    return np.concatenate([[rho * un], m * un + p * np.asarray(n), [(E + p) * un]])


def _kminus(u, n):
    m = u.size
    h = 1e-30
    A = np.zeros((m, m))
    for j in range(m):
        du = np.zeros(m, dtype=complex)
        du[j] = 1j * h
        A[:, j] = _flux_n(u.astype(complex) + du, n).imag / h
    lam, R = np.linalg.eig(A)
    return (R @ np.diag(np.minimum(lam.real, 0)) @ np.linalg.inv(R)).real, np.abs(lam).max()


def _boundary_part(mesh, u, bcs):
    """Sum over the nodes of each boundary of V_dual * (res with BCs - res without)."""
    P0 = oracle_problem(mesh, O.EULER2D if mesh.dim == 2 else O.EULER3D, O.LDA, (GAMMA, 287.05))
    P1 = oracle_problem(mesh, O.EULER2D if mesh.dim == 2 else O.EULER3D, O.LDA, (GAMMA, 287.05))
    for kind, nodes, par in bcs:
        add_oracle_weak(P1, mesh, kind, nodes, par)
    U = np.tile(u, (mesh.nb_nodes, 1))
    r0, w0, rc0 = P0.assemble_rhs(U)
    r1, w1, rc1 = P1.assemble_rhs(U)
    assert rc0 == 0 and rc1 == 0
    vol = P0.dual_volume()
    return (r1 - r0) * vol[:, None], w1 - w0, vol


def _length(mesh, nodes):
    fn, fv, _ = boundary_facets(mesh, nodes)
    x = mesh.coords[fv]
    return np.linalg.norm(x[:, 1] - x[:, 0], axis=1).sum()


@pytest.mark.parametrize("order", [1, 2])
def test_wall(order):
    mesh = rectangle(6, 5, split=2, quad_rows=2 if order == 1 else 0, order=order, jitter=0.1, seed=3)
    b = mesh.boundaries
    u = _state(1.2, (0.5, 0.2), 0.9)
    dres, dws, vol = _boundary_part(mesh, u, [("WeakWall", b["ymin"], ())])
    n = np.array([0.0, -1.0])
    v = u[1:3] / u[0]
    g = _state(u[0], v - (v @ n) * n, (GAMMA - 1) * (u[3] - 0.5 * u[0] * v @ v))
    dF = _flux_n(g, n) - _flux_n(u, n)
    L = _length(mesh, b["ymin"])
    assert abs(L - 1.0) < 1e-14
    assert np.allclose(dres.sum(0), L * dF, rtol=1e-12, atol=1e-13)
    fn, _, _ = boundary_facets(mesh, b["ymin"])
    on = np.unique(fn)
    _, eig = _kminus(u, n)
    # every wall node: V_dual * max |eig| per facet holding it
    counts = np.bincount(fn.ravel(), minlength=mesh.nb_nodes)
    assert np.allclose(dws[on], counts[on] * vol[on] * eig, rtol=1e-12)
    assert np.abs(dres[np.setdiff1d(np.arange(mesh.nb_nodes), on)]).max() == 0.0
    # flow along the wall: no residual
    dres2, _, _ = _boundary_part(mesh, _state(1.0, (0.6, 0.0), 0.8), [("WeakWall", b["ymin"], ())])
    assert np.abs(dres2).max() < 1e-14


def test_subsonic_outlet_and_inlet():
    mesh = rectangle(5, 6, split=2, jitter=0.1, seed=5)
    b = mesh.boundaries
    u = _state(1.1, (0.45, 0.05), 1.0 / GAMMA)
    # outlet at x = 1, exterior pressure 0.8
    dres, _, _ = _boundary_part(mesh, u, [("WeakSubOutlet", b["xmax"], (0.8,))])
    n = np.array([1.0, 0.0])
    Km, _ = _kminus(u, n)
    g = u.copy()
    g[3] = 0.8 / (GAMMA - 1) + 0.5 * (u[1] ** 2 + u[2] ** 2) / u[0]
    assert np.allclose(dres.sum(0), 1.0 * Km @ (g - u), rtol=1e-11, atol=1e-13)
    p = (GAMMA - 1) * (u[3] - 0.5 * (u[1] ** 2 + u[2] ** 2) / u[0])
    dres0, _, _ = _boundary_part(mesh, u, [("WeakSubOutlet", b["xmax"], (p,))])
    assert np.abs(dres0).max() < 1e-13
    # inlet at x = 0 from total temperature / pressure at the state's Mach number
    Tt, pt = 1.2 / R_GAS, 1.1
    dres, _, _ = _boundary_part(mesh, u, [("WeakSubInlet", b["xmin"], (Tt, pt))])
    n = np.array([-1.0, 0.0])
    Km, _ = _kminus(u, n)
    v = u[1:3] / u[0]
    pu = (GAMMA - 1) * (u[3] - 0.5 * u[0] * v @ v)
    Ma = np.sqrt(v @ v) / np.sqrt(GAMMA * pu / u[0])
    fac = 1 + 0.5 * (GAMMA - 1) * Ma ** 2
    Tg, pg = Tt / fac, pt / fac ** (GAMMA / (GAMMA - 1))
    rg = pg / (R_GAS * Tg)
    ug = rg * Ma * np.sqrt(GAMMA * R_GAS * Tg)
    g = np.array([rg, ug, 0.0, pg / (GAMMA - 1) + 0.5 * ug ** 2 / rg])
    assert np.allclose(dres.sum(0), 1.0 * Km @ (g - u), rtol=1e-11, atol=1e-13)


def test_wall_3d():
    mesh = box_tetra(3, 3, 2, extent=(1.0, 1.0, 0.5), jitter=0.1, seed=4)
    b = mesh.boundaries
    name = "zmin" if "zmin" in b else sorted(b)[0]
    u = _state(1.0, (0.4, 0.1, -0.15), 0.7)
    fn, fv, _ = boundary_facets(mesh, b[name])
    x = mesh.coords[fv]
    nrm = np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0])
    area = 0.5 * np.linalg.norm(nrm, axis=1)
    n = np.array([0.0, 0.0, -1.0])
    assert np.allclose(np.abs(nrm / np.linalg.norm(nrm, axis=1)[:, None]) @ np.abs(n), 1.0)
    dres, _, _ = _boundary_part(mesh, u, [("WeakWall", b[name], ())])
    v = u[1:4] / u[0]
    g = _state(u[0], v - (v @ n) * n, (GAMMA - 1) * (u[4] - 0.5 * u[0] * v @ v))
    dF = _flux_n(g, n) - _flux_n(u, n)
    assert np.allclose(dres.sum(0), area.sum() * dF, rtol=1e-12, atol=1e-13)


def test_farfield_symmetry_superinlet_dirichlet():
    mesh = rectangle(5, 4, split=2, jitter=0.1, seed=7)
    b = mesh.boundaries
    u = _state(1.05, (0.4, 0.12), 0.75)
    ref = _state(1.0, (0.5, 0.0), 1.0 / GAMMA)
    n = np.array([0.0, 1.0])  # ymax
    Km, eig = _kminus(u, n)
    for kind, par, g in (("WeakFarfield", tuple(ref), ref), ("WeakDirichlet", tuple(ref), ref)):
        dres, _, _ = _boundary_part(mesh, u, [(kind, b["ymax"], par)])
        assert np.allclose(dres.sum(0), 1.0 * Km @ (g - u), rtol=1e-11, atol=1e-13)
    v = u[1:3] / u[0]
    g = np.concatenate([[u[0]], u[0] * (v - 2 * (v @ n) * n), [u[3]]])
    dres, _, _ = _boundary_part(mesh, u, [("WeakSymmetry", b["ymax"], ())])
    assert np.allclose(dres.sum(0), 1.0 * Km @ (g - u), rtol=1e-11, atol=1e-13)
    # the supersonic inlet only adds to the wave speed
    dres, dws, vol = _boundary_part(mesh, u, [("WeakSuperInlet", b["ymax"], tuple(ref))])
    assert np.abs(dres).max() == 0.0 and dws.max() > 0


def test_scalar_weak_dirichlet():
    """Linear advection a = (0.8, 0.6): on the inflow side x = 0 (n = (-1, 0),
    a.n = -0.8) the correction is -0.8 (g - u) per unit length, on the
    outflow side x = 1 it vanishes."""
    mesh = rectangle(6, 5, split=2, jitter=0.1, seed=2)
    b = mesh.boundaries
    P0 = oracle_problem(mesh, O.ADVECTION2D, O.N, (0.8, 0.6))
    P1 = oracle_problem(mesh, O.ADVECTION2D, O.N, (0.8, 0.6))
    add_oracle_weak(P1, mesh, "WeakDirichlet", b["xmin"], (1.0,))
    add_oracle_weak(P1, mesh, "WeakDirichlet", b["xmax"], (1.0,))
    U = np.full((mesh.nb_nodes, 1), 0.25)
    r0, w0, _ = P0.assemble_rhs(U)
    r1, w1, _ = P1.assemble_rhs(U)
    vol = P0.dual_volume()
    d = (r1 - r0)[:, 0] * vol
    fn_in, _, _ = boundary_facets(mesh, b["xmin"])
    fn_out, _, _ = boundary_facets(mesh, b["xmax"])
    assert abs(d[np.unique(fn_in)].sum() - (-0.8) * (1.0 - 0.25)) < 1e-13
    assert np.abs(d[np.setdiff1d(np.unique(fn_out), np.unique(fn_in))]).max() < 1e-15


def test_rotation_weak_dirichlet():
    """The rotating advection, velocity (y, -x) of the point, on [0, 1]^2: the
    edges x = 0 (n = (-1, 0), a.n = -y) and y = 1 (n = (0, 1), a.n = -x) are
    inflow, so with uniform u and g the correction integrates to
    -(g - u) * 1/2 over each; the edges x = 1 (a.n = y) and y = 0 (a.n = x)
    are outflow and add nothing."""
    mesh = rectangle(6, 5, split=2, jitter=0.1, seed=2)
    b = mesh.boundaries
    P0 = oracle_problem(mesh, O.ROTATION2D, O.N, ())
    P1 = oracle_problem(mesh, O.ROTATION2D, O.N, ())
    for side in ("xmin", "ymax", "xmax", "ymin"):
        add_oracle_weak(P1, mesh, "WeakDirichlet", b[side], (1.0,))
    U = np.full((mesh.nb_nodes, 1), 0.25)
    r0, w0, rc0 = P0.assemble_rhs(U)
    r1, w1, rc1 = P1.assemble_rhs(U)
    assert rc0 == 0 and rc1 == 0
    vol = P0.dual_volume()
    d = (r1 - r0)[:, 0] * vol
    corner_in = np.intersect1d(b["xmin"], b["ymax"])
    x_in = np.unique(boundary_facets(mesh, b["xmin"])[0])
    y_in = np.unique(boundary_facets(mesh, b["ymax"])[0])
    assert abs(d[np.union1d(x_in, y_in)].sum() - 2 * (-0.5) * (1.0 - 0.25)) < 1e-13
    # each inflow edge alone (the corner node is shared; its share of either edge is separated by adding one edge)
    P2 = oracle_problem(mesh, O.ROTATION2D, O.N, ())
    add_oracle_weak(P2, mesh, "WeakDirichlet", b["xmin"], (1.0,))
    r2, _, _ = P2.assemble_rhs(U)
    assert abs(((r2 - r0)[:, 0] * vol).sum() - (-0.5) * (1.0 - 0.25)) < 1e-13
    out = np.setdiff1d(np.union1d(b["xmax"], b["ymin"]), np.union1d(x_in, y_in))
    assert corner_in.size == 1 and np.abs(d[out]).max() < 1e-15
