"""Pins of the CPU oracle (oracle/rd_oracle.cpp) by mathematics the reference
algorithm must satisfy -- none of these checks retypes the oracle's formula:

* P1 scalar schemes reduce to the textbook LDA / N distributions
  (beta_i = k_i+ / sum k+, N: k_i+ (u_i - u_in)), B = LDA + theta (N - LDA)
  with theta = |Phi_T| / (sum |Phi_N| + 1e-6), the PSI limiter of
  LFResidualLimiter.hpp:84-110 applied to N, and the PSI / LDA blend
  (1 - theta) LDA + theta PSI; k_i = |T| a . grad(phi_i).
* conservation: sum_i Phi_i = integral of div F for every scheme; with the
  P+1 flux element the divergence of a linear advection flux and of the
  Burgers flux of a P1 state is exact, so the total equals a closed form.
* linearity preservation of LDA and B for an exact linear steady state.
* Euler: R diag(D) L = A g_x + B g_y (+ C g_z) with A, B, C the flux Jacobians
  obtained by complex-step differentiation of the flux, L R = I (as
  test/utest-physics/utest-euler-model-2D.cpp:29-108 at state (1,1,2,5),
  direction (7,2)); free-stream preservation; small inverses A A^-1 = I
  (test/utest-math/utest-matrix-inverter.cpp:31-260).
* dual volumes sum to (nodes per cell) x area (RDTimeUpdate.hpp:162-175).
* end to end: the make_unit_quad advection problem converges.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from pdekit_amd import _lib as L
from pdekit_amd.mesh import box_tetra, rectangle, unit_quad_triag
from tests.helpers import euler_state, oracle_problem

A = (0.8, 0.6)


def _p1_grads(x):
    """Gradients of the P1 basis of triangles x[c][3][2] and their areas."""
    e1 = x[:, 1] - x[:, 0]
    e2 = x[:, 2] - x[:, 0]
    area = 0.5 * (e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    Jt = np.stack([e1, e2], 1)  # rows: dx/dl1, dx/dl2
    inv = np.linalg.inv(Jt)      # grad(l1), grad(l2) are the columns
    g1 = inv[:, :, 0]
    g2 = inv[:, :, 1]
    g0 = -g1 - g2
    return np.stack([g0, g1, g2], 1), area


def _textbook(scheme, k, u):
    kp, km = np.maximum(k, 0), np.minimum(k, 0)
    phiT = (k * u).sum(1)
    S = kp.sum(1)
    lda = kp / S[:, None] * phiT[:, None]
    uin = (km * u).sum(1) / km.sum(1)
    n = kp * (u - uin[:, None])
    if scheme == O.LDA:
        return lda
    if scheme == O.N:
        return n
    th = np.abs(phiT) / (np.abs(n).sum(1) + 1e-6)
    if scheme == O.B:
        return lda + th[:, None] * (n - lda)
    x = np.maximum(n / phiT[:, None], 0) + 1e-12
    psi = x / x.sum(1)[:, None] * phiT[:, None]
    if scheme == O.PSI:
        return psi
    # the PSI / LDA blend with the same cell theta
    return (1 - th[:, None]) * lda + th[:, None] * psi


@pytest.mark.parametrize("scheme", [O.LDA, O.N, O.B, O.PSI, O.BPSI])
def test_p1_advection_textbook_schemes(scheme):
    mesh = rectangle(7, 6, split=2, jitter=0.2, seed=11)
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, mesh.nb_nodes)
    P = oracle_problem(mesh, O.ADVECTION2D, scheme, A)
    phi, wave, rc = P.cell_residuals(0, u)
    assert rc == 0
    conn = mesh.blocks[0].sol_conn
    g, area = _p1_grads(mesh.coords[mesh.blocks[0].geo_conn])
    k = area[:, None] * (A[0] * g[:, :, 0] + A[1] * g[:, :, 1])
    ref = _textbook(scheme, k, u[conn])
    assert np.abs(phi[:, :, 0] - ref).max() <= 1e-13 * np.abs(ref).max()
    lam = np.maximum(np.abs(A[0] * g[:, :, 0] + A[1] * g[:, :, 1]), 1e-6).max(1)
    assert np.allclose(wave, area * lam, rtol=1e-13, atol=0)


@pytest.mark.parametrize("scheme", [O.LDA, O.N, O.B, O.PSI, O.BPSI])
def test_burgers_p1_conservation_is_exact(scheme):
    """u linear in each cell: F = (u^2/2, u) is quadratic, interpolated exactly
    by the P2 flux element, so sum_i Phi_i = |T| (u_x mean(u) + u_y)."""
    mesh = rectangle(6, 5, split=4, jitter=0.2, seed=3)
    rng = np.random.default_rng(2)
    u = rng.uniform(0.2, 1.5, mesh.nb_nodes)
    P = oracle_problem(mesh, O.BURGERS2D, scheme, ())
    phi, _, rc = P.cell_residuals(0, u)
    g, area = _p1_grads(mesh.coords[mesh.blocks[0].geo_conn])
    uc = u[mesh.blocks[0].sol_conn]
    grad = np.einsum("cn,cnd->cd", uc, g)
    exact = area * (grad[:, 0] * uc.mean(1) + grad[:, 1])
    assert np.abs(phi[:, :, 0].sum(1) - exact).max() <= 1e-13 * np.abs(exact).max()


@pytest.mark.parametrize("name", ["p2", "q1", "q2", "euler", "euler3d_p2"])
def test_schemes_share_the_element_total(name):
    """The N / B dissipation, the PSI limiter and the PSI / LDA blend
    redistribute without changing the element total, which equals the LDA
    total."""
    if name == "q2":
        mesh = rectangle(4, 3, split=1, order=2, jitter=0.15, seed=5)
        phys, par = O.BURGERS2D, ()
        u = np.random.default_rng(3).uniform(0.2, 1.5, (mesh.nb_nodes, 1))
    elif name == "euler3d_p2":
        mesh = box_tetra(2, 2, 2, order=2, jitter=0.1, seed=4)
        phys, par = O.EULER3D, (1.4, 287.05)
        u = euler_state(mesh, dim=3)
    elif name == "p2":
        mesh = rectangle(5, 4, split=2, order=2, jitter=0.15, seed=5)
        phys, par = O.BURGERS2D, ()
        u = np.random.default_rng(3).uniform(0.2, 1.5, (mesh.nb_nodes, 1))
    elif name == "q1":
        mesh = rectangle(5, 4, split=1, jitter=0.15, seed=5)
        phys, par = O.BURGERS2D, ()
        u = np.random.default_rng(3).uniform(0.2, 1.5, (mesh.nb_nodes, 1))
    else:
        mesh = rectangle(5, 4, split=2, jitter=0.15, seed=5)
        phys, par = O.EULER2D, (1.4, 287.05)
        u = euler_state(mesh)
    tot = []
    for sch in (O.LDA, O.N, O.B, O.PSI, O.BPSI):
        P = oracle_problem(mesh, phys, sch, par)
        phi, _, rc = P.cell_residuals(0, u)
        assert rc == 0
        tot.append(phi.sum(1))
    scale = np.abs(tot[0]).max()
    for t in tot[1:]:
        assert np.abs(t - tot[0]).max() <= 1e-12 * scale


@pytest.mark.parametrize("quads", [False, True])
def test_linearity_preservation(quads):
    """P2 triangles, or Q2 quadrilaterals on bilinear cells: a linear exact
    solution of a . grad u = 0 is represented exactly and LDA / B and the
    PSI / LDA blend (theta = 0 where the element total vanishes) leave it."""
    mesh = rectangle(8, 7, split=1 if quads else 2, jitter=0.2, seed=8, order=2)
    x, y = mesh.node_coords.T
    u = 0.3 + (-A[1] * x + A[0] * y)  # a . grad u = 0
    for sch in (O.LDA, O.B, O.BPSI):
        P = oracle_problem(mesh, O.ADVECTION2D, sch, A)
        res, _, rc = P.assemble_rhs(u)
        assert np.abs(res).max() < 1e-13
    res_n, _, _ = oracle_problem(mesh, O.ADVECTION2D, O.N, A).assemble_rhs(u)
    assert np.abs(res_n).max() > 1e-6  # N is not linearity preserving


def _euler_flux(dim, gamma, u):
    rho = u[0]
    m = u[1:1 + dim]
    E = u[1 + dim]
    v = m / rho
    p = (gamma - 1) * (E - 0.5 * rho * np.dot(v, v))
    H = (E + p) / rho
    F = []
    for d in range(dim):
        f = [m[d]] + [m[d] * v[k] + (p if k == d else 0) for k in range(dim)] + [m[d] * H]
        F.append(np.array(f))
    return F


def _jacobians(dim, gamma, u):
    m = dim + 2
    h = 1e-30
    Jac = [np.zeros((m, m)) for _ in range(dim)]
    for j in range(m):
        du = np.zeros(m, dtype=complex)
        du[j] = 1j * h
        F = _euler_flux(dim, gamma, u.astype(complex) + du)
        for d in range(dim):
            Jac[d][:, j] = F[d].imag / h
    return Jac


@pytest.mark.parametrize("dim,u,g", [(2, [1.0, 1.0, 2.0, 5.0], [7.0, 2.0]),
                                     (2, [1.2, 0.3, -0.4, 2.5], [-0.3, 0.9]),
                                     (3, [1.0, 0.3, 0.2, 0.1, 2.5], [7.0, 2.0, 3.0]),
                                     (3, [0.8, -0.2, 0.1, 0.4, 2.0], [0.1, -1.0, 0.5])])
def test_euler_eigenstructure(dim, u, g):
    u = np.array(u)
    g = np.array(g)
    R, Lm, D, rc = O.euler_eigen(dim, 1.4, u, g)
    assert rc == 0
    Jac = _jacobians(dim, 1.4, u)
    K = sum(Jac[d] * g[d] for d in range(dim))
    assert np.abs(R @ np.diag(D) @ Lm - K).max() <= 1e-12 * np.abs(K).max()
    assert np.abs(Lm @ R - np.eye(dim + 2)).max() <= 1e-12
    Kp, eig, _ = O.euler_kplus(dim, 1.4, u, g)
    Km = R @ np.diag(np.minimum(D, 0)) @ Lm
    assert np.abs(Kp + Km - K).max() <= 1e-12 * np.abs(K).max()
    assert np.allclose(np.sort(np.linalg.eigvals(Kp).real), np.sort(np.maximum(eig, 0)), atol=1e-10)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_small_inverses(n):
    rng = np.random.default_rng(n)
    for _ in range(20):
        a = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
        inv = O.invert(a)
        assert np.abs(a @ inv - np.eye(n)).max() < 1e-12


@pytest.mark.parametrize("kind", ["mixed_p1", "mixed_p2_q2", "tet_p2"])
@pytest.mark.parametrize("scheme", [O.LDA, O.N, O.B, O.PSI])
def test_euler_free_stream_is_preserved(scheme, kind):
    if kind == "tet_p2":
        mesh = box_tetra(3, 3, 2, extent=(1.0, 1.0, 0.5), order=2, jitter=0.1, seed=4,
                         bump=(0.5, 0.08, 25.0, 0.5, 0.5))
        u = euler_state(mesh, pert=0.0, dim=3)
        phys = O.EULER3D
    else:
        mesh = rectangle(8, 6, extent=(-0.5, 1.5, 0, 1), split=2, quad_rows=3, bump=(1.0, 0.1, 25.0, 0.5),
                         order=2 if kind == "mixed_p2_q2" else 1)
        u = euler_state(mesh, pert=0.0)
        phys = O.EULER2D
    P = oracle_problem(mesh, phys, scheme, (1.4, 287.05))
    res, ws, rc = P.assemble_rhs(u)
    assert rc == 0
    assert np.abs(res).max() < 1e-12 * ws.max()


def test_dual_volumes():
    for order, nd in ((1, 3), (2, 6)):
        mesh = rectangle(6, 5, split=2, order=order, jitter=0.2, seed=1)
        P = oracle_problem(mesh, O.ADVECTION2D, O.LDA, A)
        assert abs(P.dual_volume().sum() - nd * 1.0) < 1e-13
    mesh = rectangle(6, 5, split=1, jitter=0.2, seed=1)
    assert abs(oracle_problem(mesh, O.ADVECTION2D, O.LDA, A).dual_volume().sum() - 4.0) < 1e-13
    mesh = rectangle(6, 5, split=1, jitter=0.2, seed=1, order=2)
    assert abs(oracle_problem(mesh, O.ADVECTION2D, O.LDA, A).dual_volume().sum() - 9.0) < 1e-13
    mesh = box_tetra(3, 3, 2, extent=(1.0, 1.0, 0.5), order=2, jitter=0.1, seed=4)
    assert abs(oracle_problem(mesh, O.EULER3D, O.LDA, (1.4, 287.05)).dual_volume().sum() - 10 * 0.5) < 1e-13


@pytest.mark.parametrize("scheme", [O.N, O.LDA])
def test_advection_converges(scheme):
    """make_unit_quad(17) on [0,1]^2, a = (0.8, 0.6), u = 1 on the inflow for
    y < 0.5 else 0: the residual of the non-Dirichlet rows drops below 1e-12
    and the N solution stays within the inflow bounds."""
    mesh = unit_quad_triag(17)
    bn = np.union1d(mesh.boundaries["Edge0"], mesh.boundaries["Edge3"])
    g = np.where(mesh.node_coords[bn, 1] < 0.5, 1.0, 0.0)[:, None]
    P = oracle_problem(mesh, O.ADVECTION2D, scheme, A)
    u, res, dt, norms, rc = P.iterate(np.zeros(mesh.nb_nodes), 0.9, 400, bn, g)
    mask = np.ones(mesh.nb_nodes, bool)
    mask[bn] = False
    assert np.sqrt(np.mean(res[mask] ** 2)) < 1e-12
    if scheme == O.N:
        assert u.min() > -1e-12 and u.max() < 1 + 1e-12


def test_threaded_iterations_match():
    """The multi-core timing baseline computes the same iterations."""
    mesh = rectangle(12, 10, split=4, order=2, jitter=0.15, seed=2)
    x = mesh.node_coords[:, 0]
    u = (1.5 - 2.0 * x)[:, None]
    P = oracle_problem(mesh, O.BURGERS2D, O.B, ())
    a, ra, _, _, rc = P.iterate(u, 0.5, 4)
    b, rb, _, rc2 = P.iterate_threads(u, 0.5, 4, 3)
    assert rc == rc2 == 0
    assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()
    assert np.abs(ra - rb).max() <= 1e-12 * np.abs(ra).max()
