"""Shared builders for the parity tests: meshes, states, oracle problems."""
from __future__ import annotations

import numpy as np

from oracle.pyoracle import Problem
from pdekit_amd import _lib as L
from pdekit_amd.fe import element_tables
from pdekit_amd.mesh import box_tetra, rectangle, unit_quad_triag

QUAD_ORDER = {L.TRIAG: 4, L.QUAD: 4, L.TETRA: 2}
P3_QUAD_ORDER = 6


def quad_orders(mesh):
    q = dict(QUAD_ORDER)
    if mesh.order == 3:
        q[L.TRIAG] = P3_QUAD_ORDER
    return {b.shape: q[b.shape] for b in mesh.blocks}


def oracle_problem(mesh, physics, scheme, params, qorders=None):
    qorders = qorders or quad_orders(mesh)
    tables = {b.shape: element_tables(b.shape, mesh.order, qorders[b.shape]) for b in mesh.blocks}
    vol = {b.shape: element_tables(b.shape, mesh.order, mesh.order) for b in mesh.blocks}
    return Problem(mesh, tables, vol, physics, scheme, params)


def euler_state(mesh, mach=0.5, gamma=1.4, seed=3, pert=0.01, dim=2):
    """Free stream rho=1, p=1/gamma (a=1), velocity mach along x, with a seeded
    relative perturbation of every conserved variable."""
    rng = np.random.default_rng(seed)
    n = mesh.nb_nodes
    rho = np.ones(n)
    vel = np.zeros((n, dim))
    vel[:, 0] = mach
    p = np.full(n, 1.0 / gamma)
    E = p / (gamma - 1.0) + 0.5 * rho * (vel ** 2).sum(1)
    u = np.concatenate([rho[:, None], rho[:, None] * vel, E[:, None]], axis=1)
    return u * (1.0 + pert * rng.uniform(-1.0, 1.0, u.shape))


def burgers_state(mesh, seed=7, pert=0.01):
    rng = np.random.default_rng(seed)
    x = mesh.node_coords[:, 0]
    return ((1.5 - 2.0 * x) * (1.0 + pert * rng.uniform(-1, 1, x.size)))[:, None]


def advection_state(mesh, seed=5):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, (mesh.nb_nodes, 1))


def rel_err(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


CASES = {
    # name: (mesh builder, physics, params, state builder)
    "adv_p1": (lambda: unit_quad_triag(33), L.ADVECTION2D, (0.8, 0.6), advection_state),
    "adv_p2": (lambda: rectangle(12, 10, split=2, order=2, jitter=0.15, seed=2), L.ADVECTION2D, (0.8, 0.6),
               advection_state),
    "adv_p3": (lambda: rectangle(8, 8, split=2, order=3, jitter=0.1, seed=4), L.ADVECTION2D, (0.8, 0.6),
               advection_state),
    "burgers_p1": (lambda: rectangle(16, 16, split=4, order=1, jitter=0.15, seed=2), L.BURGERS2D, (),
                   burgers_state),
    "burgers_p2": (lambda: rectangle(16, 16, split=4, order=2, jitter=0.15, seed=2), L.BURGERS2D, (),
                   burgers_state),
    "burgers_q1": (lambda: rectangle(12, 12, split=2, quad_rows=6, order=1, jitter=0.1, seed=2), L.BURGERS2D, (),
                   burgers_state),
    "euler2d_p1": (lambda: rectangle(16, 8, extent=(-0.5, 1.5, 0.0, 1.0), split=2, order=1,
                                     bump=(1.0, 0.1, 25.0, 0.5)), L.EULER2D, (1.4, 287.05), euler_state),
    "euler2d_mixed": (lambda: rectangle(16, 8, extent=(-0.5, 1.5, 0.0, 1.0), split=2, quad_rows=4, order=1,
                                        bump=(1.0, 0.1, 25.0, 0.5)), L.EULER2D, (1.4, 287.05), euler_state),
    "euler2d_p2": (lambda: rectangle(8, 6, extent=(-0.5, 1.5, 0.0, 1.0), split=2, order=2,
                                     bump=(1.0, 0.1, 25.0, 0.5)), L.EULER2D, (1.4, 287.05), euler_state),
    "euler3d_p1": (lambda: box_tetra(5, 5, 4, extent=(1.0, 1.0, 0.5), jitter=0.1, seed=4,
                                     bump=(0.5, 0.08, 25.0, 0.5, 0.5)), L.EULER3D, (1.4, 287.05),
                   lambda m: euler_state(m, dim=3)),
    "adv_q2": (lambda: rectangle(8, 7, split=1, order=2, jitter=0.1, seed=3), L.ADVECTION2D, (0.8, 0.6),
               advection_state),
    "burgers_q2_mixed": (lambda: rectangle(10, 10, split=2, quad_rows=4, order=2, jitter=0.1, seed=6), L.BURGERS2D,
                         (), burgers_state),
    "euler2d_q2_mixed": (lambda: rectangle(12, 6, extent=(-0.5, 1.5, 0.0, 1.0), split=2, quad_rows=3, order=2,
                                           bump=(1.0, 0.1, 25.0, 0.5)), L.EULER2D, (1.4, 287.05), euler_state),
    "euler3d_p2": (lambda: box_tetra(3, 3, 2, extent=(1.0, 1.0, 0.5), order=2, jitter=0.1, seed=4,
                                     bump=(0.5, 0.08, 25.0, 0.5, 0.5)), L.EULER3D, (1.4, 287.05),
                   lambda m: euler_state(m, dim=3)),
}


# ---------------------------------------------------------------------------
# Weak boundary conditions: boundary facets and facet rules restated in numpy
# (independent of the library's facet search) for the oracle.
# ---------------------------------------------------------------------------
_LOCAL_FACETS = {L.TRIAG: [(0, 1), (1, 2), (2, 0)], L.QUAD: [(0, 1), (1, 2), (2, 3), (3, 0)],
                 L.TETRA: [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]}


_TET_EDGES = [(0, 1), (1, 2), (2, 0), (3, 0), (3, 2), (3, 1)]


def boundary_facets(mesh, nodes):
    """Facets on the domain boundary whose solution nodes all lie in `nodes`:
    (facet nodes [nf][nfn] (vertices, then the P2 mid-edge node), facet
    vertices [nf][nfv], centroid of each facet's cell [nf][dim])."""
    in_set = np.zeros(mesh.nb_nodes, bool)
    in_set[np.asarray(nodes, dtype=np.int64)] = True
    count = {}
    for b in mesh.blocks:
        for cv in b.geo_conn:
            for lf in _LOCAL_FACETS[b.shape]:
                k = tuple(sorted(int(cv[i]) for i in lf))
                count[k] = count.get(k, 0) + 1
    fn, fv, ins = [], [], []
    for b in mesh.blocks:
        for cv, cs in zip(b.geo_conn, b.sol_conn):
            for e, lf in enumerate(_LOCAL_FACETS[b.shape]):
                if count[tuple(sorted(int(cv[i]) for i in lf))] != 1:
                    continue
                nd = [int(cs[i]) for i in lf]
                if mesh.order == 2 and b.shape == L.TRIAG:
                    nd.append(int(cs[3 + e]))
                if mesh.order == 2 and b.shape == L.QUAD:
                    nd.append(int(cs[4 + e]))
                if mesh.order == 2 and b.shape == L.TETRA:
                    for k in range(3):
                        pair = {lf[k], lf[(k + 1) % 3]}
                        nd.append(int(cs[4 + [set(t) for t in _TET_EDGES].index(pair)]))
                if not in_set[nd].all():
                    continue
                fn.append(nd)
                fv.append([int(cv[i]) for i in lf])
                ins.append(mesh.coords[cv].mean(0))
    return np.array(fn, dtype=np.uint32), np.array(fv, dtype=np.uint32), np.array(ins)


def facet_rule(dim, nfn, order):
    """Gauss-Legendre on [-1, 1] with max(2, order // 2 + 1) points (edges) or
    the triangle rule of the given order (faces), and the facet nodes' values."""
    if dim == 2:
        x, w = np.polynomial.legendre.leggauss(max(2, order // 2 + 1))
        if nfn == 2:
            V = np.stack([0.5 * (1 - x), 0.5 * (1 + x)], 1)
        else:
            V = np.stack([0.5 * x * (x - 1), 0.5 * x * (x + 1), 1 - x * x], 1)
        return w, V
    from pdekit_amd.fe import quadrature
    xi, w = quadrature(L.TRIAG, order)
    a, b = xi[:, 0], xi[:, 1]
    lam = np.stack([-0.5 * (a + b), 0.5 * (1 + a), 0.5 * (1 + b)], 1)
    if nfn == 3:
        return w, lam
    # P2 face: the quadratic Lagrange basis by solving on its six nodes
    nodes = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [.5, .5, 0], [0, .5, .5], [.5, 0, .5]])
    mono = lambda l: np.stack([l[:, 0] ** 2, l[:, 1] ** 2, l[:, 2] ** 2, l[:, 0] * l[:, 1], l[:, 1] * l[:, 2],
                               l[:, 2] * l[:, 0]], 1)
    return w, mono(lam) @ np.linalg.inv(mono(nodes))


def add_oracle_weak(P, mesh, bc_type, nodes, params=(), quad_order=4):
    from oracle import pyoracle as O
    fn, fv, ins = boundary_facets(mesh, nodes)
    w, V = facet_rule(mesh.dim, fn.shape[1], quad_order + 1)
    # the facet points: the vertices' linear functions at the same rule
    _, V1 = facet_rule(mesh.dim, fv.shape[1], quad_order + 1)
    qpoints = np.einsum("qv,fvd->fqd", V1, mesh.coords[fv])
    kinds = {"WeakWall": O.WEAK_WALL, "WeakSubInlet": O.WEAK_SUBINLET, "WeakSubOutlet": O.WEAK_SUBOUTLET,
             "WeakFarfield": O.WEAK_FARFIELD, "WeakSymmetry": O.WEAK_SYMMETRY, "WeakSuperInlet": O.WEAK_SUPERINLET,
             "WeakDirichlet": O.WEAK_DIRICHLET}
    P.add_weak(kinds[bc_type], params, mesh.order, fn, fv, ins, w, V, qpoints=qpoints)
    return fn
