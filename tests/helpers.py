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
}
