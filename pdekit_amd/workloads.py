"""Synthetic workloads of the benchmark: meshes, states and boundary data.

c1  2D linear advection a=(0.8, 0.6), MeshCreator::make_unit_quad(129) mapped
    to [0,1]^2 (32 768 P1 triangles, 16 641 nodes), strong Dirichlet inflow
    u = 1 for y < 0.5 else 0 on x = 0 and y = 0, u0 = 0, CFL 0.9.
c2  2D Burgers, P2 triangles: a 512 x 512 grid on [0,1]^2 with every cell
    split into 4 triangles around a centre vertex (1 048 576 triangles,
    2 099 201 nodes), interior vertices jittered U(-0.15 h, 0.15 h) with
    std::mt19937_64(2); B scheme; inflow u(x,0) = 1.5 - 2x, u = 1.5 on x = 0,
    u = -0.5 on x = 1 (the boundary values of examples/testcases-rdm/
    Burgers-RDS-2D.cpp); timing state 1.5 - 2x with a seeded 1% perturbation.
c3  2D Euler channel [-0.5,1.5] x [0,1] over y = 0.1 exp(-25 (x-0.5)^2):
    1024 x 512 cells, the lower 256 rows Q1 quadrilaterals, the upper rows
    split triangles (262 144 + 524 288 = 786 432 P1 elements), M = 0.5 free
    stream with a seeded 1% perturbation.
c4  3D Euler box [0,1]^2 x [0,0.5] over z = 0.08 exp(-25((x-.5)^2+(y-.5)^2)):
    96^3 hexahedra split into 6 tetrahedra (5 308 416 P1 tetrahedra,
    912 673 nodes), interior vertices jittered 10% with seed 4, M = 0.5, the
    PSI / LDA blend (PGB's cell theta).
c3p2 the c3 channel with P2 triangles and Q2 quadrilaterals (786 432
    elements, 2 100 225 nodes), rule of order 4.
c4p2 the c4 box with 48^3 hexahedra split into 6 P2 tetrahedra (663 552
    elements, 912 673 nodes), rule of order 2 as examples/testcases-rdm/
    Bump-RDS-3D-Explicit.cpp:53 uses, the PSI / LDA blend.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .mesh import Mesh, box_tetra, rectangle, unit_quad_triag


_TAG = {L.LDA: "lda", L.N: "n", L.B: "b", L.PSI: "psi", L.BPSI: "bpsi"}


@dataclass
class Workload:
    name: str
    mesh: Mesh
    physics: int
    scheme: int
    params: tuple
    quad_order: dict
    cfl: float
    state: np.ndarray
    dirichlet_nodes: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.uint32))
    dirichlet_values: np.ndarray = field(default_factory=lambda: np.zeros((0, 1)))

    @property
    def nb_elements(self) -> int:
        return self.mesh.nb_cells


def _free_stream(mesh: Mesh, dim: int, mach: float = 0.5, gamma: float = 1.4, seed: int = 3, pert: float = 0.01):
    rng = np.random.default_rng(seed)
    n = mesh.nb_nodes
    rho = np.ones(n)
    vel = np.zeros((n, dim))
    vel[:, 0] = mach
    p = np.full(n, 1.0 / gamma)
    E = p / (gamma - 1.0) + 0.5 * rho * (vel ** 2).sum(1)
    u = np.concatenate([rho[:, None], rho[:, None] * vel, E[:, None]], axis=1)
    return u * (1.0 + pert * rng.uniform(-1.0, 1.0, u.shape))


def c1(scheme: int = L.N, N: int = 129) -> Workload:
    mesh = unit_quad_triag(N)
    bn = np.union1d(mesh.boundaries["Edge0"], mesh.boundaries["Edge3"]).astype(np.uint32)
    g = np.where(mesh.node_coords[bn, 1] < 0.5, 1.0, 0.0)[:, None]
    return Workload("advection2d_p1_c1", mesh, L.ADVECTION2D, scheme, (0.8, 0.6), {L.TRIAG: 4}, 0.9,
                    np.zeros((mesh.nb_nodes, 1)), bn, g)


def c2(n: int = 512, scheme: int = L.B, seed: int = 2) -> Workload:
    mesh = rectangle(n, n, split=4, order=2, jitter=0.15, seed=seed)
    x = mesh.node_coords[:, 0]
    rng = np.random.default_rng(7)
    u = ((1.5 - 2.0 * x) * (1.0 + 0.01 * rng.uniform(-1.0, 1.0, x.size)))[:, None]
    b = mesh.boundaries
    ymin, xmin, xmax = b["ymin"], b["xmin"], b["xmax"]
    xmin = np.setdiff1d(xmin, ymin)
    xmax = np.setdiff1d(xmax, ymin)
    nodes = np.concatenate([ymin, xmin, xmax]).astype(np.uint32)
    vals = np.concatenate([1.5 - 2.0 * mesh.node_coords[ymin, 0], np.full(xmin.size, 1.5),
                           np.full(xmax.size, -0.5)])[:, None]
    return Workload(f"burgers2d_p2_{_TAG[scheme]}_{mesh.nb_cells}", mesh, L.BURGERS2D, scheme, (), {L.TRIAG: 4}, 0.5, u,
                    nodes, vals)


def c3(nx: int = 1024, ny: int = 512, scheme: int = L.LDA) -> Workload:
    mesh = rectangle(nx, ny, extent=(-0.5, 1.5, 0.0, 1.0), split=2, quad_rows=ny // 2, order=1,
                     bump=(1.0, 0.1, 25.0, 0.5))
    return Workload(f"euler2d_bump_p1q1_{mesh.nb_cells}", mesh, L.EULER2D, scheme, (1.4, 287.05),
                    {L.TRIAG: 4, L.QUAD: 4}, 0.5, _free_stream(mesh, 2))


def c4(n: int = 96, scheme: int = L.BPSI) -> Workload:
    mesh = box_tetra(n, n, n, extent=(1.0, 1.0, 0.5), jitter=0.1, seed=4, bump=(0.5, 0.08, 25.0, 0.5, 0.5))
    return Workload(f"euler3d_bump_p1_{_TAG[scheme]}_{mesh.nb_cells}", mesh, L.EULER3D, scheme, (1.4, 287.05), {L.TETRA: 2},
                    0.5, _free_stream(mesh, 3, seed=4))


def c3p2(nx: int = 1024, ny: int = 512, scheme: int = L.LDA) -> Workload:
    mesh = rectangle(nx, ny, extent=(-0.5, 1.5, 0.0, 1.0), split=2, quad_rows=ny // 2, order=2,
                     bump=(1.0, 0.1, 25.0, 0.5))
    return Workload(f"euler2d_bump_p2q2_{mesh.nb_cells}", mesh, L.EULER2D, scheme, (1.4, 287.05),
                    {L.TRIAG: 4, L.QUAD: 4}, 0.5, _free_stream(mesh, 2))


def c4p2(n: int = 48, scheme: int = L.BPSI) -> Workload:
    mesh = box_tetra(n, n, n, extent=(1.0, 1.0, 0.5), order=2, jitter=0.1, seed=4,
                     bump=(0.5, 0.08, 25.0, 0.5, 0.5))
    return Workload(f"euler3d_bump_p2_{_TAG[scheme]}_{mesh.nb_cells}", mesh, L.EULER3D, scheme, (1.4, 287.05), {L.TETRA: 2},
                    0.5, _free_stream(mesh, 3, seed=4))


BUILDERS = {"c1": c1, "c2": c2, "c3": c3, "c4": c4, "c3p2": c3p2, "c4p2": c4p2}
