"""Serving loop: steady solves of differing sizes on one GPU with continuous
batching.

Requests arrive on a seeded Poisson stream. Each request is a steady explicit
RD solve run to a residual norm tolerance: 2D advection P1-P3 or 2D Burgers P2
with strong Dirichlet inflow, or the 2D Euler channel over a bump (P1 / Q1
elements as examples/testcases-rdm/Bump-RDS-2D-Explicit.cpp:56-70 uses, LDA,
free stream at the request's Mach number, weak walls, subsonic inlet and
outlet as there, :138-149, 235-243). The server

  * admits requests in arrival order as soon as their device footprint fits
    the HBM budget next to the requests already resident,
  * keeps every admitted mesh and state resident on the GPU,
  * advances all resident requests together: one pdrd_batch_iterate per chunk
    of iterations, whose launches cover the patches of every resident request
    of one element type / physics / scheme,
  * retires a request when its norm (every `chunk` iterations, counted from
    its own admission) reaches the tolerance, freeing its memory.

A request therefore performs the same iterations with the same kernels as
when it is solved alone, and its result is the same bits. The sequential
mode solves one request at a time in arrival order (the plain back-to-back
baseline). Times are wall-clock seconds from the start of the stream;
time-to-converge = retirement - arrival.

The scheduling is independent of the GPU: `Server` drives an executor with
start / step / finish, so the admission and retirement rules are testable
with a CPU stand-in.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .mesh import Mesh, rectangle, unit_quad_triag

FAMILIES = ("advection", "burgers", "euler")
GAMMA, GAS_R = 1.4, 287.05


@dataclass
class Request:
    rid: int
    arrival: float   # seconds after the start of the stream
    family: str      # 'advection' (P1-P3, N scheme), 'burgers' (P2, N scheme) or 'euler' (P1 / Q1, LDA)
    order: int
    cells: int       # target number of elements
    tol: float = 1e-10
    mach: float = 0.0  # euler: free-stream Mach number


def request_stream(n: int, rate: float = 20.0, seed: int = 5, adv_cells=(10_000, 1_000_000),
                   burgers_cells=(50_000, 500_000), tol: float = 1e-10, euler_cells=None,
                   mach=(0.3, 0.6)) -> list[Request]:
    """Seeded Poisson arrivals (mean `rate` per second); family uniform over
    advection / Burgers (and the Euler bump when euler_cells is given),
    advection order uniform in 1..3, Mach number uniform in `mach`, sizes
    uniform in the given ranges of elements. The Euler bump runs at P1 / Q1:
    with P2 / Q2 the LDA iteration of this channel diverges (no shock
    capturing; the reference's bump example is P1)."""
    fams = FAMILIES if euler_cells is not None else FAMILIES[:2]
    rng = np.random.Generator(np.random.MT19937(seed))
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        fam = fams[int(rng.integers(0, len(fams)))]
        m = 0.0
        if fam == "advection":
            order = int(rng.integers(1, 4))
            cells = int(rng.integers(adv_cells[0], adv_cells[1] + 1))
        elif fam == "burgers":
            order = 2
            cells = int(rng.integers(burgers_cells[0], burgers_cells[1] + 1))
        else:
            order = 1
            cells = int(rng.integers(euler_cells[0], euler_cells[1] + 1))
            m = float(rng.uniform(mach[0], mach[1]))
        out.append(Request(i, float(t[i]), fam, order, cells, tol, m))
    return out


@dataclass
class Problem:
    """Host data of one request: mesh, physics, scheme, initial state, the
    strong Dirichlet inflow (may be empty) and the weak boundary conditions
    (type, nodes, parameters)."""
    mesh: Mesh
    physics: int
    scheme: int
    params: tuple
    cfl: float
    u0: np.ndarray
    bc_nodes: np.ndarray
    bc_values: np.ndarray
    weak: list = field(default_factory=list)


def bump_problem(cells: int, order: int, mach: float) -> Problem:
    """The 2D channel [-0.5, 1.5] x [0, 1] over the bump y = 0.1 exp(-25 (x -
    0.5)^2) with 2n x n cells, the lower half quadrilaterals, the upper half
    split triangles (3 n^2 elements), free stream rho = 1, p = 1 / gamma at
    `mach`; weak walls below and above, subsonic inlet (total temperature and
    pressure of the free stream) and outlet (its pressure)."""
    n = max(2, 2 * int(round(np.sqrt(cells / 3.0) / 2)))
    mesh = rectangle(2 * n, n, extent=(-0.5, 1.5, 0.0, 1.0), split=2, quad_rows=n // 2, order=order,
                     bump=(1.0, 0.1, 25.0, 0.5))
    p = 1.0 / GAMMA
    fac = 1.0 + 0.5 * (GAMMA - 1.0) * mach ** 2
    u0 = np.zeros((mesh.nb_nodes, 4))
    u0[:, 0] = 1.0
    u0[:, 1] = mach
    u0[:, 3] = p / (GAMMA - 1.0) + 0.5 * mach ** 2
    b = mesh.boundaries
    weak = [("WeakWall", b["ymin"], {}), ("WeakWall", b["ymax"], {}),
            ("WeakSubInlet", b["xmin"], {"total_temperature": p / GAS_R * fac,
                                         "total_pressure": p * fac ** (GAMMA / (GAMMA - 1.0))}),
            ("WeakSubOutlet", b["xmax"], {"total_pressure": p})]
    empty = np.zeros(0, dtype=np.uint32)
    return Problem(mesh, L.EULER2D, L.LDA, (GAMMA, GAS_R), 0.5, u0, empty, np.zeros((0, 4)), weak)


def build_problem(req: Request) -> Problem:
    if req.family == "euler":
        return bump_problem(req.cells, req.order, req.mach)
    if req.family == "advection":
        # make_unit_quad(N) on [0,1]^2: 2 (N-1)^2 triangles
        N = max(3, int(round(np.sqrt(req.cells / 2.0))) + 1)
        mesh = unit_quad_triag(N, order=req.order)
        bn = np.union1d(mesh.boundaries["Edge0"], mesh.boundaries["Edge3"]).astype(np.uint32)
        g = np.where(mesh.node_coords[bn, 1] < 0.5, 1.0, 0.0)[:, None]
        return Problem(mesh, L.ADVECTION2D, L.N, (0.8, 0.6), 0.9, np.zeros((mesh.nb_nodes, 1)), bn, g)
    # Burgers: n x n cells split in 4 triangles, inflow 1.5 - 2x, 1.5 / -0.5 on the sides
    n = max(2, int(round(np.sqrt(req.cells / 4.0))))
    mesh = rectangle(n, n, split=4, order=req.order)
    b = mesh.boundaries
    ymin = b["ymin"]
    xmin = np.setdiff1d(b["xmin"], ymin)
    xmax = np.setdiff1d(b["xmax"], ymin)
    nodes = np.concatenate([ymin, xmin, xmax]).astype(np.uint32)
    vals = np.concatenate([1.5 - 2.0 * mesh.node_coords[ymin, 0], np.full(xmin.size, 1.5),
                           np.full(xmax.size, -0.5)])[:, None]
    x = mesh.node_coords[:, 0]
    # the N scheme: the B blend of this shocked flow does not settle below ~1e-3
    return Problem(mesh, L.BURGERS2D, L.N, (), 0.5, (1.5 - 2.0 * x)[:, None], nodes, vals)


def footprint_bytes(p: Problem) -> int:
    """Device bytes of one resident request (a bound on what pdrd_initialize
    allocates: nodal fields and accumulators, coordinates, connectivity in
    three layouts and the patch plan with its two buffers of shared-node
    rows and states)."""
    m = p.mesh
    neq = p.u0.shape[1]
    N, E = m.nb_nodes, m.nb_cells
    nd = max(b.sol_conn.shape[1] for b in m.blocks)
    ng = max(b.geo_conn.shape[1] for b in m.blocks)
    nodal = 8 * N * (3 * neq + 6) + 4 * N
    slots = 2 * N  # nodes plus their extra slots on patch borders, bounded
    # per slot: node id, slot row, pair offset, shared index, row range, two rows
    patch = slots * (4 + 4 + 2 + 4 + 8 + 2 * 8 * (neq + 1)) + 2 * 8 * neq * N + E * (2 * nd + 4 * ng)
    # weak facets: per boundary node and facet, two extra rows, its row and
    # node ids, the facet normal / Jacobian / rule
    nb = sum(int(np.asarray(w[1]).size) for w in p.weak)
    weak = nb * (2 * 2 * 8 * (neq + 1) + 2 * 4 * 2 + 8 * 8)
    return int(nodal + 8 * m.coords.size + E * 4 * (2 * nd + 2 * ng) + patch + weak + (1 << 20))


@dataclass
class Result:
    rid: int
    family: str
    order: int
    cells: int
    arrival: float
    admitted: float = 0.0
    finished: float = 0.0
    iterations: int = 0
    norm: float = float("nan")
    converged: bool = False
    solution: np.ndarray | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class GpuExecutor:
    """Requests as PGRDMethodExplicit handles on the current GPU, advanced by
    pdrd_batch_iterate."""

    def __init__(self):
        from .rdm import Batch
        self.batch = Batch()

    def start(self, p: Problem):
        from .rdm import PGRDMethodExplicit
        m = PGRDMethodExplicit(p.physics, p.scheme, params=p.params if p.params else None)
        m.configure_mesh_data(p.mesh).initialize_work_data(quad_order=6 if p.mesh.order == 3 else 4)
        for kind, nodes, par in p.weak:  # before the method's first device use
            m.add_boundary_condition(kind, kind, nodes, params=par)
        m.set_vec_function("solution", p.u0)
        if p.bc_nodes.size:
            m.add_boundary_condition("StrongDirichlet", "inflow", p.bc_nodes, p.bc_values)
        m.cfl = p.cfl
        return m

    def step(self, handles, n_iter: int):
        if len(handles) == 1:  # the plain path: pdrd_iterate of one method
            h = handles[0]
            return [float(np.max(h.iterate(h.cfl, n_iter, exclude_dirichlet=True)))]
        norms = self.batch.iterate(handles, [h.cfl for h in handles], n_iter, exclude_dirichlet=True)
        return [float(np.max(x)) for x in norms]

    def finish(self, h):
        u = h.vec_function("solution")
        h.destroy()
        return u


class Server:
    """Continuous batching under an HBM budget (mode 'batched') or one
    request at a time (mode 'sequential')."""

    def __init__(self, executor, hbm_budget: int = 288 * 10**9, chunk: int = 50, max_iters: int = 200_000,
                 max_active: int = 64, clock=time.perf_counter, sleep=time.sleep):
        self.ex = executor
        self.budget = hbm_budget
        self.chunk = chunk
        self.max_iters = max_iters
        self.max_active = max_active
        self.clock = clock
        self.sleep = sleep

    def run(self, requests: list[Request], problems: dict, mode: str = "batched", keep_solutions: bool = False,
            on_finish=None):
        """problems: rid -> Problem (host data prepared by the clients before
        the stream starts); on_finish(result), if given, is called as each
        request leaves the device. Returns the Results in request order and
        the largest device footprint held at once."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        results = {r.rid: Result(r.rid, r.family, r.order, r.cells, r.arrival) for r in reqs}
        pending = list(reqs)
        active = []  # [req, handle, footprint, iterations]
        used = 0
        peak = 0
        cap = 1 if mode == "sequential" else self.max_active
        t0 = self.clock()
        while pending or active:
            now = self.clock() - t0
            # admission in arrival order while the footprint fits
            while pending and pending[0].arrival <= now and len(active) < cap:
                r = pending[0]
                fp = footprint_bytes(problems[r.rid])
                if fp > self.budget:
                    raise ValueError(f"request {r.rid} needs {fp} bytes, more than the HBM budget")
                if active and used + fp > self.budget:
                    break
                pending.pop(0)
                h = self.ex.start(problems[r.rid])
                results[r.rid].admitted = self.clock() - t0
                active.append([r, h, fp, 0])
                used += fp
                peak = max(peak, used)
            if not active:
                # idle until the next arrival
                wait = pending[0].arrival - (self.clock() - t0)
                if wait > 0:
                    self.sleep(min(wait, 0.05))
                continue
            norms = self.ex.step([a[1] for a in active], self.chunk)
            keep = []
            for a, nrm in zip(active, norms):
                a[3] += self.chunk
                res = results[a[0].rid]
                res.iterations = a[3]
                res.norm = nrm
                if nrm <= a[0].tol or a[3] >= self.max_iters or not np.isfinite(nrm):
                    res.converged = bool(nrm <= a[0].tol)
                    u = self.ex.finish(a[1])
                    res.finished = self.clock() - t0
                    if keep_solutions:
                        res.solution = u
                    used -= a[2]
                    if on_finish is not None:
                        on_finish(res)
                else:
                    keep.append(a)
            active = keep
        return [results[r.rid] for r in requests], peak


def summary(results: list[Result]) -> dict:
    lat = np.array([r.latency for r in results])
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    return {"requests": len(results), "converged": int(sum(r.converged for r in results)),
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "throughput_req_s": len(results) / span,
            "iterations": int(sum(r.iterations for r in results))}
