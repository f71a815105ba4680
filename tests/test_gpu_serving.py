"""Serving loop on the GPU: a small seeded stream of advection (P1-P3),
Burgers (P2) and Euler bump (P1 / Q1, weak walls, subsonic inlet / outlet)
steady solves, run with continuous batching and one at a time; every request
gives the same solution bits and iteration count alone and in the mix, and
converges to the tolerance. One request of each family is checked against
the oracle: its converged state is a steady state of the oracle's discrete
operator too (the oracle's residual there is at the tolerance's level)."""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd import serving as S
from tests.helpers import add_oracle_weak, oracle_problem, quad_orders

pytestmark = pytest.mark.gpu


def _stream():
    return S.request_stream(9, rate=400.0, seed=5, adv_cells=(1_000, 6_000), burgers_cells=(1_000, 6_000),
                            euler_cells=(600, 1_500), tol=1e-10)


def _oracle_residual(p: S.Problem, u):
    """norm_L2 of the oracle's residual at a state, without the strong
    Dirichlet rows. The state is taken as the iteration leaves it: its
    Dirichlet nodes hold g - dt r (the reference applies the value before the
    update and never zeroes those rows, StrongDirichletBC.hpp:202-233), which
    is what the cells see in the next sweep."""
    P = oracle_problem(p.mesh, p.physics, p.scheme, p.params, quad_orders(p.mesh))
    for kind, nodes, par in p.weak:
        keys = {"WeakSubInlet": ("total_temperature", "total_pressure"), "WeakSubOutlet": ("total_pressure",)}
        add_oracle_weak(P, p.mesh, kind, nodes, tuple(par[k] for k in keys.get(kind, ())),
                        quad_order=max(quad_orders(p.mesh).values()))
    res, _, rc = P.assemble_rhs(np.array(u))
    assert rc == 0
    keep = np.ones(p.mesh.nb_nodes, bool)
    keep[p.bc_nodes] = False
    return float(np.sqrt((res[keep] ** 2).mean(0)).max())


def test_results_alone_and_in_the_mix_are_identical():
    reqs = _stream()
    assert {r.family for r in reqs} == {"advection", "burgers", "euler"}
    probs = {r.rid: S.build_problem(r) for r in reqs}
    out = {}
    for mode in ("batched", "sequential"):
        print(f"serving mode {mode}: {len(reqs)} requests", flush=True)
        srv = S.Server(S.GpuExecutor(), chunk=50, max_iters=120_000)
        res, peak = srv.run(reqs, probs, mode=mode, keep_solutions=True)
        print(mode, S.summary(res), flush=True)
        out[mode] = res
    checked = set()
    for a, b in zip(out["batched"], out["sequential"]):
        assert a.rid == b.rid
        assert a.converged and b.converged, (a.rid, a.family, a.norm, b.norm)
        assert a.iterations == b.iterations
        assert np.array_equal(a.solution, b.solution)
        if a.family not in checked:
            checked.add(a.family)
            r = _oracle_residual(probs[a.rid], a.solution)
            print(f"request {a.rid} {a.family}: {a.iterations} iterations, oracle residual {r:.2e}", flush=True)
            assert r <= 1e-9


def test_footprint_bounds_the_device_allocation():
    """The admission footprint is an upper bound of what a request's method
    allocates on the device."""
    reqs = S.request_stream(9, rate=400.0, seed=11, adv_cells=(500, 20_000), burgers_cells=(500, 20_000),
                            euler_cells=(500, 20_000))
    ex = S.GpuExecutor()
    for r in reqs:
        p = S.build_problem(r)
        m = ex.start(p)
        used = m.info()["device_bytes"]
        print(f"request {r.rid} {r.family} P{r.order} {p.mesh.nb_cells} cells: {used} of {S.footprint_bytes(p)} B",
              flush=True)
        assert 0 < used <= S.footprint_bytes(p)
        ex.finish(m)
