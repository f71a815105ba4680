"""Serving loop on the GPU: a small seeded stream of advection (P1-P3) and
Burgers (P2) steady solves, run with continuous batching and one at a time;
every request gives the same solution bits and iteration count alone and in
the mix, and converges to the tolerance."""
import numpy as np
import pytest

from pdekit_amd import serving as S

pytestmark = pytest.mark.gpu


def test_results_alone_and_in_the_mix_are_identical():
    reqs = S.request_stream(8, rate=400.0, seed=5, adv_cells=(1_000, 6_000), burgers_cells=(1_000, 6_000),
                            tol=1e-10)
    probs = {r.rid: S.build_problem(r) for r in reqs}
    out = {}
    for mode in ("batched", "sequential"):
        print(f"serving mode {mode}: {len(reqs)} requests", flush=True)
        srv = S.Server(S.GpuExecutor(), chunk=50, max_iters=40_000)
        res, peak = srv.run(reqs, probs, mode=mode, keep_solutions=True)
        print(mode, S.summary(res), flush=True)
        out[mode] = res
    for a, b in zip(out["batched"], out["sequential"]):
        assert a.rid == b.rid
        assert a.converged and b.converged, (a.rid, a.norm, b.norm)
        assert a.iterations == b.iterations
        assert np.array_equal(a.solution, b.solution)


def test_footprint_bounds_the_device_allocation():
    """The admission footprint is an upper bound of what a request's method
    allocates on the device."""
    reqs = S.request_stream(6, rate=400.0, seed=11, adv_cells=(500, 20_000), burgers_cells=(500, 20_000))
    ex = S.GpuExecutor()
    for r in reqs:
        p = S.build_problem(r)
        m = ex.start(p)
        used = m.info()["device_bytes"]
        print(f"request {r.rid} {r.family} P{r.order} {p.mesh.nb_cells} cells: {used} of {S.footprint_bytes(p)} B",
              flush=True)
        assert 0 < used <= S.footprint_bytes(p)
        ex.finish(m)
