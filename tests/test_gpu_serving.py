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
