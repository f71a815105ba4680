"""Scheduling of the serving loop (pdekit_amd/serving.py) with a CPU stand-in
for the GPU executor and a simulated clock: admission in arrival order under
the HBM budget, retirement at the tolerance checked every chunk counted from
admission, identical per-request iteration counts alone and in the mix."""
import numpy as np
import pytest

from pdekit_amd import serving as S


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class FakeExecutor:
    """A request converges geometrically, rate set by its size; one step of
    the batch costs time proportional to the cells it advances."""

    def __init__(self, clock, cost_per_cell_iter=1e-9):
        self.clock = clock
        self.cost = cost_per_cell_iter
        self.live = {}
        self.max_live = 0

    def start(self, p):
        h = object()
        self.live[h] = [p, 1.0]
        self.max_live = max(self.max_live, len(self.live))
        return h

    def step(self, handles, n):
        out = []
        for h in handles:
            p, nrm = self.live[h]
            rate = 0.9 + 0.09 * min(1.0, p.mesh.nb_cells / 1e5)
            nrm *= rate ** n
            self.live[h][1] = nrm
            out.append(nrm)
            self.clock.t += self.cost * p.mesh.nb_cells * n
        return out

    def finish(self, h):
        p, _ = self.live.pop(h)
        return np.zeros((p.mesh.nb_nodes, 1))


def _problems(reqs):
    return {r.rid: S.build_problem(r) for r in reqs}


def test_request_stream_is_seeded_poisson():
    a = S.request_stream(400, rate=20.0, seed=5)
    b = S.request_stream(400, rate=20.0, seed=5)
    assert [(r.arrival, r.family, r.order, r.cells) for r in a] == [(r.arrival, r.family, r.order, r.cells) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 20.0) < 0.15 / 20.0  # exponential gaps of mean 1/rate
    assert {r.family for r in a} == {"advection", "burgers"}
    assert all(1 <= r.order <= 3 for r in a if r.family == "advection")
    assert all(r.order == 2 and 50_000 <= r.cells <= 500_000 for r in a if r.family == "burgers")


def test_problems_have_the_requested_sizes():
    for r in S.request_stream(12, seed=9, adv_cells=(2_000, 20_000), burgers_cells=(2_000, 20_000)):
        p = S.build_problem(r)
        assert abs(p.mesh.nb_cells - r.cells) / r.cells < 0.1
        assert p.mesh.order == r.order
        assert p.bc_nodes.size == p.bc_values.shape[0] > 0
        assert S.footprint_bytes(p) > 8 * p.mesh.nb_nodes * 4


@pytest.mark.parametrize("mode", ["batched", "sequential"])
def test_admission_respects_budget_and_order(mode):
    reqs = S.request_stream(16, rate=200.0, seed=3, adv_cells=(2_000, 20_000), burgers_cells=(2_000, 20_000))
    probs = _problems(reqs)
    fps = sorted(S.footprint_bytes(p) for p in probs.values())
    budget = fps[-1] + fps[-2]  # at most two of the largest at once
    clk = Clock()
    ex = FakeExecutor(clk)
    srv = S.Server(ex, hbm_budget=budget, chunk=10, clock=clk, sleep=clk.sleep)
    res, peak = srv.run(reqs, probs, mode=mode)
    assert peak <= budget
    assert all(r.converged for r in res)
    adm = [r.admitted for r in sorted(res, key=lambda r: r.arrival)]
    assert adm == sorted(adm)  # first come, first admitted
    assert all(r.admitted >= r.arrival and r.finished > r.admitted for r in res)
    assert all(r.iterations % 10 == 0 for r in res)
    if mode == "sequential":
        assert ex.max_live == 1


def test_same_iterations_alone_and_in_the_mix():
    reqs = S.request_stream(10, rate=100.0, seed=4, adv_cells=(2_000, 20_000), burgers_cells=(2_000, 20_000))
    probs = _problems(reqs)
    out = {}
    for mode in ("batched", "sequential"):
        clk = Clock()
        srv = S.Server(FakeExecutor(clk), chunk=25, clock=clk, sleep=clk.sleep)
        res, _ = srv.run(reqs, probs, mode=mode)
        out[mode] = {r.rid: (r.iterations, r.norm) for r in res}
        summ = S.summary(res)
        assert summ["p99_s"] >= summ["p50_s"] > 0
    assert out["batched"] == out["sequential"]


def test_oversized_request_is_rejected():
    reqs = S.request_stream(2, seed=1, adv_cells=(2_000, 3_000), burgers_cells=(2_000, 3_000))
    probs = _problems(reqs)
    clk = Clock()
    srv = S.Server(FakeExecutor(clk), hbm_budget=1000, clock=clk, sleep=clk.sleep)
    with pytest.raises(ValueError):
        srv.run(reqs, probs)
