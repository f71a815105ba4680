"""Cells per patch (PDRD_OPT_PATCH_CELLS) and the autotuner on the GPU: every
patch size gives the oracle's residual, and the tuner returns a size in use."""
import pytest

from pdekit_amd import _lib as L
from pdekit_amd import autotune as A
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import CASES, oracle_problem, quad_orders, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["burgers_p2", "euler2d_mixed", "euler3d_p1", "adv_p3"])
@pytest.mark.parametrize("cells", [64, 128, 256])
def test_patch_sizes_match_oracle(name, cells):
    build, physics, params, state = CASES[name]
    mesh = build()
    u = state(mesh)
    print(f"patch cells {cells} case {name}", flush=True)
    m = PGRDMethodExplicit(physics, L.B, params=params if params else None)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh), patch_cells=cells)
    assert m.get_option(L.OPT_SCATTER) == 2
    used = m.get_option(L.OPT_PATCH_CELLS)
    assert used <= cells and used in (64, 128, 256)
    m.set_vec_function("solution", u)
    m.assemble_lhs_and_rhs(0.5)
    ref, ref_ws, rc = oracle_problem(mesh, physics, L.B, params).assemble_rhs(u)
    assert rel_err(m.vec_function("residuals"), ref) <= 1e-12
    assert rel_err(m.vec_function("wave_speed"), ref_ws) <= 1e-12


def test_tuner_picks_a_candidate(tmp_path):
    build, physics, params, state = CASES["burgers_p2"]
    mesh = build()
    path = str(tmp_path / "cells.json")
    best = A.tune(mesh, physics, L.B, params, quad_orders(mesh), state(mesh), 0.5, cache_path=path, iters=3)
    assert best in (64, 128, 256)
    key = A.cache_key(mesh, physics, L.B, 0, quad_orders(mesh))
    entry = A.load_cache(path)[key]
    assert entry["cells"] == best and len(entry["ms_per_iteration"]) >= 2
    assert A.tune(mesh, physics, L.B, params, quad_orders(mesh), state(mesh), 0.5, cache_path=path) == best
