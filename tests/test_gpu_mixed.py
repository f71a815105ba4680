"""Mixed-precision cells (PDRD_OPT_PRECISION = 1): FP32 interpolation, flux
differences, divergence and K+, FP64 sums of K+, solves, residuals and
everything after them.

* one sweep: residuals within 2e-5 normwise of the FP64 oracle on every
  case and scheme (FP32 rounding of O(h |grad u|) flux differences);
* steady state: the FP64 iteration is run to a residual norm of 1e-12; the
  mixed iteration, whose residual norm stalls at the FP32 rounding level
  (about 1e-8 here), is run for the same number of iterations plus a margin,
  with and without the defect correction (every 50 iterations
  R64(u) - Rmixed(u) recomputed and added to the mixed residual); its
  solution must lie within 1e-6 of the FP64 steady state
  (max |u_mixed - u_plain| / max |u_plain|). The figures are printed per case.
"""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.mesh import rectangle, unit_quad_triag
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import CASES, oracle_problem, quad_orders, rel_err

pytestmark = pytest.mark.gpu

SWEEP_TOL = 2e-5


def _method(mesh, physics, scheme, params, precision):
    m = PGRDMethodExplicit(physics, scheme, params=params if params else None)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    m.set_option(L.OPT_PRECISION, precision)
    assert m.get_option(L.OPT_PRECISION) == precision and m.get_option(L.OPT_SCATTER) == 2
    return m


_cache = {}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("scheme", [L.LDA, L.N, L.B, L.PSI])
def test_mixed_sweep_close_to_oracle(name, scheme):
    if name not in _cache:
        build, physics, params, state = CASES[name]
        mesh = build()
        _cache[name] = (mesh, physics, params, state(mesh))
    mesh, physics, params, u = _cache[name]
    print(f"mixed sweep {name} scheme {scheme}", flush=True)
    m = _method(mesh, physics, scheme, params, 1)
    m.set_vec_function("solution", u)
    m.assemble_lhs_and_rhs(0.5)
    m.check_status()
    ref, ref_ws, rc = oracle_problem(mesh, physics, scheme, params).assemble_rhs(u)
    assert rc == 0
    e = rel_err(m.vec_function("residuals"), ref)
    print(f"  residual rel err {e:.3e}", flush=True)
    assert e <= SWEEP_TOL
    assert rel_err(m.vec_function("wave_speed"), ref_ws) <= SWEEP_TOL
    m.destroy()


def _steady(mesh, physics, scheme, params, u0, bn, g, cfl, precision, tol=1e-12, chunk=200, cap=40_000, refresh=50):
    m = _method(mesh, physics, scheme, params, precision)
    m.set_option(L.OPT_REFRESH, refresh)
    m.set_vec_function("solution", u0)
    m.add_boundary_condition("StrongDirichlet", "inflow", bn, g)
    it = 0
    while True:
        nrm = m.iterate(cfl, chunk, exclude_dirichlet=True)
        it += chunk
        if nrm.max() <= tol or it >= cap:
            break
    return m.vec_function("solution"), nrm.max(), it


@pytest.mark.parametrize("case", ["advection_p1_n", "advection_p2_lda", "burgers_p2_n"])
def test_mixed_steady_state_matches_plain(case):
    if case.startswith("advection"):
        order = 1 if case.endswith("p1_n") else 2
        scheme = L.N if order == 1 else L.LDA
        mesh = unit_quad_triag(33, order=order)
        bn = np.union1d(mesh.boundaries["Edge0"], mesh.boundaries["Edge3"]).astype(np.uint32)
        g = np.where(mesh.node_coords[bn, 1] < 0.5, 1.0, 0.0)[:, None]
        args = (mesh, L.ADVECTION2D, scheme, (0.8, 0.6), np.zeros((mesh.nb_nodes, 1)), bn, g, 0.9)
    else:
        mesh = rectangle(16, 16, split=4, order=2, jitter=0.15, seed=2)
        b = mesh.boundaries
        ymin = b["ymin"]
        xmin = np.setdiff1d(b["xmin"], ymin)
        xmax = np.setdiff1d(b["xmax"], ymin)
        bn = np.concatenate([ymin, xmin, xmax]).astype(np.uint32)
        g = np.concatenate([1.5 - 2.0 * mesh.node_coords[ymin, 0], np.full(xmin.size, 1.5),
                            np.full(xmax.size, -0.5)])[:, None]
        u0 = (1.5 - 2.0 * mesh.node_coords[:, 0])[:, None]
        args = (mesh, L.BURGERS2D, L.N, (), u0, bn, g, 0.5)
    print(f"steady {case}", flush=True)
    up, np_, ip = _steady(*args, precision=0)
    um, nm, im = _steady(*args, precision=1, cap=2 * ip + 2000)
    u0m, n0m, i0m = _steady(*args, precision=1, refresh=0, cap=2 * ip + 2000)
    d = float(np.abs(um - up).max() / np.abs(up).max())
    d0 = float(np.abs(u0m - up).max() / np.abs(up).max())
    print(f"  plain: norm {np_:.2e} after {ip} it; mixed + correction: norm {nm:.2e} after {im} it, "
          f"state difference {d:.3e}; mixed alone: norm {n0m:.2e} after {i0m} it, state difference {d0:.3e}",
          flush=True)
    assert np_ <= 1e-12 and nm <= 1e-6 and n0m <= 1e-6
    assert d <= 1e-6 and d0 <= 1e-6


def test_mixed_needs_the_patch_assembly():
    build, physics, params, state = CASES["adv_p1"]
    mesh = build()
    m = PGRDMethodExplicit(physics, L.N, params=params)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh), scatter=1)
    m.set_option(L.OPT_PRECISION, 1)
    m.set_vec_function("solution", state(mesh))
    with pytest.raises(L.PdrdError):
        m.assemble_lhs_and_rhs(0.5)
