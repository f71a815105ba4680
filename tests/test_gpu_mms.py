"""The reference's end-to-end Euler test (utest-Euler-continuous-rdm.cpp:
454-623, Euler2D_explicit_utest) on the GPU: a manufactured smooth solution
of the 2D Euler equations on make_unit_quad(20) mapped to [0,1]^2 (P1, rule
of order 4, PGLDA), its source term as a nodal source field, WeakDirichlet
on the four edges with the exact state at every facet point, the constant
initial state (1, 77.2, 82.6, p = 2.94e5), CFL 0.5, 2 500 iterations; the
L2 norm of the density error must stay below 5e-3, as the reference checks.

The manufactured fields are the test's table (utest-Euler-continuous-rdm.cpp:
156-172: rho = 1 + 0.15 sin(pi x) - 0.1 cos(pi y / 2), u = 70 + 5 sin(1.5 pi
x) - 7 cos(0.6 pi y), v = 90 - 15 cos(pi x / 2) + 8.5 sin(2 pi y / 3),
p = 1e5 + 2e4 cos(2 pi x) + 5e4 sin(pi y)); the source term, div F of the
exact state, is differentiated here by complex steps of the fluxes.

Also: one sweep and 10 iterations of the same set-up against the oracle,
<= 1e-12 normwise relative."""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.mesh import unit_quad_triag
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import rel_err
from tests.mms import GAMMA, QORDER, density_l2_error, exact_state, initial, oracle, source

pytestmark = pytest.mark.gpu


def _gpu(mesh):
    m = PGRDMethodExplicit(L.EULER2D, L.LDA, params=(GAMMA, 287.05))
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=QORDER)
    edges = ["Edge0", "Edge1", "Edge2", "Edge3"]
    for e in edges:
        m.add_boundary_condition("WeakDirichlet", e, mesh.boundaries[e], values=initial(1)[0])
    m.set_vec_function("solution", initial(mesh.nb_nodes))
    for k in range(len(edges)):
        xq = m.weak_bc_points(k)
        m.set_weak_bc_state(k, exact_state(xq[..., 0], xq[..., 1]))
    x, y = mesh.node_coords[:, 0], mesh.node_coords[:, 1]
    m.set_vec_function("sources", source(x, y))
    return m


def test_mms_sweep_and_iterations_match_oracle():
    mesh = unit_quad_triag(20)
    m = _gpu(mesh)
    P = oracle(mesh)
    u0 = initial(mesh.nb_nodes)
    m.assemble_rhs()
    ref, ref_ws, rc = P.assemble_rhs(u0)
    assert rc == 0
    assert rel_err(m.vec_function("residuals"), ref) <= 1e-12
    assert rel_err(m.vec_function("wave_speed"), ref_ws) <= 1e-12
    m.set_vec_function("solution", u0)
    m.iterate(0.5, 10)
    uo, res, dt, norms, rc = P.iterate(u0, 0.5, 10)
    assert rc == 0
    assert rel_err(m.vec_function("solution"), uo) <= 1e-12
    m.destroy()


def test_mms_converges_as_the_reference_requires():
    mesh = unit_quad_triag(20)
    m = _gpu(mesh)
    for _ in range(25):
        norm = m.iterate(0.5, 100)
    m.check_status()
    u = m.vec_function("solution")
    m.destroy()
    l2 = density_l2_error(mesh, u)
    print(f"MMS Euler 2D, make_unit_quad(20), LDA, 2500 iterations: residual norm {norm.max():.3e}, "
          f"density L2 error {l2:.3e}", flush=True)
    assert np.isfinite(l2) and l2 <= 5e-3
