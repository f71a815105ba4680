"""Element residuals and Jacobians of the implicit forms on the GPU
(pdrd_element_jacobians) against the oracle, ≤ 1e-12 normwise relative per
block: LDA, N and B on the parity meshes (scalar P1-P3 / Q1 / Q2, Euler 2D
P1 / P2 / P1 + Q1 / P2 + Q2, Euler 3D P1 / P2), B with an external blending
coefficient and LDA with an artificial viscosity; PSI is refused."""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import CASES, oracle_problem, quad_orders, rel_err

pytestmark = pytest.mark.gpu

NAMES = sorted(n for n in CASES if n not in ("burgers_p1",))


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("scheme", [L.LDA, L.N, L.B])
def test_element_jacobians_match_oracle(name, scheme):
    build, physics, params, state = CASES[name]
    mesh = build()
    u = state(mesh)
    print(f"jacobians {name} scheme {scheme} cells {mesh.nb_cells}", flush=True)
    m = PGRDMethodExplicit(physics, scheme, params=params if params else None)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    m.set_vec_function("solution", u)
    P = oracle_problem(mesh, physics, scheme, params)
    for blk in range(len(mesh.blocks)):
        phi, jac, wave = m.element_jacobians(blk)
        rphi, rjac, rwave, rc = P.cell_jacobians(blk, u)
        assert rc == 0
        assert rel_err(phi, rphi) <= 1e-12
        assert rel_err(jac, rjac) <= 1e-12
        assert rel_err(wave, rwave) <= 1e-12
    m.check_status()
    m.destroy()


@pytest.mark.parametrize("scheme", [L.LDA, L.B])
def test_element_jacobians_with_nodal_inputs(scheme):
    build, physics, params, state = CASES["euler2d_mixed"]
    mesh = build()
    u = state(mesh)
    v = np.random.default_rng(4).uniform(0.0, 1.0, mesh.nb_nodes) * (0.05 if scheme == L.LDA else 1.0)
    m = PGRDMethodExplicit(physics, scheme, params=params)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    m.set_vec_function("solution", u)
    P = oracle_problem(mesh, physics, scheme, params)
    if scheme == L.LDA:
        m.set_artificial_viscosity(v)
        P.set_artificial_viscosity(v)
    else:
        m.set_blending_coeff(v)
        P.set_blending_coeff(v)
    for blk in range(len(mesh.blocks)):
        phi, jac, wave = m.element_jacobians(blk)
        rphi, rjac, rwave, rc = P.cell_jacobians(blk, u)
        assert rel_err(phi, rphi) <= 1e-12
        assert rel_err(jac, rjac) <= 1e-12


def test_psi_has_no_implicit_form():
    build, physics, params, state = CASES["adv_p1"]
    mesh = build()
    m = PGRDMethodExplicit(physics, L.PSI, params=params)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    m.set_vec_function("solution", state(mesh))
    with pytest.raises(L.PdrdError):
        m.element_jacobians(0)
