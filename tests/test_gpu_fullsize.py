"""Parity at the benchmark's full sizes: the six workloads of bench.py (C1
advection P1 32 768 triangles; C2 Burgers P2 B 1 048 576 triangles; C3 Euler
P1 + Q1 channel 786 432 elements; C4 Euler box 5 308 416 P1 tetrahedra; the
P2 / Q2 channel and the P2 tetrahedron box), each with its own state and
strong Dirichlet data. One GPU iteration from the workload's state against
one iteration of the oracle (its cell loop split over host threads, as the
CPU baseline runs it): residual, time step and updated state, <= 1e-12
normwise relative. Size-independent properties besides: a second sweep gives
the same bits, and a uniform state (free stream for Euler, a constant for the
scalar equations) has residuals at the rounding level of its fluxes."""
import os

import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd import workloads as W
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import oracle_problem, rel_err

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _method(wl, state):
    m = PGRDMethodExplicit(wl.physics, wl.scheme, params=wl.params if wl.params else None)
    m.configure_mesh_data(wl.mesh).initialize_work_data(quad_order=wl.quad_order)
    m.set_vec_function("solution", state)
    if wl.dirichlet_nodes.size:
        m.add_boundary_condition("StrongDirichlet", "inflow", wl.dirichlet_nodes, wl.dirichlet_values)
    return m


@pytest.mark.parametrize("name", ["c1", "c2", "c3", "c4", "c3p2", "c4p2"])
def test_full_size_iteration_matches_oracle(name):
    wl = getattr(W, name)()
    print(f"full size {wl.name}: {wl.mesh.nb_cells} elements, {wl.mesh.nb_nodes} nodes", flush=True)
    m = _method(wl, wl.state)
    m.assemble_lhs_and_rhs(wl.cfl)
    m.check_status()
    res, dt = m.vec_function("residuals"), m.vec_function("time_step")
    m.set_vec_function("solution", wl.state)  # the sweep applied the Dirichlet values
    m.assemble_lhs_and_rhs(wl.cfl)
    assert np.array_equal(res, m.vec_function("residuals"))
    m.set_vec_function("solution", wl.state)
    m.iterate(wl.cfl, 1)
    u1 = m.vec_function("solution")
    m.destroy()

    P = oracle_problem(wl.mesh, wl.physics, wl.scheme, wl.params, wl.quad_order)
    dn = wl.dirichlet_nodes if wl.dirichlet_nodes.size else None
    dv = wl.dirichlet_values if wl.dirichlet_nodes.size else None
    uo, ro, dto, rc = P.iterate_threads(wl.state, wl.cfl, 1, THREADS, dn, dv)
    assert rc == 0
    e_res, e_dt, e_u = rel_err(res, ro), rel_err(dt, dto), rel_err(u1, uo)
    print(f"  residual {e_res:.2e}, time step {e_dt:.2e}, state {e_u:.2e}", flush=True)
    assert e_res <= 1e-12 and e_dt <= 1e-12 and e_u <= 1e-12


@pytest.mark.parametrize("name", ["c2", "c3", "c4"])
def test_full_size_uniform_state_has_no_residual(name):
    wl = getattr(W, name)()
    if wl.physics in (L.EULER2D, L.EULER3D):
        dim = 2 if wl.physics == L.EULER2D else 3
        u = np.zeros_like(wl.state)
        rho, mach, gamma = 1.0, 0.5, 1.4
        u[:, 0] = rho
        u[:, 1] = rho * mach
        u[:, -1] = (1.0 / gamma) / (gamma - 1.0) + 0.5 * rho * mach ** 2
        scale = rho * mach ** 2 + 1.0 / gamma  # momentum flux
        assert u.shape[1] == dim + 2
    else:
        u = np.full_like(wl.state, 0.7)
        scale = 0.7 ** 2
    m = PGRDMethodExplicit(wl.physics, wl.scheme, params=wl.params if wl.params else None)
    m.configure_mesh_data(wl.mesh).initialize_work_data(quad_order=wl.quad_order)
    m.set_vec_function("solution", u)
    m.assemble_rhs()
    m.check_status()
    r = m.vec_function("residuals")
    ws = m.vec_function("wave_speed")
    m.destroy()
    # residual = sum Phi / V with Phi ~ |K| F rounding: relative to F / h
    h = 1.0 / np.sqrt(wl.mesh.nb_cells) if wl.mesh.dim == 2 else wl.mesh.nb_cells ** (-1.0 / 3.0)
    worst = float(np.abs(r).max()) * h / scale
    print(f"uniform state {wl.name}: max |residual| h / |F| = {worst:.2e}", flush=True)
    assert np.isfinite(ws).all() and ws.min() > 0
    assert worst <= 1e-10
