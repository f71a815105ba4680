"""Chained iterations of the patch assembly: within one iterate call the
shared nodes of a sweep are finished in the next sweep's gather (no fix-up
launch between sweeps, the row buffers alternating).
The state, residual, time step and norm after iterate(n) must equal, bit for
bit, those of n calls of iterate(1), each of which closes its sweep with the
fix-up kernel -- for scalar and Euler physics, mixed element types, strong
Dirichlet and weak boundary conditions, the eager timing mode, and small
patches (many shared nodes); in the mixed-precision form (linear sweeps
between FP64 ones) a chained call equals the same call with a fix-up after
every sweep. The chained state is also checked against the oracle."""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.mesh import box_tetra, rectangle, unit_quad_triag
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import burgers_state, euler_state, oracle_problem, quad_orders, rel_err

pytestmark = pytest.mark.gpu

GAMMA, RG = 1.4, 287.046
P_INF = 1.0 / GAMMA
T_INF = P_INF / RG
FAC = 1 + 0.5 * (GAMMA - 1) * 0.25


def _euler_bcs(mesh):
    b = mesh.boundaries
    walls = ("ymin", "ymax") if mesh.dim == 2 else ("zmin", "zmax", "ymin", "ymax")
    out = [("WeakWall", b[w], {}) for w in walls]
    out.append(("WeakSubInlet", b["xmin"], {"total_temperature": T_INF * FAC,
                                           "total_pressure": P_INF * FAC ** (GAMMA / (GAMMA - 1))}))
    out.append(("WeakSubOutlet", b["xmax"], {"total_pressure": P_INF}))
    return out


def _advection(options):
    mesh = unit_quad_triag(33)
    m = PGRDMethodExplicit(L.ADVECTION2D, L.N, params=(0.8, 0.6))
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=4, **options)
    bn = np.union1d(mesh.boundaries["Edge0"], mesh.boundaries["Edge3"]).astype(np.uint32)
    g = np.where(mesh.node_coords[bn, 1] < 0.5, 1.0, 0.0)[:, None]
    m.add_boundary_condition("StrongDirichlet", "inflow", bn, g)
    m.set_vec_function("solution", np.zeros((mesh.nb_nodes, 1)))
    return m


def _burgers(options):
    mesh = rectangle(12, 12, extent=(0.0, 1.0, 0.0, 1.0), split=2, order=2)
    m = PGRDMethodExplicit(L.BURGERS2D, L.B)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh), **options)
    m.set_vec_function("solution", burgers_state(mesh))
    return m


def _euler2d(options):
    mesh = rectangle(16, 8, extent=(-0.5, 1.5, 0.0, 1.0), split=2, quad_rows=4, bump=(1.0, 0.1, 25.0, 0.5))
    m = PGRDMethodExplicit(L.EULER2D, L.LDA, params=(GAMMA, 287.05))
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh), **options)
    for kind, nodes, par in _euler_bcs(mesh):
        m.add_boundary_condition(kind, kind, nodes, params=par)
    m.set_vec_function("solution", euler_state(mesh))
    return m


def _euler3d(options):
    mesh = box_tetra(5, 5, 3, extent=(1.0, 1.0, 0.5), jitter=0.1, seed=4, bump=(0.5, 0.08, 25.0, 0.5, 0.5))
    m = PGRDMethodExplicit(L.EULER3D, L.N, params=(GAMMA, 287.05))
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh), **options)
    for kind, nodes, par in _euler_bcs(mesh):
        m.add_boundary_condition(kind, kind, nodes, params=par)
    m.set_vec_function("solution", euler_state(mesh, dim=3))
    return m


CASES = {"advection_p1_dirichlet": (_advection, 0.9), "burgers_p2_b": (_burgers, 0.5),
         "euler2d_mixed_weak": (_euler2d, 0.3), "euler3d_weak": (_euler3d, 0.3)}
FIELDS = ("solution", "residuals", "time_step", "wave_speed")


def _run(make, cfl, n, chained, options=None, precision=0, refresh=0, timing=0, chain=1):
    m = make(options or {})
    m.set_option(L.OPT_CHAIN, chain)
    assert m.get_option(L.OPT_CHAIN) == chain
    if precision:
        m.set_option(L.OPT_PRECISION, precision)
        m.set_option(L.OPT_REFRESH, refresh)
    if timing:
        m.set_option(L.OPT_TIMING, 1)
    if chained:
        norm = m.iterate(cfl, n)
    else:
        for _ in range(n):
            norm = m.iterate(cfl, 1)
    m.check_status()
    out = {f: m.vec_function(f).copy() for f in FIELDS}
    m.destroy()
    return out, norm


def _same(a, b):
    (fa, na), (fb, nb) = a, b
    for f in FIELDS:
        assert np.array_equal(fa[f], fb[f]), f
    assert np.array_equal(na, nb)


@pytest.mark.parametrize("name", sorted(CASES))
def test_chained_iterations_equal_closed_sweeps(name):
    print(f"chain {name}", flush=True)
    make, cfl = CASES[name]
    # 7 sweeps: the opening sweep, deferred sweeps of both buffer parities,
    # the closing sweep
    _same(_run(make, cfl, 7, True), _run(make, cfl, 7, False))


@pytest.mark.parametrize("name", ["advection_p1_dirichlet", "euler2d_mixed_weak"])
def test_chain_option_off_gives_the_same_iterations(name):
    make, cfl = CASES[name]
    _same(_run(make, cfl, 7, True, chain=1), _run(make, cfl, 7, True, chain=0))


def test_chain_option_default_and_range():
    make, _ = CASES["advection_p1_dirichlet"]
    m = make({})
    assert m.get_option(L.OPT_CHAIN) == 1  # scalar physics
    m.set_option(L.OPT_CHAIN, 0)
    assert m.get_option(L.OPT_CHAIN) == 0
    with pytest.raises(RuntimeError):
        m.set_option(L.OPT_CHAIN, 2)
    m.destroy()
    m = CASES["euler2d_mixed_weak"][0]({})
    assert m.get_option(L.OPT_CHAIN) == 0  # Euler systems
    m.destroy()


@pytest.mark.parametrize("name", ["burgers_p2_b", "euler3d_weak"])
def test_chained_iterations_in_timing_mode(name):
    print(f"chain timing {name}", flush=True)
    make, cfl = CASES[name]
    _same(_run(make, cfl, 6, True, timing=1), _run(make, cfl, 6, False))


@pytest.mark.parametrize("name", ["advection_p1_dirichlet", "euler2d_mixed_weak"])
def test_chained_iterations_with_small_patches(name):
    print(f"chain patch_cells 64 {name}", flush=True)
    make, cfl = CASES[name]
    opts = {"patch_cells": 64}
    _same(_run(make, cfl, 9, True, opts), _run(make, cfl, 9, False, opts))


@pytest.mark.parametrize("refresh", [1, 3, 4])
def test_chained_mixed_precision_with_refreshes(refresh):
    """Mixed precision: chains run within runs of linear sweeps and are closed
    before every FP64 sweep (every `refresh` iterations and the last)."""
    print(f"chain mixed refresh {refresh}", flush=True)
    make, cfl = CASES["burgers_p2_b"]
    _same(_run(make, cfl, 10, True, precision=1, refresh=refresh),
          _run(make, cfl, 10, True, precision=1, refresh=refresh, chain=0))


def test_chained_iterations_match_oracle():
    mesh = rectangle(12, 12, extent=(0.0, 1.0, 0.0, 1.0), split=2, order=2)
    u0 = burgers_state(mesh)
    out, norm = _run(CASES["burgers_p2_b"][0], 0.5, 12, True)
    P = oracle_problem(mesh, L.BURGERS2D, L.B, (), quad_orders(mesh))
    uo, res, dt, norms, rc = P.iterate(u0, 0.5, 12)
    assert rc == 0
    assert rel_err(out["solution"], uo) <= 1e-12
    assert rel_err(out["residuals"], res) <= 1e-11
    assert abs(norm[0] - norms[-1, 0]) <= 1e-11 * norms[-1, 0]


def test_calls_resume_after_a_closed_chain():
    """Two calls of iterate(4) equal one iterate(8): every call ends complete."""
    make, cfl = CASES["euler2d_mixed_weak"]
    m = make({})
    m.set_option(L.OPT_CHAIN, 1)
    m.iterate(cfl, 4)
    norm = m.iterate(cfl, 4)
    a = {f: m.vec_function(f).copy() for f in FIELDS}
    m.destroy()
    _same((a, norm), _run(make, cfl, 8, True))


@pytest.mark.parametrize("n", [2, 3, 12])
def test_chain_on_meshes_of_one_or_few_patches(n):
    """make_unit_quad(n): 2, 8 and 242 triangles -- one patch without shared
    nodes, then a few patches."""
    def make(options):
        mesh = unit_quad_triag(n)
        m = PGRDMethodExplicit(L.ADVECTION2D, L.N, params=(0.8, 0.6))
        m.configure_mesh_data(mesh).initialize_work_data(quad_order=4, **options)
        bn = np.union1d(mesh.boundaries["Edge0"], mesh.boundaries["Edge3"]).astype(np.uint32)
        m.add_boundary_condition("StrongDirichlet", "inflow", bn, np.ones((bn.size, 1)))
        m.set_vec_function("solution", np.zeros((mesh.nb_nodes, 1)))
        return m
    _same(_run(make, 0.9, 6, True), _run(make, 0.9, 6, False))
    _same(_run(make, 0.9, 6, True, {"patch_cells": 64}), _run(make, 0.9, 6, True, {"patch_cells": 64}, chain=0))
