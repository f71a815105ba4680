"""Mixed-precision form (PDRD_OPT_PRECISION = 1): linearised defect correction.

An FP64 sweep sets the linearisation point u_ref, its residual R(u_ref) and
time step dt(u_ref); every PDRD_OPT_REFRESH iterations it also recomputes the
element Jacobians J_e (FP32). The sweeps in between are linear,
u <- u0 - dt(u_ref) (R(u_ref) + sum_e J_e (u - u_ref)_e / V), with J delta in
FP32; the last iteration of every pdrd_iterate call is an FP64 sweep.

* refresh <= 1: every sweep is FP64, bit for bit the plain iteration;
* the first iteration of a call is the plain FP64 iteration, bit for bit;
* FP64 / linear / FP64 (three iterations) against a numpy restatement with
  the oracle's residual and Jacobians: the linear step agrees to the FP32
  rounding of J delta (scalar, Euler 2D / 3D, mixed element types, weak BCs
  -- whose facet residuals the linear sweeps evaluate at the current state --,
  external blending coefficient, source field);
* steady states: the FP64 iteration run to a residual norm of 1e-12 and the
  mixed one to the same norm (its FP64 sweeps measure the true residual);
  the states agree within 1e-8 (max |u_mixed - u_plain| / max |u_plain|):
  advection P1 N and P2 LDA, Burgers P2 with the B scheme (C2's scheme on a
  16 x 16 criss-cross grid with C2's jitter; a shock-free inflow, since the
  B blend with C2's inflow 1.5 - 2x does not reach a steady state in FP64
  either), and a supersonic Euler 2D channel over the bump with LDA.
"""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.mesh import rectangle, unit_quad_triag
from pdekit_amd.rdm import Batch, PGRDMethodExplicit
from tests.helpers import CASES, add_oracle_weak, euler_state, oracle_problem, quad_orders, rel_err

pytestmark = pytest.mark.gpu


def _method(mesh, physics, scheme, params, precision, refresh=50, weak=()):
    m = PGRDMethodExplicit(physics, scheme, params=params if params else None)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    for kind, nodes, par in weak:  # before the method's first device use
        m.add_boundary_condition(kind, kind, nodes, params=par)
    m.set_option(L.OPT_PRECISION, precision)
    m.set_option(L.OPT_REFRESH, refresh)
    assert m.get_option(L.OPT_PRECISION) == precision and m.get_option(L.OPT_SCATTER) == 2
    return m


def _dirichlet(mesh):
    """Strong Dirichlet values on the xmin / Edge3 boundary (the state there)."""
    b = mesh.boundaries
    return np.asarray(b["xmin"] if "xmin" in b else b["Edge3"], dtype=np.uint32)


_cache = {}


def _case(name):
    if name not in _cache:
        build, physics, params, state = CASES[name]
        mesh = build()
        _cache[name] = (mesh, physics, params, state(mesh))
    return _cache[name]


@pytest.mark.parametrize("name", ["adv_p1", "burgers_p2", "euler2d_mixed", "euler3d_p1"])
@pytest.mark.parametrize("scheme", [L.LDA, L.N, L.B])
def test_refresh_one_and_first_iteration_are_fp64(name, scheme):
    mesh, physics, params, u = _case(name)
    print(f"mixed fp64 bits {name} scheme {scheme}", flush=True)
    bn = _dirichlet(mesh)
    runs = []
    for precision, refresh, n in ((0, 50, 7), (1, 1, 7), (0, 50, 1), (1, 50, 1)):
        m = _method(mesh, physics, scheme, params, precision, refresh)
        m.set_vec_function("solution", u)
        m.add_boundary_condition("StrongDirichlet", "in", bn, u[bn])
        m.iterate(0.3, n)
        runs.append((m.vec_function("solution"), m.vec_function("residuals"), m.vec_function("time_step")))
        m.destroy()
    for a, b in ((0, 1), (2, 3)):
        for x, y in zip(runs[a], runs[b]):
            assert np.array_equal(x, y)


def _emulate(P, mesh, u0, bn, g, cfl, m_eq, Pc=None):
    """FP64 iteration, linear iteration about its start state, FP64
    iteration, restated with the oracle (FP64 Jacobians). Pc: the same
    problem without its weak BCs, whose facet part the linear iteration takes
    at the current state (R_all - R_cells)."""
    vol = P.dual_volume()
    r0, ws0, rc = P.assemble_rhs(u0)
    assert rc == 0
    dt0 = cfl * vol / ws0
    u1, _, _, _, rc = P.iterate(u0, cfl, 1, bn, g)
    assert rc == 0
    if Pc is not None:
        facet = lambda u: P.assemble_rhs(u)[0] - Pc.assemble_rhs(u)[0]
        r0 = r0 - facet(u0) + facet(u1)
    acc = np.zeros((mesh.nb_nodes, m_eq))
    d = u1 - u0
    for bi, b in enumerate(mesh.blocks):
        _, J, _, rc = P.cell_jacobians(bi, u0)
        assert rc == 0
        c = b.sol_conn.astype(np.int64)
        ne, nd = c.shape
        y = np.einsum("eij,ej->ei", J, d[c].reshape(ne, nd * m_eq)).reshape(ne, nd, m_eq)
        for n in range(nd):
            np.add.at(acc, c[:, n], y[:, n])
    u1d = u1.copy()
    if bn.size:
        u1d[bn] = g
    u2 = u1d - dt0[:, None] * (r0 + acc / vol[:, None])
    u3, r3, dt3, _, rc = P.iterate(u2, cfl, 1, bn, g)
    assert rc == 0
    return u1, u2, u3, r3


EMULATED = {
    # name: (case, scheme, weak BCs, nodal input)
    "adv_p2_n": ("adv_p2", L.N, False, None),
    "burgers_p2_b": ("burgers_p2", L.B, False, None),
    "burgers_q2_mixed_lda": ("burgers_q2_mixed", L.LDA, False, None),
    "euler2d_mixed_lda": ("euler2d_mixed", L.LDA, False, None),
    "euler2d_p2_n": ("euler2d_p2", L.N, False, None),
    "euler3d_p1_b": ("euler3d_p1", L.B, False, None),
    "euler2d_mixed_weak_lda": ("euler2d_mixed", L.LDA, True, None),
    "burgers_p2_b_theta": ("burgers_p2", L.B, False, "theta"),
    "adv_p2_lda_source": ("adv_p2", L.LDA, False, "source"),
}


@pytest.mark.parametrize("name", sorted(EMULATED))
def test_linear_sweep_against_oracle_restatement(name):
    case, scheme, weak, extra = EMULATED[name]
    mesh, physics, params, u0 = _case(case)
    print(f"mixed F/L/F {name}", flush=True)
    m_eq = u0.shape[1]
    cfl = 0.3
    P = oracle_problem(mesh, physics, scheme, params)
    Pc = oracle_problem(mesh, physics, scheme, params) if weak else None
    bcs = []
    if weak:
        bn = np.zeros(0, dtype=np.uint32)
        b = mesh.boundaries
        gamma = 1.4
        for kind, nodes, par, opar in (
                ("WeakWall", b["ymin"], {}, ()), ("WeakWall", b["ymax"], {}, ()),
                ("WeakSubOutlet", b["xmax"], {"total_pressure": 1.0 / gamma}, (1.0 / gamma,))):
            bcs.append((kind, nodes, par))
            add_oracle_weak(P, mesh, kind, nodes, opar, quad_order=max(quad_orders(mesh).values()))
    else:
        bn = _dirichlet(mesh)
    m = _method(mesh, physics, scheme, params, 1, 50, weak=bcs)
    g = u0[bn]
    m.set_vec_function("solution", u0)
    if bn.size:
        m.add_boundary_condition("StrongDirichlet", "in", bn, g)
    if extra == "theta":
        theta = np.random.default_rng(3).uniform(0.0, 1.0, mesh.nb_nodes)
        m.set_blending_coeff(theta)
        P.set_blending_coeff(theta)
    elif extra == "source":
        src = np.random.default_rng(4).uniform(-1.0, 1.0, (mesh.nb_nodes, m_eq))
        m.set_vec_function("sources", src)
        P.set_sources(src)
    m.iterate(cfl, 3)
    m.check_status()
    u1, u2, u3, r3 = _emulate(P, mesh, u0, bn, g, cfl, m_eq, Pc)
    ug = m.vec_function("solution")
    # the linear step's J delta carries FP32 rounding relative to |J| |u1 - u0|
    tol = 1e-4 * np.abs(u2 - u1).max() + 1e-13 * np.abs(u3).max()
    err = np.abs(ug - u3).max()
    print(f"  |u_gpu - u_emulated| = {err:.3e} (tolerance {tol:.3e}; step {np.abs(u2 - u1).max():.3e})",
          flush=True)
    assert err <= tol
    assert rel_err(m.vec_function("residuals"), r3) <= 1e-4
    m.destroy()


def _steady_case(case):
    if case.startswith("advection"):
        order = 1 if case.endswith("p1_n") else 2
        scheme = L.N if order == 1 else L.LDA
        mesh = unit_quad_triag(33, order=order)
        bn = np.union1d(mesh.boundaries["Edge0"], mesh.boundaries["Edge3"]).astype(np.uint32)
        g = np.where(mesh.node_coords[bn, 1] < 0.5, 1.0, 0.0)[:, None]
        return mesh, L.ADVECTION2D, scheme, (0.8, 0.6), np.zeros((mesh.nb_nodes, 1)), bn, g, 0.9
    if case == "burgers_p2_b":
        mesh = rectangle(16, 16, split=4, order=2, jitter=0.15, seed=2)
        b = mesh.boundaries
        ymin = b["ymin"]
        xmin = np.setdiff1d(b["xmin"], ymin)
        bn = np.concatenate([ymin, xmin]).astype(np.uint32)
        g = np.concatenate([0.5 + 0.5 * mesh.node_coords[ymin, 0], np.full(xmin.size, 0.5)])[:, None]
        u0 = (0.5 + 0.5 * mesh.node_coords[:, 0])[:, None]
        return mesh, L.BURGERS2D, L.B, (), u0, bn, g, 0.5
    # supersonic (M 2) channel over the bump, density varying across the inflow
    mesh = rectangle(24, 12, extent=(-0.5, 1.5, 0.0, 1.0), split=2, order=1, bump=(1.0, 0.1, 25.0, 0.5))
    u0 = euler_state(mesh, mach=2.0, pert=0.0)
    b = mesh.boundaries
    bn = np.union1d(np.union1d(b["xmin"], b["ymin"]), b["ymax"]).astype(np.uint32)
    g = u0[bn] * (1.0 + 0.1 * np.sin(2 * np.pi * mesh.node_coords[bn, 1]))[:, None]
    return mesh, L.EULER2D, L.LDA, (1.4, 287.05), u0, bn, g, 0.5


def _steady(mesh, physics, scheme, params, u0, bn, g, cfl, precision, tol=1e-12, chunk=100, cap=60_000):
    m = _method(mesh, physics, scheme, params, precision)
    m.set_vec_function("solution", u0)
    m.add_boundary_condition("StrongDirichlet", "inflow", bn, g)
    it = 0
    while True:
        nrm = m.iterate(cfl, chunk, exclude_dirichlet=True)
        it += chunk
        if nrm.max() <= tol or it >= cap:
            break
    u = m.vec_function("solution")
    m.destroy()
    return u, nrm.max(), it


@pytest.mark.parametrize("case", ["advection_p1_n", "advection_p2_lda", "burgers_p2_b", "euler2d_supersonic_lda"])
def test_mixed_steady_state_matches_plain(case):
    args = _steady_case(case)
    print(f"steady {case}", flush=True)
    up, np_, ip = _steady(*args, precision=0)
    um, nm, im = _steady(*args, precision=1, cap=2 * ip + 2000)
    d = float(np.abs(um - up).max() / np.abs(up).max())
    print(f"  plain: norm {np_:.2e} after {ip} it; mixed: norm {nm:.2e} after {im} it; "
          f"state difference {d:.3e}", flush=True)
    assert np_ <= 1e-12 and nm <= 1e-12
    assert d <= 1e-8


def test_mixed_refusals():
    mesh, physics, params, u = _case("adv_p1")
    m = PGRDMethodExplicit(physics, L.PSI, params=params)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    with pytest.raises(L.PdrdError):
        m.set_option(L.OPT_PRECISION, 1)  # PSI has no element Jacobians
    m.destroy()
    m = PGRDMethodExplicit(physics, L.N, params=params)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh), scatter=1)
    m.set_option(L.OPT_PRECISION, 1)
    m.set_vec_function("solution", u)
    with pytest.raises(L.PdrdError):
        m.iterate(0.5, 3)
    m.destroy()
    m = _method(mesh, physics, L.N, params, 1)
    m.set_vec_function("solution", u)
    with pytest.raises(L.PdrdError):
        Batch().iterate([m], [0.5], 3)
    m.destroy()
