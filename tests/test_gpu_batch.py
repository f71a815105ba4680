"""Batched iterations (pdrd_batch_iterate): several independent methods of
different element types, orders, physics and schemes advanced in shared
launches (direct, or replayed from a captured HIP graph) give bit for bit
what each method's own pdrd_iterate gives, and what the oracle's iterations
give to 1e-12."""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.rdm import Batch, PGRDMethodExplicit
from tests.helpers import CASES, oracle_problem, quad_orders, rel_err

pytestmark = pytest.mark.gpu

MIX = [("adv_p1", L.N), ("adv_p2", L.LDA), ("adv_p3", L.B), ("burgers_p2", L.B), ("burgers_q1", L.PSI),
       ("euler2d_mixed", L.LDA), ("euler2d_p2", L.N), ("euler3d_p1", L.LDA), ("burgers_p1", L.B)]
CFL = {L.ADVECTION2D: 0.9, L.BURGERS2D: 0.5, L.EULER2D: 0.3, L.EULER3D: 0.3}


def _make(name, scheme):
    build, physics, params, state = CASES[name]
    mesh = build()
    u = state(mesh)
    m = PGRDMethodExplicit(physics, scheme, params=params if params else None)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    assert m.get_option(L.OPT_SCATTER) == 2
    m.set_vec_function("solution", u)
    bn = mesh.boundaries.get("xmin", mesh.boundaries.get("Edge3"))
    if bn is not None and physics in (L.ADVECTION2D, L.BURGERS2D):
        m.add_boundary_condition("StrongDirichlet", "inflow", bn, u[bn])
    return m, CFL[physics]


@pytest.mark.parametrize("chain", [None, 0, 1])
@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("n_iter", [1, 2, 7, 8])
def test_batch_matches_single_methods(n_iter, graphs, chain):
    """chain: the methods' PDRD_OPT_CHAIN (None: each method's default, which
    chains the scalar methods and not the Euler ones)."""
    print(f"batch of {len(MIX)} methods, {n_iter} iterations, graphs {graphs}, chain {chain}", flush=True)
    alone = [_make(n, s) for n, s in MIX]
    mixed = [_make(n, s) for n, s in MIX]
    if chain is not None:
        for m, _ in alone + mixed:
            m.set_option(L.OPT_CHAIN, chain)
    ref_norms = [m.iterate(c, n_iter) for m, c in alone]
    b = Batch(graphs=graphs)
    norms = b.iterate([m for m, _ in mixed], [c for _, c in mixed], n_iter)
    for (ma, _), (mb, _), na, nb, (name, _) in zip(alone, mixed, ref_norms, norms, MIX):
        for f in ("solution", "residuals", "time_step", "wave_speed"):
            a, c = ma.vec_function(f), mb.vec_function(f)
            assert np.array_equal(a, c), (name, f, np.abs(a - c).max())
        assert np.array_equal(na, nb), name
    # a second call on a different subset and order reuses the batch buffers
    sub = [mixed[5], mixed[0], mixed[3]]
    for m, c in [alone[5], alone[0], alone[3]]:
        m.iterate(c, 3)
    b.iterate([m for m, _ in sub], [c for _, c in sub], 3)
    for (ma, _), (mb, _) in zip([alone[5], alone[0], alone[3]], sub):
        assert np.array_equal(ma.vec_function("solution"), mb.vec_function("solution"))


def test_batch_rejects_bad_input():
    m, c = _make("adv_p1", L.N)
    b = Batch()
    with pytest.raises(L.PdrdError):
        b.iterate([m, m], [c, c], 2)
    m2 = PGRDMethodExplicit(L.ADVECTION2D, L.N, params=(0.8, 0.6))
    build = CASES["adv_p1"][0]
    mesh = build()
    m2.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh), scatter=1)
    m2.set_vec_function("solution", np.zeros((mesh.nb_nodes, 1)))
    with pytest.raises(L.PdrdError):
        b.iterate([m, m2], [c, c], 2)


def _with_inputs(kind):
    """Methods whose cells take a nodal input or a source field."""
    rng = np.random.default_rng(9)
    if kind == "visc":
        m, c = _make("euler2d_mixed", L.LDA)
        m.set_artificial_viscosity(rng.uniform(0.0, 0.05, m.mesh.nb_nodes))
    elif kind == "theta":
        m, c = _make("burgers_p2", L.B)
        m.set_blending_coeff(rng.uniform(0.0, 1.0, m.mesh.nb_nodes))
    elif kind == "src_adv":
        m, c = _make("adv_p2", L.LDA)
        m.set_sources(rng.uniform(-0.5, 0.5, (m.mesh.nb_nodes, 1)))
    elif kind == "src_euler":
        m, c = _make("euler2d_p2", L.N)
        m.set_sources(rng.uniform(-0.01, 0.01, (m.mesh.nb_nodes, 4)))
    else:
        m, c = _make("adv_p1", L.N)
    return m, c


@pytest.mark.parametrize("n_iter", [1, 7])
def test_batch_with_nodal_inputs_and_sources(n_iter):
    """Methods with an artificial viscosity, an external blending coefficient
    or a source field, beside a plain one, in one batch: the same bits as their
    own pdrd_iterate."""
    kinds = ["visc", "theta", "src_adv", "src_euler", "plain"]
    print(f"batch with nodal inputs and sources, {n_iter} iterations", flush=True)
    alone = [_with_inputs(k) for k in kinds]
    batched = [_with_inputs(k) for k in kinds]
    ref = [m.iterate(c, n_iter) for m, c in alone]
    norms = Batch().iterate([m for m, _ in batched], [c for _, c in batched], n_iter)
    for (ma, _), (mb, _), na, nb, k in zip(alone, batched, ref, norms, kinds):
        for f in ("solution", "residuals", "time_step", "wave_speed"):
            a, c = ma.vec_function(f), mb.vec_function(f)
            assert np.array_equal(a, c), (k, f, np.abs(a - c).max())
        assert np.array_equal(na, nb), k


@pytest.mark.parametrize("graphs", [False, True])
def test_batch_matches_oracle(graphs):
    """A batch of methods of every element type and physics (the P2 / Q2 Euler
    cells with LDA, PSI and the PSI / LDA blend among them), iterated together,
    against the oracle's explicit iterations (PGRDMethodExplicit.hpp:427-463,
    pseudo-time steps of every method from its own CFL): solutions <= 1e-12 and
    residuals <= 1e-11 normwise relative, the Dirichlet inflow nodes of the
    scalar methods kept at their values."""
    chosen = [("adv_p2", L.LDA), ("burgers_p2", L.B), ("euler2d_mixed", L.LDA), ("euler3d_p1", L.N),
              ("euler2d_q2_mixed", L.LDA), ("euler2d_p2", L.PSI), ("euler3d_p2", L.BPSI)]
    n_iter = 5
    print(f"batch against the oracle, {n_iter} iterations, graphs {graphs}", flush=True)
    made = [_make(n, s) for n, s in chosen]
    Batch(graphs=graphs).iterate([m for m, _ in made], [c for _, c in made], n_iter)
    for (m, cfl), (name, scheme) in zip(made, chosen):
        build, physics, params, state = CASES[name]
        mesh = build()
        u0 = state(mesh)
        P = oracle_problem(mesh, physics, scheme, params, quad_orders(mesh))
        bn = mesh.boundaries.get("xmin", mesh.boundaries.get("Edge3"))
        if bn is not None and physics in (L.ADVECTION2D, L.BURGERS2D):
            uo, res, dt, norms, rc = P.iterate(u0, cfl, n_iter, np.asarray(bn, np.uint32), u0[bn])
        else:
            uo, res, dt, norms, rc = P.iterate(u0, cfl, n_iter)
        assert rc == 0, name
        assert rel_err(m.vec_function("solution"), uo) <= 1e-12, name
        assert rel_err(m.vec_function("residuals"), res) <= 1e-11, name
        m.destroy()
