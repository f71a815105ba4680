"""The rotating advection (RotationAdvection2D with AroundOrigin2D, velocity
(y, -x) at every point) on the GPU against the oracle, <= 1e-12 normwise
relative: one sweep on P1 / P2 / P3 triangles and Q1 / Q2 / mixed meshes in
every scheme and assembly, 20 iterations with strong Dirichlet inflow values
(chained and unchained), the element Jacobians, a radial steady state,
WeakDirichlet (the velocity of every facet point) for one sweep and 10
iterations, and the mixed-precision form's steady state within 1e-8 of the
FP64 one."""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.mesh import rectangle, unit_quad_triag
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import add_oracle_weak, advection_state, oracle_problem, quad_orders, rel_err

pytestmark = pytest.mark.gpu

ROT = L.ROTATION2D
MESHES = {
    "p1": lambda: unit_quad_triag(17),
    "p2": lambda: rectangle(10, 8, extent=(-1.0, 1.0, -0.5, 1.0), split=2, order=2, jitter=0.15, seed=2),
    "p3": lambda: rectangle(6, 6, split=2, order=3, jitter=0.1, seed=4),
    "q1_mixed": lambda: rectangle(10, 10, extent=(-1.0, 1.0, -1.0, 1.0), split=2, quad_rows=5, order=1, jitter=0.1,
                                  seed=2),
    "q2": lambda: rectangle(7, 6, split=1, order=2, jitter=0.1, seed=3),
}


def _method(mesh, scheme, scatter=None):
    m = PGRDMethodExplicit(ROT, scheme)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh), scatter=scatter)
    return m


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("scheme", [L.LDA, L.N, L.B, L.PSI])
@pytest.mark.parametrize("scatter", [2, 1, 0])
def test_rotation_sweep_matches_oracle(name, scheme, scatter):
    print(f"rotation {name} scheme {scheme} scatter {scatter}", flush=True)
    mesh = MESHES[name]()
    u = advection_state(mesh)
    m = _method(mesh, scheme, scatter)
    m.set_vec_function("solution", u)
    m.assemble_rhs()
    P = oracle_problem(mesh, ROT, scheme, (), quad_orders(mesh))
    ref, ref_ws, rc = P.assemble_rhs(u)
    assert rc == 0
    assert rel_err(m.vec_function("residuals"), ref) <= 1e-12
    assert rel_err(m.vec_function("wave_speed"), ref_ws) <= 1e-12
    m.destroy()


@pytest.mark.parametrize("chain", [0, 1])
def test_rotation_iterations_match_oracle(chain):
    mesh = unit_quad_triag(33)
    x, y = mesh.node_coords[:, 0], mesh.node_coords[:, 1]
    # inflow of (y, -x) on [0,1]^2: the edge y = 1 (a.n = -x < 0 on it) and x = 0 (a.n = -y)
    bn = np.union1d(mesh.boundaries["Edge2"], mesh.boundaries["Edge3"]).astype(np.uint32)
    g = np.exp(-20 * (mesh.node_coords[bn, 0] + mesh.node_coords[bn, 1] - 0.8) ** 2)[:, None]
    m = _method(mesh, L.N)
    m.set_option(L.OPT_CHAIN, chain)
    m.set_vec_function("solution", np.zeros((mesh.nb_nodes, 1)))
    m.add_boundary_condition("StrongDirichlet", "inflow", bn, g)
    norm = m.iterate(0.9, 20)
    P = oracle_problem(mesh, ROT, L.N, ())
    uo, res, dt, norms, rc = P.iterate(np.zeros(mesh.nb_nodes), 0.9, 20, bn, g)
    assert rc == 0
    assert rel_err(m.vec_function("solution"), uo) <= 1e-12
    assert abs(norm[0] - norms[-1, 0]) <= 1e-12 * norms[-1, 0]
    m.destroy()


@pytest.mark.parametrize("scheme", [L.LDA, L.N, L.B])
def test_rotation_jacobians_match_oracle(scheme):
    mesh = MESHES["p2"]()
    u = advection_state(mesh)
    m = _method(mesh, scheme)
    m.set_vec_function("solution", u)
    phi, jac, wave = m.element_jacobians(0)
    P = oracle_problem(mesh, ROT, scheme, (), quad_orders(mesh))
    po, jo, wo, rc = P.cell_jacobians(0, u)
    assert rc == 0
    assert rel_err(phi, po) <= 1e-12 and rel_err(jac, jo) <= 1e-12 and rel_err(wave, wo) <= 1e-12
    m.destroy()


def test_radial_state_is_steady_on_the_gpu():
    mesh = MESHES["p2"]()
    x, y = mesh.node_coords[:, 0], mesh.node_coords[:, 1]
    m = _method(mesh, L.LDA)
    m.set_vec_function("solution", x[:, None])
    m.assemble_rhs()
    scale = np.abs(m.vec_function("residuals")).max()
    m.set_vec_function("solution", (x * x + y * y)[:, None])
    m.assemble_rhs()
    assert scale > 1e-2 and np.abs(m.vec_function("residuals")).max() <= 1e-12 * scale
    m.destroy()


@pytest.mark.parametrize("name", ["p1", "p2", "q1_mixed"])
@pytest.mark.parametrize("scheme", [L.LDA, L.N])
def test_rotation_weak_dirichlet_matches_oracle(name, scheme):
    print(f"rotation WeakDirichlet {name} scheme {scheme}", flush=True)
    mesh = MESHES[name]()
    b = mesh.boundaries
    sides = [s for s in ("xmin", "xmax", "ymin", "ymax") if s in b] or ["Edge0", "Edge1", "Edge2", "Edge3"]
    u = advection_state(mesh)
    q = quad_orders(mesh)
    m = _method(mesh, scheme)
    P = oracle_problem(mesh, ROT, scheme, (), q)
    for k, side in enumerate(sides):
        val = 0.3 + 0.2 * k
        m.add_boundary_condition("WeakDirichlet", side, b[side], values=[val])
        add_oracle_weak(P, mesh, "WeakDirichlet", b[side], (val,), quad_order=max(q.values()))
    m.set_vec_function("solution", u)
    m.assemble_rhs()
    m.check_status()
    ref, ref_ws, rc = P.assemble_rhs(u)
    assert rc == 0
    assert rel_err(m.vec_function("residuals"), ref) <= 1e-12
    assert rel_err(m.vec_function("wave_speed"), ref_ws) <= 1e-12
    m.iterate(0.5, 10)
    uo, _, _, _, rc = P.iterate(u, 0.5, 10)
    assert rc == 0
    assert rel_err(m.vec_function("solution"), uo) <= 1e-12
    m.destroy()


def test_rotation_mixed_steady_state_matches_plain():
    """Steady inflow problem on make_unit_quad(33) (strong Dirichlet on x = 0
    and y = 1, N scheme, CFL 0.9): the FP64 iteration and the mixed form each
    run to a residual norm <= 1e-12; the states agree within 1e-8."""
    mesh = unit_quad_triag(33)
    bn = np.union1d(mesh.boundaries["Edge2"], mesh.boundaries["Edge3"]).astype(np.uint32)
    g = np.exp(-20 * (mesh.node_coords[bn, 0] + mesh.node_coords[bn, 1] - 0.8) ** 2)[:, None]
    out = {}
    for prec in (0, 1):
        m = _method(mesh, L.N)
        m.set_option(L.OPT_PRECISION, prec)
        m.set_vec_function("solution", np.zeros((mesh.nb_nodes, 1)))
        m.add_boundary_condition("StrongDirichlet", "inflow", bn, g)
        it = 0
        while True:
            nrm = m.iterate(0.9, 100, exclude_dirichlet=True)
            it += 100
            if nrm.max() <= 1e-12 or it >= 40_000:
                break
        out[prec] = (m.vec_function("solution"), float(nrm.max()), it)
        m.destroy()
    (up, np_, ip), (um, nm, im) = out[0], out[1]
    d = float(np.abs(um - up).max() / np.abs(up).max())
    print(f"rotation steady: plain {np_:.2e} after {ip}, mixed {nm:.2e} after {im}, difference {d:.3e}", flush=True)
    assert np_ <= 1e-12 and nm <= 1e-12
    assert d <= 1e-8
