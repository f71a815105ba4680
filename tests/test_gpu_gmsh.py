"""The GPU cell loop on the reference's Gmsh meshes (tests/golden/gmsh) against
the oracle on the same mesh, ≤ 1e-12 normwise relative: advection on P1/P2/P3
triangles, on quadrilaterals and on the mixed triangle + quadrilateral
rectangle (LDA, N, B; the P2 files read by their corners), Euler with weak
walls / subsonic inlet / outlet on the files' named boundaries (mixed P1,
triangles P2), the cube's tetrahedra; and a Gmsh $NodeData checkpoint in the
middle of a run giving the same state, bit for bit, as the uninterrupted run."""
import os

import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.gmsh import GmshReader, GmshWriter
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import add_oracle_weak, euler_state, oracle_problem, quad_orders, rel_err

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(__file__), "golden", "gmsh")
GAMMA = 1.4


def _read(name, order):
    return GmshReader().read_mesh_from_file(os.path.join(HERE, name + ".msh"), order)


def _adv_state(mesh):
    x = mesh.node_coords
    return (np.sin(2.0 * x[:, 0]) * np.cos(3.0 * x[:, 1]) + 0.1 * x[:, 0])[:, None].copy()


@pytest.mark.parametrize("name,order", [("test_p1_tri", 1), ("test_p2_tri", 2), ("test_p3_tri", 3),
                                        ("test_p2_quad", 1), ("rectangle_mixed_elem_p2", 1)])
@pytest.mark.parametrize("scheme", [L.LDA, L.N, L.B])
def test_advection_on_gmsh_mesh(name, order, scheme):
    print(f"gmsh {name} P{order} scheme {scheme}", flush=True)
    mesh = _read(name, order)
    u = _adv_state(mesh)
    m = PGRDMethodExplicit(L.ADVECTION2D, scheme, params=(0.8, 0.6))
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    m.set_vec_function("solution", u)
    m.assemble_lhs_and_rhs(0.5)
    m.check_status()
    P = oracle_problem(mesh, L.ADVECTION2D, scheme, (0.8, 0.6))
    ref, ref_ws, rc = P.assemble_rhs(u)
    assert rc == 0
    assert rel_err(m.vec_function("residuals"), ref) <= 1e-12
    assert rel_err(m.vec_function("wave_speed"), ref_ws) <= 1e-12


def _sides(mesh):
    """The named boundaries by position: bottom, top, left, right."""
    x = mesh.node_coords
    lo, hi = x.min(0), x.max(0)
    out = {}
    for nm, ids in mesh.boundaries.items():
        p = x[ids]
        for side, (d, v) in {"bottom": (1, lo[1]), "top": (1, hi[1]), "left": (0, lo[0]), "right": (0, hi[0])}.items():
            if np.abs(p[:, d] - v).max() < 1e-12:
                out[side] = ids
    return out


def _euler_bcs(mesh):
    pin, tin = 1.0 / GAMMA, 1.0 / GAMMA / 287.046
    fac = 1 + 0.5 * (GAMMA - 1) * 0.25
    b = _sides(mesh)
    return [("WeakWall", b["bottom"], {}, ()), ("WeakWall", b["top"], {}, ()),
            ("WeakSubInlet", b["left"], {"total_temperature": tin * fac,
                                         "total_pressure": pin * fac ** (GAMMA / (GAMMA - 1))},
             (tin * fac, pin * fac ** (GAMMA / (GAMMA - 1)))),
            ("WeakSubOutlet", b["right"], {"total_pressure": pin}, (pin,))]


def _euler(mesh, scheme):
    m = PGRDMethodExplicit(L.EULER2D, scheme, params=(GAMMA, 287.05))
    q = quad_orders(mesh)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=q)
    for kind, nodes, par, _ in _euler_bcs(mesh):
        m.add_boundary_condition(kind, kind, nodes, params=par)
    return m


@pytest.mark.parametrize("name,order", [("rectangle_mixed_elem_p1", 1), ("test_p2_tri", 2)])
@pytest.mark.parametrize("scheme", [L.LDA, L.N])
def test_euler_gmsh_with_weak_bcs(name, order, scheme):
    print(f"gmsh euler {name} P{order} scheme {scheme}", flush=True)
    mesh = _read(name, order)
    u = euler_state(mesh)
    m = _euler(mesh, scheme)
    P = oracle_problem(mesh, L.EULER2D, scheme, (GAMMA, 287.05))
    for kind, nodes, _, opar in _euler_bcs(mesh):
        add_oracle_weak(P, mesh, kind, nodes, opar, quad_order=max(quad_orders(mesh).values()))
    m.set_vec_function("solution", u)
    m.iterate(0.3, 10)
    uo, res, _, _, rc = P.iterate(u, 0.3, 10)
    assert rc == 0
    assert rel_err(m.vec_function("solution"), uo) <= 1e-12


def test_euler3d_on_cube():
    print("gmsh euler3d cube", flush=True)
    mesh = _read("cube", 1)
    u = euler_state(mesh, dim=3)
    m = PGRDMethodExplicit(L.EULER3D, L.LDA, params=(GAMMA, 287.05))
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    m.set_vec_function("solution", u)
    m.assemble_lhs_and_rhs(0.5)
    m.check_status()
    ref, ref_ws, rc = oracle_problem(mesh, L.EULER3D, L.LDA, (GAMMA, 287.05)).assemble_rhs(u)
    assert rc == 0
    assert rel_err(m.vec_function("residuals"), ref) <= 1e-12


def test_checkpoint_resume_through_gmsh(tmp_path):
    mesh = _read("test_p2_tri", 2)
    u0 = euler_state(mesh)
    a = _euler(mesh, L.LDA)
    a.set_vec_function("solution", u0)
    a.iterate(0.4, 60)
    ref = a.vec_function("solution")
    b = _euler(mesh, L.LDA)
    b.set_vec_function("solution", u0)
    b.iterate(0.4, 25)
    ck = str(tmp_path / "checkpoint.msh")
    GmshWriter().write_mesh_to_file(mesh, ck)
    GmshWriter().append_nodal_function_to_file(mesh, ck, b.vec_function("solution"), "u", timestep=25)
    mesh2 = GmshReader().read_mesh_from_file(ck, 2)
    assert np.array_equal(mesh2.node_coords, mesh.node_coords)
    c = _euler(mesh2, L.LDA)
    c.set_vec_function("solution", GmshReader().read_nodal_function_from_file(mesh2, ck, [f"u_{k}" for k in range(4)]))
    c.iterate(0.4, 35)
    assert np.array_equal(c.vec_function("solution"), ref)
