"""Features used together on the GPU:

* a batch mixing methods with weak boundary conditions (Euler 2D on P1 + Q1 and
  P2 + Q2, Euler 3D on P2 tetrahedra) with plain ones gives bit for bit what
  each method's own pdrd_iterate gives;
* a method read from a Gmsh file, with weak BCs on its named boundaries, in a
  batch;
* the mixed-precision form with weak BCs and every sweep FP64 is the plain
  iteration bit for bit;
* a method with a nodal input batches with one with weak BCs.
"""
import os

import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.gmsh import GmshReader
from pdekit_amd.mesh import box_tetra, rectangle
from pdekit_amd.rdm import Batch, PGRDMethodExplicit
from tests.helpers import euler_state, quad_orders

pytestmark = pytest.mark.gpu

GAMMA = 1.4
HERE = os.path.join(os.path.dirname(__file__), "golden", "gmsh")


def _bcs(mesh, names=("xmin", "xmax", "ymin", "ymax")):
    pin, tin = 1.0 / GAMMA, 1.0 / GAMMA / 287.046
    fac = 1 + 0.5 * (GAMMA - 1) * 0.25
    b = mesh.boundaries
    out = [("WeakWall", b[names[2]], {}), ("WeakWall", b[names[3]], {}),
           ("WeakSubInlet", b[names[0]], {"total_temperature": tin * fac,
                                          "total_pressure": pin * fac ** (GAMMA / (GAMMA - 1))}),
           ("WeakSubOutlet", b[names[1]], {"total_pressure": pin})]
    if mesh.dim == 3:
        out += [("WeakWall", b["zmin"], {}), ("WeakWall", b["zmax"], {})]
    return out


def _method(mesh, scheme=L.LDA, weak=True, names=("xmin", "xmax", "ymin", "ymax")):
    phys = L.EULER2D if mesh.dim == 2 else L.EULER3D
    m = PGRDMethodExplicit(phys, scheme, params=(GAMMA, 287.05))
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    if weak:
        for kind, nodes, par in _bcs(mesh, names):
            m.add_boundary_condition(kind, kind, nodes, params=par)
    m.set_vec_function("solution", euler_state(mesh, dim=mesh.dim, pert=0.02))
    return m


MESHES = [
    lambda: rectangle(12, 6, extent=(-0.5, 1.5, 0.0, 1.0), split=2, quad_rows=3, bump=(1.0, 0.1, 25.0, 0.5)),
    lambda: rectangle(8, 4, extent=(-0.5, 1.5, 0.0, 1.0), split=2, quad_rows=2, order=2,
                      bump=(1.0, 0.1, 25.0, 0.5)),
    lambda: box_tetra(3, 3, 2, extent=(1.0, 1.0, 0.5), order=2, jitter=0.1, seed=4),
]


def test_batch_with_weak_bcs_matches_single_methods():
    specs = [(mk, s, w) for mk in MESHES for s, w in ((L.LDA, True), (L.N, False))]
    alone = [_method(mk(), s, w) for mk, s, w in specs]
    mixed = [_method(mk(), s, w) for mk, s, w in specs]
    for m in alone:
        m.iterate(0.3, 6)
    Batch().iterate(mixed, [0.3] * len(mixed), 6)
    for a, b in zip(alone, mixed):
        for f in ("solution", "residuals", "time_step", "wave_speed"):
            assert np.array_equal(a.vec_function(f), b.vec_function(f)), f


def test_gmsh_method_in_a_batch():
    mesh = GmshReader().read_mesh_from_file(os.path.join(HERE, "rectangle_mixed_elem_p1.msh"), 1)
    names = ("Left", "Right", "Bottom", "Top")
    a = _method(mesh, names=names)
    b = _method(mesh, names=names)
    other = _method(MESHES[0]())
    a.iterate(0.3, 5)
    Batch().iterate([other, b], [0.3, 0.3], 5)
    assert np.array_equal(a.vec_function("solution"), b.vec_function("solution"))


def test_mixed_precision_with_weak_bcs():
    """The mixed-precision form with weak BCs: refresh 1 (every sweep FP64)
    gives the plain bits, chained or not."""
    mesh = rectangle(24, 12, extent=(-0.5, 1.5, 0.0, 1.0), split=2, quad_rows=6, bump=(1.0, 0.1, 25.0, 0.5))
    plain = _method(mesh)
    one = _method(mesh)
    one.set_option(L.OPT_PRECISION, 1)
    one.set_option(L.OPT_REFRESH, 1)
    for m in (plain, one):
        m.iterate(0.5, 40)
    assert np.array_equal(plain.vec_function("solution"), one.vec_function("solution"))
    assert np.array_equal(plain.vec_function("residuals"), one.vec_function("residuals"))


def test_nodal_inputs_in_a_batch_with_weak_bcs():
    """A method with an external blending coefficient and one with weak BCs in
    one batch: the same bits as their own pdrd_iterate."""
    mesh = MESHES[0]()
    theta = np.random.default_rng(3).uniform(0.0, 1.0, mesh.nb_nodes)
    alone = [_method(mesh, L.B, weak=False), _method(mesh)]
    batched = [_method(mesh, L.B, weak=False), _method(mesh)]
    for m in (alone[0], batched[0]):
        m.set_blending_coeff(theta)
    for m in alone:
        m.iterate(0.3, 6)
    Batch().iterate(batched, [0.3, 0.3], 6)
    for a, b in zip(alone, batched):
        assert np.array_equal(a.vec_function("solution"), b.vec_function("solution"))
        assert np.array_equal(a.vec_function("residuals"), b.vec_function("residuals"))
