"""Persistent iteration (PDRD_OPT_PERSISTENT): the chained sweeps of one
iterate call in one cooperative launch, the workgroups meeting at a grid
barrier between sweeps. State, residual, time step, wave speed and norm must
equal those of the chained launches bit for bit (advection P1 with strong
Dirichlet inflow, Burgers P2 B, small patches so that many workgroups share
nodes). Combinations it does not cover are refused, and a barrier that runs
out of polls (forced from the host with a poll budget of zero) makes the call
fail instead of waiting."""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.mesh import rectangle, unit_quad_triag
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import burgers_state, euler_state, quad_orders

pytestmark = pytest.mark.gpu


def _advection(patch_cells=0):
    mesh = unit_quad_triag(33)
    m = PGRDMethodExplicit(L.ADVECTION2D, L.N, params=(0.8, 0.6))
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=4, patch_cells=patch_cells)
    bn = np.union1d(mesh.boundaries["Edge0"], mesh.boundaries["Edge3"]).astype(np.uint32)
    g = np.where(mesh.node_coords[bn, 1] < 0.5, 1.0, 0.0)[:, None]
    m.add_boundary_condition("StrongDirichlet", "inflow", bn, g)
    m.set_vec_function("solution", np.zeros((mesh.nb_nodes, 1)))
    return m


def _burgers(patch_cells=0):
    mesh = rectangle(20, 20, split=4, order=2, jitter=0.15, seed=2)
    m = PGRDMethodExplicit(L.BURGERS2D, L.B)
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh), patch_cells=patch_cells)
    m.set_vec_function("solution", burgers_state(mesh))
    return m


def _fields(m):
    return [m.vec_function(f) for f in ("solution", "residuals")] + \
           [m.vec_function(f) for f in ("time_step", "wave_speed")]


@pytest.mark.parametrize("build,cfl", [(_advection, 0.9), (_burgers, 0.5)])
@pytest.mark.parametrize("patch_cells", [0, 64])
@pytest.mark.parametrize("calls", [(57,), (1, 2, 40)])
def test_persistent_equals_chained_launches(build, cfl, patch_cells, calls):
    print(f"persistent {build.__name__} patch_cells {patch_cells} calls {calls}", flush=True)
    ref, per = build(patch_cells), build(patch_cells)
    per.set_option(L.OPT_PERSISTENT, 1)
    for n in calls:
        nr = ref.iterate(cfl, n)
        npe = per.iterate(cfl, n)
        assert np.array_equal(nr, npe)
    print(f"  {per.get_option(L.OPT_PERSIST_GRID)} workgroups", flush=True)
    assert per.get_option(L.OPT_PERSIST_GRID) >= 1
    for a, b in zip(_fields(ref), _fields(per)):
        assert np.array_equal(a, b)
    per.check_status()
    ref.destroy()
    per.destroy()


def test_persistent_refusals():
    mesh = rectangle(8, 4, extent=(-0.5, 1.5, 0.0, 1.0), split=2, bump=(1.0, 0.1, 25.0, 0.5))
    m = PGRDMethodExplicit(L.EULER2D, L.LDA, params=(1.4, 287.05))
    m.configure_mesh_data(mesh).initialize_work_data(quad_order=quad_orders(mesh))
    m.set_vec_function("solution", euler_state(mesh))
    m.set_option(L.OPT_PERSISTENT, 1)
    with pytest.raises(RuntimeError, match="scalar 2D physics"):
        m.iterate(0.5, 3)
    m.destroy()
    a = _advection()
    a.set_option(L.OPT_PERSISTENT, 1)
    a.set_option(L.OPT_CHAIN, 0)
    with pytest.raises(RuntimeError, match="chained sweeps"):
        a.iterate(0.9, 3)
    a.set_option(L.OPT_CHAIN, -1)
    a.set_option(L.OPT_PRECISION, 1)
    with pytest.raises(RuntimeError, match="FP64"):
        a.iterate(0.9, 3)
    a.destroy()


def test_barrier_gives_up_when_out_of_polls():
    m = _advection(patch_cells=64)
    m.set_option(L.OPT_PERSISTENT, 1)
    m.set_option(L.OPT_BARRIER_POLLS, 0)
    with pytest.raises(RuntimeError, match="timed out"):
        m.iterate(0.9, 20)
    print(f"  {m.get_option(L.OPT_PERSIST_GRID)} workgroups gave up", flush=True)
    # the status word is clean again and a normal call runs
    m.set_option(L.OPT_BARRIER_POLLS, 1 << 20)
    m.set_vec_function("solution", np.zeros((m.mesh.nb_nodes, 1)))
    ref = _advection(patch_cells=64)
    assert np.array_equal(m.iterate(0.9, 10), ref.iterate(0.9, 10))
    m.check_status()
    m.destroy()
    ref.destroy()
