"""pdrd_assemble_rhs_noalloc (the cell loop over caller-owned device arrays,
scratch in a static device array) against pdrd_assemble_rhs of a method on
the same mesh and state: residual and wave speed within 1e-12 normwise, on
two calls in a row (the second with a smaller mesh, so that the scratch
keeps the first call's larger contents), for scalar and Euler physics."""
import ctypes as C

import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.fe import element_tables
from pdekit_amd.mesh import box_tetra, rectangle
from pdekit_amd.rdm import PGRDMethodExplicit
from tests.helpers import burgers_state, euler_state, rel_err

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _reference(mesh, physics, scheme, params, state, qorder):
    m = PGRDMethodExplicit(physics, scheme, params=params).configure_mesh_data(mesh)
    m.initialize_work_data(quad_order=qorder)
    m.set_vec_function("solution", state)
    m.assemble_rhs()
    out = (m.vec_function("residuals"), m.vec_function("wave_speed"), m.vec_function("dual_volume"))
    m.destroy()
    return out


def _noalloc(mesh, physics, scheme, params, state, qorder, vol):
    """One call through the C ABI with every array in device memory (torch
    tensors on cuda:0); returns residual, wave speed, status."""
    assert len(mesh.blocks) == 1
    blk = mesh.blocks[0]
    t = element_tables(blk.shape, mesh.order, qorder)
    dev = torch.device("cuda:0")
    tab = torch.tensor(np.asarray(t.packed), dtype=torch.float64, device=dev)
    gconn = torch.tensor(np.ascontiguousarray(blk.geo_conn.T), dtype=torch.int32, device=dev)
    sconn = torch.tensor(np.ascontiguousarray(blk.sol_conn.T), dtype=torch.int32, device=dev)
    coords = torch.tensor(np.asarray(mesh.coords), dtype=torch.float64, device=dev)
    u = torch.tensor(np.asarray(state), dtype=torch.float64, device=dev)
    v = torch.tensor(np.asarray(vol), dtype=torch.float64, device=dev)
    res = torch.full(u.shape, np.nan, dtype=torch.float64, device=dev)
    ws = torch.full((mesh.nb_nodes,), np.nan, dtype=torch.float64, device=dev)
    status = torch.full((1,), -7, dtype=torch.int32, device=dev)
    p = np.zeros(4)
    p[:len(params)] = params
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream
    rc = L.lib().pdrd_assemble_rhs_noalloc(
        physics, scheme, L.ptr(p), len(params), C.byref(t.desc), tab.data_ptr(), blk.geo_conn.shape[0],
        gconn.data_ptr(), sconn.data_ptr(), mesh.coords.shape[0], coords.data_ptr(), mesh.nb_nodes, u.data_ptr(),
        v.data_ptr(), res.data_ptr(), ws.data_ptr(), status.data_ptr(), stream)
    L.check(rc, "pdrd_assemble_rhs_noalloc")
    torch.cuda.synchronize()
    return res.cpu().numpy(), ws.cpu().numpy(), int(status.cpu()[0])


CASES = {
    "burgers_p2_b": (lambda n: rectangle(n, n, split=4, order=2, jitter=0.15, seed=2), L.BURGERS2D, L.B, (),
                     burgers_state, 4, (16, 8)),
    "euler2d_p1_n": (lambda n: rectangle(n, n, split=2, order=1, jitter=0.1, seed=3), L.EULER2D, L.N,
                     (1.4, 287.05), euler_state, 4, (20, 12)),
    "euler3d_p2_lda": (lambda n: box_tetra(n, n, n, order=2, jitter=0.1, seed=4), L.EULER3D, L.LDA,
                       (1.4, 287.05), lambda m: euler_state(m, dim=3), 2, (5, 3)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_noalloc_matches_assemble_rhs(name):
    build, physics, scheme, params, state_fn, qorder, sizes = CASES[name]
    for n in sizes:  # the second call is smaller than the first
        mesh = build(n)
        state = state_fn(mesh)
        res_ref, ws_ref, vol = _reference(mesh, physics, scheme, params, state, qorder)
        print(f"[noalloc] {name} n={n}: {mesh.nb_cells} cells, {mesh.nb_nodes} nodes", flush=True)
        res, ws, status = _noalloc(mesh, physics, scheme, params, state, qorder, vol)
        assert status == 0
        assert rel_err(res, res_ref) <= 1e-12
        assert rel_err(ws, ws_ref) <= 1e-12


def test_noalloc_refuses_oversized_meshes():
    """nb_nodes * (m + 1) beyond the scratch limit is refused before any device
    work."""
    mesh = rectangle(4, 4, split=2, order=1)
    blk = mesh.blocks[0]
    t = element_tables(blk.shape, 1, 4)
    p = np.zeros(4)
    big = L.NOALLOC_MAX_ACC // 5 + 1  # Euler 2D: m + 1 = 5
    rc = L.lib().pdrd_assemble_rhs_noalloc(L.EULER2D, L.LDA, L.ptr(p), 0, C.byref(t.desc), 1, blk.geo_conn.shape[0],
                                           1, 1, mesh.coords.shape[0], 1, big, 1, 1, 1, 1, None, None)
    assert rc == -1 and b"scratch limit" in L.lib().pdrd_last_error()
