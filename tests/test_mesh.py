"""Mesh inputs: make_unit_quad numbering, continuous P1 -> Pk DOF numbering,
jitter, bump mappings and boundary node sets.

The numbering rule pinned here is the one of
src/mesh/containers/DofMap.hpp:2556-3073: P1 vertices keep their ascending
ids, every higher-order node is numbered once in order of first appearance
in a cell-order scan of the local node lists; it is restated below in plain
Python and compared entry by entry.
"""
import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.fe import reference_nodes
from pdekit_amd.mesh import MeshGen, box_tetra, rectangle, unit_quad_triag

EDGES = {L.TRIAG: [(0, 1), (1, 2), (2, 0)], L.QUAD: [(0, 1), (1, 2), (2, 3), (3, 0)],
         L.TETRA: [(0, 1), (1, 2), (2, 0), (3, 0), (3, 2), (3, 1)]}


def test_make_unit_quad_numbering():
    N = 5
    m = MeshGen(0, [N]).clone(1).build()
    assert m.nb_cells == 2 * (N - 1) ** 2 and m.nb_nodes == N * N
    g = m.blocks[0].geo_conn
    k = 0
    c = 0
    for j in range(N - 1):
        for i in range(N - 1):
            assert list(g[c]) == [k, k + 1, k + N + 1]
            assert list(g[c + 1]) == [k + N + 1, k + N, k]
            c += 2
            k += 1
        k += 1
    x = m.coords
    assert np.allclose(x[:N, 1], -1) and np.allclose(x[:N, 0], np.linspace(-1, 1, N))
    assert np.array_equal(np.sort(m.boundaries["Edge0"]), np.arange(N))
    assert np.array_equal(np.sort(m.boundaries["Edge3"]), np.arange(0, N * N, N))
    # P1 solution numbering is the vertex numbering
    assert np.array_equal(m.blocks[0].sol_conn, g)


def _reference_numbering(shapes, geo_cells, order):
    """First-appearance numbering of higher-order nodes (plain restatement)."""
    nv = max(max(c) for c in geo_cells) + 1
    next_id = nv
    edge_nodes = {}
    out = []
    for sh, cv in zip(shapes, geo_cells):
        nvert = len(cv)
        nodes = list(cv)
        for a, b in EDGES[sh]:
            va, vb = cv[a], cv[b]
            key = (min(va, vb), max(va, vb))
            ids = []
            for i in range(1, order):
                pos = i if va < vb else order - i  # position counted from the smaller vertex
                if (key, pos) not in edge_nodes:
                    edge_nodes[(key, pos)] = next_id
                    next_id += 1
                ids.append(edge_nodes[(key, pos)])
            nodes += ids
        n_total = {L.TRIAG: (order + 1) * (order + 2) // 2, L.QUAD: (order + 1) ** 2,
                   L.TETRA: (order + 1) * (order + 2) * (order + 3) // 6}[sh]
        while len(nodes) < n_total:
            nodes.append(next_id)
            next_id += 1
        assert len(nodes) == n_total and nvert <= n_total
        out.append(nodes)
    return out, next_id


@pytest.mark.parametrize("kind,order", [("tri", 2), ("tri", 3), ("mixed", 2), ("criss", 2), ("tet", 2)])
def test_clone_continuous_matches_first_appearance(kind, order):
    if kind == "tri":
        g = MeshGen(0, [6])
    elif kind == "mixed":
        g = MeshGen(1, [5, 4, 2, 2], [0, 1, 0, 1])
    elif kind == "criss":
        g = MeshGen(1, [4, 3, 4, 0], [0, 1, 0, 1])
    else:
        g = MeshGen(2, [2, 2, 2], [1, 1, 1])
    m = g.clone(order).build()
    shapes = m.cell_shape.tolist()
    # geometry cells in global order
    geo = [None] * m.nb_cells
    sol = [None] * m.nb_cells
    for b in m.blocks:
        for i, c in enumerate(b.cells):
            geo[c] = b.geo_conn[i].tolist()
            sol[c] = b.sol_conn[i].tolist()
    ref, n = _reference_numbering(shapes, geo, order)
    assert m.nb_nodes == n
    for c in range(m.nb_cells):
        assert sol[c] == ref[c], c


def test_high_order_node_coordinates():
    m = rectangle(6, 5, split=2, order=3, jitter=0.1, seed=9)
    x = m.coords
    xn = m.node_coords
    b = m.blocks[0]
    gl = 1 / np.sqrt(5)  # P3 edge nodes sit at the Gauss-Lobatto points of the edge
    for c in range(0, b.geo_conn.shape[0], 7):
        v = b.geo_conn[c]
        s = b.sol_conn[c]
        assert np.allclose(xn[s[:3]], x[v])
        for e, (i, j) in enumerate(EDGES[L.TRIAG]):
            t = np.array([(1 - gl) / 2, (1 + gl) / 2])
            pts = x[v[i]][None, :] * (1 - t[:, None]) + x[v[j]][None, :] * t[:, None]
            assert np.allclose(xn[s[3 + 2 * e:5 + 2 * e]], pts, atol=1e-13)
        assert np.allclose(xn[s[9]], x[v].mean(0), atol=1e-13)


def _volumes(coords, conn):
    if conn.shape[1] == 3:
        a, b, c = coords[conn[:, 0]], coords[conn[:, 1]], coords[conn[:, 2]]
        return 0.5 * ((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))
    if conn.shape[1] == 4 and coords.shape[1] == 3:
        a = coords[conn[:, 0]]
        e = [coords[conn[:, k]] - a for k in (1, 2, 3)]
        return np.einsum("ij,ij->i", e[0], np.cross(e[1], e[2])) / 6.0
    # quad: shoelace
    p = coords[conn]
    return 0.5 * np.sum(p[:, :, 0] * np.roll(p[:, :, 1], -1, 1) - np.roll(p[:, :, 0], -1, 1) * p[:, :, 1], 1)


def test_jitter_bump_and_orientation():
    m1 = rectangle(16, 12, split=4, order=1, jitter=0.15, seed=2)
    m2 = rectangle(16, 12, split=4, order=1, jitter=0.15, seed=2)
    m3 = rectangle(16, 12, split=4, order=1, jitter=0.15, seed=3)
    m0 = rectangle(16, 12, split=4, order=1)
    assert np.array_equal(m1.coords, m2.coords)
    assert not np.array_equal(m1.coords, m3.coords)
    d = np.abs(m1.coords - m0.coords)
    assert d[:, 0].max() <= 0.15 / 16 + 1e-15 and d[:, 1].max() <= 0.15 / 12 + 1e-15
    bnd = np.unique(np.concatenate([m0.boundaries[k] for k in ("xmin", "xmax", "ymin", "ymax")]))
    assert np.array_equal(m1.coords[bnd], m0.coords[bnd])
    assert (_volumes(m1.coords, m1.blocks[0].geo_conn) > 0).all()
    assert abs(_volumes(m1.coords, m1.blocks[0].geo_conn).sum() - 1.0) < 1e-13
    # bump channel: the lower wall follows the bump, volumes stay positive
    mb = rectangle(32, 16, extent=(-0.5, 1.5, 0, 1), split=2, quad_rows=8, bump=(1.0, 0.1, 25.0, 0.5))
    low = mb.boundaries["ymin"]
    assert np.allclose(mb.coords[low, 1], 0.1 * np.exp(-25 * (mb.coords[low, 0] - 0.5) ** 2))
    for b in mb.blocks:
        assert (_volumes(mb.coords, b.geo_conn) > 0).all()
    mt = box_tetra(4, 3, 5, extent=(1.0, 1.0, 0.5), jitter=0.1, seed=4, bump=(0.5, 0.08, 25.0, 0.5, 0.5))
    v = _volumes(mt.coords, mt.blocks[0].geo_conn)
    assert (v > 0).all() and mt.nb_cells == 6 * 4 * 3 * 5
    m_flat = box_tetra(4, 3, 5, extent=(1.0, 1.0, 0.5))
    assert abs(_volumes(m_flat.coords, m_flat.blocks[0].geo_conn).sum() - 0.5) < 1e-14


def test_boundary_nodes_are_on_the_boundary():
    m = rectangle(8, 6, split=2, order=3)
    for name, axis, val in (("xmin", 0, 0.0), ("xmax", 0, 1.0), ("ymin", 1, 0.0), ("ymax", 1, 1.0)):
        ids = m.boundaries[name]
        assert np.allclose(m.node_coords[ids, axis], val)
        on = np.isclose(m.node_coords[:, axis], val)
        assert on.sum() == ids.size
    m = unit_quad_triag(9, order=2)
    assert m.boundaries["Edge0"].size == 2 * 8 + 1


def test_sizes_of_the_benchmark_meshes():
    # 512 x 512 criss-cross P2: 1 048 576 triangles, 2 099 201 nodes (checked
    # on a 64 x 64 copy with the same formulas)
    n = 64
    m = rectangle(n, n, split=4, order=2)
    assert m.nb_cells == 4 * n * n
    nv = (n + 1) ** 2 + n * n
    ne = 2 * n * (n + 1) + 4 * n * n
    assert m.nb_nodes == nv + ne
