"""Gmsh 2.2 reader / writer (GmshReader.hpp:118-176, 487-; GmshWriter.hpp:145-264)
on the reference's own mesh files (tests/golden/gmsh):

* the continuous equidistant Pk numbering of the corner mesh puts, cell by
  cell, every solution node where the file's Pk element has it (P2 and P3
  triangles and quadrilaterals, mixed meshes), and numbers as many nodes as the
  file holds -- which pins both the reader and the local node order (Gmsh's);
* cell areas / volumes sum to the domain's (rectangles 2 x 0.5, 4 x 1, 2 x 1,
  the cube [-1, 1]^3);
* boundaries carry the $PhysicalNames names and the corner vertices of their
  facets (checked against a plain-text parse);
* write -> read round trips keep cells, node positions and boundaries;
  $NodeData round trips are bit-exact;
* curved cells, binary files and unknown element types are refused.
"""
import os

import numpy as np
import pytest

from pdekit_amd import _lib as L
from pdekit_amd.gmsh import GmshReader, GmshWriter

HERE = os.path.join(os.path.dirname(__file__), "golden", "gmsh")
CELL_ORDER = {2: 1, 9: 2, 21: 3, 3: 1, 10: 2, 36: 3, 4: 1, 11: 2, 29: 3}
FACET = {1: 2, 8: 2, 26: 2, 2: 3, 9: 3, 21: 3}


def parse(path):
    """Plain-text parse: node coordinates, elements, physical names."""
    lines = open(path).read().splitlines()
    i = lines.index("$Nodes")
    n = int(lines[i + 1])
    xyz = np.array([[float(t) for t in ln.split()[1:4]] for ln in lines[i + 2:i + 2 + n]])
    j = lines.index("$Elements")
    el = [list(map(int, ln.split())) for ln in lines[j + 2:j + 2 + int(lines[j + 1])]]
    names = {}
    if "$PhysicalNames" in lines:
        k = lines.index("$PhysicalNames")
        for ln in lines[k + 2:k + 2 + int(lines[k + 1])]:
            d, t, nm = ln.split(None, 2)
            names[(int(d), int(t))] = nm.strip().strip('"')
    return xyz, el, names


def cells_of(el):
    dims = {2: 2, 9: 2, 21: 2, 3: 2, 10: 2, 36: 2, 4: 3, 11: 3, 29: 3}
    tdim = max(dims.get(e[1], 0) for e in el)
    return [e for e in el if dims.get(e[1], 0) == tdim], tdim


def cell_measure(mesh):
    tot = 0.0
    for b in mesh.blocks:
        X = mesh.coords[b.geo_conn]
        if mesh.dim == 3:
            a, c, d = X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]
            tot += np.einsum("ij,ij->i", np.cross(a, c), d).sum() / 6
        elif X.shape[1] == 3:
            a, c = X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]
            tot += 0.5 * (a[:, 0] * c[:, 1] - a[:, 1] * c[:, 0]).sum()
        else:  # shoelace
            x, y = X[:, :, 0], X[:, :, 1]
            tot += 0.5 * (x * np.roll(y, -1, 1) - np.roll(x, -1, 1) * y).sum()
    return tot


FILES = ["test_p1_tri", "test_p2_tri", "test_p3_tri", "test_p1_quad", "test_p2_quad", "test_p3_quad",
         "rectangle_mixed_elem_p1", "rectangle_mixed_elem_p2", "rectangle_mixed_elem_p3", "cube"]
AREA = {"test": 1.0, "test_p1_quad": 4.0, "test_p2_quad": 4.0, "test_p3_quad": 4.0, "rectangle": 2.0, "cube": 8.0}


@pytest.mark.parametrize("name", FILES)
def test_read_matches_file_nodes(name):
    path = os.path.join(HERE, name + ".msh")
    xyz, el, names = parse(path)
    cells, tdim = cells_of(el)
    p = CELL_ORDER[cells[0][1]]
    mesh = GmshReader().read_mesh_from_file(path, p, L.EQUIDIST)
    assert mesh.dim == tdim and mesh.nb_cells == len(cells)
    assert mesh.nb_nodes == xyz.shape[0]
    worst = 0.0
    for b in mesh.blocks:
        for c, conn in zip(b.cells, b.sol_conn):
            e = cells[c]
            fn = np.array(e[3 + e[2]:]) - 1
            worst = max(worst, np.abs(mesh.node_coords[conn] - xyz[fn][:, :tdim]).max())
    assert worst <= 1e-12
    area = AREA.get(name, AREA.get(name.split("_")[0]))
    assert abs(cell_measure(mesh) - area) <= 1e-12 * area
    # boundaries: names and corner vertices of the facet elements
    expect = {}
    for e in el:
        if e[1] in FACET and (tdim - 1, e[3]) in names and (e[1] in (1, 8, 26)) == (tdim == 2):
            expect.setdefault(names[(tdim - 1, e[3])], set()).update(np.array(e[3 + e[2]:3 + e[2] + FACET[e[1]]]) - 1)
    assert set(mesh.boundaries) == set(expect)
    P1 = GmshReader().read_mesh_from_file(path, 1)
    for nm, verts in expect.items():
        got = P1.boundaries[nm]
        assert np.allclose(np.sort(P1.node_coords[got], axis=0), np.sort(xyz[sorted(verts)][:, :tdim], axis=0))


def test_tet20_read_by_corners():
    path = os.path.join(HERE, "gmsh_ho_output.msh")
    xyz, el, _ = parse(path)
    mesh = GmshReader().read_mesh_from_file(path, 2)
    cells, _ = cells_of(el)
    assert mesh.nb_cells == len(cells) == 24
    corners = np.unique(np.concatenate([np.array(e[3 + e[2]:3 + e[2] + 4]) - 1 for e in cells]))
    assert np.allclose(np.sort(mesh.node_coords[:corners.size], axis=0), np.sort(xyz[corners], axis=0))


@pytest.mark.parametrize("name,order", [("rectangle_mixed_elem_p1", 1), ("rectangle_mixed_elem_p1", 2),
                                        ("test_p1_quad", 3), ("test_p1_tri", 3), ("cube", 2)])
def test_write_read_round_trip(tmp_path, name, order):
    src = GmshReader().read_mesh_from_file(os.path.join(HERE, name + ".msh"), order, L.EQUIDIST)
    out = str(tmp_path / "out.msh")
    GmshWriter().write_mesh_to_file(src, out)
    back = GmshReader().read_mesh_from_file(out, order, L.EQUIDIST)
    assert back.nb_cells == src.nb_cells and back.nb_nodes == src.nb_nodes
    assert set(back.boundaries) == set(src.boundaries) | set()
    for bs, bb in zip(src.blocks, back.blocks):
        assert bs.shape == bb.shape
        assert np.array_equal(src.node_coords[bs.sol_conn], back.node_coords[bb.sol_conn])
    for nm in src.boundaries:
        assert np.array_equal(np.sort(src.node_coords[src.boundaries[nm]], axis=0),
                              np.sort(back.node_coords[back.boundaries[nm]], axis=0))


def test_node_data_round_trip(tmp_path):
    mesh = GmshReader().read_mesh_from_file(os.path.join(HERE, "test_p2_tri.msh"), 2)
    rng = np.random.default_rng(0)
    u = rng.standard_normal((mesh.nb_nodes, 4)) * 10.0 ** rng.integers(-300, 300, (mesh.nb_nodes, 4))
    s = rng.standard_normal(mesh.nb_nodes)
    out = str(tmp_path / "sol.msh")
    w = GmshWriter()
    w.write_mesh_to_file(mesh, out)
    w.append_nodal_function_to_file(mesh, out, u, "u", time=0.5, timestep=7)
    w.append_nodal_function_to_file(mesh, out, s, "rho")
    r = GmshReader()
    back = r.read_nodal_function_from_file(mesh, out, ["u_0", "u_1", "u_2", "u_3", "rho"])
    assert np.array_equal(back[:, :4], u) and np.array_equal(back[:, 4], s)
    with pytest.raises(KeyError):
        r.read_nodal_function_from_file(mesh, out, ["u_4"])
    text = open(out).read()
    assert text.count("$NodeData") == 5 and '"u_3"' in text and "\n0.5\n3\n7\n1\n" in text


def _tiny(path, mid=(0.5, 0.0), binary=False, etype=9):
    nodes = [(0, 0), (1, 0), (0, 1), mid, (0.5, 0.5), (0.0, 0.5)]
    with open(path, "w") as f:
        f.write(f"$MeshFormat\n2.2 {1 if binary else 0} 8\n$EndMeshFormat\n$Nodes\n6\n")
        for i, (x, y) in enumerate(nodes):
            f.write(f"{i + 1} {x} {y} 0\n")
        f.write(f"$EndNodes\n$Elements\n1\n1 {etype} 2 1 1 1 2 3 4 5 6\n$EndElements\n")


def test_refusals(tmp_path):
    ok = str(tmp_path / "ok.msh")
    _tiny(ok)
    m = GmshReader().read_mesh_from_file(ok, 2, L.EQUIDIST)
    assert m.nb_cells == 1 and m.nb_nodes == 6
    for kw, msg in ((dict(mid=(0.5, 0.05)), "curved"), (dict(binary=True), "binary"), (dict(etype=5), "type")):
        bad = str(tmp_path / "bad.msh")
        _tiny(bad, **kw)
        with pytest.raises(L.PdrdError, match=msg):
            GmshReader().read_mesh_from_file(bad, 1)
