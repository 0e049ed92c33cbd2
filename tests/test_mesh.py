"""Host mesh library: geometry pinned by brute force and closed forms
(reference formulas meshProperties.cu:51-137, translator.cu face/cell centres,
face numbering buildLocalMesh.cu:244-343)."""
import math

import numpy as np
import pytest

from fooskagpu_amd import Mesh, cases
from fooskagpu_amd.mesh import BC_OUTLET, BC_SLIP, BC_WALL, place_bubbles


def closure(m):
    """Per-cell sum of outward n A (owner +, neighbour -)."""
    s = np.zeros((m.n_cells, 3))
    nA = m.n * m.area[:, None]
    np.add.at(s, m.fc2el[:, 0], nA)
    inner = m.fc2el[:, 1] >= 0
    np.add.at(s, m.fc2el[inner, 1], -nA[inner])
    return s


def test_sod_box_counts_match_face_formula():
    m = cases.sod().mesh
    C, F, B = m.n_cells, m.n_faces, m.n_bfaces
    assert (C, F, B) == (12288, 25728, 2304)
    assert F == (4 * C - B) // 2 + B
    assert set(np.unique(m.bc)) == {0, BC_OUTLET, BC_SLIP}
    assert np.all((m.fc2el[:, 1] < 0) == (m.bc != 0))
    assert np.all(m.fc2el[m.fc2el[:, 1] >= 0, 0] < m.fc2el[m.fc2el[:, 1] >= 0, 1])  # owner < neighbour


@pytest.mark.parametrize("jitter", [0.0, 0.1])
def test_volumes_closure_and_outward_normals(jitter):
    m = Mesh.box(5, 4, 3, (0, 0, 0), (2.0, 1.0, 0.5), jitter, 7, bc=(BC_WALL,) * 6)
    assert m.n_cells == 5 * 4 * 3 * 6
    assert np.all(m.volume > 0)
    assert abs(m.volume.sum() - 1.0) < 1e-13
    assert np.abs(closure(m)).max() < 1e-14
    assert np.allclose(np.linalg.norm(m.n, axis=1), 1.0, atol=1e-15)
    own = m.cell_center[m.fc2el[:, 0]]
    assert np.all(np.einsum("ij,ij->i", m.n, m.fc_center - own) > 0)
    inner = m.fc2el[:, 1] >= 0
    nb = m.cell_center[m.fc2el[inner, 1]]
    assert np.all(np.einsum("ij,ij->i", m.n[inner], m.fc_center[inner] - nb) < 0)
    # boundary faces lie on the box walls
    b = ~inner
    fc = m.fc_center[b]
    on_wall = np.isclose(fc, [0, 0, 0], atol=1e-14) | np.isclose(fc, [2.0, 1.0, 0.5], atol=1e-14)
    assert np.all(on_wall.any(axis=1))


def test_unit_tet_brute_force(tmp_path):
    """One tetrahedron (0,0,0),(1,0,0),(0,1,0),(0,0,1) read from a Gambit file:
    V = 1/6, faces 1/2, 1/2, 1/2 and sqrt(3)/2, centroids, outward normals."""
    path = tmp_path / "tet.neu"
    path.write_text(
        "CONTROL INFO 1.2.1\n** GAMBIT NEUTRAL FILE\nt\nPROGRAM: x\n\n"
        "NUMNP NELEM NGRPS NBSETS NDFCD NDFVL\n 4 1 1 1 3 3\nENDOFSECTION\n"
        "NODAL COORDINATES 1.2.1\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\nENDOFSECTION\n"
        "ELEMENTS/CELLS 1.2.1\n1 6 4 1 2 3 4\nENDOFSECTION\n"
        "BOUNDARY CONDITIONS 1.2.1\nwall 1 4 0 51\n1 6 1\n1 6 2\n1 6 3\n1 6 4\nENDOFSECTION\n")
    m = Mesh.read_neu(path, rcm=False)
    assert (m.n_cells, m.n_faces, m.n_bfaces) == (1, 4, 4)
    assert m.volume[0] == pytest.approx(1 / 6, abs=1e-16)
    assert sorted(m.area) == pytest.approx([0.5, 0.5, 0.5, math.sqrt(3) / 2], abs=1e-15)
    assert np.allclose(m.cell_center[0], [0.25, 0.25, 0.25])
    assert np.all(m.bc == BC_WALL)
    expect_n = {(0, 0, -1), (0, -1, 0), (-1, 0, 0)}
    got = {tuple(np.round(v).astype(int)) for v in m.n if abs(abs(v).max() - 1) < 1e-15}
    assert got == expect_n
    slanted = [v for v in m.n if abs(abs(v).max() - 1) > 1e-6]
    assert np.allclose(slanted[0], np.ones(3) / math.sqrt(3))
    # face-node order of meshProperties.cu: face 0 = (1,0,2) -> centroid of nodes 0,1,2
    assert np.allclose(m.fc_center[0], [1 / 3, 1 / 3, 0])
    t1, t2 = m.tangents()
    assert np.allclose(np.einsum("ij,ij->i", t1, m.n), 0, atol=1e-15)
    assert np.allclose(np.cross(m.n, t1), t2, atol=1e-15)


def test_neu_round_trip_is_exact(tmp_path):
    a = Mesh.box(4, 3, 3, (0, 0, 0), (1, 1, 1), 0.1, 3, bc=(1, 2, 3, 4, 5, 2))
    path = tmp_path / "box.neu"
    a.write_neu(path)
    b = Mesh.read_neu(path, rcm=False)
    for name in ("fc2el", "bc", "n", "area", "fc_center", "cell_center", "volume"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name


def test_periodic_box_pairs_every_face():
    n = 4
    L = 2 * math.pi
    m = Mesh.box(n, n, n, (0, 0, 0), (L, L, L), 0.1, 5, periodic=(True, True, True))
    assert m.n_bfaces == 0 and m.n_faces == 2 * m.n_cells
    assert abs(m.volume.sum() - L ** 3) < 1e-11
    assert np.abs(closure(m)).max() < 1e-13
    vals = np.unique(np.round(m.shift / L, 12))
    assert set(vals) <= {-1.0, 0.0, 1.0}
    per = np.any(m.shift != 0, axis=1)
    assert per.sum() == 3 * n * n * 2  # three periodic planes, two triangles per hex face
    nb = m.cell_center[m.fc2el[:, 1]]
    assert np.all(np.einsum("ij,ij->i", m.n, (m.fc_center - m.shift) - nb) < 0)


def test_rcm_reduces_bandwidth_and_keeps_geometry():
    a = Mesh.box(6, 6, 6, (0, 0, 0), (1, 1, 1), 0.1, 9, rcm=False)
    b = Mesh.box(6, 6, 6, (0, 0, 0), (1, 1, 1), 0.1, 9, rcm=True)
    bw = lambda m: np.abs(m.fc2el[m.fc2el[:, 1] >= 0, 1] - m.fc2el[m.fc2el[:, 1] >= 0, 0]).max()
    assert bw(b) < bw(a)
    assert np.isclose(a.volume.sum(), b.volume.sum(), rtol=0, atol=1e-14)
    assert np.allclose(np.sort(a.volume), np.sort(b.volume), rtol=0, atol=0)
    assert a.n_faces == b.n_faces


def test_bubbles_do_not_overlap_and_fit():
    c, r = place_bubbles(64, (0, 0, 0), (1, 1, 1), 0.02, 0.05, 0.2, 4)
    assert np.all((r >= 0.02) & (r <= 0.05))
    assert np.all(c - r[:, None] >= 0) and np.all(c + r[:, None] <= 1) and np.all(c[:, 0] - r >= 0.2)
    d = np.linalg.norm(c[:, None] - c[None], axis=2)
    s = r[:, None] + r[None]
    np.fill_diagonal(d, 1)
    np.fill_diagonal(s, 0)
    assert np.all(d > s)


def test_bad_arguments_raise():
    from fooskagpu_amd import FkError
    with pytest.raises(FkError):
        Mesh.box(0, 1, 1)
    with pytest.raises(FkError):
        Mesh.box(2, 2, 2, periodic=(True, False, False))  # periodic axis needs >= 3 hexes
    with pytest.raises(FkError):
        Mesh.read_neu("/nonexistent/sh.neu")


# ---------------- hexahedra, prisms, pyramids (translator.cu:122-293) ----------------
def one_element_neu(path, nodes, elem_type, bc_code=51):
    """Gambit file with one element of `elem_type` (4 brick, 5 wedge, 6 tet,
    7 pyramid) on `nodes` and every face in one boundary set."""
    nf = {4: 6, 5: 5, 6: 4, 7: 5}[elem_type]
    lines = ["CONTROL INFO 1.2.1", "** GAMBIT NEUTRAL FILE", "t", "PROGRAM: x", "",
             "NUMNP NELEM NGRPS NBSETS NDFCD NDFVL", f" {len(nodes)} 1 1 1 3 3", "ENDOFSECTION",
             "NODAL COORDINATES 1.2.1"]
    lines += [f"{i + 1} {x} {y} {z}" for i, (x, y, z) in enumerate(nodes)]
    lines += ["ENDOFSECTION", "ELEMENTS/CELLS 1.2.1",
              f"1 {elem_type} {len(nodes)} " + " ".join(str(i + 1) for i in range(len(nodes))), "ENDOFSECTION",
              "BOUNDARY CONDITIONS 1.2.1", f"wall 1 {nf} 0 {bc_code}"]
    lines += [f"1 {elem_type} {j + 1}" for j in range(nf)] + ["ENDOFSECTION"]
    path.write_text("\n".join(lines) + "\n")


UNIT_HEX = [(a & 1, a >> 1 & 1, a >> 2 & 1) for a in range(8)]


@pytest.mark.parametrize("name,nodes,etype,volume,areas,centre", [
    ("hex", UNIT_HEX, 4, 1.0, [1.0] * 6, [0.5, 0.5, 0.5]),
    ("prism", [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)], 5, 0.5,
     [1.0, math.sqrt(2), 1.0, 0.5, 0.5], [1 / 3, 1 / 3, 0.5]),
    ("pyramid", [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0.5, 0.5, 1)], 7, 1 / 3,
     [1.0] + [math.sqrt(1.25) / 2] * 4, [0.5, 0.5, 0.2]),
])
def test_single_element_brute_force(tmp_path, name, nodes, etype, volume, areas, centre):
    """One element read from a Gambit file: volume from the reference's
    sub-tets, face areas, the node-mean centre, outward unit normals, closure
    sum n A = 0 for planar faces and face centres inside the element's faces."""
    path = tmp_path / f"{name}.neu"
    one_element_neu(path, nodes, etype)
    m = Mesh.read_neu(path, rcm=False)
    assert m.n_cells == 1 and m.n_faces == len(areas) and m.n_bfaces == len(areas)
    assert m.nnode[0] == len(nodes) and list(m.el2nd[0, :len(nodes)]) == list(range(len(nodes)))
    assert m.volume[0] == pytest.approx(volume, abs=1e-15)
    assert list(m.area) == pytest.approx(areas, abs=1e-15)  # local face order of the tables
    assert np.allclose(m.cell_center[0], centre, atol=1e-15)
    assert np.all(m.bc == BC_WALL)
    assert np.allclose(np.linalg.norm(m.n, axis=1), 1.0, atol=1e-15)
    assert np.all(np.einsum("ij,ij->i", m.n, m.fc_center - m.cell_center[0]) > 0)
    assert np.abs(closure(m)).max() < 1e-15


def test_hex_face_table_follows_the_reference():
    """meshProperties.cu fcNd2ElNd for hexahedra: faces 0..5 are y = 0,
    x = 1, y = 1, x = 0, z = 0, z = 1 of the Gambit brick."""
    m = Mesh.box(1, 1, 1, kinds=np.ones((1, 1, 1)), order="none")
    assert m.n_cells == 1 and m.nnode[0] == 8
    expect = [(0, -1, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, 0, -1), (0, 0, 1)]
    assert np.allclose(m.n, expect, atol=1e-15)
    assert np.allclose(m.fc_center, [[0.5, 0, 0.5], [1, 0.5, 0.5], [0.5, 1, 0.5], [0, 0.5, 0.5], [0.5, 0.5, 0],
                                      [0.5, 0.5, 1]], atol=1e-15)


@pytest.mark.parametrize("jitter", [0.0, 0.1])
def test_mixed_box_is_conforming_and_closed(jitter):
    """Tets | split hexes | hexes | split hexes | prisms: every element type,
    every interior face matched (a triangle never against a quadrilateral),
    volumes fill the box, outward normals; per-cell closure is exact for
    planar faces (jitter 0) and bounded by the quadrilaterals' warp."""
    m = Mesh.mixed_box(10, 3, 2, (0, 0, 0), (2.0, 0.6, 0.4), jitter, 5, bc=(BC_WALL,) * 6)
    counts = {k: int((m.nnode == k).sum()) for k in (4, 5, 6, 8)}
    # slabs of 2 x 3 x 2 hexes: 72 tets | 6 split hexes facing the tets (2 tets
    # + 5 pyramids) and 6 with 6 pyramids | 12 hexes | 72 pyramids | 24 prisms
    assert counts == {4: 72 + 12, 5: 30 + 36 + 72, 6: 24, 8: 12}
    assert np.all(m.volume > 0)
    # the sub-tet volumes tile the box exactly when the quadrilaterals are
    # planar; warped ones (jitter) split differently on their two sides
    assert abs(m.volume.sum() - 2.0 * 0.6 * 0.4) < (1e-13 if jitter == 0 else 1e-4 * 0.48)
    own = m.cell_center[m.fc2el[:, 0]]
    assert np.all(np.einsum("ij,ij->i", m.n, m.fc_center - own) > 0)
    inner = m.fc2el[:, 1] >= 0
    nb = m.cell_center[m.fc2el[inner, 1]]
    assert np.all(np.einsum("ij,ij->i", m.n[inner], m.fc_center[inner] - nb) < 0)
    # only the box surface is boundary: its area is the box's
    assert abs(m.area[~inner].sum() - 2 * (2.0 * 0.6 + 2.0 * 0.4 + 0.6 * 0.4)) < 1e-12
    faces = np.bincount(m.fc2el[:, 0], minlength=m.n_cells) + np.bincount(m.fc2el[inner, 1], minlength=m.n_cells)
    assert np.array_equal(faces, m.faces_per_cell)
    cl = np.abs(closure(m)).max()
    assert cl < (1e-14 if jitter == 0 else 5e-3)


def test_mixed_box_neu_round_trip_is_exact(tmp_path):
    a = Mesh.mixed_box(6, 2, 2, jitter=0.1, seed=3, bc=(1, 2, 3, 4, 5, 2), order="none")
    path = tmp_path / "mixed.neu"
    a.write_neu(path)
    b = Mesh.read_neu(path, rcm=False)
    for name in ("fc2el", "bc", "n", "area", "fc_center", "cell_center", "volume", "nnode", "el2nd"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name


def test_non_conforming_kinds_are_rejected():
    from fooskagpu_amd import FkError
    with pytest.raises(FkError, match="conform"):
        Mesh.box(2, 1, 1, kinds=np.array([[[0, 1]]]))  # tets against a hexahedron
    with pytest.raises(FkError, match="conform"):
        Mesh.box(1, 1, 2, kinds=np.array([[[1]], [[2]]]))  # hexahedron on a prism's triangles
    Mesh.box(1, 1, 2, kinds=np.array([[[0]], [[2]]]))  # tets under prisms: same diagonal
    Mesh.box(3, 1, 1, kinds=np.array([[[0, 3, 1]]]))
