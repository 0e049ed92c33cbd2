"""Field export (node averages as export_plt.cu, Tecplot bricks and VTK)."""
import numpy as np

from fooskagpu_amd import Mesh, cases, export


def test_constant_state_exports_constants():
    """A uniform state exports its own values at every node, and X, Y, Z are
    the node coordinates (sums of x / n over n cells)."""
    c = cases.bubble(n=4)
    U = np.tile(c.U[0], (c.mesh.n_cells, 1))
    cf = export.cell_fields(c.mesh, U, c.params)
    nf = export.node_fields(c.mesh, U, c.params)
    assert list(nf) == export.TWO_PHASE_VARS
    assert np.abs(cf["gradRho"]).max() < 1e-9 * cf["rho"][0]  # closed cells: sum n A = 0
    for name in export.TWO_PHASE_VARS[3:]:
        assert np.allclose(nf[name], cf[name][0], rtol=1e-12, atol=1e-9 * cf["rho"][0]), name
    for a, name in enumerate("XYZ"):
        assert np.allclose(nf[name], c.mesh.nodes[:, a], rtol=1e-14, atol=1e-15)


def test_node_average_is_the_mean_over_the_cells_of_each_node():
    """Brute force on a mixed-element box: each node's value is the mean of
    its cells' values (counted over all element types)."""
    m = Mesh.mixed_box(5, 2, 2, jitter=0.05, seed=3)
    q = np.random.default_rng(0).normal(size=m.n_cells)
    got = export.node_average(m, {"q": q})["q"]
    for node in range(m.n_nodes):
        cells = [c for c in range(m.n_cells) if node in m.el2nd[c, :m.nnode[c]]]
        assert got[node] == np.float64(sum(q[c] / len(cells) for c in cells)) or abs(
            got[node] - np.mean(q[cells])) < 1e-14


def test_green_gauss_linear_field_on_interior_cells():
    """Face averaging of centroid values is not exact per cell on Kuhn tets
    (measured up to 1.06 off a gradient of 3), but over the cells with only
    interior faces of the regular box it reproduces the gradient on average:
    the face terms telescope."""
    m = Mesh.box(6, 6, 6, order="none")
    q = 2.0 * m.cell_center[:, 0] - 3.0 * m.cell_center[:, 1] + 0.5 * m.cell_center[:, 2]
    g = export.green_gauss(m, q)
    interior = np.ones(m.n_cells, bool)
    interior[m.fc2el[m.fc2el[:, 1] < 0, 0]] = False
    assert np.allclose(g[interior].mean(0), [2.0, -3.0, 0.5], atol=1e-9)
    assert np.abs(g[interior] - [2.0, -3.0, 0.5]).max() < 1.1


def test_tecplot_brick_file_layout(tmp_path):
    """export_plt.cu:140-170: VARIABLES line, one FEPOINT brick zone, a row
    of %e values per node, then 8 1-based node ids per element with the
    TPNDS repetition of the element's nodes."""
    m = Mesh.mixed_box(5, 2, 2, jitter=0.05, seed=3, order="none")
    c = cases.bubble(n=4)
    regions = [dict(kind=0, alphaL=1 - c.params["avfMin"], p=1e7, T=cases.T_WATER),
               dict(kind=1, c=(0.5, 0.5, 0.5), r=0.3, alphaL=c.params["avfMin"], p=2340.0, T=cases.T_WATER)]
    U = m.init_regions(c.params, regions)
    nf = export.node_fields(m, U, c.params)
    path = tmp_path / export.plt_name(41)
    assert path.name == "field_0000042.dat"
    export.write_plt(path, m, nf, time=1.5e-6)
    lines = path.read_text().splitlines()
    assert lines[0] == 'VARIABLES = "X" "Y" "Z" "avfL" "avfG" "rho" "U" "V" "W" "Pressure" "Temperature" ' \
                       '"vaporMass" "gradRho"'
    assert lines[1] == (f'ZONE T="1",N={m.n_nodes},E={m.n_cells},F=FEPOINT,ET=BRICK,DT=(double,double,double),'
                        f'SOLUTIONTIME=1.500000e-06')
    assert len(lines) == 2 + m.n_nodes + m.n_cells
    vals = np.array([ln.split() for ln in lines[2:2 + m.n_nodes]], dtype=float)
    assert np.allclose(vals[:, 5], nf["rho"], rtol=1e-6)
    conn = np.array([ln.split() for ln in lines[2 + m.n_nodes:]], dtype=int)
    seen = set()
    for cidx in range(m.n_cells):
        nn = int(m.nnode[cidx])
        e = m.el2nd[cidx] + 1
        want = {4: [e[1], e[2], e[2], e[0], e[3], e[3], e[3], e[3]],
                5: [e[0], e[1], e[3], e[2], e[4], e[4], e[4], e[4]],
                6: [e[2], e[5], e[3], e[0], e[1], e[4], e[4], e[1]],
                8: [e[0], e[1], e[3], e[2], e[4], e[5], e[7], e[6]]}[nn]
        assert list(conn[cidx]) == want
        seen.add(nn)
    assert seen == {4, 5, 6, 8}


def test_vtk_file(tmp_path):
    m = Mesh.mixed_box(5, 2, 2, order="none")
    c = cases.sod(nx=8, ny=2, nz=2)
    U = m.init_sod(1.4, 0.0, 0.5, (1.0, 0, 0, 0, 1.0), (0.125, 0, 0, 0, 0.1))
    nf = export.node_fields(m, U, c.params)
    assert list(nf) == export.SINGLE_PHASE_VARS
    export.write_vtk(tmp_path / "a.vtk", m, nf)
    raw = (tmp_path / "a.vtk").read_bytes()
    assert raw.startswith(b"# vtk DataFile Version 3.0")
    i = raw.index(b"POINTS ")
    assert raw[i:].split(b"\n", 1)[0] == f"POINTS {m.n_nodes} double".encode()
    k = raw.index(b"CELL_TYPES ") + len(f"CELL_TYPES {m.n_cells}\n")
    types = np.frombuffer(raw[k:k + 4 * m.n_cells], dtype=">i4")
    assert np.array_equal(types, np.select([m.nnode == 4, m.nnode == 5, m.nnode == 6], [10, 14, 13], 12))
    j = raw.index(b"SCALARS rho double 1\nLOOKUP_TABLE default\n") + len(b"SCALARS rho double 1\nLOOKUP_TABLE default\n")
    rho = np.frombuffer(raw[j:j + 8 * m.n_nodes], dtype=">f8")
    assert np.allclose(rho, nf["rho"])
