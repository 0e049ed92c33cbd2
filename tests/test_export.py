"""Field export (node averages, Green-Gauss |grad rho|, VTK / Tecplot files)."""
import numpy as np

from fooskagpu_amd import Mesh, cases, export


def test_constant_state_exports_constants(tmp_path):
    c = cases.bubble(n=4)
    U = np.tile(c.U[0], (c.mesh.n_cells, 1))
    cf = export.cell_fields(c.mesh, U, c.params)
    nf = export.node_fields(c.mesh, U, c.params)
    assert list(nf) == export.TWO_PHASE_VARS
    assert np.abs(cf["gradRho"]).max() < 1e-9 * cf["rho"][0]  # closed cells: sum n A = 0
    for name, q in nf.items():
        assert np.allclose(q, cf[name][0], rtol=1e-12, atol=1e-9 * cf["rho"][0]), name


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


def test_vtk_and_tecplot_files(tmp_path):
    c = cases.sod(nx=8, ny=2, nz=2)
    nf = export.node_fields(c.mesh, c.U, c.params)
    assert list(nf) == export.SINGLE_PHASE_VARS
    export.write_vtk(tmp_path / "a.vtk", c.mesh, nf)
    raw = (tmp_path / "a.vtk").read_bytes()
    assert raw.startswith(b"# vtk DataFile Version 3.0")
    i = raw.index(b"POINTS ")
    assert raw[i:].split(b"\n", 1)[0] == f"POINTS {c.mesh.n_nodes} double".encode()
    j = raw.index(b"SCALARS rho double 1\nLOOKUP_TABLE default\n") + len(b"SCALARS rho double 1\nLOOKUP_TABLE default\n")
    rho = np.frombuffer(raw[j:j + 8 * c.mesh.n_nodes], dtype=">f8")
    assert np.allclose(rho, nf["rho"])
    export.write_tecplot(tmp_path / "a.dat", c.mesh, nf, time=1.5)
    lines = (tmp_path / "a.dat").read_text().splitlines()
    assert lines[1].startswith('VARIABLES = "x", "y", "z", "rho"')
    assert f"N={c.mesh.n_nodes}, E={c.mesh.n_cells}" in lines[2]
    assert len(lines) == 3 + c.mesh.n_nodes + c.mesh.n_cells
    assert np.array_equal(np.array(lines[-1].split(), dtype=int), c.mesh.el2nd[-1] + 1)
