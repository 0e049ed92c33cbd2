"""Field export for validation plots: node-averaged variables written as the
reference's Tecplot FEBRICK file (field_%07d.dat) or as legacy VTK.

Restates export_plt (two-phase mulitphase Solver/doublePrecision/
export_plt.cu:41-174, single-phase singlePhase/src/export_plt.cu:36-150):
every node gets sum over the cells c sharing it of value_c / (number of those
cells), added in ascending cell order, for
  two-phase:    X Y Z avfL avfG rho U V W Pressure Temperature vaporMass gradRho
                (Pressure = sum_k alpha_k p_k, Temperature = sum_k alpha_k T_k,
                vaporMass = consVec[9], gradRho = |Green-Gauss gradient of the
                mixture density|, the reference's grad row 9)
  single-phase: X Y Z rho U V W E Pressure Temperature Mach Rank
and each element is written as an 8-node brick with repeated nodes (TPNDS,
export_plt.cu:4-13).  The gradient comes from the solver that produced the
state (GpuSolver or the oracle, pass G of that state); the reference exports
the gradient of its last RK stage.  Host-side numpy: the export is I/O, not
part of the time step.
"""
import numpy as np

TWO_PHASE_VARS = ["X", "Y", "Z", "avfL", "avfG", "rho", "U", "V", "W", "Pressure", "Temperature", "vaporMass",
                  "gradRho"]
SINGLE_PHASE_VARS = ["X", "Y", "Z", "rho", "U", "V", "W", "E", "Pressure", "Temperature", "Mach", "Rank"]

# element node -> brick node: TPNDS[ind][nElNds] of export_plt.cu:4-13, by node count
BRICK_NODES = {4: (1, 2, 2, 0, 3, 3, 3, 3), 5: (0, 1, 3, 2, 4, 4, 4, 4), 6: (2, 5, 3, 0, 1, 4, 4, 1),
               8: (0, 1, 3, 2, 4, 5, 7, 6)}
# legacy VTK cell types and node orders (from the reference's Gambit orders)
VTK_CELLS = {4: (10, (0, 1, 2, 3)), 5: (14, (0, 1, 3, 2, 4)), 6: (13, (0, 2, 1, 3, 5, 4)),
             8: (12, (0, 1, 3, 2, 4, 5, 7, 6))}


def _eos(params, k, rho, e):
    """eqos cases 0 (p) and 1 (T) of the stiffened gas (DP eqos.h:1-23)."""
    g, pi, cp, eta = (params[n][k] for n in ("gamma", "pInf", "Cp", "eta"))
    p = (g - 1) * rho * e - g * pi - (g - 1) * eta * rho
    T = (p + pi) / ((cp / g) * (g - 1) * rho)
    return p, T


def green_gauss(mesh, q):
    """Cell gradients of the cell field q by the face-average rule
    (computeGrad.cu:51-91), boundary faces taking the cell's own value (a
    stand-alone gradient for states without a solver pass)."""
    f2 = mesh.fc2el
    o, nb = f2[:, 0], f2[:, 1]
    qn = np.where(nb >= 0, q[np.maximum(nb, 0)], q[o])
    flux = (0.5 * (q[o] + qn) * mesh.area)[:, None] * mesh.n
    g = np.zeros((mesh.n_cells, 3))
    np.add.at(g, o, flux)
    inter = nb >= 0
    np.add.at(g, nb[inter], -flux[inter])
    return g / mesh.volume[:, None]


def cell_fields(mesh, U, params, grad=None):
    """Cell values of the exported variables except X, Y, Z (dict name -> [C]).
    grad: the solver's Green-Gauss gradients [C][NEQ][3] of the primitive
    variables (two-phase: row 9 = mixture density); None computes the
    mixture-density gradient with green_gauss."""
    U = np.asarray(U, dtype=np.float64)
    if U.shape[1] == 5:
        rho = U[:, 0]
        u, v, w = (U[:, i] / rho for i in (1, 2, 3))
        E = U[:, 4] / rho
        e = E - 0.5 * (u * u + v * v + w * w)
        p, T = _eos(params, 0, rho, e)
        g, pi = params["gamma"][0], params["pInf"][0]
        mach = np.sqrt(u * u + v * v + w * w) / np.sqrt(g * (p + pi) / rho)
        return dict(rho=rho, U=u, V=v, W=w, E=E, Pressure=p, Temperature=T, Mach=mach, Rank=np.zeros_like(rho))
    aL = U[:, 0]
    aG = 1 - aL
    rL, rG = U[:, 1] / aL, U[:, 2] / aG
    rho = aL * rL + aG * rG
    u, v, w = (U[:, i] / rho for i in (3, 4, 5))
    pL, TL = _eos(params, 0, rL, U[:, 6] / (aL * rL))
    pG, TG = _eos(params, 1, rG, U[:, 7] / (aG * rG))
    gr = np.asarray(grad)[:, 9, :] if grad is not None else green_gauss(mesh, rho)
    return dict(avfL=aL, avfG=aG, rho=rho, U=u, V=v, W=w, Pressure=aL * pL + aG * pG,
                Temperature=aL * TL + aG * TG, vaporMass=U[:, 9],
                gradRho=np.sqrt(gr[:, 0] * gr[:, 0] + gr[:, 1] * gr[:, 1] + gr[:, 2] * gr[:, 2]))


def _incidence(mesh):
    """(cell, node) pairs in ascending cell order and the cell count per node."""
    nn = mesh.nnode
    cells = np.repeat(np.arange(mesh.n_cells), nn)
    mask = np.arange(8)[None, :] < nn[:, None]
    nodes = mesh.el2nd[mask]
    cnt = np.bincount(nodes, minlength=mesh.n_nodes).astype(np.float64)
    return cells, nodes, cnt


def node_average(mesh, cell_values):
    """sum over the cells sharing each node of value / count, in ascending
    cell order (computeNodeVector, export_plt.cu:41-93)."""
    cells, nodes, cnt = _incidence(mesh)
    inv = 1.0 / np.maximum(cnt, 1)
    out = {}
    for name, q in cell_values.items():
        out[name] = np.bincount(nodes, weights=np.asarray(q)[cells] * inv[nodes], minlength=mesh.n_nodes)
    return out


def node_fields(mesh, U, params, grad=None):
    """Every exported node variable, in the reference's order."""
    cells, nodes, cnt = _incidence(mesh)
    inv = 1.0 / np.maximum(cnt, 1)
    xyz = {k: np.bincount(nodes, weights=mesh.nodes[nodes, a] * inv[nodes], minlength=mesh.n_nodes)
           for a, k in enumerate("XYZ")}
    xyz.update(node_average(mesh, cell_fields(mesh, U, params, grad)))
    names = SINGLE_PHASE_VARS if np.asarray(U).shape[1] == 5 else TWO_PHASE_VARS
    return {k: xyz[k] for k in names}


def brick_connectivity(mesh):
    """[C][8] 1-based node ids of each element as a Tecplot brick."""
    out = np.empty((mesh.n_cells, 8), dtype=np.int64)
    for nn, idx in BRICK_NODES.items():
        sel = mesh.nnode == nn
        out[sel] = mesh.el2nd[sel][:, list(idx)] + 1
    return out


def write_plt(path, mesh, fields, time=0.0):
    """Tecplot ASCII in the layout of export_plt.cu:140-170: one FEPOINT
    zone of bricks, node rows of `%e` values, then 1-based brick
    connectivity."""
    names = list(fields)
    data = np.column_stack([np.asarray(fields[n]) for n in names])
    with open(path, "w") as f:
        f.write("VARIABLES = " + " ".join(f'"{n}"' for n in names) + "\n")
        f.write(f'ZONE T="1",N={mesh.n_nodes},E={mesh.n_cells},F=FEPOINT,ET=BRICK,DT=(double,double,double),'
                f"SOLUTIONTIME={time:e}\n")
        np.savetxt(f, data, fmt="%e", delimiter=" ")
        np.savetxt(f, brick_connectivity(mesh), fmt="%d", delimiter=" ")


def plt_name(step):
    """The reference's file name field_%07d.dat (iStep + 1)."""
    return f"field_{step + 1:07d}.dat"


def write_vtk(path, mesh, fields, title="fooskagpu_amd"):
    """Legacy VTK unstructured grid, binary (big endian), the cells as VTK
    tetrahedra / pyramids / wedges / hexahedra, point data = `fields`."""
    nn, nc = mesh.n_nodes, mesh.n_cells
    conn, types = [], np.empty(nc, dtype=">i4")
    for c in range(nc):
        t, order = VTK_CELLS[int(mesh.nnode[c])]
        types[c] = t
        conn.append([len(order)] + [int(mesh.el2nd[c, q]) for q in order])
    flat = np.array([x for row in conn for x in row], dtype=">i4")
    with open(path, "wb") as f:
        f.write(f"# vtk DataFile Version 3.0\n{title}\nBINARY\nDATASET UNSTRUCTURED_GRID\n".encode())
        f.write(f"POINTS {nn} double\n".encode())
        f.write(np.ascontiguousarray(mesh.nodes, dtype=">f8").tobytes())
        f.write(f"\nCELLS {nc} {len(flat)}\n".encode())
        f.write(flat.tobytes())
        f.write(f"\nCELL_TYPES {nc}\n".encode())
        f.write(types.tobytes())
        f.write(f"\nPOINT_DATA {nn}\n".encode())
        for name, q in fields.items():
            f.write(f"SCALARS {name} double 1\nLOOKUP_TABLE default\n".encode())
            f.write(np.ascontiguousarray(q, dtype=">f8").tobytes())
            f.write(b"\n")
