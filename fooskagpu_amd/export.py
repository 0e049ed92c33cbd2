"""Field export for validation plots: node-averaged variables written as
legacy VTK (binary) or Tecplot ASCII.

Follows export_plt (mulitphase Solver/doublePrecision/export_plt.cu:41-93):
every node gets the arithmetic mean over the cells that share it of
  alpha_L, alpha_G, rho, u, v, w, p = sum_k alpha_k p_k, T = sum_k alpha_k T_k,
  rho Y_v (consVec[9]) and |grad rho| (Green-Gauss gradient of the mixture
  density, grad row 9 in the reference),
and the single-phase model exports rho, u, v, w, p, T.  Host-side numpy: the
export is I/O, not part of the time step.
"""
import numpy as np

TWO_PHASE_VARS = ["alphaL", "alphaG", "rho", "u", "v", "w", "p", "T", "rhoYv", "gradRho"]
SINGLE_PHASE_VARS = ["rho", "u", "v", "w", "p", "T"]


def _eos(params, k, rho, e):
    g, pi, cp, eta = (params[n][k] for n in ("gamma", "pInf", "Cp", "eta"))
    p = (g - 1) * rho * e - g * pi - (g - 1) * eta * rho
    T = (p + pi) / ((cp / g) * (g - 1) * rho)
    return p, T


def green_gauss(mesh, q):
    """Cell gradients of the cell field q by the face-average rule
    (computeGrad.cu:51-91), boundary faces taking the cell's own value."""
    f2 = mesh.fc2el
    o, nb = f2[:, 0], f2[:, 1]
    qn = np.where(nb >= 0, q[np.maximum(nb, 0)], q[o])
    flux = (0.5 * (q[o] + qn) * mesh.area)[:, None] * mesh.n
    g = np.zeros((mesh.n_cells, 3))
    np.add.at(g, o, flux)
    inter = nb >= 0
    np.add.at(g, nb[inter], -flux[inter])
    return g / mesh.volume[:, None]


def cell_fields(mesh, U, params):
    """Cell values of the exported variables (dict name -> [C])."""
    U = np.asarray(U, dtype=np.float64)
    if U.shape[1] == 5:
        rho = U[:, 0]
        u, v, w = (U[:, i] / rho for i in (1, 2, 3))
        e = U[:, 4] / rho - 0.5 * (u * u + v * v + w * w)
        p, T = _eos(params, 0, rho, e)
        return dict(rho=rho, u=u, v=v, w=w, p=p, T=T)
    am = params["avfMin"] / 100
    aL = np.clip(U[:, 0], am, 1 - am)
    aG = 1 - aL
    rL, rG = U[:, 1] / aL, U[:, 2] / aG
    rho = aL * rL + aG * rG
    u, v, w = (U[:, i] / rho for i in (3, 4, 5))
    pL, TL = _eos(params, 0, rL, U[:, 6] / U[:, 1])
    pG, TG = _eos(params, 1, rG, U[:, 7] / U[:, 2])
    grad = green_gauss(mesh, rho)
    return dict(alphaL=aL, alphaG=aG, rho=rho, u=u, v=v, w=w, p=aL * pL + aG * pG, T=aL * TL + aG * TG,
                rhoYv=U[:, 9], gradRho=np.sqrt((grad * grad).sum(1)))


def node_average(mesh, cell_values):
    """Mean over the cells sharing each node (cells x 4 nodes, el2nd)."""
    nodes = mesh.el2nd.ravel()
    cnt = np.bincount(nodes, minlength=mesh.n_nodes).astype(np.float64)
    out = {}
    for name, q in cell_values.items():
        s = np.bincount(nodes, weights=np.repeat(q, 4), minlength=mesh.n_nodes)
        out[name] = s / np.maximum(cnt, 1)
    return out


def node_fields(mesh, U, params):
    return node_average(mesh, cell_fields(mesh, U, params))


def write_vtk(path, mesh, fields, title="fooskagpu_amd"):
    """Legacy VTK unstructured grid, binary (big endian), tetrahedra (type 10),
    point data = `fields` (name -> [n_nodes])."""
    nn, nc = mesh.n_nodes, mesh.n_cells
    with open(path, "wb") as f:
        f.write(f"# vtk DataFile Version 3.0\n{title}\nBINARY\nDATASET UNSTRUCTURED_GRID\n".encode())
        f.write(f"POINTS {nn} double\n".encode())
        f.write(np.ascontiguousarray(mesh.nodes, dtype=">f8").tobytes())
        f.write(f"\nCELLS {nc} {nc * 5}\n".encode())
        conn = np.empty((nc, 5), dtype=">i4")
        conn[:, 0] = 4
        conn[:, 1:] = mesh.el2nd
        f.write(conn.tobytes())
        f.write(f"\nCELL_TYPES {nc}\n".encode())
        f.write(np.full(nc, 10, dtype=">i4").tobytes())
        f.write(f"\nPOINT_DATA {nn}\n".encode())
        for name, q in fields.items():
            f.write(f"SCALARS {name} double 1\nLOOKUP_TABLE default\n".encode())
            f.write(np.ascontiguousarray(q, dtype=">f8").tobytes())
            f.write(b"\n")


def write_tecplot(path, mesh, fields, time=0.0, title="fooskagpu_amd"):
    """Tecplot ASCII, one FETETRAHEDRON zone in POINT packing (the reference
    writes FEBRICK with repeated nodes, export_plt.cu:4-13, 95-174)."""
    names = list(fields)
    data = np.column_stack([mesh.nodes] + [np.asarray(fields[n]) for n in names])
    with open(path, "w") as f:
        f.write(f'TITLE = "{title}"\n')
        f.write("VARIABLES = " + ", ".join(f'"{n}"' for n in ["x", "y", "z"] + names) + "\n")
        f.write(f"ZONE N={mesh.n_nodes}, E={mesh.n_cells}, DATAPACKING=POINT, ZONETYPE=FETETRAHEDRON, "
                f"SOLUTIONTIME={time:.10e}\n")
        np.savetxt(f, data, fmt="%.10e")
        np.savetxt(f, mesh.el2nd + 1, fmt="%d")
