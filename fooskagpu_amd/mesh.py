"""Host mesh: box generator and Gambit neutral file reader/writer
(libfooska_mesh.so).  Arrays follow the reference MESH semantics
(mulitphase Solver/doublePrecision/strdata.h:100-115): fc2el, boundCond,
n, area, fcCenter, cellCenter, volume, plus the periodic face shift.
Cells are linear tetrahedra, pyramids, prisms or hexahedra (4, 5, 6 or 8
nodes, el2nd[cell][8] as the reference's elNd2GlbNd, -1 past the element's
nodes)."""
import ctypes

import numpy as np

from . import _lib

BC_SYMMETRY, BC_OUTLET, BC_INLET, BC_SLIP, BC_WALL = 1, 2, 3, 4, 5


# hex element kinds of Mesh.box(kinds=...)
KIND_TETS, KIND_HEX, KIND_PRISMS, KIND_SPLIT = 0, 1, 2, 3
# faces per element, by node count
FACES_OF = {4: 4, 5: 5, 6: 5, 8: 6}


class Mesh:
    """Unstructured mesh; every array is an owned numpy copy."""

    def __init__(self, handle):
        L = _lib.mesh_lib()
        nc, nf, nn, nb = (ctypes.c_int64() for _ in range(4))
        _lib.check_mesh(L.fkm_sizes(handle, nc, nf, nn, nb))
        self.n_cells, self.n_faces, self.n_nodes, self.n_bfaces = nc.value, nf.value, nn.value, nb.value
        self._h = handle

        def arr(name, dtype, shape):
            p = ctypes.c_void_p()
            _lib.check_mesh(L.fkm_array(handle, name.encode(), ctypes.byref(p)))
            n = int(np.prod(shape))
            ct = ctypes.c_int32 if dtype == np.int32 else ctypes.c_double
            buf = (ct * n).from_address(p.value)
            return np.frombuffer(buf, dtype=dtype).reshape(shape).copy()

        F, C = self.n_faces, self.n_cells
        self.fc2el = arr("fc2el", np.int32, (F, 2))
        self.bc = arr("bc", np.int32, (F,))
        self.n = arr("n", np.float64, (F, 3))
        self.area = arr("area", np.float64, (F,))
        self.fc_center = arr("fc_center", np.float64, (F, 3))
        self.shift = arr("shift", np.float64, (F, 3))
        self.cell_center = arr("cell_center", np.float64, (C, 3))
        self.volume = arr("volume", np.float64, (C,))
        self.el2nd = arr("el2nd", np.int32, (C, 8))
        self.nnode = arr("nnode", np.int32, (C,))
        self.nodes = arr("nodes", np.float64, (self.n_nodes, 3))
        self.periodic = bool(np.any(self.shift != 0))

    def tangents(self):
        """Reference tangents t1 (first face edge) and t2 = n x t1."""
        L = _lib.mesh_lib()
        out = []
        for name in ("nt1", "nt2"):
            p = ctypes.c_void_p()
            _lib.check_mesh(L.fkm_array(self._h, name.encode(), ctypes.byref(p)))
            buf = (ctypes.c_double * (3 * self.n_faces)).from_address(p.value)
            out.append(np.frombuffer(buf, dtype=np.float64).reshape(self.n_faces, 3).copy())
        return out

    def close(self):
        if self._h:
            _lib.mesh_lib().fkm_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- construction ----
    ORDERS = {"none": 0, "rcm": 1, "morton": 2}

    @property
    def faces_per_cell(self):
        """Faces of each cell (4 tet, 5 pyramid or prism, 6 hex)."""
        return np.select([self.nnode == 8, self.nnode == 4], [6, 4], 5)

    @classmethod
    def box(cls, nx, ny, nz, lo=(0, 0, 0), hi=(1, 1, 1), jitter=0.0, seed=0, periodic=(False, False, False),
            bc=(BC_OUTLET,) * 6, rcm=True, order=None, kinds=None):
        """order: "rcm" (default), "morton" or "none" (rcm=False).
        kinds: None (every hex split into 6 tets) or an int array [nz][ny][nx]
        of KIND_TETS, KIND_HEX, KIND_PRISMS (two prisms along z) or
        KIND_SPLIT (a centre node with a pyramid on each quadrilateral face
        and two tets on each face that meets triangles); neighbouring hexes
        must present the same face shape, KIND_SPLIT adapts to both."""
        L = _lib.mesh_lib()
        lo_a = np.ascontiguousarray(lo, dtype=np.float64)
        hi_a = np.ascontiguousarray(hi, dtype=np.float64)
        bc_a = np.ascontiguousarray(bc, dtype=np.int32)
        mask = sum(1 << i for i, p in enumerate(periodic) if p)
        h = ctypes.c_void_p()
        o = cls.ORDERS[order] if order is not None else (1 if rcm else 0)
        if kinds is None:
            _lib.check_mesh(L.fkm_box(nx, ny, nz, _lib.dptr(lo_a), _lib.dptr(hi_a), jitter, seed, mask,
                                      _lib.iptr(bc_a), o, ctypes.byref(h)))
        else:
            if mask:
                raise ValueError("mixed element kinds need a box without periodic axes")
            k = np.ascontiguousarray(np.broadcast_to(np.asarray(kinds, dtype=np.int32), (nz, ny, nx)))
            _lib.check_mesh(L.fkm_box_kinds(nx, ny, nz, _lib.dptr(lo_a), _lib.dptr(hi_a), jitter, seed,
                                            _lib.iptr(bc_a), o, _lib.iptr(k), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def mixed_box(cls, nx, ny, nz, lo=(0, 0, 0), hi=(1, 1, 1), jitter=0.0, seed=0, bc=(BC_OUTLET,) * 6, rcm=True,
                  order=None):
        """Every element type of the reference in one conforming box: five
        slabs along x (nx >= 5) of tets | split hexes (pyramids + the tets
        facing the tet slab) | hexes | split hexes (pyramids) | prisms."""
        if nx < 5:
            raise ValueError("mixed_box needs nx >= 5")
        edges = np.linspace(0, nx, 6).round().astype(int)
        slab = np.searchsorted(edges, np.arange(nx), side="right") - 1
        kind_x = np.array([KIND_TETS, KIND_SPLIT, KIND_HEX, KIND_SPLIT, KIND_PRISMS])[slab]
        kinds = np.broadcast_to(kind_x, (nz, ny, nx))
        return cls.box(nx, ny, nz, lo, hi, jitter, seed, bc=bc, rcm=rcm, order=order, kinds=kinds)

    @classmethod
    def read_neu(cls, path, scale=1.0, rcm=True):
        h = ctypes.c_void_p()
        _lib.check_mesh(_lib.mesh_lib().fkm_read_neu(str(path).encode(), scale, 1 if rcm else 0, ctypes.byref(h)))
        return cls(h)

    def write_neu(self, path):
        _lib.check_mesh(_lib.mesh_lib().fkm_write_neu(self._h, str(path).encode()))

    # ---- initial conditions (reference AoS U[cell][NEQ]) ----
    def init_sod(self, gamma, pinf, x_split, left, right):
        U = np.zeros((self.n_cells, 5))
        l = np.ascontiguousarray(left, dtype=np.float64)
        r = np.ascontiguousarray(right, dtype=np.float64)
        _lib.check_mesh(_lib.mesh_lib().fkm_init_sod(self._h, gamma, pinf, x_split, _lib.dptr(l), _lib.dptr(r),
                                                     _lib.dptr(U)))
        return U

    def init_tgv(self, gamma, rho0, U0, mach):
        U = np.zeros((self.n_cells, 5))
        _lib.check_mesh(_lib.mesh_lib().fkm_init_tgv(self._h, gamma, rho0, U0, mach, _lib.dptr(U)))
        return U

    def init_regions(self, eos, regions):
        """eos: dict gamma/pInf/Cp/eta (2 each); regions: list of dicts with
        kind (0 background, 1 sphere, 2 half-space x < c[0]), c, r, alphaL, p, T, u."""
        e = _lib.FkmPhaseEos()
        for k_src, k_dst in (("gamma", "gamma"), ("pInf", "pinf"), ("Cp", "cp"), ("eta", "eta")):
            for i in range(2):
                getattr(e, k_dst)[i] = eos[k_src][i]
        arr = (_lib.FkmRegion * len(regions))()
        for i, rg in enumerate(regions):
            arr[i].kind = rg.get("kind", 0)
            for d in range(3):
                arr[i].c[d] = rg.get("c", (0, 0, 0))[d]
                arr[i].u[d] = rg.get("u", (0, 0, 0))[d]
            arr[i].r = rg.get("r", 0.0)
            arr[i].alphaL, arr[i].p, arr[i].T = rg["alphaL"], rg["p"], rg["T"]
        U = np.zeros((self.n_cells, 10))
        _lib.check_mesh(_lib.mesh_lib().fkm_init_regions(self._h, ctypes.byref(e), arr, len(regions), _lib.dptr(U)))
        return U


def place_bubbles(n, lo, hi, rmin, rmax, xmin, seed):
    c = np.zeros((n, 3))
    r = np.zeros(n)
    lo_a = np.ascontiguousarray(lo, dtype=np.float64)
    hi_a = np.ascontiguousarray(hi, dtype=np.float64)
    _lib.check_mesh(_lib.mesh_lib().fkm_place_bubbles(n, _lib.dptr(lo_a), _lib.dptr(hi_a), rmin, rmax, xmin, seed,
                                                      _lib.dptr(c), _lib.dptr(r)))
    return c, r
