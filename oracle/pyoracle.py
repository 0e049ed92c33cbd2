"""ctypes wrapper of the CPU FP64 oracle (oracle.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / CPU baseline."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_libs = {}
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def _params_type():
    from fooskagpu_amd._lib import FkParams
    return FkParams


_override = []  # a library loaded by `variant`, used in place of the built one


def lib(omp=False):
    if _override:
        return _override[-1]
    name = "liboracle_omp.so" if omp else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built; run make -C oracle")
    _libs[name] = _load(path)
    return _libs[name]


class variant:
    """Context in which every oracle call (Oracle objects created inside it
    and the pointwise helpers) uses the library at `path`: a build of
    oracle.cpp with a planted mistake, for tests that show a pin catches it."""

    def __init__(self, path):
        self.L = _load(path)

    def __enter__(self):
        _override.append(self.L)
        return self

    def __exit__(self, *exc):
        _override.pop()


def _load(path):
    L = ctypes.CDLL(path)
    P = ctypes.POINTER(_params_type())
    vp = ctypes.c_void_p
    sig = {
        "orc_create": (ctypes.c_int, [P, ctypes.c_int64, ctypes.c_int64, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp,
                                      ctypes.POINTER(vp)]),
        "orc_free": (None, [vp]),
        "orc_set_params": (ctypes.c_int, [vp, P]),
        "orc_set_state": (ctypes.c_int, [vp, _dp]),
        "orc_get_state": (ctypes.c_int, [vp, _dp]),
        "orc_init_rk": (ctypes.c_int, [vp]),
        "orc_compute_face_values": (ctypes.c_int, [vp]),
        "orc_compute_rhs": (ctypes.c_int, [vp]),
        "orc_compute_time_step": (ctypes.c_int, [vp, _dp]),
        "orc_update": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double]),
        "orc_relaxation": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int64)]),
        "orc_cons_calc": (ctypes.c_int, [vp, _dp]),
        "orc_run_steps": (ctypes.c_int, [vp, ctypes.c_int, _dp, _dp]),
        "orc_get": (ctypes.c_int, [vp, ctypes.c_char_p, vp, ctypes.c_int64]),
        "orc_set_rhs": (ctypes.c_int, [vp, _dp]),
        "orc_eqos": (ctypes.c_double, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
        "orc_cons2prim": (None, [P, _dp, _dp]),
        "orc_prim2cons": (None, [P, _dp, _dp]),
        "orc_face_flux": (None, [P, _dp, _dp, _dp, ctypes.c_double, _dp, _dp, ctypes.c_int, _dp]),
        "orc_bc_ghost": (None, [P, ctypes.c_int, _dp, _dp]),
        "orc_wale": (ctypes.c_double, [_dp, ctypes.c_double]),
        "orc_tsat": (ctypes.c_double, [P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_int)]),
        "orc_relax_cell": (ctypes.c_int, [P, _dp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "orc_log10": (ctypes.c_double, [ctypes.c_double]),
        "orc_exp": (ctypes.c_double, [ctypes.c_double]),
        "orc_fnt_press_rlx": (None, [P, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_float),
                                     ctypes.POINTER(ctypes.c_float)]),
        "orc_gibbs": (ctypes.c_double, [P, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
        "orc_threads": (ctypes.c_int, []),
        "orc_chem_trace": (ctypes.c_int, [P, _dp, _dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int)]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype, f.argtypes = r, a
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _chk(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}")


class Oracle:
    """Serial (or OpenMP) FP64 restatement of the reference step on a mesh."""

    def __init__(self, params, mesh, U=None, omp=False):
        self.L = lib(omp)
        FkParams = _params_type()
        self.p = params if isinstance(params, FkParams) else FkParams.from_dict(params)
        self.neq = 5 if self.p.model == 0 else 10
        self.n_cells = mesh.n_cells
        self._keep = [np.ascontiguousarray(a) for a in (mesh.fc2el, mesh.bc, mesh.n, mesh.area, mesh.fc_center,
                                                        mesh.cell_center, mesh.volume, mesh.shift)]
        fc2el, bc, n, area, fcc, ccc, vol, sh = self._keep
        h = ctypes.c_void_p()
        _chk(self.L.orc_create(ctypes.byref(self.p), mesh.n_cells, mesh.n_faces, fc2el.ctypes.data_as(_ip),
                               bc.ctypes.data_as(_ip), _d(n), _d(area), _d(fcc), _d(ccc), _d(vol), _d(sh),
                               ctypes.byref(h)))
        self.h = h
        if U is not None:
            self.set_state(U)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_free(self.h)
            self.h = None

    def set_params(self, params):
        FkParams = _params_type()
        self.p = params if isinstance(params, FkParams) else FkParams.from_dict(params)
        _chk(self.L.orc_set_params(self.h, ctypes.byref(self.p)))

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        _chk(self.L.orc_set_state(self.h, _d(U)))

    def state(self):
        U = np.empty((self.n_cells, self.neq))
        _chk(self.L.orc_get_state(self.h, _d(U)))
        return U

    def init_rk(self):
        _chk(self.L.orc_init_rk(self.h))

    def compute_face_values(self):
        _chk(self.L.orc_compute_face_values(self.h))

    def compute_rhs(self):
        _chk(self.L.orc_compute_rhs(self.h))

    def compute_time_step(self):
        dt = ctypes.c_double()
        _chk(self.L.orc_compute_time_step(self.h, ctypes.byref(dt)))
        return dt.value

    def update(self, irk, dt):
        _chk(self.L.orc_update(self.h, irk, dt))

    def relaxation(self):
        st = (ctypes.c_int64 * 4)()
        _chk(self.L.orc_relaxation(self.h, st))
        return dict(n_therm=st[0], n_chem=st[1], newton_iters=st[2], tsat_iters=st[3])

    def cons_calc(self):
        out = np.zeros(self.neq)
        _chk(self.L.orc_cons_calc(self.h, _d(out)))
        return out

    def run_steps(self, nsteps):
        dt, t = ctypes.c_double(), ctypes.c_double()
        _chk(self.L.orc_run_steps(self.h, nsteps, ctypes.byref(dt), ctypes.byref(t)))
        return dt.value, t.value

    def set_rhs(self, rhs):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        _chk(self.L.orc_set_rhs(self.h, _d(rhs)))

    def nslots(self):
        """Face slots per cell: 4 when every cell is a tetrahedron, else 6."""
        out = np.zeros(1, dtype=np.int32)
        _chk(self.L.orc_get(self.h, b"nslots", out.ctypes.data_as(ctypes.c_void_p), 4))
        return int(out[0])

    def get(self, name):
        C, E = self.n_cells, self.neq
        F = len(self._keep[0])
        S = self.nslots()
        shapes = {"grad": ((C, E, 3), np.float64), "theta": ((C, E), np.float64), "wmin": ((C, E), np.float64),
                  "wmax": ((C, E), np.float64), "primF": ((F, 2, E), np.float64), "face": ((C, S, E), np.float64),
                  "rhs": ((C, E), np.float64), "signal": ((C,), np.float64), "slots": ((C, S), np.int32)}
        shape, dt = shapes[name]
        out = np.empty(shape, dtype=dt)
        _chk(self.L.orc_get(self.h, name.encode(), out.ctypes.data_as(ctypes.c_void_p), out.nbytes))
        return out


# ---- pointwise helpers ----
def _P(params):
    FkParams = _params_type()
    return params if isinstance(params, FkParams) else FkParams.from_dict(params)


def eqos(params, imat, idv, v1, v2):
    return lib().orc_eqos(ctypes.byref(_P(params)), imat, idv, v1, v2)


def cons2prim(params, U):
    p = _P(params)
    neq = 5 if p.model == 0 else 10
    U = np.ascontiguousarray(U, dtype=np.float64)
    W = np.zeros(neq)
    lib().orc_cons2prim(ctypes.byref(p), _d(U), _d(W))
    return W


def prim2cons(params, W):
    p = _P(params)
    neq = 5 if p.model == 0 else 10
    W = np.ascontiguousarray(W, dtype=np.float64)
    U = np.zeros(neq)
    lib().orc_prim2cons(ctypes.byref(p), _d(W), _d(U))
    return U


def face_flux(params, WL, WR, n, A, UL, UR, bc=0):
    p = _P(params)
    neq = 5 if p.model == 0 else 10
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (WL, WR, n, UL, UR)]
    out = np.zeros(2 * neq)
    lib().orc_face_flux(ctypes.byref(p), _d(arrs[0]), _d(arrs[1]), _d(arrs[2]), A, _d(arrs[3]), _d(arrs[4]), bc,
                        _d(out))
    return out.reshape(2, neq)


def bc_ghost(params, bc, n, W):
    p = _P(params)
    W = np.array(W, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.float64)
    lib().orc_bc_ghost(ctypes.byref(p), bc, _d(n), _d(W))
    return W


def wale(g, vol):
    g = np.ascontiguousarray(g, dtype=np.float64).reshape(9)
    return lib().orc_wale(_d(g), vol)


def tsat(params, P, T0, urf=0.8):
    it = ctypes.c_int()
    T = lib().orc_tsat(ctypes.byref(_P(params)), P, T0, urf, ctypes.byref(it))
    return T, it.value


def log10(x):
    """Portable log10 shared by the oracle and the device relaxation."""
    return lib().orc_log10(float(x))


def exp(x):
    """Portable exp shared by the oracle and the device relaxation."""
    return lib().orc_exp(float(x))


def fnt_press_rlx(params, avf, rho, ener):
    """Finite-rate relaxation targets (dP, dT) as FP32 values."""
    dP, dT = ctypes.c_float(), ctypes.c_float()
    a, r, e = (np.ascontiguousarray(v, dtype=np.float64) for v in (avf, rho, ener))
    lib().orc_fnt_press_rlx(ctypes.byref(_P(params)), _d(a), _d(r), _d(e), ctypes.byref(dP), ctypes.byref(dT))
    return np.float32(dP.value), np.float32(dT.value)


def gibbs(params, k, P, T):
    return lib().orc_gibbs(ctypes.byref(_P(params)), k, P, T)


def relax_cell(params, U):
    U = np.array(U, dtype=np.float64)
    it, ts = ctypes.c_int(), ctypes.c_int()
    kind = lib().orc_relax_cell(ctypes.byref(_P(params)), _d(U), ctypes.byref(it), ctypes.byref(ts))
    return U, kind, it.value, ts.value


CHEM_TRACE_COLS = 20


def chem_trace(params, U, max_rows=256):
    """Chemical Newton solve of one cell with its iterations recorded:
    returns (relaxed U, iterations, rows[n][18], aborted); row layout as
    kChemTraceCols in oracle.cpp (x before the update, residuals, abortCheck
    flag 0 not eligible / 1 eligible / 2 new best, FP32 residual norm, x after
    the update, abort flag, the non-primary phase's volume fraction and density
    after the full step, before checkUpdate)."""
    U = np.array(U, dtype=np.float64)
    rows = np.zeros((max_rows, CHEM_TRACE_COLS))
    n, ab = ctypes.c_int(), ctypes.c_int()
    it = lib().orc_chem_trace(ctypes.byref(_P(params)), _d(U), _d(rows), max_rows, ctypes.byref(n), ctypes.byref(ab))
    return U, it, rows[:n.value].copy(), bool(ab.value)


def threads(omp=True):
    """omp_get_max_threads() of the OpenMP build (1 for the serial one)."""
    return lib(omp).orc_threads()
