"""Host driver of the relaxation surrogate (the learned part): training-set
assembly from solver-emitted states plus uniform-density states, training
on the GPU with the library's own BF16-MFMA forward/backward kernels, and
weight I/O.  See include/fooska.h (fk_mlp_*) for the device side."""
import ctypes

import numpy as np

from . import _lib

N_PARAMS = 8771
SHAPES = [("W1", (64, 3)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)), ("b3", (64,)),
          ("W4", (3, 64)), ("b4", (3,))]
P_SCALE = 1e4  # pressure target is asinh(p / P_SCALE)


def unpack(w):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(w[o:o + n], dtype=np.float32).reshape(shp)
        o += n
    return out


def init_weights(seed=0):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in SHAPES:
        if name[0] == "W":
            lim = np.sqrt(6.0 / (shp[0] + shp[1]))
            parts.append(rng.uniform(-lim, lim, shp).ravel())
        else:
            parts.append(np.zeros(shp))
    w = np.concatenate(parts).astype(np.float32)
    assert w.size == N_PARAMS
    return w


def fit_norm(x, y):
    """[x mean, x std, y mean, y std] (12 floats)."""
    xs = np.maximum(x.std(0), 1e-6)
    ys = np.maximum(y.std(0), 1e-6)
    return np.concatenate([x.mean(0), xs, y.mean(0), ys]).astype(np.float32)


def finite(x, y):
    """Rows with finite features and targets (the labelling marks states that
    are not chemical-class or did not reach equilibrium with NaN targets)."""
    ok = np.isfinite(x).all(1) & np.isfinite(y).all(1)
    return x[ok], y[ok]


def uniform_states(n, params, seed=0, rho_lo=0.02, rho_hi=1000.0):
    """Pre-relaxation two-phase states with mixture density uniform in
    [rho_lo, rho_hi]: phasic pressures drawn log-uniformly in [1e3, 5e7] Pa
    and temperatures in [280, 380] K independently per phase (out of
    equilibrium), volume fraction from the mixture density, at rest."""
    rng = np.random.default_rng(seed)
    g = np.array(params["gamma"])
    pi = np.array(params["pInf"])
    cp = np.array(params["Cp"])
    eta = np.array(params["eta"])
    cv = cp / g
    am = params["avfMin"]
    out = []
    need = n
    while need > 0:
        m = max(2 * need, 1024)
        rho = rng.uniform(rho_lo, rho_hi, m)
        pL = 10 ** rng.uniform(3, np.log10(5e7), m)
        pG = 10 ** rng.uniform(3, np.log10(5e7), m)
        TL = rng.uniform(280, 380, m)
        TG = TL + rng.uniform(-5, 5, m)
        rL = (pL + pi[0]) / (cv[0] * TL * (g[0] - 1))
        rG = (pG + pi[1]) / (cv[1] * TG * (g[1] - 1))
        a = (rho - rG) / (rL - rG)
        ok = (a > am) & (a < 1 - am) & np.isfinite(a)
        a, rL, rG, pL, pG = a[ok], rL[ok], rG[ok], pL[ok], pG[ok]
        eL = (pL + g[0] * pi[0] + (g[0] - 1) * eta[0] * rL) / (rL * (g[0] - 1))
        eG = (pG + g[1] * pi[1] + (g[1] - 1) * eta[1] * rG) / (rG * (g[1] - 1))
        U = np.zeros((len(a), 10))
        U[:, 0] = a
        U[:, 1] = a * rL
        U[:, 2] = (1 - a) * rG
        U[:, 6] = a * rL * eL
        U[:, 7] = (1 - a) * rG * eG
        U[:, 8] = U[:, 6] + U[:, 7]
        out.append(U[:need])
        need -= len(U[:need])
    return np.concatenate(out)


class Surrogate:
    """MLP state of one GpuSolver context."""

    def __init__(self, solver):
        self.s = solver
        self.L = _lib.hip_lib()

    def set_weights(self, w, norm):
        w = np.ascontiguousarray(w, dtype=np.float32)
        norm = np.ascontiguousarray(norm, dtype=np.float32)
        assert w.size == N_PARAMS and norm.size == 12
        _lib.check_hip(self.L.fk_mlp_set_weights(self.s._h, w.ctypes.data_as(ctypes.c_void_p),
                                                  norm.ctypes.data_as(ctypes.c_void_p)))

    def weights(self):
        w = np.zeros(N_PARAMS, dtype=np.float32)
        norm = np.zeros(12, dtype=np.float32)
        _lib.check_hip(self.L.fk_mlp_get_weights(self.s._h, w.ctypes.data_as(ctypes.c_void_p),
                                                  norm.ctypes.data_as(ctypes.c_void_p)))
        return w, norm

    def set_norm(self, norm):
        norm = np.ascontiguousarray(norm, dtype=np.float32)
        _lib.check_hip(self.L.fk_mlp_set_norm(self.s._h, norm.ctypes.data_as(ctypes.c_void_p)))

    def label(self, U):
        """(features, targets) of states U[n][10] by the device plain Newton."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        n = len(U)
        x = np.zeros((n, 3), dtype=np.float32)
        y = np.zeros((n, 3), dtype=np.float32)
        _lib.check_hip(self.L.fk_mlp_label(self.s._h, _lib.dptr(U), n, x.ctypes.data_as(ctypes.c_void_p),
                                            y.ctypes.data_as(ctypes.c_void_p)))
        return x, y

    def set_data(self, x, y, append=False):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        _lib.check_hip(self.L.fk_mlp_data(self.s._h, x.ctypes.data_as(ctypes.c_void_p),
                                           y.ctypes.data_as(ctypes.c_void_p), len(x), 1 if append else 0))

    def emit(self, stride, capacity):
        _lib.check_hip(self.L.fk_mlp_emit(self.s._h, stride, capacity))

    def data(self):
        n = ctypes.c_int64()
        _lib.check_hip(self.L.fk_mlp_download_data(self.s._h, None, None, 0, ctypes.byref(n)))
        x = np.zeros((n.value, 3), dtype=np.float32)
        y = np.zeros((n.value, 3), dtype=np.float32)
        _lib.check_hip(self.L.fk_mlp_download_data(self.s._h, x.ctypes.data_as(ctypes.c_void_p),
                                                    y.ctypes.data_as(ctypes.c_void_p), n.value, ctypes.byref(n)))
        return x, y

    def grad(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        g = np.zeros(N_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        _lib.check_hip(self.L.fk_mlp_grad(self.s._h, idx.ctypes.data_as(ctypes.c_void_p), len(idx),
                                           g.ctypes.data_as(ctypes.c_void_p), ctypes.byref(loss)))
        return g, loss.value

    def train(self, steps, batch=4096, lr=1e-3, seed=0):
        l0, l1 = ctypes.c_float(), ctypes.c_float()
        _lib.check_hip(self.L.fk_mlp_train(self.s._h, steps, batch, lr, seed, ctypes.byref(l0), ctypes.byref(l1)))
        return l0.value, l1.value

    def infer(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((len(x), 3), dtype=np.float32)
        _lib.check_hip(self.L.fk_mlp_infer(self.s._h, x.ctypes.data_as(ctypes.c_void_p), len(x),
                                            y.ctypes.data_as(ctypes.c_void_p)))
        return y


def train_on_solver(solver, case_U, params, n_solver=4_000_000, n_uniform=1_000_000, warm_steps=3,
                    train_steps=3000, batch=8192, lr=2e-3, seed=0, log=print):
    """Assemble the training set (solver-emitted states from `warm_steps`
    plain steps + uniform-density states), fit the normalisation, train.
    Leaves the solver state at case_U.  Returns (loss_first, loss_last)."""
    sg = Surrogate(solver)
    C = solver.n_cells
    sg.set_weights(init_weights(seed), np.array([0, 0, 0, 1, 1, 1, 0, 0, 0, 1, 1, 1], dtype=np.float32))
    stride = max(1, (C * warm_steps) // max(n_solver, 1))
    sg.set_data(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))
    sg.emit(stride, n_solver)
    solver.upload_state(case_U)
    solver.run_steps(warm_steps)
    sg.emit(0, 0)
    xs, ys = sg.data()
    xu, yu = sg.label(uniform_states(n_uniform, params, seed=seed + 1))
    log(f"surrogate: {len(xs)} solver states (stride {stride}) + {len(xu)} uniform-density states, "
        f"{int(np.isfinite(ys).all(1).sum())} + {int(np.isfinite(yu).all(1).sum())} equilibrium labels")
    x, y = finite(np.concatenate([xs, xu]), np.concatenate([ys, yu]))
    norm = fit_norm(x, y)
    sg.set_data(x, y)
    sg.set_weights(init_weights(seed), norm)
    l0, l1 = sg.train(train_steps, batch=batch, lr=lr, seed=seed)
    solver.upload_state(case_U)
    return sg, (l0, l1, len(x))
