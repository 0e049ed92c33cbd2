"""Host-side mirror of the reference solver driver over the C ABI
(libfooska_hip.so).  Method names follow the reference phase functions
(mulitphase Solver/doublePrecision/strdata.h:231-247) and explicitLoop
(explicitLoop.cu:3-70)."""
import ctypes

import numpy as np

from . import _lib


class GpuSolver:
    """One context on one HIP device: mesh, state and the phase functions."""

    def __init__(self, params, mesh=None, U=None, device=0):
        L = _lib.hip_lib()
        self._p = params if isinstance(params, _lib.FkParams) else _lib.FkParams.from_dict(params)
        h = ctypes.c_void_p()
        _lib.check_hip(L.fk_create(device, ctypes.byref(self._p), ctypes.byref(h)))
        self._h = h
        self.neq = 5 if self._p.model == _lib.MODEL_SINGLE else 10
        self.n_cells = 0
        if mesh is not None:
            self.upload_mesh(mesh)
        if U is not None:
            self.upload_state(U)

    def close(self):
        if getattr(self, "_h", None):
            _lib.hip_lib().fk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_params(self, params):
        self._p = params if isinstance(params, _lib.FkParams) else _lib.FkParams.from_dict(params)
        _lib.check_hip(_lib.hip_lib().fk_set_params(self._h, ctypes.byref(self._p)))

    def upload_mesh(self, m):
        self.mesh = m
        self.n_cells = m.n_cells
        # face slots per cell on the device: 4 on meshes of tetrahedra, else 6
        faces = np.bincount(m.fc2el[:, 0], minlength=m.n_cells) + np.bincount(
            m.fc2el[m.fc2el[:, 1] >= 0, 1], minlength=m.n_cells)
        self.n_slots = 4 if np.all(faces == 4) else 6
        shift = np.ascontiguousarray(m.shift) if m.periodic else None
        self._keep = [np.ascontiguousarray(a) for a in (m.fc2el, m.bc, m.n, m.area, m.fc_center, m.cell_center,
                                                        m.volume)]
        fc2el, bc, n, area, fcc, ccc, vol = self._keep
        _lib.check_hip(_lib.hip_lib().fk_upload_mesh(
            self._h, m.n_cells, m.n_faces, _lib.iptr(fc2el), _lib.iptr(bc), _lib.dptr(n), _lib.dptr(area),
            _lib.dptr(fcc), _lib.dptr(ccc), _lib.dptr(vol), _lib.dptr(shift) if shift is not None else None))

    def upload_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (self.n_cells, self.neq)
        _lib.check_hip(_lib.hip_lib().fk_upload_state(self._h, _lib.dptr(U)))

    def download_state(self):
        U = np.empty((self.n_cells, self.neq))
        _lib.check_hip(_lib.hip_lib().fk_download_state(self._h, _lib.dptr(U)))
        return U

    # ---- reference phase functions ----
    def init_rk(self):
        _lib.check_hip(_lib.hip_lib().fk_init_rk(self._h))

    def compute_face_values(self):
        """computePrimVec + computeGrad + computeFaceValues."""
        _lib.check_hip(_lib.hip_lib().fk_compute_face_values(self._h))

    def compute_rhs(self):
        _lib.check_hip(_lib.hip_lib().fk_compute_rhs(self._h))

    def compute_time_step(self):
        dt = ctypes.c_double()
        _lib.check_hip(_lib.hip_lib().fk_compute_time_step(self._h, ctypes.byref(dt)))
        return dt.value

    def update(self, irk, dt):
        _lib.check_hip(_lib.hip_lib().fk_update(self._h, irk, dt))

    def compute_rhs_update(self, irk, dt):
        _lib.check_hip(_lib.hip_lib().fk_compute_rhs_update(self._h, irk, dt))

    def relaxation(self, mode=_lib.RELAX_PLAIN):
        st = _lib.FkRelaxStats()
        _lib.check_hip(_lib.hip_lib().fk_relaxation(self._h, mode, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def set_tuning(self, name, value):
        _lib.check_hip(_lib.hip_lib().fk_set_tuning(self._h, name.encode(), int(value)))

    def relax_totals(self, reset=False):
        """Relaxation counts summed over fk_run_steps since the last reset."""
        st = _lib.FkRelaxStats()
        _lib.check_hip(_lib.hip_lib().fk_relax_totals(self._h, ctypes.byref(st), 1 if reset else 0))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def cons_calc(self):
        out = np.zeros(self.neq)
        _lib.check_hip(_lib.hip_lib().fk_cons_calc(self._h, _lib.dptr(out)))
        return out

    def health(self, reset=False):
        """(first step, lowest cell) of a non-physical state found by pass G
        since the last reset, or (-1, -1)."""
        st, ce = ctypes.c_int64(), ctypes.c_int64()
        _lib.check_hip(_lib.hip_lib().fk_health(self._h, ctypes.byref(st), ctypes.byref(ce), 1 if reset else 0))
        return st.value, ce.value

    def run_steps(self, nsteps):
        """explicitLoop iterations on the device; returns (last dt, time)."""
        dt, t = ctypes.c_double(), ctypes.c_double()
        _lib.check_hip(_lib.hip_lib().fk_run_steps(self._h, nsteps, ctypes.byref(dt), ctypes.byref(t)))
        return dt.value, t.value

    def clock(self):
        """(dt, physical time, step counter) of the loop."""
        dt, t, st = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        _lib.check_hip(_lib.hip_lib().fk_get_clock(self._h, ctypes.byref(dt), ctypes.byref(t), ctypes.byref(st)))
        return dt.value, t.value, st.value

    def set_clock(self, dt, time, step):
        _lib.check_hip(_lib.hip_lib().fk_set_clock(self._h, dt, time, step))

    def sync(self):
        _lib.check_hip(_lib.hip_lib().fk_sync(self._h))

    # ---- diagnostics ----
    def set_debug(self, on=True):
        _lib.check_hip(_lib.hip_lib().fk_set_debug(self._h, 1 if on else 0))

    def download(self, name):
        C, E = self.n_cells, self.neq
        f32 = self._p.precision == _lib.MIXED
        shapes = {"grad": ((C, E, 3), np.float64), "theta": ((C, E), np.float64),
                  "face": ((C, self.n_slots, E), np.float32 if f32 else np.float64), "rhs": ((C, E), np.float64),
                  "signal": ((C,), np.float64), "slots": ((C, self.n_slots), np.int32)}
        shape, dt = shapes[name]
        out = np.empty(shape, dtype=dt)
        _lib.check_hip(_lib.hip_lib().fk_download(self._h, name.encode(), out.ctypes.data_as(ctypes.c_void_p),
                                                  out.nbytes))
        return out

    def stream_copy(self, nbytes):
        """Duration (ms) of a plain device copy of nbytes (HBM reference)."""
        ms = ctypes.c_double()
        _lib.check_hip(_lib.hip_lib().fk_stream_copy(self._h, nbytes, ctypes.byref(ms)))
        return ms.value

    def set_timing(self, on=True):
        _lib.check_hip(_lib.hip_lib().fk_set_timing(self._h, 1 if on else 0))

    def kernel_times(self):
        """{kernel: (launches, total ms)} since timing was switched on."""
        L = _lib.hip_lib()
        names = (ctypes.c_char * 32 * 64)()
        counts = np.zeros(64, dtype=np.int64)
        ms = np.zeros(64)
        n = _lib.check_hip(L.fk_kernel_times(self._h, 64, ctypes.cast(names, ctypes.c_void_p),
                                             counts.ctypes.data_as(_lib.c_int64_p), _lib.dptr(ms)))
        return {bytes(names[i]).split(b"\0")[0].decode(): (int(counts[i]), float(ms[i])) for i in range(n)}

    # ---- the reference time loop, phase by phase (explicitLoop.cu:28-64) ----
    def explicit_loop_phases(self, nsteps, step0=0):
        p = self._p
        dt = p.dt
        for s in range(nsteps):
            if p.rkSteps > 1:
                self.init_rk()
            for irk in range(p.rkSteps):
                self.compute_face_values()
                self.compute_rhs()
                if p.adaptDt == 1 and irk == 0 and (step0 + s) % p.adaptDtItr == 0:
                    dt = self.compute_time_step()
                self.update(irk, dt)
            if p.model == _lib.MODEL_TWO_PHASE and p.relaxFlag == 1:
                self.relaxation(p.relaxMode)
        return dt
