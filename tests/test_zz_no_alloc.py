"""fk_cons_calc_noalloc: the volume integrals without any allocation in the
call (partials in a fixed device array), through the C ABI, against
fk_cons_calc on the same state: bitwise equal, on two calls in a row, the
second on a smaller mesh, so a stale partial of the first call would show."""
import ctypes

import numpy as np
import pytest

from fooskagpu_amd import GpuSolver, _lib, cases

pytestmark = pytest.mark.gpu


def _noalloc(g):
    out = np.full(g.neq, np.nan)
    _lib.check_hip(_lib.hip_lib().fk_cons_calc_noalloc(g._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def test_cons_calc_noalloc_equals_cons_calc(gpu):
    big = cases.shock_bubble(n=12, nbub=3)   # 10,368 cells, two-phase
    small = cases.bubble(n=6)                # 1,296 cells
    sod = cases.sod(nx=8, ny=2, nz=2)        # 192 cells, single-phase
    for c, steps in ((big, 2), (small, 1), (sod, 3)):
        with GpuSolver(c.params, c.mesh, c.U) as g:
            g.run_steps(steps)
            a = _noalloc(g)
            b = g.cons_calc()
            print(f"[no_alloc] {c.name} {c.mesh.n_cells} cells: {a}", flush=True)
            assert np.isfinite(a).all()
            assert np.array_equal(a, b)
            # a second call on the same state gives the same bits
            assert np.array_equal(_noalloc(g), b)
