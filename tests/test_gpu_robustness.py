"""Failure handling of the device path, forced from the host: states that are
not physical, a surrogate with garbage weights, Newton solves cut to one
iteration.  Every kernel still runs to completion; the library reports."""
import numpy as np
import pytest

from fooskagpu_amd import GpuSolver, _lib, cases, surrogate
from fooskagpu_amd._lib import FkError
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def rel_l1(a, b):
    den = np.abs(b).sum(0)
    return np.abs(a - b).sum(0) / np.where(den > 0, den, 1.0)


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED])
def test_healthy_run_reports_no_fault(gpu, precision):
    c = cases.shock_bubble(n=10, nbub=3)
    with GpuSolver(dict(c.params, precision=precision), c.mesh, c.U) as g:
        g.run_steps(3)
        assert g.health() == (-1, -1)


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED])
@pytest.mark.parametrize("fault", ["nan", "negative_partial_density"])
def test_non_physical_state_is_reported_with_step_and_cell(gpu, precision, fault):
    """A NaN or a negative liquid partial density planted in one cell of the
    uploaded state: the next pass G records step 0 and that cell;
    fk_run_steps completes its steps and returns FK_ERANGE naming step 0 (a
    NaN reaches the neighbours within the step, so the lowest failing cell
    of the step is at most the planted one)."""
    c = cases.bubble(n=8)
    U = c.U.copy()
    k = c.mesh.n_cells // 3
    if fault == "nan":
        U[k, 6] = np.nan  # liquid internal energy -> pressure
    else:
        U[k, 1] = -abs(U[k, 1])
    with GpuSolver(dict(c.params, precision=precision), c.mesh, U) as g:
        g.compute_face_values()
        assert g.health() == (0, k)
        with pytest.raises(FkError, match="non-physical .* from step 0 on"):
            g.run_steps(2)
        step, cell = g.health()
        assert step == 0 and (cell == k if fault != "nan" else 0 <= cell <= k)
        assert g.clock()[2] == 2  # the steps ran
        g.health(reset=True)
        assert g.health() == (-1, -1)


def test_single_phase_negative_density_is_reported(gpu):
    c = cases.sod(nx=8, ny=2, nz=2)
    U = c.U.copy()
    U[5, 0] = -1.0
    with GpuSolver(c.params, c.mesh, U) as g:
        with pytest.raises(FkError, match="non-physical"):
            g.run_steps(1)
        step, cell = g.health()
        assert step == 0 and 0 <= cell <= 5  # its NaN fluxes reach the neighbours within the step


def test_positivity_loss_of_the_bubble_collapse_is_caught(gpu):
    """The 48^3 bubble collapse loses positivity at step 54 in the reference
    scheme (a negative liquid partial density where the liquid converges on
    the collapsing vapour; the oracle shows the same): the run reports it."""
    c = cases.bubble(n=48)
    with GpuSolver(c.params, c.mesh, c.U) as g:
        g.run_steps(50)
        assert g.health() == (-1, -1)
        with pytest.raises(FkError, match="non-physical") as e:
            g.run_steps(10)
        step, cell = g.health()
        print(f"[bubble 48] {e.value}", flush=True)
        assert 50 <= step < 60 and cell >= 0


@pytest.mark.parametrize("garbage", ["nan", "huge"])
def test_garbage_surrogate_falls_back_to_the_plain_solve(gpu, garbage):
    """Weights that make every prediction useless (NaN, or outputs of order
    1e30): every chemical cell falls back to the reference solve from the cell
    state, the relaxation stays finite and agrees with the plain relaxation
    (thermal cells by the analytic Newton, to 1e-6 in every variable)."""
    from test_gpu_parity import cavitation
    c = cavitation(10)
    p = dict(c.params, relaxFlag=0)
    with GpuSolver(p, c.mesh, c.U) as g:
        g.run_steps(2)
        U = g.download_state()
        sg = surrogate.Surrogate(g)
        w = surrogate.init_weights(0)
        if garbage == "nan":
            w[:] = np.nan
        else:
            w[-3:] = 1e30  # output biases
        sg.set_weights(w, np.array([0, 0, 0, 1, 1, 1, 0, 0, 0, 1, 1, 1], dtype=np.float32))
        st_plain = g.relaxation(_lib.RELAX_PLAIN)
        Up = g.download_state()
        g.upload_state(U)
        st = g.relaxation(_lib.RELAX_MLP)
        Um = g.download_state()
    print(f"plain {st_plain}\ngarbage {garbage} {st}", flush=True)
    assert np.isfinite(Um).all()
    assert st["n_chem"] == st_plain["n_chem"] > 0
    assert st["n_warm"] == st["n_chem"] and st["n_fallback"] >= st["n_chem"]
    assert rel_l1(Um, Up).max() < 1e-6


def test_one_iteration_newton_solves_match_oracle(gpu):
    """thermItrMax = chemItrMax = 1: every thermal and chemical Newton solve
    ends on the reference's abort path (thermRlx.cu:142-170,
    chemRlx.cu:187-192, restore the best iterate); FP64 against the oracle
    (bitwise)."""
    from test_gpu_parity import cavitation
    c = cavitation(10)
    p = dict(c.params, thermItrMax=1, chemItrMax=1)
    o = Oracle(p, c.mesh, c.U)
    o.run_steps(2)
    Uo = o.state()
    with GpuSolver(p, c.mesh, c.U) as g:
        g.run_steps(2)
        Ug = g.download_state()
    err = rel_l1(Ug, Uo)
    print(f"itrMax 1: rel L1 per variable {np.array2string(err, precision=2)}", flush=True)
    assert np.isfinite(Ug).all() and err.max() <= 1e-12
