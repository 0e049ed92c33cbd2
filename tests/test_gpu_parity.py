"""GPU parity: the HIP path (through the C ABI) against the CPU FP64 oracle
on the same seeded inputs.

Bars: FP64 path <= 1e-12 relative L1 per conserved variable
after N steps; mixed path <= 1e-5 relative L1 in density.  Phase-level
intermediates of the FP64 path are expected to agree to rounding (the
arithmetic order is the same; only libm transcendentals in the relaxation
differ between glibc and the device library)."""
import numpy as np
import pytest

from fooskagpu_amd import GpuSolver, _lib, cases
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_RHO_TOL = 1e-5  # density, the bar of the mixed path (both models)


def rel_l1(a, b):
    """Relative L1 error per column: sum|a-b| / sum|b|."""
    a = np.asarray(a, dtype=np.float64).reshape(len(a), -1)
    b = np.asarray(b, dtype=np.float64).reshape(len(b), -1)
    den = np.abs(b).sum(0)
    return np.abs(a - b).sum(0) / np.where(den > 0, den, 1.0)


def cavitation(n=10):
    """Mixture cells with alpha_L = 0.5 in liquid superheated with respect to
    Tsat(1500 Pa) ~ 286 K: exercises the chemical (p-T-g) relaxation.  The
    reference chemical relaxation is not robust for this state beyond a few
    steps (the oracle itself reaches alpha_L = 1e-6 and rho Y_v ~ 1e4 at step
    4), so runs of this case are limited to STEPS["cavitation"]."""
    c = cases.bubble(n=n)
    p = c.params
    regions = [dict(kind=0, alphaL=1 - p["avfMin"], p=1e5, T=300.0),
               dict(kind=1, c=(0.5, 0.5, 0.5), r=0.3, alphaL=0.5, p=1500.0, T=300.0)]
    c.U = c.mesh.init_regions(p, regions)
    c.name = "cavitation"
    return c


def mixed_elements(model):
    """Every element type of the reference in one box (tets | split hexes:
    pyramids and tets | hexahedra | split hexes | prisms, Mesh.mixed_box),
    jittered.  Two-phase: a vapour bubble across the element slabs, WALE on;
    single-phase: a Sod tube along x with viscous terms and the heat flux."""
    from fooskagpu_amd import Mesh
    from fooskagpu_amd.mesh import BC_OUTLET, BC_SLIP, BC_WALL
    m = Mesh.mixed_box(10, 4, 4, (0, 0, 0), (1.0, 0.4, 0.4), 0.10, 11, bc=(BC_OUTLET, BC_OUTLET, BC_SLIP, BC_WALL,
                                                                        BC_SLIP, BC_SLIP))
    if model == _lib.MODEL_SINGLE:
        p = dict(cases.BASE, model=model, viscFlag=1, turbulence=1, heatFlux=1, dynVisc=[1e-3, 1e-3])
        U = m.init_sod(1.4, 0.0, 0.45, (1.0, 0.2, 0, 0, 1.0), (0.125, 0, 0, 0, 0.1))
        return cases.Case("mixed_elements_sp", p, m, U, 10)
    p = cases._two_phase_params(viscFlag=1, turbulence=1)
    am = p["avfMin"]
    regions = [dict(kind=0, alphaL=1 - am, p=1e7, T=cases.T_WATER, u=(3.0, 0, 0)),
               dict(kind=1, c=(0.5, 0.2, 0.2), r=0.15, alphaL=am, p=2340.0, T=cases.T_WATER)]
    return cases.Case("mixed_elements_tp", p, m, m.init_regions(p, regions), 10)


STEPS = {"cavitation": 2}

SMALL = {
    "sod": lambda: cases.sod(nx=16, ny=4, nz=4),
    "tgv": lambda: cases.tgv(n=8),
    "bubble": lambda: cases.bubble(n=10),
    "shock_bubble": lambda: cases.shock_bubble(n=10, nbub=3),
    "cavitation": lambda: cavitation(10),
    "mixed_elements_sp": lambda: mixed_elements(_lib.MODEL_SINGLE),
    "mixed_elements_tp": lambda: mixed_elements(_lib.MODEL_TWO_PHASE),
}


def neq_density(case):
    """Columns holding density: rho (single-phase) or alpha_k rho_k (two-phase)."""
    return [0] if case.neq == 5 else [1, 2]


@pytest.fixture(params=sorted(SMALL))
def case(request):
    c = SMALL[request.param]()
    print(f"[case] {c.name} cells={c.mesh.n_cells}", flush=True)
    return c


def test_face_values_and_rhs_match_oracle(gpu, case):
    """Pass G (grad, theta, MUSCL face states) and pass F (residual, signal)
    on the initial state, FP64."""
    o = Oracle(case.params, case.mesh, case.U)
    with GpuSolver(case.params, case.mesh, case.U) as g:
        g.set_debug(True)
        g.compute_face_values()
        o.compute_face_values()
        slots_g, slots_o = g.download("slots"), o.get("slots")
        assert np.array_equal(slots_g, slots_o)
        for name in ("grad", "theta", "face"):
            a, b = g.download(name), o.get(name)
            if name == "face":  # slots past a cell's last face hold no state
                a, b = a[slots_o >= 0], b[slots_o >= 0]
            err = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
            print(f"[{case.name}] {name}: max rel diff {err:.3e} bitwise={np.array_equal(a, b)}", flush=True)
            assert err <= 1e-13, name
        g.compute_rhs()
        o.compute_rhs()
        for name in ("rhs", "signal"):
            a, b = g.download(name), o.get(name)
            e = rel_l1(a.reshape(len(a), -1), b.reshape(len(b), -1)).max()
            print(f"[{case.name}] {name}: rel L1 {e:.3e} bitwise={np.array_equal(a, b)}", flush=True)
            assert e <= 1e-13, name
        dt_g, dt_o = g.compute_time_step(), o.compute_time_step()
        assert abs(dt_g - dt_o) <= 1e-14 * dt_o


def test_fp64_steps_match_oracle(gpu, case):
    steps = STEPS.get(case.name, 10)
    o = Oracle(case.params, case.mesh, case.U)
    o.run_steps(steps)
    with GpuSolver(case.params, case.mesh, case.U) as g:
        g.run_steps(steps)
        Ug = g.download_state()
    Uo = o.state()
    assert np.isfinite(Ug).all()
    err = rel_l1(Ug, Uo)
    print(f"[{case.name}] FP64 {steps} steps rel L1 per variable: {np.array2string(err, precision=2)}", flush=True)
    assert err.max() <= FP64_TOL


def test_phase_api_equals_fused_loop(gpu, case):
    """explicitLoop phase by phase (fk_compute_rhs + fk_update) and the fused
    device loop (fk_run_steps) give bitwise identical states."""
    k = min(4, STEPS.get(case.name, 4))
    with GpuSolver(case.params, case.mesh, case.U) as a, GpuSolver(case.params, case.mesh, case.U) as b:
        a.explicit_loop_phases(k)
        b.run_steps(k)
        assert np.array_equal(a.download_state(), b.download_state())


def test_phase_api_mixed_close_to_fused_loop(gpu, case):
    """Mixed precision: fk_compute_rhs stores the residual in FP32 and
    fk_update promotes it: the first RK stage is exactly U0 + dt (1/V) r with
    the downloaded (widened) r.  fk_run_steps keeps the residual in registers
    at the fused stages, so the two loops differ only by that rounding."""
    k = min(4, STEPS.get(case.name, 4))
    p = dict(case.params, precision=_lib.MIXED)
    dt = 1e-7
    with GpuSolver(p, case.mesh, case.U) as a, GpuSolver(p, case.mesh, case.U) as b:
        a.init_rk()
        a.compute_face_values()
        a.compute_rhs()
        r = a.download("rhs")
        a.update(0, dt)
        U1 = a.download_state()
        a.upload_state(case.U)
        a.explicit_loop_phases(k)
        b.run_steps(k)
        Ua, Ub = a.download_state(), b.download_state()
    assert np.isfinite(r).all() and np.array_equal(r, r.astype(np.float32).astype(np.float64))
    expect = case.U + (dt * (1.0 / case.mesh.volume))[:, None] * r
    if case.params.get("neglectZ", 0) == 1:
        expect[:, 3 if case.neq == 5 else 5] = 0.0
    assert np.array_equal(U1, expect)
    # density and total energy (well conditioned; momenta of flows near rest
    # are not, relative to their own size)
    ce = neq_density(case) + [4 if case.neq == 5 else 8]
    d = rel_l1(Ua[:, ce], Ub[:, ce])
    print(f"[{case.name}] mixed phase vs fused rel L1 per variable {np.array2string(rel_l1(Ua, Ub), precision=1)}",
          flush=True)
    assert d.max() <= 1e-6  # 12 FP32-rounded residuals; the mixed bar is 1e-5


def test_repeat_runs_bitwise_identical(gpu, case):
    out = []
    for _ in range(2):
        with GpuSolver(case.params, case.mesh, case.U) as g:
            g.run_steps(min(5, STEPS.get(case.name, 5)))
            out.append(g.download_state())
    assert np.array_equal(out[0], out[1])


def test_mixed_precision_density_error(gpu, case):
    steps = STEPS.get(case.name, 10)
    o = Oracle(case.params, case.mesh, case.U)
    o.run_steps(steps)
    Uo = o.state()
    p = dict(case.params, precision=_lib.MIXED)
    with GpuSolver(p, case.mesh, case.U) as g:
        g.run_steps(steps)
        Ug = g.download_state()
    assert np.isfinite(Ug).all()
    err = rel_l1(Ug, Uo)
    cols = neq_density(case)
    rho_err = rel_l1(Ug[:, cols].sum(1, keepdims=True), Uo[:, cols].sum(1, keepdims=True))[0]
    print(f"[{case.name}] mixed {steps} steps: density rel L1 {rho_err:.2e}; per variable "
          f"{np.array2string(err, precision=2)}", flush=True)
    assert rho_err <= MIXED_RHO_TOL


def test_relaxation_matches_oracle(gpu):
    """Plain Newton relaxation (partition + thermal / chemical) on a hydro-
    evolved cavitation state."""
    c = cavitation(10)
    p = dict(c.params, relaxFlag=0)
    o = Oracle(p, c.mesh, c.U)
    o.run_steps(3)
    U = o.state()
    st_o = o.relaxation()
    with GpuSolver(c.params, c.mesh, U) as g:
        st_g = g.relaxation()
        Ug = g.download_state()
    Uo = o.state()
    print(f"oracle {st_o} gpu {st_g}", flush=True)
    assert st_o["n_chem"] > 0
    assert st_g["n_chem"] == st_o["n_chem"]
    err = rel_l1(Ug, Uo)
    assert err.max() <= FP64_TOL, err


def test_cons_calc_matches_oracle(gpu):
    c = SMALL["bubble"]()
    o = Oracle(c.params, c.mesh, c.U)
    with GpuSolver(c.params, c.mesh, c.U) as g:
        a = g.cons_calc()
    b = o.cons_calc()
    assert np.allclose(a, b, rtol=1e-12, atol=0)


@pytest.mark.parametrize("name", ["shock_bubble", "tgv", "sod", "mixed_elements_tp", "mixed_elements_sp"])
@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED])
def test_tuning_options_do_not_change_results(gpu, precision, name):
    """fuse_update (RK update inside pass F or residual + separate update),
    stage_f (LDS staging of face states in pass F), prims_w (pass G reads
    FP32 primitive records written by the producers of U) and stage_g (pass G
    writes its face records through LDS) are implementation choices: bitwise
    equal states after several steps (two-phase with relaxation,
    single-phase viscous, single-phase inviscid, and both models on a mesh
    of every element type, 6 slots per cell)."""
    c = SMALL[name]()
    p = dict(c.params, precision=precision)
    out = []
    for fuse, stage_f, prims_w, stage_g in ((0, 0, 0, 0), (1, 0, 1, 1), (0, 1, 1, 0), (1, 1, 0, 1), (1, 1, 1, 1),
                                            (1, 1, 1, 0)):
        with GpuSolver(p, c.mesh, c.U) as g:
            g.set_tuning("fuse_update", fuse)
            g.set_tuning("stage_f", stage_f)
            g.set_tuning("prims_w", prims_w)
            g.set_tuning("stage_g", stage_g)
            g.run_steps(3)
            out.append(g.download_state())
    for o in out[1:]:
        assert np.array_equal(o, out[0])


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED])
@pytest.mark.parametrize("n", [10, 17])
def test_pass_g_face_records_through_lds_are_bitwise(gpu, n, precision):
    """stage_g: pass G writes the face records and gradient rows of its
    workgroup (256 cells, FP64 128) from LDS in 16-byte words; face states,
    gradients and limiter equal the direct stores bitwise (6000 and 29478
    cells: the last workgroup is partial)."""
    c = cases.shock_bubble(n=n, nbub=3)
    p = dict(c.params, precision=precision)
    out = []
    for sg in (0, 1):
        with GpuSolver(p, c.mesh, c.U) as g:
            g.set_tuning("stage_g", sg)
            g.run_steps(1)
            g.set_debug(True)
            g.compute_face_values()
            out.append([g.download(k) for k in ("face", "grad", "theta")])
    assert c.mesh.n_cells % 256 != 0 and c.mesh.n_cells % 128 != 0
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED])
@pytest.mark.parametrize("model", [_lib.MODEL_TWO_PHASE, _lib.MODEL_SINGLE])
def test_pass_g_face_records_through_lds_six_slots(gpu, model, precision):
    """stage_g on a mesh of every element type (6 slots per cell): the LDS
    store path (two-phase FP32, workgroups of 128 cells) equals the direct
    stores bitwise, including the empty slots of cells with fewer than six
    faces and the partial last workgroup; the other layouts keep the direct
    stores either way."""
    c = mixed_elements(model)
    p = dict(c.params, precision=precision)
    out = []
    for sg in (0, 1):
        with GpuSolver(p, c.mesh, c.U) as g:
            g.set_tuning("stage_g", sg)
            g.run_steps(2)
            g.set_debug(True)
            g.compute_face_values()
            out.append([g.download(k) for k in ("face", "grad", "theta")] + [g.download_state()])
    assert c.mesh.n_cells % 64 != 0
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("name,mode", [("cavitation", _lib.RELAX_PLAIN), ("cavitation", _lib.RELAX_COLD),
                                       ("shock_bubble", _lib.RELAX_COLD)])
def test_primitive_records_of_relaxation(gpu, name, mode):
    """Mixed precision: the relaxation kernels (plain: thermal + chemical
    list; surrogate screen + list) write the FP32 primitive records that the
    next pass G reads; the run is bitwise that of pass G converting U itself."""
    c = SMALL[name]()
    p = dict(c.params, precision=_lib.MIXED, relaxMode=mode)
    steps = STEPS.get(name, 3)
    out = []
    for prims_w in (0, 1):
        with GpuSolver(p, c.mesh, c.U) as g:
            g.set_tuning("prims_w", prims_w)
            g.run_steps(steps)
            out.append(g.download_state())
            st = g.relax_totals()
    print(name, mode, st)
    assert np.isfinite(out[0]).all()
    assert np.array_equal(out[1], out[0])


@pytest.mark.parametrize("bad", [0, -2])
def test_boundary_face_without_condition_is_rejected(gpu, bad):
    """A boundary face (no neighbour) needs a condition code >= 1: the slot
    link stores -bc there, and 0 or a negative code would read as an
    interior link."""
    from fooskagpu_amd._lib import FkError
    c = SMALL["sod"]()  # a fresh mesh: its host arrays can be edited
    bnd = np.flatnonzero(c.mesh.fc2el[:, 1] < 0)
    assert len(bnd) > 0
    c.mesh.bc[bnd[0]] = bad
    with pytest.raises(FkError, match="boundCond"):
        GpuSolver(c.params, c.mesh, c.U)


# ---------------- finite-rate relaxation and the inlet state ----------------
FNT = dict(thermFntRlxFlag=1, thermRlxRate=0.5, dt=1e-9)  # exp(-kappa dt) spans (0, 1) over the cells


@pytest.mark.parametrize("name", ["bubble", "shock_bubble"])
def test_finite_rate_relaxation_matches_oracle(gpu, name):
    """fntPressRlx targets in the thermal Newton solve (thermRlx.cu:42-67,
    213-214): FP64 path against the oracle after 10 steps, and the model
    changes the state well above that bar (so the comparison tests it)."""
    c = SMALL[name]()
    p = dict(c.params, **FNT)
    o = Oracle(p, c.mesh, c.U)
    o.run_steps(10)
    Uo = o.state()
    with GpuSolver(p, c.mesh, c.U) as g:
        g.run_steps(10)
        Ug = g.download_state()
    with GpuSolver(c.params, c.mesh, c.U) as g:
        g.run_steps(10)
        Ueq = g.download_state()
    err = rel_l1(Ug, Uo)
    eff = rel_l1(Ug, Ueq)
    print(f"[{name}] finite-rate FP64 rel L1 {err.max():.2e}; change from equilibrium {eff.max():.2e}", flush=True)
    assert np.isfinite(Ug).all() and err.max() <= FP64_TOL
    assert eff.max() > 1e-6


def test_finite_rate_relaxation_mixed_precision(gpu):
    c = SMALL["shock_bubble"]()
    p = dict(c.params, **FNT)
    o = Oracle(p, c.mesh, c.U)
    o.run_steps(10)
    Uo = o.state()
    with GpuSolver(dict(p, precision=_lib.MIXED), c.mesh, c.U) as g:
        g.run_steps(10)
        Ug = g.download_state()
    rho_err = rel_l1(Ug[:, [1, 2]].sum(1, keepdims=True), Uo[:, [1, 2]].sum(1, keepdims=True))[0]
    assert rho_err <= MIXED_RHO_TOL


def inlet_case(model, **inlet):
    """Box with an inlet (boundCond 3) at x = 0, outlet at x = 1, slip walls."""
    from fooskagpu_amd import Mesh
    from fooskagpu_amd.mesh import BC_INLET, BC_OUTLET, BC_SLIP
    bc = (BC_INLET, BC_OUTLET) + (BC_SLIP,) * 4
    if model == _lib.MODEL_SINGLE:
        c = cases.sod(nx=16, ny=4, nz=4)
        m = cases.Mesh.box(16, 4, 4, (0, 0, 0), (1.0, 0.25, 0.25), 0.0, 0, bc=bc)
        U = m.init_sod(1.4, 0.0, 0.5, (1.0, 0, 0, 0, 1.0), (0.125, 0, 0, 0, 0.1))
        return cases.Case("sod_inlet", dict(c.params, **inlet), m, U, 10)
    p = cases._two_phase_params(**inlet)
    m = Mesh.box(10, 6, 6, (0, 0, 0), (1, 0.6, 0.6), 0.10, 3, bc=bc)
    am = p["avfMin"]
    regions = [dict(kind=0, alphaL=1 - am, p=1e7, T=cases.T_WATER, u=(5.0, 0, 0)),
               dict(kind=1, c=(0.5, 0.3, 0.3), r=0.15, alphaL=am, p=2340.0, T=cases.T_WATER)]
    return cases.Case("bubble_inlet", p, m, m.init_regions(p, regions), 10)


@pytest.mark.parametrize("model,inlet", [(_lib.MODEL_TWO_PHASE, {}),
                                         (_lib.MODEL_TWO_PHASE, dict(inletP=1.2e7, inletT=300.0)),
                                         (_lib.MODEL_SINGLE, dict(inletP=1.2, inletT=0.004))])
def test_inlet_boundary_matches_oracle(gpu, model, inlet):
    """Inlet ghost (bc.h:2-25) with the reference's constants or a configured
    pressure and temperature: FP64 against the oracle after 10 steps; mixed
    precision within its density bar."""
    c = inlet_case(model, **inlet)
    o = Oracle(c.params, c.mesh, c.U)
    o.run_steps(10)
    Uo = o.state()
    with GpuSolver(c.params, c.mesh, c.U) as g:
        g.run_steps(10)
        Ug = g.download_state()
    with GpuSolver(dict(c.params, precision=_lib.MIXED), c.mesh, c.U) as g:
        g.run_steps(10)
        Um = g.download_state()
    err = rel_l1(Ug, Uo)
    cols = neq_density(c)
    rho_err = rel_l1(Um[:, cols].sum(1, keepdims=True), Uo[:, cols].sum(1, keepdims=True))[0]
    print(f"[{c.name} {inlet}] FP64 rel L1 {err.max():.2e}; mixed density {rho_err:.2e}", flush=True)
    assert np.isfinite(Ug).all() and err.max() <= FP64_TOL
    assert rho_err <= MIXED_RHO_TOL
    # the inlet state reaches the solution: another inlet pressure changes it
    o2 = Oracle(dict(c.params, inletP=2 * c.params.get("inletP", 11.4e5), inletT=c.params.get("inletT", 115.3)),
                c.mesh, c.U)
    o2.run_steps(10)
    assert rel_l1(o2.state(), Uo).max() > 1e-3


def test_finite_rate_with_huge_rate_is_the_equilibrium_solve(gpu):
    """exp(-kappa dt) = 0 makes both finite-rate targets 0: the device run is
    bitwise the equilibrium run."""
    c = SMALL["bubble"]()
    out = []
    for extra in ({}, dict(thermFntRlxFlag=1, thermRlxRate=1e30, dt=1e-6)):
        with GpuSolver(dict(c.params, **extra), c.mesh, c.U) as g:
            g.run_steps(5)
            out.append(g.download_state())
    assert np.array_equal(out[0], out[1])


def test_finite_rate_refuses_the_surrogate_relaxation(gpu):
    from fooskagpu_amd._lib import FkError
    c = SMALL["bubble"]()
    with GpuSolver(dict(c.params, **FNT), c.mesh, c.U) as g:
        with pytest.raises(FkError, match="finite-rate"):
            g.relaxation(_lib.RELAX_MLP)
        g.relaxation(_lib.RELAX_PLAIN)


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.MIXED])
@pytest.mark.parametrize("mode", [_lib.RELAX_COLD, _lib.RELAX_MLP, _lib.RELAX_PLAIN])
def test_fused_relaxation_stage_equals_separate_kernel(gpu, precision, mode):
    """fuse_relax: the per-cell relaxation pass (stage 1 of the surrogate
    relaxation, or the plain thermal solve with its Tsat classification)
    inside the last RK stage's fused pass F, or in its own kernel after it:
    bitwise equal states and equal counts (cavitation: thermal and chemical
    cells; shock-bubble)."""
    from fooskagpu_amd import surrogate
    rng = np.random.default_rng(0)
    x = (rng.normal(size=(1000, 3)) * [2.0, 3e5, 0.3] + [5.0, 1.4e6, 0.5]).astype(np.float32)
    y = (rng.normal(size=(1000, 3)) * [20.0, 2.0, 0.3] + [300.0, 5.0, 0.5]).astype(np.float32)
    for name in ("cavitation", "shock_bubble"):
        c = SMALL[name]()
        p = dict(c.params, precision=precision, relaxMode=mode)
        out, tot = [], []
        for fused in (0, 1):
            with GpuSolver(p, c.mesh, c.U) as g:
                if mode == _lib.RELAX_MLP:
                    surrogate.Surrogate(g).set_weights(surrogate.init_weights(1), surrogate.fit_norm(x, y))
                g.set_tuning("fuse_relax", fused)
                g.run_steps(STEPS.get(name, 3))
                out.append(g.download_state())
                tot.append(g.relax_totals())
        print(name, tot[1], flush=True)
        assert np.isfinite(out[0]).all()
        assert np.array_equal(out[0], out[1]) and tot[0] == tot[1]


@pytest.mark.parametrize("model", [_lib.MODEL_SINGLE, _lib.MODEL_TWO_PHASE])
def test_all_boundary_conditions_match_oracle(gpu, model):
    """Every boundary code on one box (bc.h): inlet and outlet in x, symmetry
    and no-slip wall in y, slip and no-slip wall in z, with viscous + WALE
    terms (the wall ghost enters the viscous face gradient): FP64 against
    the oracle after 10 steps, mixed within its density bar."""
    from fooskagpu_amd import Mesh
    from fooskagpu_amd.mesh import BC_INLET, BC_OUTLET, BC_SLIP, BC_SYMMETRY, BC_WALL
    bc = (BC_INLET, BC_OUTLET, BC_SYMMETRY, BC_WALL, BC_SLIP, BC_WALL)
    m = Mesh.box(10, 5, 5, (0, 0, 0), (1, 0.5, 0.5), 0.10, 7, bc=bc)
    if model == _lib.MODEL_SINGLE:
        p = dict(cases.BASE, model=model, viscFlag=1, turbulence=1, dynVisc=[1e-3, 1e-3], inletP=1.2, inletT=0.004)
        U = m.init_sod(1.4, 0.0, 0.5, (1.0, 0.3, 0.1, 0, 1.0), (0.125, 0, 0, 0.2, 0.1))
    else:
        p = cases._two_phase_params(viscFlag=1, turbulence=1, inletP=1.2e7, inletT=300.0)
        am = p["avfMin"]
        regions = [dict(kind=0, alphaL=1 - am, p=1e7, T=cases.T_WATER, u=(5.0, 2.0, -1.0)),
                   dict(kind=1, c=(0.5, 0.25, 0.25), r=0.12, alphaL=am, p=2340.0, T=cases.T_WATER)]
        U = m.init_regions(p, regions)
    o = Oracle(p, m, U)
    o.run_steps(10)
    Uo = o.state()
    with GpuSolver(p, m, U) as g:
        g.run_steps(10)
        Ug = g.download_state()
    with GpuSolver(dict(p, precision=_lib.MIXED), m, U) as g:
        g.run_steps(10)
        Um = g.download_state()
    cols = [0] if model == _lib.MODEL_SINGLE else [1, 2]
    err = rel_l1(Ug, Uo)
    rho_err = rel_l1(Um[:, cols].sum(1, keepdims=True), Uo[:, cols].sum(1, keepdims=True))[0]
    print(f"[bcs model {model}] FP64 rel L1 {err.max():.2e}; mixed density {rho_err:.2e}", flush=True)
    assert np.isfinite(Ug).all() and err.max() <= FP64_TOL
    assert rho_err <= MIXED_RHO_TOL


def test_neglect_flags_match_oracle(gpu):
    """neglectZ zeroes the z momentum in every RK update (update.cu:26-32);
    neglectX and neglectY are inert there, because the reference assigns its
    neglect flag three times and only the z test survives.  FP64 against the
    oracle, fused and split RK updates."""
    c = SMALL["shock_bubble"]()
    p = dict(c.params, neglectX=1, neglectY=1, neglectZ=1)
    o = Oracle(p, c.mesh, c.U)
    o.run_steps(5)
    Uo = o.state()
    for fuse in (1, 0):
        with GpuSolver(p, c.mesh, c.U) as g:
            g.set_tuning("fuse_update", fuse)
            g.run_steps(5)
            Ug = g.download_state()
        assert np.all(Ug[:, 5] == 0.0) and np.abs(Ug[:, 3]).max() > 0 and np.abs(Ug[:, 4]).max() > 0
        assert rel_l1(Ug, Uo).max() <= FP64_TOL


@pytest.mark.parametrize("flags", [dict(adaptDt=0, dt=1.5e-6), dict(rkSteps=1), dict(hllcEnergy=1),
                                   dict(woodsFlag=0), dict(adaptDtItr=3)],
                         ids=["fixed_dt", "euler", "toro_energy", "frozen_sound_speed", "dt_every_3"])
def test_step_control_flags_match_oracle(gpu, flags):
    """The INPUT switches of the step beyond the defaults: FP64 against the
    oracle after 6 steps, mixed within its density bar (shock-bubble, WALE)."""
    c = SMALL["shock_bubble"]()
    p = dict(c.params, **flags)
    o = Oracle(p, c.mesh, c.U)
    dto, to = o.run_steps(6)
    Uo = o.state()
    with GpuSolver(p, c.mesh, c.U) as g:
        dtg, tg = g.run_steps(6)
        Ug = g.download_state()
    with GpuSolver(dict(p, precision=_lib.MIXED), c.mesh, c.U) as g:
        g.run_steps(6)
        Um = g.download_state()
    rho_err = rel_l1(Um[:, [1, 2]].sum(1, keepdims=True), Uo[:, [1, 2]].sum(1, keepdims=True))[0]
    assert np.isfinite(Ug).all() and rel_l1(Ug, Uo).max() <= FP64_TOL
    assert dtg == dto and tg == pytest.approx(to, rel=1e-14)
    assert rho_err <= MIXED_RHO_TOL
