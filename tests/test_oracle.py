"""Pins of the CPU FP64 oracle against what the mathematics fixes: closed
forms, invariants, textbook solutions and brute force on tiny inputs.  The
reference ships no tests, fixtures or golden vectors for this path, so these
are the checks that make the oracle trustworthy; each is chosen so that a
plausible mistake (a dropped term, a wrong sign or index, a transposed
operand) fails it."""
import math

import numpy as np
import pytest
from scipy.optimize import brentq, fsolve

from fooskagpu_amd import Mesh, cases
from fooskagpu_amd.mesh import BC_INLET, BC_OUTLET, BC_SLIP, BC_SYMMETRY, BC_WALL
from oracle import pyoracle as po
from oracle.pyoracle import Oracle

SP = dict(cases.BASE, model=0, gamma=[1.4, 1.4], pInf=[0.0, 0.0], Cp=[1006.0, 1006.0])
TP = cases._two_phase_params()
RNG = np.random.default_rng(1234)


# ---------------- EOS and conversions ----------------
@pytest.mark.parametrize("k", [0, 1])
def test_two_phase_eos_round_trips(k):
    for p, T in [(1e7, 293.15), (2340.0, 300.0), (5e7, 320.0), (1e5, 400.0)]:
        rho = po.eqos(TP, k, 2, p, T)
        e = po.eqos(TP, k, 3, p, rho)
        assert po.eqos(TP, k, 0, rho, e) == pytest.approx(p, rel=1e-9, abs=1e-6)
        assert po.eqos(TP, k, 1, rho, e) == pytest.approx(T, rel=1e-12)
        # sound speed of the stiffened gas, c^2 = gamma (p + pInf) / rho
        g, pi = TP["gamma"][k], TP["pInf"][k]
        assert po.eqos(TP, k, 4, rho, p) == pytest.approx(math.sqrt(g * (p + pi) / rho), rel=1e-15)


def test_water_density_and_saturation_points():
    assert po.eqos(TP, 0, 2, 1e7, cases.T_WATER) == pytest.approx(998.0, rel=1e-12)
    for P, Tref in [(2339.0, 293.15), (101325.0, 373.15)]:
        T, _ = po.tsat(TP, P, 300.0)
        assert T == pytest.approx(Tref, abs=0.1)  # Newton stops at |f| <= 1e-2 (log10 units)


def test_single_phase_eos_ideal_gas():
    rho, p = 1.3, 2.0e5
    e = po.eqos(SP, 0, 3, p, rho)
    assert e == pytest.approx(p / (0.4 * rho), rel=1e-15)
    assert po.eqos(SP, 0, 0, rho, e) == pytest.approx(p, rel=1e-15)
    assert po.eqos(SP, 0, 1, rho, e) == pytest.approx(p / (rho * (1006.0 - 1006.0 / 1.4)), rel=1e-14)


@pytest.mark.parametrize("params", [SP, TP], ids=["single", "two_phase"])
def test_prim_cons_round_trip(params):
    for _ in range(20):
        if params["model"] == 0:
            W = np.array([RNG.uniform(0.1, 2), *RNG.uniform(-50, 50, 3), RNG.uniform(1e4, 1e6)])
        else:
            a = RNG.uniform(0.01, 0.99)
            p = RNG.uniform(1e4, 1e7)
            W = np.array([a, po.eqos(params, 0, 2, p, 300.0), po.eqos(params, 1, 2, p, 300.0),
                          *RNG.uniform(-20, 20, 3), p, p, RNG.uniform(0, 1e-3), 0.0])
            W[9] = W[0] * W[1] + (1 - W[0]) * W[2]
        U = po.prim2cons(params, W)
        W2 = po.cons2prim(params, U)
        assert np.allclose(W2, W, rtol=1e-9, atol=1e-9 * np.abs(W).max())


# ---------------- boundary ghosts ----------------
@pytest.mark.parametrize("params", [SP, TP], ids=["single", "two_phase"])
def test_boundary_ghosts(params):
    iu = 1 if params["model"] == 0 else 3
    n = RNG.normal(size=3)
    n /= np.linalg.norm(n)
    W = po.cons2prim(params, cases.sod().U[0]) if params["model"] == 0 else po.cons2prim(params, cases.bubble(n=2).U[0])
    W[iu:iu + 3] = [3.0, -4.0, 5.0]
    u = W[iu:iu + 3].copy()
    un = u @ n
    for bc in (BC_SYMMETRY, BC_SLIP):
        g = po.bc_ghost(params, bc, n, W)
        ug = g[iu:iu + 3]
        assert ug @ n == pytest.approx(-un, abs=1e-13)
        assert np.allclose(ug - (ug @ n) * n, u - un * n, atol=1e-13)  # tangential part kept
        assert np.linalg.norm(ug) == pytest.approx(np.linalg.norm(u), rel=1e-14)
        assert np.array_equal(np.delete(g, range(iu, iu + 3)), np.delete(W, range(iu, iu + 3)))
    assert np.array_equal(po.bc_ghost(params, BC_WALL, n, W)[iu:iu + 3], -u)
    assert np.array_equal(po.bc_ghost(params, BC_OUTLET, n, W), W)
    gi = po.bc_ghost(params, BC_INLET, n, W)
    assert np.array_equal(gi[iu:iu + 3], u)
    if params["model"] == 1:
        assert gi[6] == gi[7] == 11.4e5 and gi[0] == 1 - params["avfMin"]


# ---------------- HLLC ----------------
def euler_flux(W, n):
    """Physical flux of the Euler equations for W = (rho, u, v, w, p), gamma 1.4."""
    rho, u, p = W[0], W[1:4], W[4]
    un = u @ n
    E = p / 0.4 + 0.5 * rho * (u @ u)
    return np.array([rho * un, *(rho * u * un + p * n), (E + p) * un])


def test_single_phase_flux_consistency_and_upwinding():
    n = np.array([0.48, 0.6, 0.64])
    A = 0.37
    for _ in range(10):
        W = np.array([RNG.uniform(0.2, 2), *RNG.uniform(-30, 30, 3), RNG.uniform(1e4, 2e5)])
        out = po.face_flux(SP, W, W, n, A, W, W)
        assert np.allclose(out[0], -euler_flux(W, n) * A, rtol=1e-12, atol=1e-9)
        assert np.allclose(out[1], euler_flux(W, n) * A, rtol=1e-12, atol=1e-9)
    # supersonic to the right: the flux is F(W_L)
    WL = np.array([1.0, 800.0 * n[0], 800 * n[1], 800 * n[2], 1e5])
    WR = np.array([0.5, 700.0 * n[0], 700 * n[1], 700 * n[2], 5e4])
    assert np.allclose(po.face_flux(SP, WL, WR, n, 1.0, WL, WR)[1], euler_flux(WL, n), rtol=1e-12)
    assert np.allclose(po.face_flux(SP, WR * [1, -1, -1, -1, 1], WL * [1, -1, -1, -1, 1], n, 1.0, WL, WR)[1],
                       euler_flux(WL * [1, -1, -1, -1, 1], n), rtol=1e-12)


def test_two_phase_flux_uniform_state_balance():
    """Uniform mixture: convective and non-conservative terms cancel in
    alpha (F0 - alpha v_n = 0) and the pressure correction vanishes, so the
    momentum contribution is the momentum flux rho u u_n."""
    c = cases.bubble(n=2)
    W = po.cons2prim(TP, c.U[0])
    W[3:6] = [12.0, -3.0, 4.0]
    U = po.prim2cons(TP, W)
    n = np.array([0.0, 0.6, 0.8])
    out = po.face_flux(TP, W, W, n, 2.0, U, U)
    assert out[0, 0] == pytest.approx(0.0, abs=1e-12)
    rho, u = W[9], W[3:6]
    assert np.allclose(out[0, 3:6], -(rho * u * (u @ n)) * 2.0, rtol=1e-10)
    assert np.allclose(out[0, 1], -W[0] * W[1] * (u @ n) * 2.0, rtol=1e-12)
    assert np.allclose(out[0], -out[1], rtol=1e-12, atol=1e-9)


# ---------------- WALE closed forms ----------------
def test_wale_closed_forms():
    vol = 2.7e-4
    Ls = 0.325 * vol ** 0.3333333333333
    assert po.wale(np.zeros((3, 3)), vol) == 0.0
    shear = np.zeros((3, 3))
    shear[0, 1] = 5.0  # g^2 = 0 -> S^d = 0 -> nu_t = 0
    assert po.wale(shear, vol) == 0.0
    w = 3.0
    rot = np.array([[0, -w, 0], [w, 0, 0], [0, 0, 0]])  # S = 0, S^d:S^d = 2/3 w^4
    s1 = 2.0 / 3.0 * w ** 4
    expect = Ls * Ls * s1 ** 1.5 / (s1 ** 1.25 + 1e-12)
    assert po.wale(rot, vol) == pytest.approx(expect, rel=1e-13)
    assert expect == pytest.approx(Ls * Ls * (2 / 3) ** 0.25 * w, rel=1e-12)


# ---------------- RK3 ----------------
def test_ssp_rk3_is_third_order_on_linear_ode():
    """dU/dt = lam U: one Shu-Osher step equals 1 + z + z^2/2 + z^3/6."""
    c = cases.sod(nx=2, ny=1, nz=1)
    o = Oracle(c.params, c.mesh, c.U)
    V = c.mesh.volume[:, None]
    lam, dt = -3.0, 0.1
    U = o.state()
    o.init_rk()
    for irk in range(3):
        o.set_rhs(lam * o.state() * V)
        o.update(irk, dt)
    z = lam * dt
    assert np.allclose(o.state(), U * (1 + z + z * z / 2 + z ** 3 / 6), rtol=1e-14)


# ---------------- time step brute force ----------------
def test_time_step_single_tet(tmp_path):
    path = tmp_path / "tet.neu"
    path.write_text(
        "CONTROL INFO 1.2.1\nNUMNP NELEM NGRPS NBSETS NDFCD NDFVL\n 4 1 1 1 3 3\nENDOFSECTION\n"
        "NODAL COORDINATES 1.2.1\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\nENDOFSECTION\n"
        "ELEMENTS/CELLS 1.2.1\n1 6 4 1 2 3 4\nENDOFSECTION\n"
        "BOUNDARY CONDITIONS 1.2.1\nout 1 4 0 26\n1 6 1\n1 6 2\n1 6 3\n1 6 4\nENDOFSECTION\n")
    m = Mesh.read_neu(path, rcm=False)
    W = np.array([1.2, 30.0, -20.0, 10.0, 1e5])
    U = po.prim2cons(SP, W)[None]
    o = Oracle(SP, m, U)
    o.compute_face_values()
    o.compute_rhs()
    dt = o.compute_time_step()
    c = math.sqrt(1.4 * W[4] / W[0])
    smax = max(abs(W[1:4] @ nv) + c for nv in m.n)
    assert dt == pytest.approx(0.5 * m.volume[0] ** 0.3333 / smax, rel=1e-13)


@pytest.mark.parametrize("rho_yv,woods", [(1.0, 1), (0.0, 1), (1.0, 0)])
def test_two_phase_time_step_uses_wood_or_frozen_sound_speed(tmp_path, rho_yv, woods):
    """Single tet at rest, alpha_L = 0.5: the signal speed is the mixture sound
    speed, Wood's 1/(rho c^2) = sum_k alpha_k / (rho_k c_k^2) where
    rho Y_v > chemLimit and woodsFlag == 1, else the frozen
    c^2 = sum_k Y_k c_k^2, with c_k^2 = gamma_k (p_k + pinf_k) / rho_k."""
    path = tmp_path / "tet.neu"
    path.write_text(
        "CONTROL INFO 1.2.1\nNUMNP NELEM NGRPS NBSETS NDFCD NDFVL\n 4 1 1 1 3 3\nENDOFSECTION\n"
        "NODAL COORDINATES 1.2.1\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\nENDOFSECTION\n"
        "ELEMENTS/CELLS 1.2.1\n1 6 4 1 2 3 4\nENDOFSECTION\n"
        "BOUNDARY CONDITIONS 1.2.1\nout 1 4 0 26\n1 6 1\n1 6 2\n1 6 3\n1 6 4\nENDOFSECTION\n")
    m = Mesh.read_neu(path, rcm=False)
    prm = dict(TP, woodsFlag=woods)
    a, p = 0.5, 2e5
    rL, rG = po.eqos(TP, 0, 2, p, 300.0), po.eqos(TP, 1, 2, p, 300.0)
    W = np.array([a, rL, rG, 0.0, 0.0, 0.0, p, p, rho_yv, 0.0])
    o = Oracle(prm, m, po.prim2cons(prm, W)[None])
    o.compute_face_values()
    o.compute_rhs()
    dt = o.compute_time_step()
    g, pi = TP["gamma"], TP["pInf"]
    c2 = [g[k] * (p + pi[k]) / r for k, r in ((0, rL), (1, rG))]
    rho = a * rL + (1 - a) * rG
    if rho_yv > TP["chemLimit"] and woods == 1:
        c = 1 / math.sqrt(rho * (a / (rL * c2[0]) + (1 - a) / (rG * c2[1])))
    else:
        c = math.sqrt((a * rL * c2[0] + (1 - a) * rG * c2[1]) / rho)
    assert dt == pytest.approx(TP["cfl"] * m.volume[0] ** 0.3333 / c, rel=1e-12)


# ---------------- whole-step invariants ----------------
def periodic_box(n=4, jitter=0.1):
    return Mesh.box(n, n, n, (0, 0, 0), (1, 1, 1), jitter, 11, periodic=(True, True, True))


@pytest.mark.parametrize("rk", [1, 3])
def test_interface_advection_keeps_pressure_and_velocity_uniform(rk):
    """A smooth two-phase interface (alpha_L from 0.1 to 0.9) advected at
    uniform u = 10 m/s and p = 1e5 Pa, equal phase temperatures, periodic
    jittered box, no relaxation: alpha moves while u and both phase pressures
    stay uniform to round-off.  This pins the consistency of the phase mass
    and energy fluxes with the volume-fraction transport; the
    non-conservative terms vanish at uniform u and p."""
    prm = dict(TP, relaxFlag=0, chemFlag=0, adaptDt=0, dt=2e-6, rkSteps=rk)
    m = Mesh.box(8, 3, 3, (0, 0, 0), (1, 0.375, 0.375), 0.1, 5, periodic=(True, True, True))
    p, u = 1e5, 10.0
    aL = 0.1 + 0.8 * np.exp(-((m.cell_center[:, 0] - 0.5) / 0.15) ** 2)
    rL, rG = po.eqos(prm, 0, 2, p, 300.0), po.eqos(prm, 1, 2, p, 300.0)
    U = np.array([po.prim2cons(prm, np.array([a, rL, rG, u, 0, 0, p, p, 0.0, 0.0])) for a in aL])
    o = Oracle(prm, m, U)
    o.run_steps(3)
    W = np.array([po.cons2prim(prm, x) for x in o.state()])
    assert np.abs(W[:, 0] - aL).max() > 1e-4  # the interface moved
    assert np.abs(W[:, 3] - u).max() < 1e-10 and np.abs(W[:, 4:6]).max() < 1e-10
    assert np.abs(W[:, 7] - p).max() < 1e-9 * p
    assert np.abs(W[:, 6] - p).max() < 1e-12 * (p + prm["pInf"][0])


@pytest.mark.parametrize("rk", [1, 3])
def test_compression_keeps_uniform_volume_fraction(rk):
    """Uniform alpha_L = 0.5, p and T, velocity u = 10 sin(2 pi x): the
    volume-fraction equation d alpha/dt + div(alpha u) - alpha div u = 0
    leaves a uniform alpha unchanged while the densities change.  A missing
    or mis-signed non-conservative term alpha (u . n) changes alpha by about
    alpha div(u) dt."""
    prm = dict(TP, relaxFlag=0, chemFlag=0, adaptDt=0, dt=2e-6, rkSteps=rk)
    m = Mesh.box(8, 3, 3, (0, 0, 0), (1, 0.375, 0.375), 0.1, 5, periodic=(True, True, True))
    p = 1e5
    rL, rG = po.eqos(prm, 0, 2, p, 300.0), po.eqos(prm, 1, 2, p, 300.0)
    ux = 10 * np.sin(2 * np.pi * m.cell_center[:, 0])
    U = np.array([po.prim2cons(prm, np.array([0.5, rL, rG, v, 0, 0, p, p, 0.0, 0.0])) for v in ux])
    o = Oracle(prm, m, U)
    o.run_steps(3)
    U2 = o.state()
    assert np.abs(U2[:, 2] / U[:, 2] - 1).max() > 1e-6  # the gas density changed
    assert np.abs(U2[:, 0] - 0.5).max() < 1e-13


def test_two_phase_pure_liquid_limit_matches_single_phase_stiffened_gas():
    """At alpha_L = 1 - avfMin/100 the two-phase model is the single-phase
    stiffened-gas Euler system of the liquid: a 2 MPa pressure pulse in
    water (periodic box, 10 SSP-RK3 steps, no relaxation, frozen sound speed)
    evolves the same in both models up to the vapour's 1e-6 volume fraction.
    Observed differences are 3.5e-5 (momentum) and 4.2e-5 (pressure
    perturbation); an oracle without the face-pressure momentum term differs
    by 1.0 and 0.11, one without the liquid energy term alpha_L p_L (u . n) by
    2.8e-3 and 2.9e-3.  Without the HLLC energy term the change of total
    energy differs by 196 %."""
    base = dict(relaxFlag=0, chemFlag=0, adaptDt=0, dt=2e-6, rkSteps=3, woodsFlag=0)
    tp = dict(TP, eta=[0.0, TP["eta"][1]], **base)
    sp = dict(cases.BASE, model=0, gamma=[TP["gamma"][0]] * 2, pInf=[TP["pInf"][0]] * 2, Cp=[TP["Cp"][0]] * 2,
              eta=[0.0, 0.0], **base)
    m = Mesh.box(16, 3, 3, (0, 0, 0), (1, 0.1875, 0.1875), 0.1, 5, periodic=(True, True, True))
    p = 1e7 + 2e6 * np.exp(-((m.cell_center[:, 0] - 0.5) / 0.1) ** 2)
    rl = np.array([po.eqos(tp, 0, 2, q, 300.0) for q in p])
    rg = np.array([po.eqos(tp, 1, 2, q, 300.0) for q in p])
    am = 1 - tp["avfMin"] / 100
    U2 = np.array([po.prim2cons(tp, np.array([am, a, b, 0, 0, 0, q, q, 0.0, 0.0])) for a, b, q in zip(rl, rg, p)])
    U1 = np.array([po.prim2cons(sp, np.array([a, 0, 0, 0, q])) for a, q in zip(rl, p)])
    o2, o1 = Oracle(tp, m, U2), Oracle(sp, m, U1)
    o2.run_steps(10)
    o1.run_steps(10)
    A, B = o2.state(), o1.state()
    W2 = np.array([po.cons2prim(tp, u) for u in A])
    W1 = np.array([po.cons2prim(sp, u) for u in B])
    assert np.abs(W1[:, 4] - p).max() > 1e5  # the pulse spread
    assert np.abs((A[:, 1] + A[:, 2]) - B[:, 0]).max() / B[:, 0].max() < 1e-5
    assert np.abs(A[:, 3] - B[:, 1]).max() / np.abs(B[:, 1]).max() < 3e-4
    assert np.abs(W2[:, 6] - W1[:, 4]).max() / (W1[:, 4].max() - 1e7) < 3e-4
    # total energy: with Toro's star-state energy (hllcEnergy 1) the change
    # of E matches the single-phase HLLC (observed 2.1e-4); the reference's
    # form (hllcEnergy 0, no rho_K in the energy term) differs by 5.1 %
    dE1 = B[:, 4] - U1[:, 4]
    for h, lo, hi in ((0, 1e-2, 0.2), (1, 0.0, 1e-3)):
        o3 = Oracle(dict(tp, hllcEnergy=h), m, U2)
        o3.run_steps(10)
        d = np.abs((o3.state()[:, 8] - U2[:, 8]) - dE1).max() / np.abs(dE1).max()
        assert lo <= d < hi, (h, d)


def test_two_phase_pure_vapour_limit_matches_single_phase_ideal_gas():
    """The vapour-side counterpart of the pure-liquid limit: alpha_L = 1e-10
    (avfMin 1e-8), a 20 kPa pulse in vapour at 400 K against the single-phase
    ideal gas (gamma 1.327).  Observed differences are 7e-4 (momentum) and
    1.1e-3 (pressure perturbation); an oracle without the gas energy term
    alpha_G p_G (u . n) differs by 0.16 and 0.21."""
    base = dict(relaxFlag=0, chemFlag=0, adaptDt=0, dt=2e-5, rkSteps=3, woodsFlag=0, avfMin=1e-8)
    tp = dict(TP, eta=[TP["eta"][0], 0.0], **base)
    sp = dict(cases.BASE, model=0, gamma=[TP["gamma"][1]] * 2, pInf=[0.0, 0.0], Cp=[TP["Cp"][1]] * 2,
              eta=[0.0, 0.0], **base)
    m = Mesh.box(16, 3, 3, (0, 0, 0), (1, 0.1875, 0.1875), 0.1, 5, periodic=(True, True, True))
    p = 1e5 + 2e4 * np.exp(-((m.cell_center[:, 0] - 0.5) / 0.1) ** 2)
    rl = np.array([po.eqos(tp, 0, 2, q, 400.0) for q in p])
    rg = np.array([po.eqos(tp, 1, 2, q, 400.0) for q in p])
    am = tp["avfMin"] / 100
    U2 = np.array([po.prim2cons(tp, np.array([am, a, b, 0, 0, 0, q, q, 0.0, 0.0])) for a, b, q in zip(rl, rg, p)])
    U1 = np.array([po.prim2cons(sp, np.array([b, 0, 0, 0, q])) for b, q in zip(rg, p)])
    o2, o1 = Oracle(tp, m, U2), Oracle(sp, m, U1)
    o2.run_steps(10)
    o1.run_steps(10)
    A, B = o2.state(), o1.state()
    W2 = np.array([po.cons2prim(tp, u) for u in A])
    W1 = np.array([po.cons2prim(sp, u) for u in B])
    assert np.abs(W1[:, 4] - p).max() > 5e3  # the pulse spread
    assert np.abs((A[:, 1] + A[:, 2]) - B[:, 0]).max() / B[:, 0].max() < 1e-3
    assert np.abs(A[:, 3] - B[:, 1]).max() / np.abs(B[:, 1]).max() < 5e-3
    assert np.abs(W2[:, 7] - W1[:, 4]).max() / (W1[:, 4].max() - 1e5) < 5e-3


def vapour_pulse_energy_change(T_base, amp, hllc_energy):
    """Pure-vapour limit (alpha_L = 1e-10) of a pressure pulse of amplitude
    `amp` on 1e5 Pa at temperature T_base, 10 SSP-RK3 steps: the largest
    difference of the change of total energy between the two-phase model
    (energy star state form `hllc_energy`) and the single-phase ideal gas,
    relative to the largest change, and the vapour density range."""
    base = dict(relaxFlag=0, chemFlag=0, adaptDt=0, dt=2e-5, rkSteps=3, woodsFlag=0, avfMin=1e-8)
    tp = dict(TP, eta=[TP["eta"][0], 0.0], hllcEnergy=hllc_energy, **base)
    sp = dict(cases.BASE, model=0, gamma=[TP["gamma"][1]] * 2, pInf=[0.0, 0.0], Cp=[TP["Cp"][1]] * 2,
              eta=[0.0, 0.0], **base)
    m = Mesh.box(16, 3, 3, (0, 0, 0), (1, 0.1875, 0.1875), 0.1, 5, periodic=(True, True, True))
    p = 1e5 + amp * np.exp(-((m.cell_center[:, 0] - 0.5) / 0.1) ** 2)
    rl = np.array([po.eqos(tp, 0, 2, q, T_base) for q in p])
    rg = np.array([po.eqos(tp, 1, 2, q, T_base) for q in p])
    am = tp["avfMin"] / 100
    U2 = np.array([po.prim2cons(tp, np.array([am, a, b, 0, 0, 0, q, q, 0.0, 0.0])) for a, b, q in zip(rl, rg, p)])
    U1 = np.array([po.prim2cons(sp, np.array([b, 0, 0, 0, q])) for b, q in zip(rg, p)])
    o2, o1 = Oracle(tp, m, U2), Oracle(sp, m, U1)
    o2.run_steps(10)
    o1.run_steps(10)
    dE1 = o1.state()[:, 4] - U1[:, 4]
    dE2 = o2.state()[:, 8] - U2[:, 8]
    return np.abs(dE2 - dE1).max() / np.abs(dE1).max(), rg.min(), rg.max()


# vapour at 1e5 Pa and 215.8 K has rho_G = 1.0003 kg/m^3
T_UNIT_VAPOUR = 215.8


def test_reference_energy_star_state_at_unit_density_matches_ideal_gas():
    """The reference's two-phase energy star state (starRegion.h:32) has
    (S* + p / (rho_K (S_K - u_K))) where Toro's form has
    (rho_K S* + p / (S_K - u_K)): the two coincide at rho_K = 1.  In the
    pure-vapour limit at rho_G ~ 1 (1.0003 .. 1.02 kg/m^3 over a 2 kPa pulse)
    the reference form must therefore reproduce the single-phase ideal-gas
    HLLC energy change (observed 3.4e-3, bar 5e-3), while at rho_G ~ 0.54
    (400 K) it departs from it (observed 0.18): that is the missing rho_K
    (Q21), and Toro's form matches at both densities (observed 2e-4).  An
    oracle with the sign of the pressure term of the reference form flipped
    misses the bar at unit density (tests/test_oracle_plants.py)."""
    d, lo, hi = vapour_pulse_energy_change(T_UNIT_VAPOUR, 2e3, 0)
    assert 1.0 <= lo and hi < 1.03
    assert d < 5e-3, d
    d_toro, _, _ = vapour_pulse_energy_change(T_UNIT_VAPOUR, 2e3, 1)
    assert d_toro < 1e-3
    d_far, lo, hi = vapour_pulse_energy_change(400.0, 2e3, 0)
    assert hi < 0.56 and d_far > 0.1, d_far
    assert vapour_pulse_energy_change(400.0, 2e3, 1)[0] < 1e-3


def test_free_stream_preserved_on_jittered_tets():
    m = periodic_box()
    W = np.array([1.1, 40.0, -10.0, 25.0, 1e5])
    U = np.tile(po.prim2cons(SP, W), (m.n_cells, 1))
    o = Oracle(dict(SP, viscFlag=1, turbulence=1, heatFlux=1, dynVisc=[1e-3, 1e-3]), m, U)
    o.run_steps(5)
    assert np.allclose(o.state(), U, rtol=1e-11, atol=0)


def test_two_phase_free_stream_and_conservation_periodic():
    m = periodic_box()
    c = cases.bubble(n=2)
    W = po.cons2prim(TP, c.U[0])
    W[3:6] = [5.0, 2.0, -1.0]
    U = np.tile(po.prim2cons(TP, W), (m.n_cells, 1))
    o = Oracle(TP, m, U)
    o.run_steps(3)
    assert np.allclose(o.state(), U, rtol=1e-9, atol=1e-9)


def test_conservation_on_periodic_box():
    """Mass, momentum and energy integrals are conserved to round-off with
    the deterministic gather (owner and neighbour contributions cancel)."""
    m = periodic_box(5)
    x = m.cell_center
    W = np.stack([1 + 0.2 * np.sin(2 * np.pi * x[:, 0]), 10 * np.cos(2 * np.pi * x[:, 1]), 5 * np.sin(2 * np.pi * x[:, 2]),
                  np.zeros(len(x)), 1e5 * (1 + 0.3 * (x[:, 0] > 0.5))], axis=1)
    U = np.array([po.prim2cons(SP, w) for w in W])
    p = dict(SP, viscFlag=1, turbulence=1, heatFlux=1, dynVisc=[1e-2, 1e-2])
    o = Oracle(p, m, U)
    I0 = o.cons_calc()
    o.run_steps(5)
    I1 = o.cons_calc()
    assert np.allclose(I1, I0, rtol=1e-12, atol=1e-12 * np.abs(I0).max())
    # two-phase on the same box: partial masses, momentum, total energy and
    # vapour mass are conserved (alpha and phasic energies are not)
    c = cases.bubble(n=2)
    regions = [dict(kind=0, alphaL=0.9, p=1e6, T=300.0), dict(kind=1, c=(0.5, 0.5, 0.5), r=0.3, alphaL=0.2, p=1e5,
                                                                T=300.0)]
    U2 = m.init_regions(TP, regions)
    o2 = Oracle(TP, m, U2)
    J0 = o2.cons_calc()
    o2.run_steps(3)
    J1 = o2.cons_calc()
    for i in (1, 2, 3, 4, 5, 8, 9):
        assert J1[i] == pytest.approx(J0[i], rel=1e-12, abs=1e-9), i


def sod_exact(x, t, x0=0.5, gamma=1.4, WL=(1.0, 0.0, 1.0), WR=(0.125, 0.0, 0.1)):
    """Exact Riemann solution (Toro, ch. 4) for the density."""
    rl, ul, pl = WL
    rr, ur, pr = WR
    cl, cr = math.sqrt(gamma * pl / rl), math.sqrt(gamma * pr / rr)

    def fk(p, r, pk, ck):
        if p > pk:
            A, B = 2 / ((gamma + 1) * r), (gamma - 1) / (gamma + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B))
        return 2 * ck / (gamma - 1) * ((p / pk) ** ((gamma - 1) / (2 * gamma)) - 1)

    ps = brentq(lambda p: fk(p, rl, pl, cl) + fk(p, rr, pr, cr) + ur - ul, 1e-8, 10)
    us = 0.5 * (ul + ur) + 0.5 * (fk(ps, rr, pr, cr) - fk(ps, rl, pl, cl))
    g = gamma
    rsl = rl * (ps / pl) ** (1 / g)
    rsr = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
    csl = cl * (ps / pl) ** ((g - 1) / (2 * g))
    S = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
    out = np.empty_like(x)
    for i, xi in enumerate(x):
        s = (xi - x0) / t
        if s < ul - cl:
            out[i] = rl
        elif s < us - csl:
            out[i] = rl * (2 / (g + 1) + (g - 1) / ((g + 1) * cl) * (ul - s)) ** (2 / (g - 1))
        elif s < us:
            out[i] = rsl
        elif s < S:
            out[i] = rsr
        else:
            out[i] = rr
    return out


def test_sod_converges_to_exact_riemann_solution():
    errs = []
    for nx in (8, 16, 32):  # cubic hexes, as in the 32x8x8 configuration
        c = cases.sod(nx=nx, ny=nx // 4, nz=nx // 4)
        o = Oracle(c.params, c.mesh, c.U)
        t = 0.0
        while t < 0.15:
            _, t = o.run_steps(1)
        rho = o.state()[:, 0]
        x = c.mesh.cell_center[:, 0]
        errs.append(np.abs(rho - sod_exact(x, t)).mean())
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 0.03


# ---------------- limiter ----------------
def test_limiter_bounds_and_uniform_field():
    c = cases.sod(nx=8, ny=3, nz=3)
    m = c.mesh
    U = c.U.copy()
    U[:, 0] *= 1 + 0.3 * RNG.random(m.n_cells)
    o = Oracle(c.params, m, U)
    o.compute_face_values()
    th, wmin, wmax = o.get("theta"), o.get("wmin"), o.get("wmax")
    face, slots = o.get("face"), o.get("slots")
    W = np.array([po.cons2prim(SP, u) for u in U])
    assert np.all((th >= 0) & (th <= 1))
    d = face - W[:, None, :]  # theta * Delta per slot
    lo = np.minimum(0, (wmin - W) / 2)[:, None, :]
    hi = np.maximum(0, (wmax - W) / 2)[:, None, :]
    assert np.all(d >= lo - 1e-12) and np.all(d <= hi + 1e-12)
    # the bound is reached: where 0 < theta < 1, some face of the cell sits on
    # W + (wmax - W) / 2 or W + (wmin - W) / 2 (a factor other than 1/2 in the
    # limiter misses it)
    scale = np.abs(W)[:, None, :] + 1e-300
    on_bound = (np.abs(d - lo) <= 1e-11 * scale) | (np.abs(d - hi) <= 1e-11 * scale)
    active = (th > 1e-9) & (th < 1 - 1e-9)
    assert active.sum() > 20
    assert np.all(on_bound.any(axis=1)[active])
    # uniform field: zero gradients, face states equal the cell state
    o2 = Oracle(c.params, m, np.tile(c.U[0], (m.n_cells, 1)))
    o2.compute_face_values()
    assert np.abs(o2.get("grad")).max() < 1e-10
    assert np.allclose(o2.get("face"), np.tile(W[0] * 0 + po.cons2prim(SP, c.U[0]), (m.n_cells, 4, 1)), rtol=1e-13)


# ---------------- relaxation ----------------
def mixture_state(alphaL, pL, pG, T, u=(3.0, 0, 0)):
    rhoL = po.eqos(TP, 0, 2, pL, T)
    rhoG = po.eqos(TP, 1, 2, pG, T)
    W = np.array([alphaL, rhoL, rhoG, *u, pL, pG, 0.0, 0.0])
    return po.prim2cons(TP, W)


def test_thermal_relaxation_reaches_pt_equilibrium_and_conserves():
    for aL, pL, pG in [(0.7, 2e6, 1e6), (0.3, 1e5, 3e5), (0.999, 1e7, 5e6), (0.5, 5e5, 4e5)]:
        U = mixture_state(aL, pL, pG, 350.0)
        U2, kind, it, _ = po.relax_cell(dict(TP, chemFlag=0), U)
        assert kind == 0 and it <= TP["thermItrMax"] + 1
        W = po.cons2prim(TP, U2)
        assert abs(W[6] - W[7]) < 1e-2 * max(1.0, 1e-9 * W[6]) * 10
        TLq = po.eqos(TP, 0, 1, W[1], U2[6] / (W[0] * W[1]))
        TGs = po.eqos(TP, 1, 1, W[2], U2[7] / ((1 - W[0]) * W[2]))
        assert abs(TLq - TGs) < 1e-2
        assert np.array_equal(U2[[1, 2, 3, 4, 5, 8, 9]], U[[1, 2, 3, 4, 5, 8, 9]])
        # mixture internal energy is the total minus kinetic
        rho = U[1] + U[2]
        ke = 0.5 * (U[3:6] @ U[3:6]) / rho
        assert U2[6] + U2[7] == pytest.approx(U[8] - ke, rel=1e-12)


def test_chemical_relaxation_reaches_gibbs_equilibrium():
    U = mixture_state(0.5, 1500.0, 1500.0, 300.0, u=(0, 0, 0))
    U2, kind, it, _ = po.relax_cell(TP, U)
    assert kind == 1
    W = po.cons2prim(TP, U2)
    TL = po.eqos(TP, 0, 1, W[1], U2[6] / (W[0] * W[1]))
    TG = po.eqos(TP, 1, 1, W[2], U2[7] / ((1 - W[0]) * W[2]))
    assert abs(W[6] - W[7]) < 0.2 and abs(TL - TG) < 0.2
    # equilibrium lies on the saturation curve
    Ts, _ = po.tsat(TP, W[6], TL)
    assert TL == pytest.approx(Ts, abs=0.1)
    assert U2[1] + U2[2] == pytest.approx(U[1] + U[2], rel=1e-12)
    assert U2[9] == pytest.approx(U2[2], rel=1e-15)
    assert U2[6] + U2[7] == pytest.approx(U[8], rel=1e-12)  # at rest: internal = total
    # brute-force root of p, T, g balance from a different start agrees
    rhoMix, rE = U[1] + U[2], U[8]

    def res(z):
        aG, rG, eG = z
        aL = 1 - aG
        rL = (rhoMix - aG * rG) / aL
        eL = (rE - aG * rG * eG) / (aL * rL)
        pl, pg = po.eqos(TP, 0, 0, rL, eL), po.eqos(TP, 1, 0, rG, eG)
        tl, tg = po.eqos(TP, 0, 1, rL, eL), po.eqos(TP, 1, 1, rG, eG)
        gl = TP["Cp"][0] * tl + TP["eta"][0] - tl * (TP["Cp"][0] / TP["gamma"][0] * math.log10(tl ** TP["gamma"][0] / (pl + TP["pInf"][0]) ** (TP["gamma"][0] - 1)))
        gg = TP["Cp"][1] * tg + TP["eta"][1] - tg * (TP["Cp"][1] / TP["gamma"][1] * math.log10(tg ** TP["gamma"][1] / pg ** (TP["gamma"][1] - 1)) + TP["satConst"][0] * math.log10(pg) + TP["satConst"][1])
        return [(pl - pg) / 1e3, tl - tg, (gl - gg) / 1e3]

    z0 = [W[0] * 0 + (1 - W[0]) * 1.1, W[2] * 0.9, U2[7] / U2[2] * 1.01]
    z = fsolve(res, z0, xtol=1e-13)
    assert z[0] == pytest.approx(1 - U2[0], rel=1e-4)
    assert z[1] == pytest.approx(W[2], rel=1e-4)


def test_tsat_agrees_with_bracketing_root():
    g1, g2 = TP["gamma"]
    Cp1, Cp2 = TP["Cp"]
    Cv1, Cv2 = Cp1 / g1, Cp2 / g2
    for P in (3000.0, 5e4, 2e6):
        eta2p = TP["satConst"][0] * math.log10(P) + TP["satConst"][1]
        A = (Cp1 - Cp2 + eta2p) / (Cp2 - Cv2)
        B = (TP["eta"][0] - TP["eta"][1]) / (Cp2 - Cv2)
        C = (Cp2 - Cp1) / (Cp2 - Cv2)
        D = (Cp1 - Cv1) / (Cp2 - Cv2)
        f = lambda T: A + B / T + C * math.log10(T) + D * math.log10(P + TP["pInf"][0]) - math.log10(P)
        root = brentq(f, 200, 900)
        T, it = po.tsat(TP, P, 300.0)
        assert abs(f(T)) <= 1e-2 and T == pytest.approx(root, abs=0.5) and it < 50


def test_tsat_fit_root_is_the_gibbs_equality():
    """The saturation temperature of relaxation.cu:18-31 (a Newton iteration on
    a log10 fit in A, B, C, D) must be the temperature at which the two
    phases' Gibbs energies of chemRlx.cu:38-50 are equal: the fit is that
    equality rearranged.  The bracketing root of g_L(p, T) - g_G(p, T), from
    the oracle's Gibbs routine, and the oracle's Tsat agree to the Newton
    stop (|fun| <= 1e-2 in log10 units, observed <= 0.1 K) from 1.5 kPa to
    5 MPa.  An oracle without the D log10(p + pInf_L) term is off by tens of
    kelvin (tests/test_oracle_plants.py)."""
    assert tsat_gibbs_gap() <= 0.15


def tsat_gibbs_gap():
    gap = 0.0
    for P in (1500.0, 2340.0, 1e4, 1e5, 1e6, 5e6):
        root = brentq(lambda T: po.gibbs(TP, 0, P, T) - po.gibbs(TP, 1, P, T), 200, 900, xtol=1e-12)
        T, it = po.tsat(TP, P, 300.0)
        if not (np.isfinite(T) and it < 50):
            return np.inf
        gap = max(gap, abs(T - root))
    return gap


def chem_states(n, seed):
    """Chemical-class cells (superheated liquid in a mixture) far from
    equilibrium: volume fractions over the whole range, including nearly pure
    phases, phasic pressures 300 Pa .. 1 MPa and temperatures 290 .. 450 K
    drawn independently."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        k = rng.integers(3)
        aL = [rng.uniform(0.01, 0.99), 10 ** rng.uniform(-2.6, -0.5), 1 - 10 ** rng.uniform(-3, -0.5)][k]
        pL, pG = 10 ** rng.uniform(2.5, 6, 2)
        TL = rng.uniform(290, 450)
        TG = TL + rng.uniform(-40, 40)
        W = np.array([aL, po.eqos(TP, 0, 2, pL, TL), po.eqos(TP, 1, 2, pG, TG), 3.0, 0, 0, pL, pG, 0.0, 0.0])
        U = po.prim2cons(TP, W)
        if po.relax_cell(TP, U)[1] == 1:
            out.append(U)
    return out


CHEM_STATES = []


def load_chem_states():
    """900 chemical cells; the second seed's draw holds states whose full
    Newton step at urf = 1 leaves the volume-fraction range, so that the
    bound on it is exercised."""
    if not CHEM_STATES:
        CHEM_STATES.extend(chem_states(600, 11) + chem_states(300, 35))


def chem_newton_violations(params, states):
    """The bounds the reference's chemical Newton iteration keeps
    (checkUpdate, chemRlx.cu:139-162) on every iterate that does not abort,
    for the non-primary phase k: 0 <= alpha_k <= 1, rho_k >= 0, and
    alpha_k rho_k <= rho_mix when the full step left rho_k >= 0 (after a
    step to rho_k < 0 the reference halves back to rho_k >= 0 but tests
    alpha_k rho_k <= rho with the product of the negative step, which passes:
    a quirk kept as is).  Returns (bound violations, abort-path results that
    are not the best eligible iterate, aborted solves)."""
    viol = bad_best = aborted = 0
    for U in states:
        rho = U[1] + U[2]
        U2, _, rows, ab = po.chem_trace(params, U)
        k = 1 - (0 if rows[0][0] >= rows[0][1] else 1)  # non-primary phase
        for r in rows:
            if r[17] != 0:
                continue
            a, rk = r[11 + k], r[13 + k]
            ok = 0 <= a <= 1 and rk >= 0
            if r[19] >= 0:
                ok = ok and a * rk <= rho
            viol += not ok
        if ab:
            aborted += 1
            elig = [(np.float32(r[10]), i) for i, r in enumerate(rows) if r[9] >= 1]
            if elig:
                best = rows[min(elig)[1]]
                # the abort path restores the best eligible iterate (first
                # minimum of the FP32 residual norm, chemRlx.cu:125-137, 187-192)
                bad_best += U2[1 + k] != best[k] * best[2 + k]
    return viol, bad_best, aborted


def test_chemical_newton_keeps_its_bounds_and_aborts_to_the_best_iterate():
    """The FD-Newton chemical relaxation (chemRlx.cu:78-193) on 900
    far-from-equilibrium chemical cells, at the default under-relaxation and
    at urf = 1, and with at most 3 iterations so that every solve ends on
    the abort path: no iterate that the solve accepts leaves the bounds
    checkUpdate enforces, and every aborted solve returns the eligible
    iterate with the smallest residual norm.  Oracles without the density
    bound, without the volume-fraction bound or returning the last iterate
    on abort fail this (tests/test_oracle_plants.py)."""
    load_chem_states()
    n_ab = 0
    for extra in ({}, dict(chemUrf=1.0), dict(chemItrMax=3)):
        viol, bad_best, ab = chem_newton_violations(dict(TP, **extra), CHEM_STATES)
        assert viol == 0 and bad_best == 0, (extra, viol, bad_best)
        n_ab += ab
    assert n_ab > 900  # the abort path is exercised (all 900 with 3 iterations)


def test_portable_log10_within_two_ulp_of_libm():
    """fkmath::log10 (shared by the oracle and the device relaxation) against
    the C library: log-uniform samples over the double range, values near 1,
    exact powers of ten and the special cases."""
    rng = np.random.default_rng(11)
    xs = np.concatenate([10.0 ** rng.uniform(-307, 308, 20000), 1 + rng.uniform(-0.3, 0.5, 5000),
                         1 + rng.uniform(-1e-8, 1e-8, 1000), [5e-324, 2.2250738585072014e-308, 0.5, 2.0, 1e5]])
    worst = 0.0
    for x in xs:
        a, b = po.log10(x), math.log10(x)
        if b != 0:
            worst = max(worst, abs(a - b) / math.ulp(b))
    assert worst <= 2.0, worst
    assert po.log10(1.0) == 0.0
    for k in range(-20, 21):
        assert abs(po.log10(10.0 ** k) - k) <= 2 * math.ulp(k if k else 1.0)
    assert po.log10(0.0) == -math.inf and po.log10(math.inf) == math.inf
    assert math.isnan(po.log10(-1.0)) and math.isnan(po.log10(math.nan))


def test_gibbs_energy_closed_form():
    """g_k = Cp T + eta - T [(Cp/gamma) log10(T^gamma / (p + pinf)^(gamma-1)) + eta'_k],
    eta'_liquid = 0, eta'_vapour = A log10(p) + B (chemRlx.cu)."""
    p = cases.WATER_VAPOUR
    for k in (0, 1):
        for P, T in ((1e5, 300.0), (3e3, 290.0), (2.3e6, 450.0)):
            et = 0.0 if k == 0 else p["satConst"][0] * math.log10(P) + p["satConst"][1]
            s = (p["Cp"][k] / p["gamma"][k]) * math.log10(T ** p["gamma"][k] / (P + p["pInf"][k]) ** (p["gamma"][k] - 1))
            g = p["Cp"][k] * T + p["eta"][k] - T * (s + et)
            assert po.gibbs(p, k, P, T) == pytest.approx(g, rel=1e-12, abs=1e-6)


def test_portable_exp_within_one_ulp_of_libm():
    """fkmath::exp (the finite-rate relaxation factor, shared by the oracle and
    the device) against the C library over the range where exp is a normal
    double, near 0, and at the special cases."""
    rng = np.random.default_rng(12)
    xs = np.concatenate([rng.uniform(-708, 709.7, 20000), rng.uniform(-1, 1, 5000), rng.uniform(-1e-8, 1e-8, 1000),
                         [0.5, 1.0, -1.0, 2.0, 100.0, -100.0, 709.0, -707.9]])
    worst = max(abs(po.exp(x) - math.exp(x)) / math.ulp(math.exp(x)) for x in xs)
    assert worst <= 1.0, worst
    assert po.exp(0.0) == 1.0 and po.exp(-0.0) == 1.0
    assert po.exp(-800.0) == 0.0 and po.exp(-math.inf) == 0.0
    assert po.exp(710.0) == math.inf and po.exp(math.inf) == math.inf
    assert math.isnan(po.exp(math.nan))


def fnt_closed_form(params, avf, rho, ener):
    """fntPressRlx (thermRlx.cu:42-67) in closed form: with de/drho =
    -(gamma pinf + p) / ((gamma - 1) rho^2) and de/dp = 1 / ((gamma - 1) rho),
    c_k = (p / rho^2 - de/drho) / (de/dp) = gamma (p + pinf) / rho."""
    g, pi = params["gamma"], params["pInf"]
    p = [po.eqos(params, k, 0, rho[k], ener[k]) for k in (0, 1)]
    T = [po.eqos(params, k, 1, rho[k], ener[k]) for k in (0, 1)]
    kappa = avf[1] * params["thermRlxRate"] * sum(rho[k] * g[k] * (p[k] + pi[k]) / rho[k] / avf[k] for k in (0, 1))
    dP = (p[0] - p[1]) * math.exp(-abs(kappa) * params["dt"])
    theta = dP / (p[0] - p[1]) if p[0] - p[1] > 1e-5 else 0.0
    return dP, theta * (T[0] - T[1]), kappa


@pytest.mark.parametrize("aL,pL,pG,TL,TG", [(0.7, 2e6, 1e6, 350.0, 340.0), (0.3, 3e5, 1e5, 300.0, 320.0),
                                            (0.999, 1e7, 2340.0, 293.15, 293.15), (0.5, 5e5, 4e5, 400.0, 390.0)])
def test_finite_rate_targets_closed_form(aL, pL, pG, TL, TG):
    avf = [aL, 1 - aL]
    rho = [po.eqos(TP, 0, 2, pL, TL), po.eqos(TP, 1, 2, pG, TG)]
    ener = [po.eqos(TP, 0, 3, pL, rho[0]), po.eqos(TP, 1, 3, pG, rho[1])]
    base = dict(TP, thermFntRlxFlag=1, thermRlxRate=0.5, dt=1.0)
    _, _, kappa = fnt_closed_form(base, avf, rho, ener)
    for kdt in (0.05, 1.0, 3.0):  # relaxation factor exp(-kappa dt) from 0.95 to 0.05
        prm = dict(base, dt=kdt / kappa)
        dP, dT = po.fnt_press_rlx(prm, avf, rho, ener)
        dP_ref, dT_ref, _ = fnt_closed_form(prm, avf, rho, ener)
        assert dP == pytest.approx(dP_ref, rel=1e-5), kdt
        assert dT == pytest.approx(dT_ref, rel=1e-5, abs=1e-5), kdt
    # dt = 0: no relaxation, the targets are the current differences (rounded to FP32)
    dP, dT = po.fnt_press_rlx(dict(base, dt=0.0), avf, rho, ener)
    assert dP == np.float32(np.float32(pL) - np.float32(pG))
    assert dT == pytest.approx(TL - TG, rel=1e-5, abs=1e-4)


def test_finite_rate_temperature_target_needs_liquid_overpressure():
    """theta is dP / (p_L - p_G) only where p_L - p_G > 1e-5; otherwise the
    temperature target is 0 (thermRlx.cu:65-66)."""
    avf = [0.6, 0.4]
    rho = [po.eqos(TP, 0, 2, 1e5, 320.0), po.eqos(TP, 1, 2, 3e5, 300.0)]
    ener = [po.eqos(TP, 0, 3, 1e5, rho[0]), po.eqos(TP, 1, 3, 3e5, rho[1])]
    prm = dict(TP, thermFntRlxFlag=1, thermRlxRate=0.5, dt=1e-9)
    dP, dT = po.fnt_press_rlx(prm, avf, rho, ener)
    dP_ref, _, _ = fnt_closed_form(prm, avf, rho, ener)
    assert dP < 0 and dP == pytest.approx(dP_ref, rel=1e-5)
    assert dT == 0.0


def test_finite_rate_relaxation_reaches_its_targets():
    """With the model on, the thermal Newton solve ends at p_L - p_G = dP and
    T_L - T_G = dT (to the solver's tolerance) and conserves what the
    equilibrium solve conserves; a rate so large that exp(-kappa dt) = 0 gives
    the equilibrium solve bit for bit."""
    for aL, pL, pG in [(0.7, 2e6, 1e6), (0.3, 1e5, 3e5), (0.5, 5e5, 4e5)]:
        U = mixture_state(aL, pL, pG, 350.0)
        off = dict(TP, chemFlag=0)
        U_eq, _, _, _ = po.relax_cell(off, U)
        U_huge, kind, _, _ = po.relax_cell(dict(off, thermFntRlxFlag=1, thermRlxRate=1e30, dt=1e-6), U)
        assert kind == 0 and np.array_equal(U_huge, U_eq)
        # relaxation factor exp(-kappa dt) = 0.5
        a = np.clip(U[0], TP["avfMin"] / 100, 1 - TP["avfMin"] / 100)
        avf, rho = [a, 1 - a], [U[1] / a, U[2] / (1 - a)]
        rhoMix = U[1] + U[2]
        de = U[8] / rhoMix - (U[6] + U[7]) / rhoMix - 0.5 * (U[3:6] @ U[3:6]) / rhoMix ** 2
        ener = [U[6] / U[1] + de, U[7] / U[2] + de]
        _, _, kappa = fnt_closed_form(dict(off, thermRlxRate=0.5, dt=1.0), avf, rho, ener)
        prm = dict(off, thermFntRlxFlag=1, thermRlxRate=0.5, dt=math.log(2) / kappa)
        dP, dT = po.fnt_press_rlx(prm, avf, rho, ener)
        U2, kind, _, _ = po.relax_cell(prm, U)
        assert kind == 0
        W = po.cons2prim(TP, U2)
        TLq = po.eqos(TP, 0, 1, W[1], U2[6] / (W[0] * W[1]))
        TGs = po.eqos(TP, 1, 1, W[2], U2[7] / ((1 - W[0]) * W[2]))
        assert abs((W[6] - W[7]) - dP) < 0.1 and abs((TLq - TGs) - dT) < 1e-2
        assert abs(dP) > 1e3 and abs(dP - 0.5 * (pL - pG)) < 0.05 * abs(pL - pG)
        assert np.array_equal(U2[[1, 2, 3, 4, 5, 8, 9]], U[[1, 2, 3, 4, 5, 8, 9]])
        ke = 0.5 * (U[3:6] @ U[3:6]) / rhoMix
        assert U2[6] + U2[7] == pytest.approx(U[8] - ke, rel=1e-12)


@pytest.mark.parametrize("model", [0, 1])
def test_inlet_ghost_takes_configured_state(model):
    """Inlet (boundCond 3): fk_params inletP / inletT set the ghost's pressure
    and temperature, rho_k = (p + pinf_k) / (Cv_k T (gamma_k - 1)); both 0
    keep the constants of bc.h:5-7."""
    params = SP if model == 0 else TP
    n = np.array([1.0, 0.0, 0.0])
    W = po.cons2prim(params, cases.sod().U[0]) if model == 0 else po.cons2prim(params, cases.bubble(n=2).U[0])
    P, T = 3.3e6, 310.0
    g = po.bc_ghost(dict(params, inletP=P, inletT=T), BC_INLET, n, W)
    gam, pi, cp = params["gamma"], params["pInf"], params["Cp"]
    rho = [(P + pi[k]) / (cp[k] / gam[k] * T * (gam[k] - 1)) for k in (0, 1)]
    if model == 0:
        assert g[4] == P and g[0] == pytest.approx(rho[0], rel=1e-14)
        d = po.bc_ghost(params, BC_INLET, n, W)
        assert d[4] == 4.2 * 101300 and d[0] == pytest.approx(4.2 * 101300 / (1006 / 1.4 * 288 * 0.4), rel=1e-14)
    else:
        assert g[6] == g[7] == P and g[0] == 1 - params["avfMin"] and g[8] == g[9] == 0
        assert g[1] == pytest.approx(rho[0], rel=1e-14) and g[2] == pytest.approx(rho[1], rel=1e-14)
        d = po.bc_ghost(params, BC_INLET, n, W)
        assert d[6] == 11.4e5 and d[1] == pytest.approx((11.4e5 + pi[0]) / (cp[0] / gam[0] * 115.3 * (gam[0] - 1)),
                                                         rel=1e-14)
