"""Case definitions: solver constants (the reference INPUT struct, see
mulitphase Solver/doublePrecision/buildInput.cu:10-69) and initial states of
the five workloads, each at a configurable resolution.

  sod            3D Sod tube, 32x8x8 hexes x 6 tets, single-phase FP64
  tgv            WALE LES Taylor-Green vortex, 64^3 periodic jittered hexes x 6
  bubble         water/vapour bubble collapse, 128^3 jittered hexes x 6
  shock_bubble   5e7 Pa liquid shock into 64 vapour bubbles, 200^3 hexes x 6

Water/vapour constants.  gamma_L = 6.12, pInf_L = 3.43e8 (stiffened-gas
water) and an ideal-gas vapour (gamma 1.327, Cp 1880) are given; the
remaining constants are chosen here: Cp_L makes rho_L(1e7 Pa, 293.15 K) =
998 kg/m^3, and (eta_L - eta_G, satConst[1]) fit the saturation relation of
relaxation.cu:18-31 (with satConst[0] = 0) through Tsat(2339 Pa) = 293.15 K
and Tsat(101325 Pa) = 373.15 K; eta_G = 2e6 J/kg fixes the energy origin.
"""
import math

import numpy as np

from . import _lib
from .mesh import BC_OUTLET, BC_SLIP, Mesh, place_bubbles

WATER_VAPOUR = dict(
    gamma=[6.12, 1.327],
    pInf=[3.43e8, 0.0],
    Cp=[1442.2339904096102, 1880.0],
    eta=[1026161.9362866953, 2.0e6],
    satConst=[0.0, -6057.965655177897],
    dynVisc=[1.0e-3, 1.0e-5],
    Pr=[7.0, 1.0],
)

BASE = dict(
    model=_lib.MODEL_SINGLE, precision=_lib.FP64, rkSteps=3, adaptDt=1, adaptDtItr=1, viscFlag=0,
    turbulence=0, heatFlux=0, woodsFlag=1, neglectX=0, neglectY=0, neglectZ=0, chemFlag=1,
    thermItrMax=25, chemItrMax=40, hllcEnergy=0, relaxFlag=1, relaxMode=0,
    dt=0.0, cfl=0.5, maxDt=1e30, avfMin=1e-4, thermUrf=0.8, chemUrf=0.5, chemLimit=20e-4,
    satTempUrf=0.8, polishTol=1e-10,
    gamma=[1.4, 1.4], pInf=[0.0, 0.0], Cp=[1006.0, 1006.0], eta=[0.0, 0.0], satConst=[0.0, 0.0],
    dynVisc=[0.0, 0.0], Pr=[0.71, 0.71],
)


class Case:
    """A workload: constants, mesh and initial conserved state."""

    def __init__(self, name, params, mesh, U, steps):
        self.name, self.params, self.mesh, self.U, self.steps = name, params, mesh, U, steps

    @property
    def neq(self):
        return 5 if self.params["model"] == _lib.MODEL_SINGLE else 10

    def fk_params(self, **over):
        d = dict(self.params)
        d.update(over)
        return _lib.FkParams.from_dict(d)


def sod(nx=32, ny=8, nz=8, steps=100, rcm=True):
    """3D Sod tube on [0,1]x[0,0.25]x[0,0.25]: transmissive x ends (outlet,
    code 2), slip walls (code 4), gamma = 1.4, x<0.5 (1,0,1) | (0.125,0,0.1),
    HLLC + MUSCL, SSP-RK3, CFL 0.5."""
    p = dict(BASE, model=_lib.MODEL_SINGLE, gamma=[1.4, 1.4], pInf=[0.0, 0.0], Cp=[1006.0, 1006.0])
    m = Mesh.box(nx, ny, nz, (0, 0, 0), (1.0, 0.25, 0.25), 0.0, 0,
                 bc=(BC_OUTLET, BC_OUTLET, BC_SLIP, BC_SLIP, BC_SLIP, BC_SLIP), rcm=rcm)
    U = m.init_sod(1.4, 0.0, 0.5, (1.0, 0, 0, 0, 1.0), (0.125, 0, 0, 0, 0.1))
    return Case("sod", p, m, U, steps)


def tgv(n=64, steps=200, seed=1, rcm=True):
    """Taylor-Green vortex on the periodic box [0,2pi]^3, interior nodes
    jittered +-10 % of the spacing, Mach 0.1, Re 1600 (U0 = rho0 = L = 1),
    Pr 0.71, WALE (Cw = 0.325) and the Fourier heat flux."""
    gamma, mach, re = 1.4, 0.1, 1600.0
    p0 = 1.0 / (gamma * mach * mach)  # background pressure rho0 U0^2 / (gamma Ma^2)
    # pRef = p0: the mixed path carries the FP32 face pressures as p - p0
    p = dict(BASE, model=_lib.MODEL_SINGLE, gamma=[gamma, gamma], pInf=[0.0, 0.0], Cp=[1006.0, 1006.0],
             viscFlag=1, turbulence=1, heatFlux=1, dynVisc=[1.0 / re, 1.0 / re], Pr=[0.71, 0.71], pRef=p0, cfl=0.3)
    two_pi = 2 * math.pi
    m = Mesh.box(n, n, n, (0, 0, 0), (two_pi,) * 3, 0.10, seed, periodic=(True, True, True), rcm=rcm)
    U = m.init_tgv(gamma, 1.0, 1.0, mach)
    return Case("tgv", p, m, U, steps)


def _two_phase_params(**over):
    # cfl = 0.1 as in the reference (buildInput.cu:16): at 0.5 the liquid
    # pressure in interface cells overshoots to -1e8 Pa within one stage and
    # the chemical relaxation then converges to non-physical roots.
    p = dict(BASE, model=_lib.MODEL_TWO_PHASE, cfl=0.1, **WATER_VAPOUR)
    p.update(over)
    return p


T_WATER = 293.15


def bubble(n=128, steps=100, seed=2, rcm=True):
    """Vapour sphere r = 0.15 at rest at 2340 Pa in the centre of the unit
    box, liquid at 1e7 Pa and 998 kg/m^3 (293.15 K), transmissive walls."""
    p = _two_phase_params()
    m = Mesh.box(n, n, n, (0, 0, 0), (1, 1, 1), 0.10, seed, bc=(BC_OUTLET,) * 6, rcm=rcm)
    am = p["avfMin"]
    regions = [dict(kind=0, alphaL=1 - am, p=1e7, T=T_WATER),
               dict(kind=1, c=(0.5, 0.5, 0.5), r=0.15, alphaL=am, p=2340.0, T=T_WATER)]
    U = m.init_regions(p, regions)
    return Case("bubble", p, m, U, steps)


def shock_bubble(n=200, steps=50, seed=4, nbub=64, rcm=True, order=None):
    """Planar 5e7 Pa liquid shock (rho 1005) entering from x = 0 into liquid
    at 1e5 Pa holding nbub vapour bubbles (radii U[0.02, 0.05], seeded,
    non-overlapping, x >= 0.2); WALE on."""
    p = _two_phase_params(viscFlag=1, turbulence=1)
    m = Mesh.box(n, n, n, (0, 0, 0), (1, 1, 1), 0.10, seed, bc=(BC_OUTLET,) * 6, rcm=rcm, order=order)
    am = p["avfMin"]
    gL, piL, cpL = p["gamma"][0], p["pInf"][0], p["Cp"][0]
    cvL = cpL / gL
    p0, p1, rho1 = 1e5, 5e7, 1005.0
    rho0 = (p0 + piL) / (cvL * T_WATER * (gL - 1))
    T1 = (p1 + piL) / (cvL * rho1 * (gL - 1))  # post-shock temperature from rho = 1005
    u1 = math.sqrt(max((p1 - p0) * (rho1 - rho0) / (rho0 * rho1), 0.0))
    regions = [dict(kind=0, alphaL=1 - am, p=p0, T=T_WATER)]
    centres, radii = place_bubbles(nbub, (0, 0, 0), (1, 1, 1), 0.02, 0.05, 0.2, seed)
    for c, r in zip(centres, radii):
        regions.append(dict(kind=1, c=tuple(c), r=float(r), alphaL=am, p=2340.0, T=T_WATER))
    regions.append(dict(kind=2, c=(0.1, 0, 0), alphaL=1 - am, p=p1, T=T1, u=(u1, 0, 0)))
    U = m.init_regions(p, regions)
    return Case("shock_bubble", p, m, U, steps)


CASES = {"sod": sod, "tgv": tgv, "bubble": bubble, "shock_bubble": shock_bubble}
