"""The oracle's pins catch the mistakes they are meant to catch: oracle.cpp
is rebuilt with one planted mistake at a time (an ORC_PLANT_* switch, never
set in the real build) and each pin of tests/test_oracle.py that targets that
part of the reference must then fail.  CPU only."""
import os
import subprocess

import pytest

import test_oracle as t
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

PLANTS = ["ENERGY_P_SIGN", "TSAT_NO_D", "NO_DENSITY_BOUND", "NO_ALPHA_BOUND", "ABORT_KEEPS_LAST"]


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    d = tmp_path_factory.mktemp("plants")
    out = {}
    for name in PLANTS:
        so = str(d / f"liboracle_{name.lower()}.so")
        subprocess.run(["g++", "-I" + os.path.join(ROOT, "fooskagpu_amd", "csrc"), "-O1", "-std=c++17",
                        "-ffp-contract=off", "-fPIC", "-shared", f"-DORC_PLANT_{name}",
                        os.path.join(ORACLE, "oracle.cpp"), "-o", so], check=True)
        out[name] = so
    return out


def test_energy_star_state_with_a_wrong_sign_is_caught(planted):
    """(S* - p / (rho_K (S_K - u_K))) for (S* + p / (rho_K (S_K - u_K))): the
    unit-density vapour pulse no longer matches the ideal gas."""
    with po.variant(planted["ENERGY_P_SIGN"]):
        d, _, _ = t.vapour_pulse_energy_change(t.T_UNIT_VAPOUR, 2e3, 0)
    assert d > 5e-3, d


def test_tsat_fit_without_its_pinf_term_is_caught(planted):
    with po.variant(planted["TSAT_NO_D"]):
        gap = t.tsat_gibbs_gap()
    assert gap > 0.15, gap


@pytest.mark.parametrize("name,extra,what", [
    ("NO_DENSITY_BOUND", {}, "viol"), ("NO_DENSITY_BOUND", dict(chemUrf=1.0), "viol"),
    ("NO_ALPHA_BOUND", dict(chemUrf=1.0), "viol"), ("ABORT_KEEPS_LAST", {}, "best"),
    ("ABORT_KEEPS_LAST", dict(chemItrMax=3), "best")])
def test_newton_bound_and_abort_mistakes_are_caught(planted, name, extra, what):
    t.load_chem_states()
    with po.variant(planted[name]):
        viol, bad_best, _ = t.chem_newton_violations(dict(t.TP, **extra), t.CHEM_STATES)
    assert (viol if what == "viol" else bad_best) > 0, (viol, bad_best)
