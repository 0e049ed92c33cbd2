"""Field export anchored on the oracle: the node fields (export_plt.cu:41-93)
and the Tecplot brick file of a state the GPU advanced, with the GPU's own
Green-Gauss gradient, against the same export of the oracle's state and
gradient after the same steps."""
import numpy as np
import pytest

from fooskagpu_amd import GpuSolver, _lib, export
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def exports(c, steps, precision=_lib.FP64):
    o = Oracle(c.params, c.mesh, c.U)
    o.run_steps(steps)
    o.compute_face_values()
    Uo, go = o.state(), o.get("grad")
    with GpuSolver(dict(c.params, precision=precision), c.mesh, c.U) as g:
        g.run_steps(steps)
        g.set_debug(True)
        g.compute_face_values()
        Ug, gg = g.download_state(), g.download("grad")
    return export.node_fields(c.mesh, Ug, c.params, grad=gg), export.node_fields(c.mesh, Uo, c.params, grad=go)


@pytest.mark.parametrize("name", ["shock_bubble", "sod", "tgv", "mixed_elements_tp"])
def test_export_of_the_gpu_state_equals_the_oracle_export(gpu, tmp_path, name):
    """FP64: every node field bitwise equal, and the two field_%07d.dat files
    identical byte for byte."""
    from test_gpu_parity import SMALL
    c = SMALL[name]()
    ng, no = exports(c, 5)
    for k in no:
        assert np.array_equal(ng[k], no[k]), k
    export.write_plt(tmp_path / "g.dat", c.mesh, ng, time=1.0)
    export.write_plt(tmp_path / "o.dat", c.mesh, no, time=1.0)
    assert (tmp_path / "g.dat").read_bytes() == (tmp_path / "o.dat").read_bytes()


def test_mixed_precision_export_within_the_density_bar(gpu):
    """Mixed precision: node density within the 1e-5 bar of the state; the
    gradient magnitude (FP32 Green-Gauss) to 1e-3."""
    from test_gpu_parity import SMALL
    c = SMALL["shock_bubble"]()
    ng, no = exports(c, 5, _lib.MIXED)
    rel = lambda k: np.abs(ng[k] - no[k]).sum() / np.abs(no[k]).sum()
    print(f"mixed export: rho {rel('rho'):.2e}, Pressure {rel('Pressure'):.2e}, gradRho {rel('gradRho'):.2e}")
    assert rel("rho") <= 1e-5 and rel("gradRho") <= 1e-3
