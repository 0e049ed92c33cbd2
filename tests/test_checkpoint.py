"""Checkpoint file format (CPU) and bit-exact restart of the device loop (GPU)."""
import numpy as np
import pytest

from fooskagpu_amd import checkpoint


class FakeSolver:
    """Stands in for GpuSolver in the file-format tests (no device)."""

    def __init__(self, U, clock):
        self.U, self._clock = U.copy(), clock
        self.n_cells, self.neq = U.shape

    def download_state(self):
        return self.U.copy()

    def upload_state(self, U):
        self.U = np.array(U)

    def clock(self):
        return self._clock

    def set_clock(self, dt, time, step):
        self._clock = (dt, time, step)


def test_round_trip_is_bit_exact(tmp_path):
    rng = np.random.default_rng(0)
    U = rng.normal(size=(1000, 10)) * 10.0 ** rng.integers(-300, 300, size=(1000, 10))
    src = FakeSolver(U, (1.25e-7, 3.5e-5, 42))
    checkpoint.save(tmp_path / "a.ck", src)
    dst = FakeSolver(np.zeros_like(U), (0.0, 0.0, 0))
    assert checkpoint.load(tmp_path / "a.ck", dst) == (1.25e-7, 3.5e-5, 42)
    assert np.array_equal(dst.U.view(np.uint64), U.view(np.uint64))


def test_corruption_and_mismatch_are_detected(tmp_path):
    U = np.arange(50.0).reshape(10, 5)
    checkpoint.save(tmp_path / "a.ck", FakeSolver(U, (1.0, 2.0, 3)))
    raw = bytearray((tmp_path / "a.ck").read_bytes())
    flipped = bytearray(raw)
    flipped[-3] ^= 0x10
    (tmp_path / "b.ck").write_bytes(bytes(flipped))
    with pytest.raises(checkpoint.CheckpointError, match="checksum"):
        checkpoint.read(tmp_path / "b.ck")
    # two words swapped: same plain sum, different weighted sum
    V = U.copy()
    V[0, 0], V[0, 1] = V[0, 1], V[0, 0]
    assert checkpoint.checksum(V) != checkpoint.checksum(U)
    (tmp_path / "c.ck").write_bytes(bytes(raw[:-8]))
    with pytest.raises(checkpoint.CheckpointError, match="truncated"):
        checkpoint.read(tmp_path / "c.ck")
    (tmp_path / "d.ck").write_bytes(b"NOTACKPT" + bytes(raw[8:]))
    with pytest.raises(checkpoint.CheckpointError, match="not a checkpoint"):
        checkpoint.read(tmp_path / "d.ck")
    with pytest.raises(checkpoint.CheckpointError, match="does not match"):
        checkpoint.load(tmp_path / "a.ck", FakeSolver(np.zeros((10, 10)), (0, 0, 0)))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["shock_bubble", "tgv"])
def test_restart_continues_bit_exactly(gpu, tmp_path, name):
    """6 steps in one go == 3 steps, checkpoint, a new context, 3 steps
    (adaptive dt recomputed every 2 steps, so the step counter matters)."""
    from fooskagpu_amd import GpuSolver
    from test_gpu_parity import SMALL
    c = SMALL[name]()
    p = dict(c.params, adaptDtItr=2)
    with GpuSolver(p, c.mesh, c.U) as g:
        g.run_steps(6)
        ref = g.download_state()
        ref_clock = g.clock()
    with GpuSolver(p, c.mesh, c.U) as g:
        g.run_steps(3)
        checkpoint.save(tmp_path / "mid.ck", g)
    with GpuSolver(p, c.mesh, c.U) as g:
        checkpoint.load(tmp_path / "mid.ck", g)
        g.run_steps(3)
        assert np.array_equal(g.download_state(), ref)
        assert g.clock() == ref_clock


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["shock_bubble", "tgv", "mixed_elements_tp"])
def test_restart_equals_the_oracles_uninterrupted_run(gpu, tmp_path, name):
    """The restart anchored on the oracle: 3 steps on the device, checkpoint,
    a new context loads it and runs 3 more steps; the state and the clock
    equal the CPU oracle's 6 uninterrupted steps bit for bit (FP64, adaptive
    dt every 2 steps, so the restored step counter decides the dt steps).
    Replaces writeFieldRestart / readField (restart.cu:113-152, 424-497),
    whose %.10e text is not bit-exact."""
    from fooskagpu_amd import GpuSolver
    from oracle.pyoracle import Oracle
    from test_gpu_parity import SMALL
    c = SMALL[name]()
    p = dict(c.params, adaptDtItr=2)
    o = Oracle(p, c.mesh, c.U)
    dto, to = o.run_steps(6)
    Uo = o.state()
    with GpuSolver(p, c.mesh, c.U) as g:
        g.run_steps(3)
        checkpoint.save(tmp_path / "mid.ck", g)
    with GpuSolver(p, c.mesh, c.U) as g:
        checkpoint.load(tmp_path / "mid.ck", g)
        dtg, tg = g.run_steps(3)
        Ug = g.download_state()
        step = g.clock()[2]
    assert np.isfinite(Ug).all()
    assert np.array_equal(Ug, Uo)
    assert dtg == dto and tg == to and step == 6
