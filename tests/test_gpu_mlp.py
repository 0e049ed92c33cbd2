"""GPU tests of the relaxation surrogate: BF16-MFMA forward and backward
against the CPU reference network, training, and the warm-started FP64
Newton polish against the plain reference relaxation."""
import numpy as np
import pytest

from fooskagpu_amd import GpuSolver, _lib, cases, surrogate
from oracle import mlp

pytestmark = pytest.mark.gpu


def small_solver(n=10):
    c = cases.bubble(n=n)
    return c, GpuSolver(c.params, c.mesh, c.U)


def random_setup(sg, n=1000, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3)).astype(np.float32) * [2.0, 3e5, 0.3] + [5.0, 1.4e6, 0.5]
    y = rng.normal(size=(n, 3)).astype(np.float32) * [20.0, 2.0, 0.3] + [300.0, 5.0, 0.5]
    norm = surrogate.fit_norm(x, y)
    w = surrogate.init_weights(seed + 1)
    w[4352:4416] = rng.normal(scale=0.1, size=64)  # non-zero b2
    sg.set_weights(w, norm)
    sg.set_data(x, y)
    return x, y, w, norm


def test_forward_matches_cpu_reference(gpu):
    c, g = small_solver()
    with g:
        sg = surrogate.Surrogate(g)
        x, _, w, norm = random_setup(sg)
        yg = sg.infer(x)
    xn = (x - norm[0:3]) / norm[3:6]
    yn, _ = mlp.forward(w, xn, bf16_round=True)
    yc = yn * norm[9:12] + norm[6:9]
    err = np.abs(yg - yc) / norm[9:12]
    print(f"forward: max |dy|/std {err.max():.2e}, mean {err.mean():.2e}")
    assert err.max() < 5e-3 and err.mean() < 5e-4


def test_backward_matches_cpu_reference(gpu):
    c, g = small_solver()
    with g:
        sg = surrogate.Surrogate(g)
        x, y, w, norm = random_setup(sg, n=300, seed=4)
        idx = np.arange(0, 300, 1)
        gg, loss = sg.grad(idx)
    xn = (x - norm[0:3]) / norm[3:6]
    tn = (y - norm[6:9]) / norm[9:12]
    gc = mlp.backward(w, xn[idx], tn[idx], bf16_round=True)
    lc = mlp.loss(w, xn[idx], tn[idx], bf16_round=True)
    rel = np.linalg.norm(gg - gc) / np.linalg.norm(gc)
    cos = gg @ gc / (np.linalg.norm(gg) * np.linalg.norm(gc))
    print(f"backward: relative L2 {rel:.2e}, cosine {cos:.6f}, loss gpu {loss:.6f} cpu {lc:.6f}")
    assert rel < 1e-5 and cos > 0.9999999
    assert loss == pytest.approx(lc, rel=1e-3)
    for name, (a, b) in {"W1": (0, 192), "W2": (256, 4352), "W3": (4416, 8512), "W4": (8576, 8768)}.items():
        r = np.linalg.norm(gg[a:b] - gc[a:b]) / np.linalg.norm(gc[a:b])
        assert r < 1e-5, (name, r)


def test_training_reduces_loss_and_is_deterministic(gpu):
    out = []
    for _ in range(2):
        c, g = small_solver()
        with g:
            sg = surrogate.Surrogate(g)
            U = surrogate.uniform_states(40000, c.params, seed=2)
            x, y = surrogate.finite(*sg.label(U))
            norm = surrogate.fit_norm(x, y)
            sg.set_data(x, y)
            sg.set_weights(surrogate.init_weights(0), norm)
            l0, l1 = sg.train(500, batch=2048, lr=2e-3, seed=7)
            out.append((l0, l1, sg.weights()[0]))
    print(f"loss {out[0][0]:.4f} -> {out[0][1]:.5f}")
    assert out[0][1] < 0.3 * out[0][0]
    assert np.array_equal(out[0][2], out[1][2])


def test_surrogate_relaxation_agrees_with_plain_relaxation(gpu):
    """The surrogate path changes the Newton iterations, not the converged
    equilibrium: exact classification + cold-start screen + MLP warm start +
    FP64 polish vs the plain reference solve, on hydro-evolved states of the
    cavitation case (thermal and chemical cells)."""
    from test_gpu_parity import cavitation
    c = cavitation(10)
    p = dict(c.params, relaxFlag=0)
    with GpuSolver(p, c.mesh, c.U) as g:
        g.run_steps(2)
        U = g.download_state()
        sg = surrogate.Surrogate(g)
        xs, ys = sg.label(U)
        xu, yu = sg.label(surrogate.uniform_states(100000, c.params, seed=3))
        x, y = surrogate.finite(np.concatenate([xs, xu]), np.concatenate([ys, yu]))
        sg.set_data(x, y)
        sg.set_weights(surrogate.init_weights(0), surrogate.fit_norm(x, y))
        sg.train(600, batch=2048, lr=2e-3, seed=1)
        g.upload_state(U)
        st_plain = g.relaxation(_lib.RELAX_PLAIN)
        Up = g.download_state()
        g.upload_state(U)
        st_mlp = g.relaxation(_lib.RELAX_MLP)
        Um = g.download_state()
    print(f"plain {st_plain}\nmlp   {st_mlp}")
    err = np.abs(Um - Up).sum(0) / np.maximum(np.abs(Up).sum(0), 1e-300)
    print("rel L1 per variable", err)
    assert st_mlp["n_chem"] == st_plain["n_chem"] and st_plain["n_chem"] > 0
    assert st_mlp["newton_iters"] < st_plain["newton_iters"]
    assert st_mlp["n_fallback"] < 0.05 * len(U)
    assert err.max() < 1e-6


def test_network_warm_start_beats_cold_start(gpu):
    """What the network adds: on states far from equilibrium (phasic
    pressures and temperatures drawn independently), the same polish of the
    chemical cells from the MLP prediction needs fewer Newton iterations and
    fallbacks than from the cell state (FK_RELAX_COLD), and both reach the
    plain solution."""
    c, g = small_solver(n=20)
    with g:
        sg = surrogate.Surrogate(g)
        x, y = surrogate.finite(*sg.label(surrogate.uniform_states(300000, c.params, seed=6)))
        print(f"{len(x)} training states")
        sg.set_data(x, y)
        sg.set_weights(surrogate.init_weights(0), surrogate.fit_norm(x, y))
        l0, l1 = sg.train(1500, batch=4096, lr=2e-3, seed=2)
        U = surrogate.uniform_states(g.n_cells, c.params, seed=5)
        out = {}
        for name, mode in (("plain", _lib.RELAX_PLAIN), ("cold", _lib.RELAX_COLD), ("mlp", _lib.RELAX_MLP)):
            g.upload_state(U)
            out[name] = (g.relaxation(mode), g.download_state())
    for k, (st, _) in out.items():
        print(f"{k:5s} {st}")
    print(f"training loss {l0:.4f} -> {l1:.5f}")
    cold, mlp = out["cold"][0], out["mlp"][0]
    assert cold["n_warm"] == 0 and mlp["n_warm"] == mlp["n_chem"] > 0.1 * len(U)
    assert mlp["newton_iters"] < 0.9 * cold["newton_iters"]
    assert mlp["n_fallback"] <= cold["n_fallback"]
    # far from equilibrium the plain solve often stops short (absolute
    # tolerances, abort to its best iterate): compare with it where it reached
    # p-T equilibrium, and require the polished paths to reach it at least as often
    Up = out["plain"][1]
    ok_p = at_equilibrium(Up, c.params)
    scale = np.maximum(np.abs(Up).mean(0), 1e-300)
    for k in ("cold", "mlp"):
        ok_k = at_equilibrium(out[k][1], c.params)
        both = ok_p & ok_k
        d = (np.abs(out[k][1] - Up) / scale)[both].max()
        print(f"{k}: at equilibrium {ok_k.mean():.4f} (plain {ok_p.mean():.4f}); max scaled diff where both are {d:.2e}")
        assert ok_k.mean() >= ok_p.mean()
        assert d < 1e-5


def at_equilibrium(U, p):
    """Cells whose phases are at p-T equilibrium (the relaxation targets)."""
    g, pi, cp, eta = (np.array(p[k]) for k in ("gamma", "pInf", "Cp", "eta"))
    am = p["avfMin"] / 100
    a = np.clip(U[:, 0], am, 1 - am)
    r = np.stack([U[:, 1] / a, U[:, 2] / (1 - a)], 1)
    e = np.stack([U[:, 6] / U[:, 1], U[:, 7] / U[:, 2]], 1)
    pk = (g - 1) * r * (e - eta) - g * pi
    T = (e - eta - pi / r) * g / cp
    return (np.abs(pk[:, 0] - pk[:, 1]) <= 1e-6 * np.maximum(np.abs(pk[:, 0]), 1e3)) & \
        (np.abs(T[:, 0] - T[:, 1]) <= 1e-3)


def test_surrogate_steps_without_weights_fail_before_any_stage(gpu):
    from fooskagpu_amd._lib import FkError
    c = cases.bubble(n=8)
    with GpuSolver(dict(c.params, relaxMode=_lib.RELAX_MLP), c.mesh, c.U) as g:
        with pytest.raises(FkError, match="weights"):
            g.run_steps(2)
        assert np.array_equal(g.download_state(), c.U) and g.clock()[2] == 0


def test_surrogate_classification_check_fails_before_any_stage(gpu):
    """The exact classification of the surrogate relaxation needs
    eta_L <= eta_G and Cp_G >= Cp_L: with Cp_G < Cp_L fk_run_steps refuses
    before the first RK stage, leaving the state and the clock untouched."""
    from fooskagpu_amd._lib import FkError
    c = cases.bubble(n=8)
    p = dict(c.params, relaxMode=_lib.RELAX_COLD, Cp=[c.params["Cp"][0], c.params["Cp"][0] * 0.5])
    with GpuSolver(p, c.mesh, c.U) as g:
        with pytest.raises(FkError, match="Cp_G"):
            g.run_steps(2)
        assert np.array_equal(g.download_state(), c.U) and g.clock()[2] == 0


def test_network_halves_the_newton_work_on_chemical_cells(gpu):
    """What the network adds where it is used: on chemical-class states
    whose equilibrium exists (those the labelling gives a finite target: the
    (ln p, T) solve from the cell state reached the p-T-g equilibrium), the polish
    from the MLP prediction needs at most 0.6x the Newton iterations of the
    same polish from the cell state (FK_RELAX_COLD), falls back to the plain
    solve for fewer than 5 % of the cells, and leaves at least 95 % of them
    at p-T equilibrium."""
    c, g = small_solver(n=20)
    with g:
        sg = surrogate.Surrogate(g)
        x, y = surrogate.finite(*sg.label(surrogate.uniform_states(300000, c.params, seed=6)))
        sg.set_data(x, y)
        sg.set_weights(surrogate.init_weights(0), surrogate.fit_norm(x, y))
        l0, l1 = sg.train(3000, batch=4096, lr=2e-3, seed=2)
        # chemical states with an equilibrium: label, keep the finite ones
        cand = surrogate.uniform_states(200000, c.params, seed=5)
        _, yc = sg.label(cand)
        U = cand[np.isfinite(yc).all(1)][:g.n_cells]
        U = np.concatenate([U, np.tile(U[:1], (g.n_cells - len(U), 1))]) if len(U) < g.n_cells else U
        out = {}
        for name, mode in (("cold", _lib.RELAX_COLD), ("mlp", _lib.RELAX_MLP)):
            g.upload_state(U)
            out[name] = (g.relaxation(mode), g.download_state())
    cold, mlp = out["cold"][0], out["mlp"][0]
    print(f"training loss {l0:.4f} -> {l1:.6f}; cold {cold}\nmlp {mlp}")
    assert mlp["n_chem"] == cold["n_chem"] and mlp["n_chem"] >= 0.99 * g.n_cells
    assert mlp["newton_iters"] <= 0.6 * cold["newton_iters"]
    assert mlp["n_fallback"] < 0.05 * mlp["n_chem"]
    assert at_equilibrium(out["mlp"][1], c.params).mean() >= 0.95
