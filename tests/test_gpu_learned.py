"""Learned preconditioner on the GPU against its CPU oracle (oracle/learned.py).

The device network stores BF16 weights and activations and accumulates every
product in FP32 on MFMA; the oracle rounds at the same points and accumulates
in float64.  Most values then agree to FP32 accumulation-order error (~1e-7),
but an activation that lands within that error of a BF16 rounding boundary
rounds one BF16 ulp (2^-8 relative) apart and carries the difference forward,
so the bounds are a few BF16 ulps: 5e-3 relative for P and y, 1e-3 for the
loss, 2e-2 and cosine > 0.999 for each layer's gradient.
"""
import numpy as np
import pytest

from andosim_amd import scenes
from oracle import learned as L
from nn_systems import grid_system

pytestmark = pytest.mark.gpu

native = pytest.importorskip("andosim_amd.native")


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def trained_like_weights(seed):
    w = L.init_weights(seed)
    w[L.WO:] = np.random.default_rng(seed).normal(size=len(w) - L.WO).astype(np.float32) * 0.2
    return w


def test_init_weights_match_oracle():
    tr = native.Trainer()
    tr.init_weights(11)
    assert np.array_equal(tr.get_weights(), L.init_weights(11))


@pytest.mark.parametrize("s,seed", [(6, 3), (11, 4)])
def test_forward_matches_oracle(s, seed):
    rp, ci, A, b = grid_system(s, seed)
    tr = native.Trainer()
    i = tr.add_system(rp, ci, A, b)
    w = trained_like_weights(seed)
    tr.set_weights(w)
    P, y = tr.forward(i)
    F = L.forward(w, rp, ci, A, b, bf=True)
    assert rel(P, F["P"]) < 5e-3 and rel(y, F["y"]) < 5e-3


@pytest.mark.parametrize("s,seed", [(6, 3), (11, 4)])
def test_gradient_matches_oracle(s, seed):
    rp, ci, A, b = grid_system(s, seed)
    tr = native.Trainer()
    i = tr.add_system(rp, ci, A, b)
    w = trained_like_weights(seed)
    tr.set_weights(w)
    loss_g, g = tr.loss_grad(i)
    loss_o, go, _ = L.loss_grad(w, rp, ci, A, b, bf=True)
    assert abs(loss_g - loss_o) <= 1e-3 * abs(loss_o)
    blocks = [(0, L.B0 + L.KH)] + [(L.LAYER0 + l * L.LAYER_SIZE, L.LAYER0 + (l + 1) * L.LAYER_SIZE)
                                   for l in range(L.KL)] + [(L.WO, L.NUM_WEIGHTS)]
    for a, bb in blocks:
        x, z = g[a:bb].astype(np.float64), go[a:bb]
        cos = float(x @ z / (np.linalg.norm(x) * np.linalg.norm(z)))
        assert rel(x, z) < 2e-2 and cos > 0.999, (a, bb, rel(x, z), cos)


def test_untrained_network_is_block_jacobi_with_warm_start():
    """Zero output layer: P = D^-1, so PCG from the warm start converges like
    plain PCG (never more than one extra iteration) on the same system."""
    rp, ci, A, b = grid_system(9, 5)
    tr = native.Trainer()
    i = tr.add_system(rp, ci, A, b)
    tr.init_weights(2)
    P, _ = tr.forward(i)
    F = L.forward(L.init_weights(2), rp, ci, A, b, bf=True)
    assert rel(P, F["Dinv"]) < 1e-6
    ip, cp = tr.solve(i, False)
    il, cl = tr.solve(i, True)
    assert cp and cl and il <= ip + 1


def test_training_lowers_the_loss():
    systems = [grid_system(s, k) for k, s in enumerate([5, 6, 7, 8])]
    tr = native.Trainer()
    ids = [tr.add_system(*sy) for sy in systems]
    tr.init_weights(3)
    before = np.mean([tr.loss_grad(i)[0] for i in ids])
    losses = tr.train(300, lr=1e-3, seed=9)
    after = np.mean([tr.loss_grad(i)[0] for i in ids])
    assert np.isfinite(losses).all()
    assert after < before


def test_step_in_learned_mode():
    """ando_step with ANDO_MODE_LEARNED_PRECOND: every solve uses the network,
    failed ones fall back to plain PCG, and the state stays finite."""
    sc = scenes.small_sheet(9, pins=2, wall=True)
    g = native.GpuSim(sc)
    with pytest.raises(RuntimeError):
        g.set_mode(learned=True)          # no weights yet
    g.set_precond_weights(trained_like_weights(4))
    g.set_mode(learned=True)
    for _ in range(10):
        g.substep()
        st = g.stats()
        assert st["learned_solves"] == st["pcg_solves"]
        assert st["fallback_iterations"] <= st["pcg_iterations"]
    assert np.isfinite(g.get("x")).all()


def test_learned_solves_stay_within_the_plain_cap():
    """Never lose more than the cap: with weights far from trained, every
    solve of a corrected-physics substep (self contacts, coupling blocks) uses
    at most pcg_max_iters iterations in all -- half for the learned
    preconditioner, the rest for block-Jacobi continuing from its iterate."""
    sc = scenes.stack_resting(s=9, sheets=2, gap=8e-5)
    sc.params["pcg_max_iters"] = 40
    g = native.GpuSim(sc)
    g.set_precond_weights(trained_like_weights(7))
    g.set_mode(learned=True, corrected_physics=True)
    for _ in range(5):
        g.substep()
        st = g.stats()
        assert max(st["pcg_per_solve"]) <= 40, st["pcg_per_solve"]
        assert st["learned_solves"] == st["pcg_solves"]
    assert np.isfinite(g.get("x")).all()


def test_warm_start_counts_the_coupling_blocks():
    """A warm-started solve's starting residual includes the self-contact
    coupling blocks' share of A x0: a substep in warm-start-only mode lands
    where the plain solve does (to the PCG tolerance's reach), on the resting
    stack whose coupling blocks are active."""
    sc = scenes.stack_resting(s=9, sheets=2, gap=8e-5)
    plain = native.GpuSim(sc)
    plain.set_mode(corrected_physics=True)
    ws = native.GpuSim(sc)
    ws.set_precond_weights(trained_like_weights(4))
    ws.set_mode(learned=True, warm_start_only=True, corrected_physics=True)
    for k in range(3):
        plain.set_state(ws.get("x"), ws.get("v"))
        x0 = ws.get("x").astype(np.float64)
        plain.substep()
        ws.substep()
        dp = plain.get("x").astype(np.float64) - x0
        dw = ws.get("x").astype(np.float64) - x0
        assert np.max(np.abs(dw - dp)) <= 1e-2 * np.max(np.abs(dp)), k
