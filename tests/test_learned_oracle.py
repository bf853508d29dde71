"""CPU checks of the learned-preconditioner oracle (oracle/learned.py).

The oracle's backward pass is hand-written, so it is pinned here against
central finite differences in exact float64 arithmetic; its P is checked to be
SPD and to reduce to block-Jacobi when the output layer is zero.
"""
import numpy as np
import pytest

from oracle import learned as L
from nn_systems import grid_system


def small_system(s=5, seed=0):
    return grid_system(s, seed)


def test_untrained_network_is_block_jacobi():
    rp, ci, A, b = small_system()
    w = L.init_weights(3)
    F = L.forward(w, rp, ci, A, b, bf=False)
    for i in range(len(rp) - 1):
        np.testing.assert_allclose(F["P"][i], F["Dinv"][i], rtol=1e-12, atol=1e-14)


def test_precond_blocks_spd_for_any_output():
    rng = np.random.default_rng(5)
    o = rng.normal(size=(200, 10)) * 5
    D = rng.normal(size=(200, 3, 3))
    Dinv = np.einsum("nij,nkj->nik", D, D) + 0.1 * np.eye(3)
    P, R, _, _ = L.precond_blocks(o, Dinv)
    for k in range(200):
        assert np.linalg.det(R[k]) != 0
        assert np.all(np.linalg.eigvalsh(0.5 * (P[k] + P[k].T)) > 0)


@pytest.mark.parametrize("seed", [1, 2])
def test_backward_matches_finite_differences(seed):
    rp, ci, A, b = small_system(4, seed)
    rng = np.random.default_rng(seed)
    w = L.init_weights(seed).astype(np.float64)
    w[L.WO:L.BO + L.KOUT] = rng.normal(size=L.BO + L.KOUT - L.WO) * 0.2
    loss, g, _ = L.loss_grad(w, rp, ci, A, b, bf=False)
    idx = np.concatenate([rng.choice(np.arange(L.W0, L.B0 + L.KH), 12, replace=False),
                          rng.choice(np.arange(L.LAYER0, L.WO), 24, replace=False),
                          rng.choice(np.arange(L.WO, L.NUM_WEIGHTS), 12, replace=False)])
    for k in idx:
        e = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[k] += e
        wm[k] -= e
        lp = L.loss_grad(wp, rp, ci, A, b, bf=False)[0]
        lm = L.loss_grad(wm, rp, ci, A, b, bf=False)[0]
        fd = (lp - lm) / (2 * e)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (k, fd, g[k])


def test_loss_is_energy_error():
    """With x0 = alpha y (alpha = b^T y / y^T A y), L + b^T A^-1 b / N is the
    normalised A-norm error of x0: >= 0, and 0 when y is parallel to A^-1 b."""
    rp, ci, A, b = small_system(4, 7)
    n = len(rp) - 1
    Ad = np.zeros((3 * n, 3 * n))
    Ab = np.asarray(A).reshape(-1, 3, 3)
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            Ad[3 * i:3 * i + 3, 3 * ci[k]:3 * ci[k] + 3] = Ab[k]
    xs = np.linalg.solve(Ad, b.ravel())
    w = L.init_weights(1)
    loss, _, F = L.loss_grad(w, rp, ci, A, b, bf=False)
    N = float(np.einsum("ni,nij,nj->", F["b"], F["Dinv"], F["b"]))
    y = F["y"].ravel()
    alpha = (b.ravel() @ y) / (y @ Ad @ y)
    e = alpha * y - xs
    np.testing.assert_allclose(loss + b.ravel() @ xs / N, e @ Ad @ e / N, rtol=1e-9, atol=1e-12)
    assert e @ Ad @ e <= xs @ Ad @ xs   # never worse than x0 = 0


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5e-3], np.float32)
    r = L.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even, 1.
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 and r[3] == 1.0078125
    assert abs(r[4] - -2.5e-3) / 2.5e-3 < 2 ** -8
