"""CPU oracle of the learned block preconditioner (TEST INFRASTRUCTURE ONLY).

A numpy restatement of the network, loss and gradients that
andosim_amd/csrc/learned.hip computes on the GPU.  Only tests, the smoke
check and the benchmark's CPU leg use it, and only as the checker.  The
reference has no learned preconditioner (its PCG is block-Jacobi from x0 = 0,
pcg_solver.cpp:8-75, integrator.cpp:187), so there is no reference line to
follow; this file fixes the definitions:

  features f_i (16 used of 32) from the BSR system matrix,
  h0 = relu(f W0 + b0); h_l = relu([h_{l-1} | mean_nbr h_{l-1}] W_l + b_l), l = 1..4,
  o = h4 Wo + bo; R = I + 0.3 tanh(o[0:9]); s = 2 tanh(o[9]);
  P_i = exp(s) R D_i^-1 R^T;  y = P b;  warm start x0 = alpha y,
  alpha = b^T y / y^T A y;
  L = -(b^T y)^2 / ((y^T A y)(b^T D^-1 b))
(the A-norm error the warm start leaves, normalised, constant terms dropped).

`bf16=True` rounds at the points where the device stores BF16 (features,
weights, hidden activations, the neighbour means, and the gradient operands of
the weight products); `bf16=False` is exact float64 arithmetic, which the
finite-difference test pins.  The backward pass is written by hand (no
autodiff), so it is checked here against central differences.
"""
from __future__ import annotations

import numpy as np

KF, KF_USED, KH, KL, KOUT = 32, 16, 64, 4, 10
DELTA, LOGS = 0.3, 2.0
W0 = 0
B0 = W0 + KF * KH
LAYER0 = B0 + KH
LAYER_SIZE = 2 * KH * KH + KH
WO = LAYER0 + KL * LAYER_SIZE
BO = WO + KH * KOUT
NUM_WEIGHTS = BO + KOUT


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(w):
    w = np.asarray(w, np.float64)
    W = {"W0": w[W0:B0].reshape(KF, KH), "b0": w[B0:B0 + KH]}
    for l in range(KL):
        o = LAYER0 + l * LAYER_SIZE
        W[f"W{l + 1}"] = w[o:o + 2 * KH * KH].reshape(2 * KH, KH)
        W[f"b{l + 1}"] = w[o + 2 * KH * KH:o + LAYER_SIZE]
    W["Wo"] = w[WO:BO].reshape(KH, KOUT)
    W["bo"] = w[BO:BO + KOUT]
    return W


def block_jacobi_inverse(B):
    """Block inverse as pcg_solver.cpp:100-112: cofactor inverse if |det| > 1e-16, else I."""
    B = np.asarray(B, np.float64)
    det = np.linalg.det(B)
    if abs(det) > 1e-16:
        return np.linalg.inv(B)
    return np.eye(3)


def graph(rp, ci):
    n = len(rp) - 1
    nbr = []
    diag = np.zeros(n, np.int64)
    for i in range(n):
        row = []
        for b in range(rp[i], rp[i + 1]):
            if ci[b] == i:
                diag[i] = b
            else:
                row.append(ci[b])
        nbr.append(np.array(row, np.int64))
    return nbr, diag


def features(rp, ci, A):
    """Per-vertex features and D^-1 (definitions as in k_nn_features)."""
    A = np.asarray(A, np.float64).reshape(-1, 3, 3)
    n = len(rp) - 1
    nbr, diag = graph(rp, ci)
    f = np.zeros((n, KF))
    Dinv = np.zeros((n, 3, 3))
    for i in range(n):
        D = A[diag[i]]
        Dinv[i] = block_jacobi_inverse(D)
        t = np.trace(D) / 3.0
        if not t > 0:
            t = 1.0
        S = np.zeros((3, 3))
        cpl = 0.0
        for b in range(rp[i], rp[i + 1]):
            if ci[b] == i:
                continue
            S += A[b]
            cpl += np.sqrt(np.sum(A[b] ** 2))
        sym = lambda M: [M[0, 0], M[1, 1], M[2, 2], 0.5 * (M[0, 1] + M[1, 0]), 0.5 * (M[0, 2] + M[2, 0]),
                         0.5 * (M[1, 2] + M[2, 1])]
        f[i, 0:6] = np.array(sym(D)) / t
        f[i, 6:12] = np.array(sym(S)) / t
        f[i, 12] = cpl / t / 3.0
        f[i, 13] = np.log1p(f[i, 12])
        f[i, 14] = len(nbr[i]) / 8.0
        f[i, 15] = 1.0
    return f, Dinv, nbr


def precond_blocks(o, Dinv):
    th = np.tanh(o[:, :9]).reshape(-1, 3, 3)
    R = np.eye(3)[None] + DELTA * th
    sc = np.exp(LOGS * np.tanh(o[:, 9]))
    P = sc[:, None, None] * np.einsum("nij,njk,nlk->nil", R, Dinv, R)
    return P, R, th, sc


def spmv(rp, ci, A, x):
    A = np.asarray(A, np.float64).reshape(-1, 3, 3)
    n = len(rp) - 1
    out = np.zeros((n, 3))
    for i in range(n):
        for b in range(rp[i], rp[i + 1]):
            out[i] += A[b] @ x[ci[b]]
    return out


def forward(w, rp, ci, A, b, bf=True):
    """Returns a dict with P, y and the intermediates the backward pass needs."""
    rnd = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    W = {k: rnd(v) for k, v in unpack(w).items()}
    f, Dinv, nbr = features(rp, ci, A)
    f = rnd(f)
    h = [rnd(np.maximum(f @ W["W0"] + W["b0"], 0.0))]
    X = []
    for l in range(KL):
        hp = h[-1]
        mean = np.zeros_like(hp)
        for i, nb in enumerate(nbr):
            if len(nb):
                mean[i] = hp[nb].sum(axis=0) / len(nb)
        Xl = np.concatenate([hp, rnd(mean)], axis=1)
        X.append(Xl)
        h.append(rnd(np.maximum(Xl @ W[f"W{l + 1}"] + W[f"b{l + 1}"], 0.0)))
    o = h[-1] @ W["Wo"] + W["bo"]
    P, R, th, sc = precond_blocks(o, Dinv)
    bb = np.asarray(b, np.float64).reshape(-1, 3)
    y = np.einsum("nij,nj->ni", P, bb)
    return dict(W=W, f=f, Dinv=Dinv, nbr=nbr, h=h, X=X, o=o, P=P, R=R, th=th, sc=sc, y=y, b=bb)


def loss_grad(w, rp, ci, A, b, bf=True):
    """Loss and gradient (flat, in the weight layout), by the chain rule."""
    F = forward(w, rp, ci, A, b, bf)
    rnd = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    W, y, bb, Dinv = F["W"], F["y"], F["b"], F["Dinv"]
    Ay = spmv(rp, ci, A, y)
    N = float(np.einsum("ni,nij,nj->", bb, Dinv, bb))
    u = float(np.sum(bb * y))
    v = float(np.sum(y * Ay))
    loss = -u * u / (v * N)
    gy = 2.0 * u / (v * N) * (u / v * Ay - bb)
    G = np.einsum("ni,nj->nij", gy, bb)                      # dL/dP
    P, R, th, sc = F["P"], F["R"], F["th"], F["sc"]
    dls = np.einsum("nij,nij->n", G, P)
    RM = np.einsum("nij,njk->nik", R, Dinv)
    RMt = np.einsum("nij,nkj->nik", R, Dinv)                  # R M^T
    dR = sc[:, None, None] * (np.einsum("nik,nkj->nij", G, RMt) + np.einsum("nki,nkj->nij", G, RM))
    do = np.zeros_like(F["o"])
    do[:, :9] = (dR * DELTA * (1.0 - th ** 2)).reshape(-1, 9)
    do[:, 9] = dls * LOGS * (1.0 - np.tanh(F["o"][:, 9]) ** 2)
    grad = np.zeros(NUM_WEIGHTS)
    h = F["h"]
    grad[WO:BO] = (h[-1].T @ do).ravel()
    grad[BO:BO + KOUT] = do.sum(axis=0)
    dh = do @ W["Wo"].T
    nbr = F["nbr"]
    for l in range(KL, -1, -1):
        dz = rnd(dh * (h[l] > 0))
        Xin = F["X"][l - 1] if l > 0 else F["f"]
        gW = Xin.T @ dz
        gb = dz.sum(axis=0)
        if l > 0:
            o = LAYER0 + (l - 1) * LAYER_SIZE
            grad[o:o + 2 * KH * KH] = gW.ravel()
            grad[o + 2 * KH * KH:o + LAYER_SIZE] = gb
            dX = dz @ W[f"W{l}"].T
            dh = dX[:, :KH].copy()
            for j, nb in enumerate(nbr):
                if len(nb):
                    dh[nb] += dX[j, KH:] / len(nb)
        else:
            grad[W0:B0] = gW.ravel()
            grad[B0:B0 + KH] = gb
    return loss, grad, F


def init_weights(seed: int) -> np.ndarray:
    """Same stream as the library's initialiser (SplitMix64 + Box-Muller, He
    scaling, zero biases and output layer)."""
    M = (1 << 64) - 1
    st = [seed & M]

    def nxt():
        st[0] = (st[0] + 0x9E3779B97F4A7C15) & M
        z = st[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    def unif():
        return ((nxt() >> 11) + 0.5) * (1.0 / 9007199254740992.0)

    def gauss():
        u1 = unif()
        u2 = unif()
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)

    w = np.zeros(NUM_WEIGHTS, np.float32)
    for k in range(KF_USED):
        for j in range(KH):
            w[W0 + k * KH + j] = gauss() * np.sqrt(2.0 / KF_USED)
    for l in range(KL):
        o = LAYER0 + l * LAYER_SIZE
        for k in range(2 * KH * KH):
            w[o + k] = gauss() * np.sqrt(2.0 / (2 * KH))
    return w
