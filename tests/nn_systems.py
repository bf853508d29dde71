"""Small SPD block systems for the learned-preconditioner tests (helper, not a test module)."""
import numpy as np


def grid_system(s=5, seed=0):
    """SPD BSR system on an s x s triangulated grid graph: negative
    semidefinite off-diagonal blocks and diagonally dominant diagonal blocks."""
    rng = np.random.default_rng(seed)
    n = s * s
    nb = [[i] for i in range(n)]
    for r in range(s):
        for c in range(s):
            i = r * s + c
            for dr, dc in ((0, 1), (1, 0), (1, 1)):
                rr, cc = r + dr, c + dc
                if rr < s and cc < s:
                    j = rr * s + cc
                    nb[i].append(j)
                    nb[j].append(i)
    rp = [0]
    ci = []
    for i in range(n):
        ci += sorted(nb[i])
        rp.append(len(ci))
    rp, ci = np.array(rp), np.array(ci)
    A = np.zeros((len(ci), 3, 3))
    for i in range(n):
        for b in range(rp[i], rp[i + 1]):
            j = ci[b]
            if j > i:
                M = rng.normal(size=(3, 3)) * 0.3
                K = -(M @ M.T + 0.1 * np.eye(3))
                A[b] = K
                bt = rp[j] + np.searchsorted(ci[rp[j]:rp[j + 1]], i)
                A[bt] = K.T
    for i in range(n):
        d = rp[i] + np.searchsorted(ci[rp[i]:rp[i + 1]], i)
        off = sum(np.abs(A[b]).sum(axis=1).max() for b in range(rp[i], rp[i + 1]) if ci[b] != i)
        A[d] = np.diag(off + 1.0 + rng.random(3))
    b = rng.normal(size=(n, 3))
    return rp, ci, A.reshape(-1, 9), b
