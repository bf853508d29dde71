"""CPU model of the two-level Schwarz preconditioned CG (test infrastructure).

The reference solves every Newton system with block-Jacobi PCG
(pcg_solver.cpp:8-146).  The HIP library's opt-in two-level solver keeps that
loop (same stopping test: max-norm relative residual against max|b|, the
|p.Ap| < 1e-16 breakdown, the iteration cap) and replaces the preconditioner:

    M = sum_p R_p^T C_p R_p + P Ac^-1 P^T

* rows renumbered in Morton order of the rest positions (30-bit codes, 10 bits
  per axis over the bounding cube, ties by index) and cut into patches of
  patch_rows(n) consecutive rows (256 to 512, at most 256 patches while that
  allows);
* C_p: `degree` steps of Chebyshev iteration for A_pp y = f from y = 0,
  preconditioned by the block-Jacobi blocks D, on [lmax / ratio, lmax], where
  lmax = 1.2 x a 10-step power estimate of the largest eigenvalue of
  D_p^-1 A_pp from a fixed start vector;
* P: three coarse unknowns per patch (the patch's rows summed), Ac = P^T A P,
  solved exactly.

Nothing here is imported by the library; tests compare the GPU solver with it
(iteration counts) and check the GPU solutions against float64 residuals.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

POWER_ITERS = 10
SAFETY = 1.2


def morton_order(verts) -> np.ndarray:
    """Vertex order by Morton code of the positions (new -> old)."""
    v = np.asarray(verts, dtype=np.float32).astype(np.float64)
    lo, hi = v.min(axis=0), v.max(axis=0)
    ext = max(float((hi - lo).max()), 1e-30)
    q = np.clip((v - lo) / ext * 1023.0, 0.0, 1023.0).astype(np.uint64)

    def spread(x):
        x = x & 0x3FF
        x = (x | (x << 16)) & 0x030000FF
        x = (x | (x << 8)) & 0x0300F00F
        x = (x | (x << 4)) & 0x030C30C3
        x = (x | (x << 2)) & 0x09249249
        return x

    code = spread(q[:, 0]) | (spread(q[:, 1]) << 1) | (spread(q[:, 2]) << 2)
    return np.lexsort((np.arange(len(v)), code)).astype(np.int64)


def patch_rows(n: int) -> int:
    """Rows per patch: at least 256, at most 512, no more than 256 patches while that allows."""
    return min(512, max(256, (n + 255) // 256))


def bsr_to_csr(rp, ci, vals, n):
    return sp.bsr_matrix((np.asarray(vals, np.float64).reshape(-1, 3, 3), ci, rp), shape=(3 * n, 3 * n)).tocsr()


class TwoLevel:
    """The preconditioner for one system (A: 3n x 3n CSR, float64)."""

    def __init__(self, A, order, degree: int = 9, ratio: float = 5.0, coarse: bool = True):
        n = A.shape[0] // 3
        self.n, self.degree, self.ratio = n, degree, ratio
        self.order = np.asarray(order)
        dof = (3 * self.order[:, None] + np.arange(3)).ravel()
        self.dof = dof
        Ap = A[dof][:, dof].tocsr()  # renumbered
        self.A = Ap
        pr = patch_rows(n)
        self.npatch = (n + pr - 1) // pr
        pid = np.arange(n) // pr
        self.pid = pid
        dp = np.repeat(pid, 3)
        C = Ap.tocoo()
        keep = dp[C.row] == dp[C.col]
        self.Aloc = sp.csr_matrix((C.data[keep], (C.row[keep], C.col[keep])), shape=Ap.shape)
        D = np.zeros((n, 3, 3))
        diag = (C.row // 3) == (C.col // 3)
        D[C.row[diag] // 3, C.row[diag] % 3, C.col[diag] % 3] = C.data[diag]
        self.Dinv = np.linalg.inv(D)
        self.lmax = self._power()
        self.coarse = coarse
        if coarse:
            P = sp.csr_matrix((np.ones(3 * n), (np.arange(3 * n), 3 * np.repeat(pid, 3) + np.tile(np.arange(3), n))),
                              shape=(3 * n, 3 * self.npatch))
            self.P = P
            Ac = (P.T @ Ap @ P).toarray()
            self.Acho = sla.cho_factor(Ac)

    def _dj(self, r):
        return np.einsum("nij,nj->ni", self.Dinv, r.reshape(-1, 3)).ravel()

    def _patch_norms(self, w):
        return np.sqrt(np.bincount(np.repeat(self.pid, 3), weights=w * w, minlength=self.npatch))

    def _power(self):
        n = self.n
        k = (np.arange(3 * n, dtype=np.uint64) * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)
        v = 0.5 + (k >> np.uint64(8)).astype(np.float64) / 16777216.0
        lam = np.zeros(self.npatch)
        for _ in range(POWER_ITERS):
            w = self._dj(self.Aloc @ v)
            nw, nv = self._patch_norms(w), self._patch_norms(v)
            lam = np.where(nv > 0, nw / np.where(nv > 0, nv, 1), 0.0)
            scale = np.where(nw > 0, 1.0 / np.where(nw > 0, nw, 1.0), 0.0)
            v = w * np.repeat(scale[self.pid], 3)
        return np.where((lam > 0) & np.isfinite(lam), SAFETY * lam, 1.0)

    def local(self, f):
        """C_p f for every patch (renumbered vectors)."""
        lmx = np.repeat(self.lmax[self.pid], 3)
        lmn = lmx / self.ratio
        th, de = 0.5 * (lmx + lmn), 0.5 * (lmx - lmn)
        sigma = th / de
        rho = 1.0 / sigma
        d = self._dj(f) / th
        y = d.copy()
        for _ in range(1, self.degree):
            res = f - self.Aloc @ y
            rho_n = 1.0 / (2.0 * sigma - rho)
            d = rho_n * rho * d + (2.0 * rho_n / de) * self._dj(res)
            y = y + d
            rho = rho_n
        return y

    def apply(self, r):
        z = self.local(r)
        if self.coarse:
            z = z + self.P @ sla.cho_solve(self.Acho, self.P.T @ r)
        return z


def solve(A, b, order, tol=1e-6, maxit=1000, degree=9, ratio=5.0, coarse=True, precision=64):
    """Two-level PCG on A x = b (z = M r every iteration, as the GPU forms it).
    precision=32 rounds the CG vectors and the step sizes to float32 after
    every operation, as the GPU stores them (dot products and M stay float64).
    Returns (x in the original numbering, iterations, converged)."""
    M = TwoLevel(A, order, degree, ratio, coarse)
    rnd = (lambda v: np.asarray(v, np.float32).astype(np.float64)) if precision == 32 else (lambda v: v)
    bp = rnd(np.asarray(b, np.float64)[M.dof])
    Ap_ = M.A
    x = np.zeros_like(bp)
    r = bp.copy()
    bmax = np.abs(bp).max()
    z = rnd(M.apply(r))
    p = z.copy()
    rz = r @ z
    rel = np.abs(r).max() / bmax if bmax >= 1e-16 else np.abs(r).max()
    it, ok = 0, rel < tol
    while not ok and it < maxit:
        q = rnd(Ap_ @ p)
        pAp = p @ q
        if abs(pAp) < 1e-16:
            it += 1
            break
        a = float(rnd(rz / pAp))
        x = rnd(x + a * p)
        r = rnd(r - a * q)
        z = rnd(M.apply(r))
        it += 1
        rel = np.abs(r).max() / bmax if bmax >= 1e-16 else np.abs(r).max()
        if rel < tol:
            ok = True
            break
        rzn = r @ z
        p = rnd(z + float(rnd(rzn / rz)) * p)
        rz = rzn
    out = np.empty_like(x)
    out[M.dof] = x
    return out, it, ok
