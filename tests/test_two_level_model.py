"""Pins the CPU model of the two-level Schwarz solver (oracle/two_level.py),
which the GPU tests compare the HIP two-level PCG against (no GPU).

Expected values come from the mathematics, not from the model's formulas:
  * Morton order: codes built bit by bit (x bit b at 3b, y at 3b+1, z at 3b+2)
    are non-decreasing along the order, ties by index;
  * the patch size rule at its three regimes;
  * degree-1 Chebyshev is the block-Jacobi step D^-1 f / theta;
  * M = sum_p R_p^T C_p R_p + P Ac^-1 P^T is symmetric positive definite
    (odd degree) as a dense matrix;
  * the patch polynomial converges to A_pp^-1 f as the degree grows when the
    interval covers the spectrum;
  * the coarse term is the A-orthogonal projection onto the patch
    translations: P Ac^-1 P^T A P c = P c exactly;
  * the solve meets its residual test and agrees with a dense solve, and the
    coarse level cuts iterations on a grid Laplacian with a mass term.
"""
import numpy as np
import scipy.sparse as sp

from oracle import two_level as tlm


def grid_system(s, mass=1e-3, seed=0):
    """3 DOF per vertex of an s x s grid: a vector Laplacian (edge springs,
    weight 1) plus mass * I, as a BSR (row_ptr, cols, 3x3 blocks) and the
    vertex positions."""
    rng = np.random.default_rng(seed)
    n = s * s
    edges = []
    for i in range(s):
        for j in range(s):
            v = i * s + j
            if j + 1 < s:
                edges.append((v, v + 1))
            if i + 1 < s:
                edges.append((v, v + s))
            if i + 1 < s and j + 1 < s:
                edges.append((v, v + s + 1))
    rows = [[] for _ in range(n)]
    for v in range(n):
        rows[v].append(v)
    for a, b in edges:
        rows[a].append(b)
        rows[b].append(a)
    rp = [0]
    ci = []
    for v in range(n):
        cs = sorted(rows[v])
        ci += cs
        rp.append(len(ci))
    rp = np.array(rp, np.int32)
    ci = np.array(ci, np.int32)
    deg = np.array([len(r) - 1 for r in rows], np.float64)
    vals = np.zeros((len(ci), 3, 3))
    for v in range(n):
        for k in range(rp[v], rp[v + 1]):
            j = ci[k]
            if j == v:
                w = deg[v] + mass * (1.0 + 0.5 * rng.random())
                vals[k] = w * np.eye(3)
            else:
                vals[k] = -np.eye(3)
    xy = np.array([[j / (s - 1), i / (s - 1), 0.0] for i in range(s) for j in range(s)])
    return rp, ci, vals, xy


def dense(M_apply, N):
    return np.column_stack([M_apply(e) for e in np.eye(N)])


def test_morton_order_bit_by_bit():
    rng = np.random.default_rng(3)
    v = rng.random((200, 3)).astype(np.float32)
    v[17] = v[5]  # a tie
    order = tlm.morton_order(v)
    vd = v.astype(np.float64)
    lo, hi = vd.min(axis=0), vd.max(axis=0)
    ext = float((hi - lo).max())
    q = np.clip((vd - lo) / ext * 1023.0, 0.0, 1023.0).astype(np.int64)
    codes = []
    for p in q:
        c = 0
        for b in range(10):
            for ax in range(3):
                c |= ((int(p[ax]) >> b) & 1) << (3 * b + ax)
        codes.append(c)
    codes = np.array(codes)
    assert sorted(order.tolist()) == list(range(200))
    seq = [(codes[i], i) for i in order]
    assert seq == sorted(seq)


def test_patch_rows_rule():
    assert tlm.patch_rows(1000) == 256
    assert tlm.patch_rows(66049) == 259
    assert tlm.patch_rows(10 ** 6) == 512


def test_degree_one_is_scaled_block_jacobi():
    rp, ci, vals, xy = grid_system(12)
    n = len(rp) - 1
    A = tlm.bsr_to_csr(rp, ci, vals, n)
    M = tlm.TwoLevel(A, tlm.morton_order(xy), degree=1, coarse=False)
    f = np.random.default_rng(1).standard_normal(3 * n)
    lmx = np.repeat(M.lmax[M.pid], 3)
    th = 0.5 * (lmx + lmx / M.ratio)
    D = vals[[k for v in range(n) for k in range(rp[v], rp[v + 1]) if ci[k] == v]]
    Dp = D[M.order]
    expect = np.einsum("nij,nj->ni", np.linalg.inv(Dp), f.reshape(-1, 3)).ravel() / th
    assert np.allclose(M.local(f), expect, rtol=1e-12, atol=0)


def test_preconditioner_is_symmetric_positive_definite():
    rp, ci, vals, xy = grid_system(24)  # 576 vertices: three patches
    n = len(rp) - 1
    A = tlm.bsr_to_csr(rp, ci, vals, n)
    for degree in (1, 3, 5):
        M = tlm.TwoLevel(A, tlm.morton_order(xy), degree=degree)
        assert M.npatch == 3
        Md = dense(M.apply, 3 * n)
        assert np.abs(Md - Md.T).max() <= 1e-10 * np.abs(Md).max(), degree
        assert np.linalg.eigvalsh(0.5 * (Md + Md.T)).min() > 0.0, degree


def test_patch_polynomial_converges_to_the_patch_inverse():
    rp, ci, vals, xy = grid_system(10, mass=2.0)  # one patch, well conditioned
    n = len(rp) - 1
    A = tlm.bsr_to_csr(rp, ci, vals, n)
    f = np.random.default_rng(2).standard_normal(3 * n)
    errs = []
    for degree in (1, 3, 5, 9, 15):
        M = tlm.TwoLevel(A, tlm.morton_order(xy), degree=degree, ratio=4.0, coarse=False)
        exact = np.linalg.solve(M.Aloc.toarray(), f)
        errs.append(np.abs(M.local(f) - exact).max() / np.abs(exact).max())
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-3, errs


def test_coarse_term_is_the_projection_onto_patch_translations():
    rp, ci, vals, xy = grid_system(24)
    n = len(rp) - 1
    A = tlm.bsr_to_csr(rp, ci, vals, n)
    M = tlm.TwoLevel(A, tlm.morton_order(xy), degree=1)
    c = np.random.default_rng(4).standard_normal(3 * M.npatch)
    Pc = M.P @ c
    import scipy.linalg as sla
    back = M.P @ sla.cho_solve(M.Acho, M.P.T @ (M.A @ Pc))
    assert np.allclose(back, Pc, rtol=1e-10, atol=1e-12)
    # P sums each patch's rows: column 3p + k has ones on component k of patch p.
    cols = M.P.toarray()
    assert np.all(cols.sum(axis=0) == np.bincount(M.pid, minlength=M.npatch).repeat(3))


def test_solve_matches_dense_and_coarse_level_cuts_iterations():
    rp, ci, vals, xy = grid_system(64, mass=1e-4)  # 4,096 vertices, 16 patches, stiff
    n = len(rp) - 1
    A = tlm.bsr_to_csr(rp, ci, vals, n)
    b = np.random.default_rng(5).standard_normal(3 * n)
    order = tlm.morton_order(xy)
    x, it, ok = tlm.solve(A, b, order, tol=1e-8, maxit=2000)
    assert ok
    exact = sp.linalg.spsolve(A.tocsc(), b)
    assert np.abs(x - exact).max() <= 1e-6 * np.abs(exact).max()
    assert np.abs(A @ x - b).max() < 1e-8 * np.abs(b).max()
    _, it_local, ok_local = tlm.solve(A, b, order, tol=1e-8, maxit=2000, coarse=False)
    assert 1.3 * it <= it_local or not ok_local, (it, it_local)
    _, it32, ok32 = tlm.solve(A, b, order, tol=1e-6, maxit=2000, precision=32)
    assert ok32
