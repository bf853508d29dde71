"""The sliced-ELL SpMV (default) against the BSR row SpMV (ANDO_MODE_PCG_ROW_SPMV).

Both kernels sum a row's blocks in the same order, so their products agree bit
for bit; the p.Ap partials are grouped differently (per workgroup, in double),
which can move alpha by a rounding step, so whole solves are compared to 1e-6
with equal iteration counts.  Irregular patterns (slices of different widths,
rows longer than one chunk of eight blocks) go through ando_pcg.
"""
import numpy as np
import pytest

from andosim_amd import scenes

pytestmark = pytest.mark.gpu

native = pytest.importorskip("andosim_amd.native")


def random_spd_bsr(n, seed, extra_rows=()):
    """Symmetric block pattern with random neighbours (rows in extra_rows get
    many), diagonally dominant so it is SPD."""
    rng = np.random.default_rng(seed)
    nbr = [set() for _ in range(n)]
    for i in range(n):
        for j in rng.choice(n, size=rng.integers(0, 5), replace=False):
            if j != i:
                nbr[i].add(int(j))
                nbr[int(j)].add(i)
    for i in extra_rows:
        for j in rng.choice(n, size=18, replace=False):
            if j != i:
                nbr[i].add(int(j))
                nbr[int(j)].add(i)
    blocks = {}
    for i in range(n):
        for j in nbr[i]:
            if (j, i) in blocks:
                blocks[(i, j)] = blocks[(j, i)].T.copy()
            else:
                blocks[(i, j)] = rng.normal(scale=0.3, size=(3, 3))
    rp, ci, vals = [0], [], []
    for i in range(n):
        cols = sorted(nbr[i] | {i})
        off = sum(np.abs(blocks[(i, j)]).sum() for j in cols if j != i)
        for j in cols:
            if j == i:
                M = rng.normal(scale=0.1, size=(3, 3))
                vals.append(M @ M.T + (off + 1.0) * np.eye(3))
            else:
                vals.append(blocks[(i, j)])
            ci.append(j)
        rp.append(len(ci))
    return (np.array(rp, np.int32), np.array(ci, np.int32), np.array(vals, np.float32),
            rng.normal(size=3 * n).astype(np.float32))


@pytest.mark.parametrize("n,seed", [(300, 1), (1000, 2), (4097, 3)])
def test_sell_pcg_matches_row_and_dense(n, seed):
    print(f"sell pcg n={n}", flush=True)
    rp, ci, vals, b = random_spd_bsr(n, seed, extra_rows=(0, n // 2, n - 1))
    assert np.max(np.diff(rp)) > 8  # rows longer than one chunk
    sell, row = native.GpuSim(), native.GpuSim()
    row.set_mode(row_spmv=True)
    xs, its, cs = sell.pcg(rp, ci, vals, b, tol=1e-6, maxit=2000)
    xr, itr, cr = row.pcg(rp, ci, vals, b, tol=1e-6, maxit=2000)
    assert cs and cr and its == itr
    assert np.max(np.abs(xs - xr)) <= 1e-6 * np.max(np.abs(xr))
    if n <= 1000:
        A = np.zeros((3 * n, 3 * n))
        for i in range(n):
            for k in range(rp[i], rp[i + 1]):
                j = ci[k]
                A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = vals[k]
        x = np.linalg.solve(A, b.astype(np.float64))
        assert np.max(np.abs(A @ xs - b)) <= 1e-5 * np.max(np.abs(b))
        assert np.max(np.abs(xs - x)) <= 1e-4 * np.max(np.abs(x))


@pytest.mark.parametrize("name,make,steps,corrected", [
    ("sphere129", lambda: scenes.sphere129(), 5, False),
    ("sheet9_rigid", lambda: scenes.small_sheet(9, pins=0, wall=False, rigid=True), 8, False),
    ("stack_resting33", lambda: scenes.stack_resting(s=33, sheets=4), 3, True),
])
def test_sell_steps_match_row(name, make, steps, corrected):
    print(f"sell steps {name}", flush=True)
    sc = make()
    a, b = native.GpuSim(sc), native.GpuSim(sc)
    a.set_mode(corrected=corrected)
    b.set_mode(corrected=corrected, row_spmv=True)
    for _ in range(steps):
        a.substep()
        b.substep()
        sa, sb = a.stats(), b.stats()
        assert sa["pcg_kernel"] == 3 and sb["pcg_kernel"] == 0
        assert sa["newton_iterations"] == sb["newton_iterations"]
        assert abs(sa["pcg_iterations"] - sb["pcg_iterations"]) <= max(2, sb["pcg_iterations"] // 100)
    xa, xb = a.get("x"), b.get("x")
    assert np.max(np.abs(xa - xb)) <= 1e-6 * max(1.0, float(np.max(np.abs(xb))))


@pytest.mark.parametrize("lanes", [2, 4])
def test_sell_lane_split_bit_identical(lanes, monkeypatch):
    """Two (the default) or four lanes per row (ANDO_SELL_LANES) on the
    one-lane kernel's grid of four slices per workgroup sum in its order and
    reduce its partials the same way: solves and steps agree bit for bit."""
    print(f"sell lanes={lanes}", flush=True)
    rp, ci, vals, b = random_spd_bsr(4097, 3, extra_rows=(0, 2048, 4096))
    monkeypatch.setenv("ANDO_SELL_WG_SLICES", "4")   # the one-lane kernel's grid
    monkeypatch.setenv("ANDO_SELL_LANES", "1")
    one = native.GpuSim()
    monkeypatch.setenv("ANDO_SELL_LANES", str(lanes))
    split = native.GpuSim()
    x1, i1, c1 = one.pcg(rp, ci, vals, b, tol=1e-6, maxit=2000)
    x2, i2, c2 = split.pcg(rp, ci, vals, b, tol=1e-6, maxit=2000)
    assert c1 and c2 and i1 == i2
    assert np.array_equal(x1, x2)
    sc = scenes.sphere129()
    monkeypatch.setenv("ANDO_SELL_LANES", "1")
    a = native.GpuSim(sc)
    monkeypatch.setenv("ANDO_SELL_LANES", str(lanes))
    s = native.GpuSim(sc)
    monkeypatch.delenv("ANDO_SELL_LANES")
    monkeypatch.delenv("ANDO_SELL_WG_SLICES")
    for _ in range(4):
        a.substep()
        s.substep()
        assert a.stats()["pcg_iterations"] == s.stats()["pcg_iterations"]
    assert np.array_equal(a.get("x"), s.get("x"))


@pytest.mark.parametrize("slices,vb", [(4, 64), (2, 256), (1, 256), (1, 64), (0, 256), (0, 128)])
def test_sell_finer_grid_matches(slices, vb, monkeypatch):
    """Fewer slices per SpMV workgroup (ANDO_SELL_WG_SLICES; 0 = one wave) and
    smaller vector-kernel workgroups (ANDO_PCG_VEC_BLOCK): the same products,
    partial sums over other workgroups; solves agree with the four-slice,
    256-thread grid to 1e-6 with iteration counts within 1%, and steps to 1e-6."""
    print(f"sell slices per workgroup={slices} vector block={vb}", flush=True)
    rp, ci, vals, b = random_spd_bsr(4097, 3, extra_rows=(0, 2048, 4096))
    monkeypatch.setenv("ANDO_SELL_WG_SLICES", "4")
    monkeypatch.setenv("ANDO_PCG_VEC_BLOCK", "256")
    ref = native.GpuSim()
    sc = scenes.sphere129()
    a = native.GpuSim(sc)
    monkeypatch.setenv("ANDO_SELL_WG_SLICES", str(slices))
    monkeypatch.setenv("ANDO_PCG_VEC_BLOCK", str(vb))
    fine = native.GpuSim()
    s = native.GpuSim(sc)
    monkeypatch.delenv("ANDO_SELL_WG_SLICES")
    monkeypatch.delenv("ANDO_PCG_VEC_BLOCK")
    x1, i1, c1 = ref.pcg(rp, ci, vals, b, tol=1e-6, maxit=2000)
    x2, i2, c2 = fine.pcg(rp, ci, vals, b, tol=1e-6, maxit=2000)
    assert c1 and c2 and abs(i1 - i2) <= max(1, i1 // 100)
    assert np.max(np.abs(x1 - x2)) <= 1e-6 * np.max(np.abs(x1))
    for _ in range(4):
        a.substep()
        s.substep()
        assert abs(a.stats()["pcg_iterations"] - s.stats()["pcg_iterations"]) <= max(2, a.stats()["pcg_iterations"] // 100)
    assert np.max(np.abs(a.get("x") - s.get("x"))) <= 1e-6 * max(1.0, float(np.max(np.abs(a.get("x")))))


# Launch shapes the autotuner may pick, as the stats report them:
# (lanes per row, slices per SpMV workgroup, threads per vector workgroup).
AUTOTUNE_SHAPES = {(2, 1, 256), (2, 0, 256), (2, 2, 256), (2, 4, 256), (4, 4, 256), (1, 4, 256), (2, 1, 128),
                   (2, 1, 64)}


def test_pcg_autotune_picks_a_tested_shape(monkeypatch):
    """ANDO_PCG_AUTOTUNE=1: the first solve times the candidate launch shapes
    on its own system and keeps the fastest.  The stats report the shape and
    that it was tuned; each candidate is a grid whose solves agree with the
    four-slice grid to 1e-6 (the tests above), so steps agree with the default
    shape to 2e-6 with iteration counts within 1%.  A second context of the
    same size takes the tuned shape from the process-wide table."""
    sc = scenes.sphere129()
    a = native.GpuSim(sc)
    monkeypatch.setenv("ANDO_PCG_AUTOTUNE", "1")
    t = native.GpuSim(sc)
    t2 = native.GpuSim(sc)
    monkeypatch.delenv("ANDO_PCG_AUTOTUNE")
    shape = None
    for k in range(4):
        a.substep()
        t.substep()
        sa, st = a.stats(), t.stats()
        print(f"substep {k}: tuned shape", (st["spmv_lanes"], st["spmv_wg_slices"], st["pcg_vec_block"]), flush=True)
        assert sa["pcg_autotuned"] == 0 and (sa["spmv_lanes"], sa["spmv_wg_slices"], sa["pcg_vec_block"]) == (2, 1, 256)
        assert st["pcg_autotuned"] == 1 and st["pcg_kernel"] == 3
        shape = (st["spmv_lanes"], st["spmv_wg_slices"], st["pcg_vec_block"])
        assert shape in AUTOTUNE_SHAPES
        assert st["pcg_solves"] == sa["pcg_solves"]
        assert abs(sa["pcg_iterations"] - st["pcg_iterations"]) <= max(2, sa["pcg_iterations"] // 100)
    assert np.max(np.abs(a.get("x") - t.get("x"))) <= 2e-6 * max(1.0, float(np.max(np.abs(a.get("x")))))
    t2.substep()
    s2 = t2.stats()
    assert s2["pcg_autotuned"] == 1 and (s2["spmv_lanes"], s2["spmv_wg_slices"], s2["pcg_vec_block"]) == shape
    assert np.array_equal(t2.get("x"), _one_step_with_shape(sc, shape, monkeypatch))


def _one_step_with_shape(sc, shape, monkeypatch):
    lanes, slices, vb = shape
    monkeypatch.setenv("ANDO_SELL_LANES", str(lanes))
    monkeypatch.setenv("ANDO_SELL_WG_SLICES", str(slices))
    monkeypatch.setenv("ANDO_PCG_VEC_BLOCK", str(vb))
    g = native.GpuSim(sc)
    for k in ("ANDO_SELL_LANES", "ANDO_SELL_WG_SLICES", "ANDO_PCG_VEC_BLOCK"):
        monkeypatch.delenv(k)
    g.substep()
    return g.get("x")


def _friction_sheet():
    sc = scenes.small_sheet(9, pins=0, wall=False, rigid=True)
    sc.params.update(enable_friction=True, friction_mu=0.5)
    return sc


def test_assembled_matrix_is_bit_symmetric():
    """The assembled system equals its block transpose bit for bit (every
    contribution is symmetric and both orientations are summed in the same
    order), on a sheet with wall and pins, a sphere drape, and friction on a
    rigid body."""
    for sc in (scenes.small_sheet(9), scenes.sphere129(), _friction_sheet()):
        g = native.GpuSim(sc, fp64_near=False)
        for _ in range(3):
            g.substep()
        x = g.get("x")
        xhat = (x + np.float32(sc.params["dt"]) * g.get("v")).astype(np.float32)
        _, H = g.linearize(xhat)
        rp, ci = g.pattern()
        for i in range(len(rp) - 1):
            for k in range(rp[i], rp[i + 1]):
                j = ci[k]
                if j <= i:
                    continue
                kk = rp[j] + int(np.searchsorted(ci[rp[j]:rp[j + 1]], i))
                assert ci[kk] == i and np.array_equal(H[k], H[kk].T), (sc.name, i, j)
