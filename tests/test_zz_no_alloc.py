"""ando_pcg_noalloc (the allocation-free PCG entry point) against ando_pcg in
its row-SpMV mode, through the C ABI, on two Newton systems in a row, the
second smaller than the first: same iteration count, same convergence
verdict, same solution bit for bit.  A warm call leaves the device's free
memory unchanged (it allocates nothing), requests beyond the fixed scratch and
a malformed pattern are refused before any solve."""
import ctypes as C

import numpy as np
import pytest

from andosim_amd import scenes

pytestmark = pytest.mark.gpu

native = pytest.importorskip("andosim_amd.native")


class Hip:
    """The caller's side: device buffers from the HIP runtime the library
    itself links (libamdhip64.so.7, already loaded with it).  torch is not
    used here: its wheel bundles a second HIP runtime."""

    def __init__(self):
        native.load()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        self.L.hipSetDevice.argtypes = [C.c_int]
        assert self.L.hipSetDevice(0) == 0
        self.bufs = []

    def put(self, arr):
        a = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p.value

    def get(self, ptr, count, dtype):
        out = np.zeros(count, dtype)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes, 2) == 0
        return out

    def free_bytes(self):
        f, t = C.c_size_t(0), C.c_size_t(0)
        assert self.L.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    def release(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def newton_system(sc, advance):
    """The first Newton system (BSR pattern, values, b = -g) of a cloth after
    `advance` substeps."""
    g = native.GpuSim(sc, fp64_near=False)
    for _ in range(advance):
        g.substep()
    x, v = g.get("x"), g.get("v")
    grad, H = g.linearize((x + np.float32(sc.params["dt"]) * v).astype(np.float32))
    rp, ci = g.pattern()
    g.close()
    return rp, ci, H.astype(np.float32), (-grad).astype(np.float32)


def existing_entry_point(rp, ci, H, b, tol, maxit):
    g = native.GpuSim()
    g.set_mode(fp64_near=False, row_spmv=True)
    x, it, ok = g.pcg(rp, ci, H, b, tol=tol, maxit=maxit)
    g.close()
    return x, it, ok


def noalloc(hip, rp, ci, H, b, tol, maxit):
    n = len(rp) - 1
    args = [hip.put(np.asarray(rp, np.int32)), hip.put(np.asarray(ci, np.int32)),
            hip.put(np.asarray(H, np.float32).reshape(-1)), hip.put(np.asarray(b, np.float32).reshape(-1)),
            hip.put(np.full(3 * n, np.nan, np.float32))]
    free0 = hip.free_bytes()
    it, ok = native.pcg_noalloc(0, n, len(ci), *args, tol, maxit)
    free1 = hip.free_bytes()
    x = hip.get(args[4], 3 * n, np.float32)
    hip.release()
    return x, it, ok, free0 - free1


def test_noalloc_pcg_matches_ando_pcg_on_two_calls():
    print("ando_pcg_noalloc: sphere129 system (16,641 rows), then a 9x9 sheet (81 rows)", flush=True)
    hip = Hip()
    warmup = newton_system(scenes.small_sheet(5, pins=0, wall=False), 1)
    noalloc(hip, *warmup, 1e-6, 100)  # first call on the device: its address lookup
    cases = [newton_system(scenes.sphere129(), 9), newton_system(scenes.small_sheet(9, pins=2, wall=True), 3)]
    assert len(cases[1][0]) < len(cases[0][0])
    for k, (rp, ci, H, b) in enumerate(cases):
        xr, itr, okr = existing_entry_point(rp, ci, H, b, 1e-6, 1000)
        xn, itn, okn, grew = noalloc(hip, rp, ci, H, b, 1e-6, 1000)
        print(f"call {k}: rows {len(rp) - 1} iterations {itn} converged {okn}", flush=True)
        assert itn == itr and okn == okr, (k, itn, itr, okn, okr)
        assert np.array_equal(xn, xr), (k, np.max(np.abs(xn - xr)))
        assert grew == 0, f"free device memory changed by {grew} bytes during the call"


def test_noalloc_refuses_what_does_not_fit_or_is_malformed():
    hip = Hip()
    rp, ci, H, b = newton_system(scenes.small_sheet(5, pins=0, wall=False), 1)
    n, nb = len(rp) - 1, len(ci)
    d_rp = hip.put(rp.astype(np.int32))
    d_v = hip.put(H.reshape(-1))
    d_b = hip.put(b.reshape(-1))
    d_x = hip.put(np.zeros(3 * n, np.float32))
    # Beyond the fixed scratch: refused before anything is read.
    with pytest.raises(RuntimeError, match="131072"):
        native.pcg_noalloc(0, 131073, 131073, d_rp, d_rp, d_v, d_b, d_x, 1e-6, 100)
    with pytest.raises(RuntimeError, match="8192"):
        native.pcg_noalloc(0, n, nb, d_rp, d_rp, d_v, d_b, d_x, 1e-6, 8193)
    # A row without its diagonal block (every index stays in range).
    bad = ci.copy()
    row0 = list(ci[rp[0]:rp[1]])
    assert 0 in row0 and n - 1 not in row0
    bad[rp[0] + row0.index(0)] = n - 1  # row 0 loses its diagonal block
    with pytest.raises(RuntimeError, match="diagonal"):
        native.pcg_noalloc(0, n, nb, d_rp, hip.put(bad.astype(np.int32)), d_v, d_b, d_x, 1e-6, 100)
    # A column out of range.
    bad2 = ci.copy()
    bad2[0] = n + 5
    with pytest.raises(RuntimeError, match="column"):
        native.pcg_noalloc(0, n, nb, d_rp, hip.put(bad2.astype(np.int32)), d_v, d_b, d_x, 1e-6, 100)
    # Row pointers that do not end at num_blocks.
    with pytest.raises(RuntimeError, match="row_ptr"):
        native.pcg_noalloc(0, n, nb - 1, d_rp, hip.put(ci.astype(np.int32)), d_v, d_b, d_x, 1e-6, 100)
    hip.release()
