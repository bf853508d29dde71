"""Batched stepping (ando_step_batch): many independent cloths stepped from
several host threads, each context on its own stream, give exactly the results
of stepping them one after another."""
import numpy as np
import pytest

from andosim_amd import scenes

pytestmark = pytest.mark.gpu

native = pytest.importorskip("andosim_amd.native")


def test_batch_equals_sequential():
    print("batch of 12 uneven cloths", flush=True)
    scs = scenes.batch256(count=12)
    seq = [native.GpuSim(sc) for sc in scs]
    bat = [native.GpuSim(sc) for sc in scs]
    for _ in range(4):
        for g in seq:
            g.substep()
        for g in bat:
            g.apply_gravity()
        native.step_batch(bat, threads=6)
    for a, b in zip(seq, bat):
        assert np.array_equal(a.get("x"), b.get("x"))
        assert np.array_equal(a.get("v"), b.get("v"))
        sa, sb = a.stats(), b.stats()
        assert sa["pcg_iterations"] == sb["pcg_iterations"] and sa["newton_iterations"] == sb["newton_iterations"]
        # The batched steps ran their solves inside batched PCG launches.
        assert sa["batched_solves"] == 0 and sb["batched_solves"] == sb["pcg_solves"] > 0


def test_batch_after_autotune(monkeypatch):
    """Contexts whose launch shape came from the autotuner (ANDO_PCG_AUTOTUNE=1,
    tuned on a single step) still join batched solves, on the batch's own
    grid: from the same state, the batched steps equal stepping contexts of the
    default shape one by one, bit for bit."""
    print("batch of 8 autotuned cloths", flush=True)
    scs = scenes.batch256(count=8)
    seq = [native.GpuSim(sc) for sc in scs]
    monkeypatch.setenv("ANDO_PCG_AUTOTUNE", "1")
    bat = [native.GpuSim(sc) for sc in scs]
    monkeypatch.delenv("ANDO_PCG_AUTOTUNE")
    start = [(g.get("x").copy(), g.get("v").copy()) for g in bat]
    for g in bat:
        g.substep()
        assert g.stats()["pcg_autotuned"] == 1
    for g, (x, v) in zip(bat, start):
        g.set_state(x, v)
    for _ in range(3):
        for g in seq:
            g.substep()
        for g in bat:
            g.apply_gravity()
        native.step_batch(bat, threads=4)
    for a, b in zip(seq, bat):
        assert np.array_equal(a.get("x"), b.get("x"))
        assert np.array_equal(a.get("v"), b.get("v"))
        sa, sb = a.stats(), b.stats()
        assert sa["pcg_iterations"] == sb["pcg_iterations"]
        assert sb["batched_solves"] == sb["pcg_solves"] > 0 and sb["pcg_autotuned"] == 0
        assert (sb["spmv_lanes"], sb["spmv_wg_slices"], sb["pcg_vec_block"]) == (2, 1, 256)


def test_batch_mixed_modes_one_thread_per_cloth():
    """Twenty cloths on twenty threads (large batched rounds), with a
    corrected-mode context among them whose solves keep their own launches
    (coupling blocks): every result equals sequential stepping."""
    print("batch of 20 with a corrected-mode context", flush=True)
    scs = scenes.batch256(count=19) + [scenes.stack_resting(s=17, sheets=2)]
    seq = [native.GpuSim(sc) for sc in scs]
    bat = [native.GpuSim(sc) for sc in scs]
    for g in (seq[-1], bat[-1]):
        g.set_mode(corrected=True)
    for _ in range(3):
        for g in seq:
            g.substep()
        for g in bat:
            g.apply_gravity()
        native.step_batch(bat, threads=len(bat))
    for a, b in zip(seq, bat):
        assert np.array_equal(a.get("x"), b.get("x"))
        assert np.array_equal(a.get("v"), b.get("v"))
    assert bat[0].stats()["batched_solves"] > 0
    # The stacked sheets touch: solves with coupling rows keep their own launches.
    sl = bat[-1].stats()
    assert sl["batched_solves"] < sl["pcg_solves"]


def test_batch_corrected_physics_equals_sequential():
    """Corrected-physics cloths (self-contact view, per-iteration detection,
    CCD line search): solves without coupling rows join the batched launches,
    and every result equals stepping the cloths one by one, bit for bit."""
    print("batch of 12 corrected-physics cloths", flush=True)
    scs = scenes.batch256(count=12)
    seq = [native.GpuSim(sc) for sc in scs]
    bat = [native.GpuSim(sc) for sc in scs]
    for g in seq + bat:
        g.set_mode(fp64_near=True, corrected_physics=True)
    batched = solves = 0
    for _ in range(4):
        for g in seq:
            g.substep()
        for g in bat:
            g.apply_gravity()
        native.step_batch(bat, threads=6)
        batched += sum(g.stats()["batched_solves"] for g in bat)
        solves += sum(g.stats()["pcg_solves"] for g in bat)
    for a, b in zip(seq, bat):
        assert np.array_equal(a.get("x"), b.get("x"))
        assert np.array_equal(a.get("v"), b.get("v"))
        sa, sb = a.stats(), b.stats()
        assert sa["pcg_iterations"] == sb["pcg_iterations"] and sa["newton_iterations"] == sb["newton_iterations"]
    assert batched > 0 and solves > 0


def test_batch_with_rigid_bodies_and_errors():
    scs = [scenes.small_sheet(9, pins=0, wall=False, rigid=True), scenes.drop21()]
    seq = [native.GpuSim(sc) for sc in scs]
    bat = [native.GpuSim(sc) for sc in scs]
    for _ in range(3):
        for g in seq:
            g.substep()
        for g in bat:
            g.apply_gravity()
        native.step_batch(bat)
    for a, b in zip(seq, bat):
        assert np.array_equal(a.get("x"), b.get("x"))
    assert np.array_equal(seq[0].rigid_state(0)["position"], bat[0].rigid_state(0)["position"])
    with pytest.raises(RuntimeError, match="twice"):
        native.step_batch([bat[0], bat[0]])
    native.step_batch([])


def test_concurrent_calls_on_one_context_are_serialised():
    """ctypes releases the GIL during a foreign call, so two threads stepping
    the same context really call ando_step at once; the context lock
    serialises them, and the result equals two steps one after the other."""
    import threading
    sc = scenes.small_sheet(17, pins=2, wall=True)
    a, b = native.GpuSim(sc), native.GpuSim(sc)
    errors = []

    def run():
        try:
            a.step()
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    ts = [threading.Thread(target=run) for _ in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    b.step()
    b.step()
    assert np.array_equal(a.get("x"), b.get("x"))
    assert np.array_equal(a.get("v"), b.get("v"))
