"""Host spill of assembled children on the GPU.

Bar: on a slab too small to hold the parents and all children at once, a runtime with a host spill budget
assembles the same children (bit for bit, by organism) as a runtime with room to spare; its slot ids and
free stack follow the oracle's bookkeeping with one spill of the assembled prefix; without a budget, or with
one smaller than the prefix, the step fails as the reference's does.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _make(pkg, slots, spill_bytes, A=48, side=8):
    rtm = pkg.runtime
    cat = pkg.synth.make_catalog(A)
    sec = pkg.synth.sample_secrets(A, 16)
    off, ids = pkg.synth.block_csr(side, side, 4, sec)
    n = side * side
    stride = rtm.slot_stride_bytes(A)
    rt = pkg.SlabRuntime(slab_bytes=slots * stride, generation_bytes=n * stride, action_count=A, grid_columns=side,
                         host_spill_bytes=spill_bytes)
    rt.upload_catalog(cat)
    rt.upload_secrets_csr(off, ids)
    rt.bootstrap(n, seed=3, variance_scale=2.0, tail_std=0.5)
    return rt, stride


def test_spilled_assembly_equals_roomy_assembly(pkg, orc):
    rtm = pkg.runtime
    n = 64
    counts = rtm.word_counts(48, 48)
    cfg = rtm.default_assembly_config(mutation_probability=0.02, mutation_sigma=0.02)
    roomy, stride = _make(pkg, 3 * n, 0)
    try:
        roomy.advance(99, counts, cfg)
        pairs = roomy.plan()
        slots, _, _, _ = roomy.generation(0)
        expect = [roomy.download_slot(s).view(np.uint32).copy() for s in slots]
        assert roomy.stage_times().spilled == 0
    finally:
        roomy.close()
    # the smallest slab on which the plan needs exactly one spill (the plan does not depend on the slab)
    extra = None
    for e in range(1, n):
        stack, fc, state, cur = orc.slab_after_bootstrap(n + e, n)
        if orc.assemble_slots(n + e, stack, fc, state, cur, pairs)["status"] == 0:
            break
        stack, fc, state, cur = orc.slab_after_bootstrap(n + e, n)
        r = orc.assemble_slots(n + e, stack, fc, state, cur, pairs, spill_capacity=n)
        if r["status"] == 0:
            extra, ref = e, r
            break
    assert extra is not None, "no slab size in range needs a single spill"
    # without a budget the step fails with slab full; with one too small for the prefix, device failure
    cases = [(0, 14)] + ([((ref["spilled"] - 1) * stride, 1)] if ref["spilled"] > 1 else [])
    for budget, status in cases:
        rt, _ = _make(pkg, n + extra, budget)
        try:
            with pytest.raises(RuntimeError) as err:
                rt.advance(99, counts, cfg)
            assert err.value.status == status
        finally:
            rt.close()
    rt, _ = _make(pkg, n + extra, n * stride)
    try:
        rt.advance(99, counts, cfg)
        assert rt.plan().tolist() == pairs.tolist()
        assert rt.stage_times().spilled == ref["spilled"] > 0
        slots, _, _, _ = rt.generation(0)
        assert slots.tolist() == ref["child_slots"].tolist()
        assert rt.info().free_slot_count == ref["free_count"]
        for i, s in enumerate(slots):
            assert np.array_equal(rt.download_slot(s).view(np.uint32), expect[i]), i
        # the next step runs on the spilled-and-restored generation
        rt.advance(100, counts, cfg)
    finally:
        rt.close()
