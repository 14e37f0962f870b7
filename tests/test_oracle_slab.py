"""Pins the oracle's slot bookkeeping against the reference's known answers
(tests/genetic_algorithm/device/genotype_slab_device_runtime_test.cu)."""
import numpy as np


def test_slot_reuse_order_known_answer(orc):
    """genotype_slab_device_runtime_test.cu:447-538: 3 slots, parents in slots 0,1,2, plan
    (0,1),(0,0): child 0 reuses the collected zero-reference parent's slot 2, child 1 the slot 1 freed
    after parent 1's last reference."""
    stack, fc, state, cur = orc.slab_after_bootstrap(3, 3)
    r = orc.assemble_slots(3, stack, fc, state, cur, np.array([[0, 1], [0, 0]], np.uint32))
    assert r["status"] == 0
    assert r["child_slots"].tolist() == [2, 1]
    assert r["batches"] == 2
    assert r["free_count"] == 1 and r["stack"][0] == 0  # parent 0 freed after its last child


def test_genuinely_full_slab(orc):
    """genotype_slab_device_runtime_test.cu:1084-1146: 2 slots, 2 parents both referenced -> slab full."""
    stack, fc, state, cur = orc.slab_after_bootstrap(2, 2)
    r = orc.assemble_slots(2, stack, fc, state, cur, np.array([[0, 1], [1, 0]], np.uint32))
    assert r["status"] == 14


def test_every_parent_freed_and_children_distinct(orc):
    rng = np.random.default_rng(5)
    n, slots = 64, 80
    stack, fc, state, cur = orc.slab_after_bootstrap(slots, n)
    pairs = rng.integers(0, n, size=(n, 2)).astype(np.uint32)
    r = orc.assemble_slots(slots, stack, fc, state, cur, pairs)
    assert r["status"] == 0
    assert len(set(r["child_slots"].tolist())) == n
    assert (r["parent_slots"] == 0xFFFFFFFF).all()
    assert r["free_count"] == slots - n
    assert r["state"].sum() == n


def test_host_spill_known_answer(orc):
    """3 slots, parents in slots 0 and 1, plan (0,1),(0,1): child 0 takes slot 2, then the slab is full with
    both parents still referenced.  A spill of one child (slab_runtime.cu:1554-1661) puts slot 2 back on
    the stack; child 1 takes it, both parents are freed (parent order: 0 then 1), and the spilled child 0
    returns to the top of the stack, slot 1."""
    pairs = np.array([[0, 1], [0, 1]], np.uint32)
    stack, fc, state, cur = orc.slab_after_bootstrap(3, 2)
    assert orc.assemble_slots(3, stack, fc, state, cur, pairs)["status"] == 14
    stack, fc, state, cur = orc.slab_after_bootstrap(3, 2)
    r = orc.assemble_slots(3, stack, fc, state, cur, pairs, spill_capacity=1)
    assert r["status"] == 0 and r["spilled"] == 1
    assert r["child_slots"].tolist() == [1, 2]
    assert r["free_count"] == 1 and r["stack"][0] == 0
    assert r["state"].tolist() == [0, 1, 1]


def test_host_spill_limits(orc):
    # nothing assembled yet: a spill frees nothing (TryCountAssembledChildPrefix returns 0)
    stack, fc, state, cur = orc.slab_after_bootstrap(2, 2)
    r = orc.assemble_slots(2, stack, fc, state, cur, np.array([[0, 1], [1, 0]], np.uint32), spill_capacity=8)
    assert r["status"] == 14
    # a spill slab smaller than the prefix: the reference reports a device failure
    stack, fc, state, cur = orc.slab_after_bootstrap(4, 2)
    r = orc.assemble_slots(4, stack, fc, state, cur, np.array([[0, 1]] * 3, np.uint32), spill_capacity=1)
    assert r["status"] == 1
    # one spill only: a second full slab stays full
    stack, fc, state, cur = orc.slab_after_bootstrap(3, 2)
    r = orc.assemble_slots(3, stack, fc, state, cur, np.array([[0, 1]] * 3, np.uint32), spill_capacity=8)
    assert r["status"] == 14


def test_host_spill_random_plans(orc):
    rng = np.random.default_rng(11)
    n, succeeded = 64, 0
    for extra in range(1, 24):
        pairs = rng.integers(0, n, size=(n, 2)).astype(np.uint32)
        stack, fc, state, cur = orc.slab_after_bootstrap(n + extra, n)
        plain = orc.assemble_slots(n + extra, stack, fc, state, cur, pairs)
        stack, fc, state, cur = orc.slab_after_bootstrap(n + extra, n)
        r = orc.assemble_slots(n + extra, stack, fc, state, cur, pairs, spill_capacity=n)
        if plain["status"] == 0:
            assert r["status"] == 0 and r["spilled"] == 0
            assert r["child_slots"].tolist() == plain["child_slots"].tolist()
        if r["status"] == 0:
            assert len(set(r["child_slots"].tolist())) == n and r["state"].sum() == n
            assert (r["parent_slots"] == 0xFFFFFFFF).all() and r["free_count"] == extra
            succeeded += r["spilled"] > 0
        else:
            assert r["status"] == 14
    assert succeeded > 0
