"""Generation telemetry and the adaptive shard release on the GPU.

Bar: the fitness row (order statistics, mean and standard deviation from the reference's ascending double
sums, distinct count, breeding metrics) equals the oracle's bit for bit; the convergence row (centroid and
pairwise distances over the reference's sample plan) equals the oracle's float arithmetic bit for bit; the
controller releases exactly when the oracle's gates say so, records the telemetry p99, and the released
words are injected and played.
"""
import numpy as np
import pytest

import telemetry as tlo

pytestmark = pytest.mark.gpu

FIELDS = ("fitness_min", "fitness_mean", "fitness_median", "fitness_p90", "fitness_p99", "fitness_max",
          "fitness_stddev")


def _runtime(pkg, side, A, catalog, offsets=None, ids=None, seed=9, gen_A=None, slack=2):
    rtm = pkg.runtime
    stride = rtm.slot_stride_bytes(gen_A or A)
    n = side * side
    rt = pkg.SlabRuntime(slab_bytes=slack * n * stride + 4 * stride, generation_bytes=n * stride, action_count=A,
                         grid_columns=side)
    rt.upload_catalog(catalog)
    if offsets is not None:
        rt.upload_secrets_csr(offsets, ids)
    rt.bootstrap(n, seed=seed, variance_scale=1.0, tail_std=0.1)
    return rt


def _check_fitness_row(row, fit, pairs=None):
    ref = tlo.fitness_telemetry(fit, pairs=pairs)
    for k in FIELDS:
        assert np.float32(getattr(row, k)).view(np.uint32) == np.float32(ref[k]).view(np.uint32), k
    assert row.population_size == fit.size and row.distinct_fitness_count == ref["distinct_fitness_count"]
    if pairs is not None:
        assert row.has_breeding_metrics == 1
        assert (row.parent_childless_count, row.parent_one_child_count, row.parent_multiple_children_count) == (
            ref["parent_childless_count"], ref["parent_one_child_count"], ref["parent_multiple_children_count"])


def test_fitness_telemetry_matches_oracle(pkg):
    rtm = pkg.runtime
    wl = pkg.synth.workload("grid16_a512")
    counts = rtm.word_counts(512, 512)
    rt = _runtime(pkg, 16, 512, wl["catalog"], wl["offsets"], wl["ids"])
    try:
        with pytest.raises(RuntimeError) as err:  # generation 0 is not evaluated yet
            rt.fitness_telemetry(counts)
        assert err.value.status == 9
        rt.evaluate(counts)
        _, fit, _, _ = rt.generation(0)
        row = rt.fitness_telemetry(counts)
        _check_fitness_row(row, fit)
        assert row.generation == 0 and row.training_word_count == 512 and row.has_breeding_metrics == 0
        with pytest.raises(RuntimeError) as err:  # no plan yet
            rt.fitness_telemetry(counts, breeding=True)
        assert err.value.status == 12
        # at the pre-recombination boundary the plan is there: breeding metrics
        cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
        rt.advance(9, counts, cfg)
        ck = rt.create_checkpoint(10, counts, cfg)
        _, fit, _, _ = rt.generation(0)
        _check_fitness_row(rt.fitness_telemetry(counts, breeding=True), fit, rt.plan())
        rt.resume(ck)
        ck.close()
    finally:
        rt.close()


@pytest.mark.parametrize("side,cols", [(12, 16), (128, 128)])
def test_fitness_telemetry_sizes(pkg, side, cols):
    """A population that is not a power of two (192: the sort pads to 256) and the largest one the LDS sort
    holds (16,384); one more grid row is refused."""
    rtm = pkg.runtime
    A = 16
    cat = pkg.synth.make_catalog(A)
    n = side * cols
    stride = rtm.slot_stride_bytes(A)
    offsets = np.arange(n + 1, dtype=np.uint32) * 2
    ids = np.tile(np.array([0, 5], np.uint32), n)
    rt = pkg.SlabRuntime(slab_bytes=n * stride + 2 * cols * stride, generation_bytes=(n + cols) * stride,
                         action_count=A, grid_columns=cols)
    try:
        rt.upload_catalog(cat)
        rt.upload_secrets_csr(offsets, ids)
        rt.bootstrap(n, seed=3, variance_scale=1.0, tail_std=0.1)
        counts = rtm.word_counts(A, A)
        rt.evaluate(counts)
        _, fit, _, _ = rt.generation(0)
        _check_fitness_row(rt.fitness_telemetry(counts), fit)
        if n == 16384:
            off2 = np.arange(n + cols + 1, dtype=np.uint32) * 2
            rt.upload_secrets_csr(off2, np.tile(np.array([0, 5], np.uint32), n + cols))
            rt.bootstrap(n + cols, seed=3, variance_scale=1.0, tail_std=0.1)
            rt.evaluate(counts)
            with pytest.raises(RuntimeError) as err:
                rt.fitness_telemetry(counts)
            assert err.value.status == 18
    finally:
        rt.close()


def test_convergence_telemetry_matches_oracle(pkg):
    rtm = pkg.runtime
    wl = pkg.synth.workload("grid16_a512")
    rt = _runtime(pkg, 16, 512, wl["catalog"], wl["offsets"], wl["ids"])
    try:
        cfg = rtm.default_assembly_config(mutation_probability=0.05, mutation_sigma=0.02)
        counts = rtm.word_counts(512, 512)
        for g in range(2):
            rt.advance(9 + g, counts, cfg)
        for seed in (7, 1234):
            row = rt.convergence_telemetry(seed)
            slots, _, _, _ = rt.generation(0)
            org, w, pairs = tlo.sample_plan(256, 132928 + 38 * 512, seed, 2)
            values = np.stack([rt.download_slot(int(slots[o]))[w] for o in org])
            ref = tlo.convergence_metrics(values, pairs)
            assert (row.generation, row.sample_organisms, row.sample_weights, row.pair_count) == (2, 256, 8192, 512)
            for k, v in ref.items():
                assert np.float32(getattr(row, k)).view(np.uint32) == np.float32(v).view(np.uint32), (seed, k)
            assert row.elapsed_ms > 0
    finally:
        rt.close()


def _catalog_with_injectable_tail(pkg, orc, n, first):
    cat = pkg.synth.make_catalog(4 * n, seed=0x1234)
    ok = np.array([orc.hint_group(w)[0] == 1 for w in cat])
    rest = [w for w, good in zip(cat[first:], ok[first:]) if good]
    return np.concatenate([cat[:first], np.stack(rest[:n - first])]).astype(np.uint8)


@pytest.mark.parametrize("release", [True, False])
def test_adaptive_shard_release_end_to_end(pkg, orc, release):
    """The reference CLI's loop: the shard schedule from word counts and the release history, the
    controller in the training-shard release callback, the released words injected into the children and
    added to the action space and the training words after the step."""
    rtm = pkg.runtime
    A0, step, total, side = 32, 16, 96, 8
    cat = _catalog_with_injectable_tail(pkg, orc, total, A0)
    rt = _runtime(pkg, side, A0, cat, gen_A=total)
    policy = rtm.default_release_policy(min_gap_generations=2, first_new_release_generation=3,
                                        centroid_distance_threshold=1e9 if release else -1.0,
                                        fitness_p99_threshold=-1.0, run_seed=7)
    hist = rtm.ReleaseHistory()
    hist.record(0, 0.0)  # the foundation shard
    words = A0
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    log = []
    try:
        for g in range(8):
            counts = rtm.word_counts(words, words, initial=A0, step=step, period=1)
            state = {}

            def on_release(_rt, counts_ptr, inj_ptr, _user, g=g, counts=counts, state=state):
                try:
                    gens, _ = hist.releases()
                    expect = tlo.release_decision(words, step, total, g, gens, dict(
                        min_gap_generations=2, first_new_release_generation=3,
                        centroid_distance_threshold=policy.centroid_distance_threshold,
                        fitness_p99_threshold=policy.fitness_p99_threshold))
                    _, fit, _, _ = rt.generation(0)
                    p99 = tlo.fitness_telemetry(fit)["fitness_p99"]
                    released, inj, nxt = rt.maybe_release_shard(counts, policy, total, hist)
                    state.update(expect=expect, released=released, nxt=nxt, p99=p99)
                    if released:
                        inj_ptr.contents.enabled = 1
                        inj_ptr.contents.first_catalog_word_index = inj.first_catalog_word_index
                        inj_ptr.contents.injection_count = inj.injection_count
                    return 0
                except Exception as e:  # a failing callback fails the step
                    state["error"] = repr(e)
                    return 1

            cbs = rtm.Callbacks(rtm.POST_FITNESS_CB(0), rtm.SHARD_RELEASE_CB(on_release), None)
            gens, _ = hist.releases()
            rt.advance(9 + g, counts, cfg, callbacks=cbs, history=gens)
            assert "error" not in state, state.get("error")
            exp = state["expect"]
            if exp == "measure":
                assert state["released"] == release, g
            else:
                assert not state["released"], g
            if state["released"]:
                assert hist.p99[hist.count.value - 1] == state["p99"]
                words = state["nxt"]
            log.append((g, state["released"], words))
        gens, p99 = hist.releases()
        if release:
            assert gens == [0, 3, 5, 7] and words == 80 and rt.info().action_count == 80
        else:
            assert gens == [0] and words == A0 and rt.info().action_count == A0
        # the last released shard is played: the evaluation covers all released words
        rt.evaluate(rtm.word_counts(words, words, initial=A0, step=step, period=1), history=gens)
        assert rt.stage_times().games > 0
    finally:
        rt.close()
