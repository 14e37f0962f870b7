"""Parity of the HIP path against the CPU oracle on the same inputs and the same counter-based draws.

Bar: bit-exact for every integer stage (feedback, episode bookkeeping, selection, crossover decisions,
slot bookkeeping) and for the draw-driven fp32 stages (bootstrap, mutation); fp32 forward within 1e-5
(normwise, against the host fp64 forward on the states the GPU visited) with argmax agreement and equal
fitness wherever the top-2 score margin exceeds the tolerance.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5        # normwise relative error of the fp32 policy/scores vs the host fp64 forward
MARGIN_TOL = 1e-4     # decisions whose fp64 top-2 margin is below this are near-ties


def _runtime(pkg, side, A, catalog, offsets, ids, seed=9, var=1.0, tail=0.1, slack=3):
    rtm = pkg.runtime
    stride = rtm.slot_stride_bytes(A)
    n = side * side
    rt = pkg.SlabRuntime(slab_bytes=slack * n * stride, generation_bytes=n * stride, action_count=A,
                         grid_columns=side)
    rt.upload_catalog(catalog)
    rt.upload_secrets_csr(offsets, ids)
    rt.bootstrap(n, seed=seed, variance_scale=var, tail_std=tail)
    return rt


@pytest.fixture(scope="module")
def grid16(pkg):
    """The 16x16 / A=512 workload (256 genomes, 16 secrets per cell)."""
    wl = pkg.synth.workload("grid16_a512")
    rt = _runtime(pkg, 16, 512, wl["catalog"], wl["offsets"], wl["ids"])
    yield wl, rt
    rt.close()


def test_bootstrap_bit_exact(pkg, orc, grid16):
    wl, rt = grid16
    slots, _, _, _ = rt.generation(0)
    assert list(slots) == list(range(256)), "generation 0 takes slots 0..N-1 in order"
    stds = orc.layer_stds(1.0)
    n_vals = orc.policy_param_count() + 38 * 512
    for organism in (0, 1, 77, 255):
        gpu = rt.download_slot(slots[organism])
        ref = orc.bootstrap_genome(512, 9, 0, organism, stds, 0.1)
        assert np.array_equal(gpu[:n_vals].view(np.uint32), ref.view(np.uint32)), organism
        assert not gpu[n_vals:].any()


def _check_evaluation(pkg, orc, rt, catalog, offsets, ids, A, n_genomes, min_same=0.99):
    """Evaluates generation 0 with traces and checks it against the oracle (see the module bar)."""
    rt.set_trace(pkg.runtime.TRACE_GAMES | pkg.runtime.TRACE_POLICY)
    counts = pkg.runtime.word_counts(A, A)
    rt.evaluate(counts)
    slots, fit, evals, has = rt.generation(0)
    assert has.all() and (evals == 1).all()
    games = 3 * int(offsets[-1])
    gtr = rt.download_trace(pkg.runtime.TRACE_GAMES, games)
    ptr_ = rt.download_trace(pkg.runtime.TRACE_POLICY, games)
    genomes = [rt.download_slot(s) for s in slots]
    ref = orc.evaluate(genomes, 0, A, A, catalog, offsets, ids, mode=0, threads=8, trace=True)
    assert ref["status"] == 0
    rtr = ref["traces"]
    assert int(rt.stage_times().decisions) == int(gtr["decisions"].sum())
    for n in range(n_genomes):  # fresh games always decide; prefilled ones may win on the prefill (W = 2, 4)
        W = int(offsets[n + 1] - offsets[n])
        assert (gtr["decisions"][3 * int(offsets[n]):3 * int(offsets[n]) + W] > 0).all()
    words = np.asarray(catalog).reshape(-1, 5)

    max_err = 0.0
    checked = near = 0
    genome_clean = np.ones(n_genomes, bool)
    for n in range(n_genomes):
        W = int(offsets[n + 1] - offsets[n])
        local = ids[offsets[n]:offsets[n + 1]]
        base = 3 * int(offsets[n])
        for slot in range(3 * W):
            g = gtr[base + slot]
            d = int(g["decisions"])
            # a guessed word (any copy of it) is never guessed again
            played = [bytes(words[a]) for a in g["guesses"][:d]]
            assert len(set(played)) == d, (n, slot, g["guesses"][:d])
            rep = orc.replay_game(genomes[n], A, catalog, local, slot, g["guesses"][:d])
            assert rep["status"] == 0
            for k in range(d):
                p64 = rep["policy"][k]
                p32 = ptr_[base + slot, k].astype(np.float64)
                err = np.max(np.abs(p32 - p64)) / max(np.max(np.abs(p64)), 1e-30)
                max_err = max(max_err, err)
                margin = rep["best"][k] - rep["second"][k]
                checked += 1
                if margin > MARGIN_TOL * max(1.0, abs(rep["best"][k])):
                    assert rep["argmax"][k] == g["guesses"][k], (n, slot, k, margin)
                else:
                    near += 1
                    genome_clean[n] = False
                # the chosen action's fp32 score against fp64
                assert abs(g["best"][k] - rep["chosen"][k]) <= FWD_TOL * max(1.0, np.max(np.abs(p64))) * 64
    assert max_err <= FWD_TOL, max_err
    # fitness identical wherever no decision of the genome was a near-tie
    assert np.array_equal(fit[genome_clean], ref["fitness"][genome_clean])
    # and the GPU's games equal the oracle's fp32 games wherever the oracle's own margins are clear
    same = np.mean(gtr["guesses"] == rtr["guesses"])
    assert same >= min_same, same
    print(f"decisions {checked}, near-ties {near}, max normwise fwd err {max_err:.2e}, guess agreement {same:.5f}")


def test_evaluation_parity_grid16(pkg, orc, grid16):
    wl, rt = grid16
    _check_evaluation(pkg, orc, rt, wl["catalog"], wl["offsets"], wl["ids"], 512, 256)


def test_evaluation_parity_duplicate_words_ragged_shards(pkg, orc):
    """Duplicate catalog words (a guess masks every copy), ragged cells of 1..70 secrets (workgroups of
    2 waves up to split workgroups of 8 waves over 14 tiles) and 23 action tiles (a half-used last chunk)."""
    A, side = 365, 4
    rng = np.random.default_rng(0xD0B1E)
    catalog = pkg.synth.make_catalog(A, seed=0xD0B1E).reshape(-1, 5).copy()
    for k in range(40, A, 9):  # every 9th word from 40 on repeats an earlier word
        catalog[k] = catalog[int(rng.integers(0, k))]
    # cells of 1 or 3 secrets make a prefilled episode repeat its solution (GuessAppendFailed)
    widths = np.array([2, 4, 5, 16, 17, 33, 70, 6, 9, 48, 2, 64, 7, 31, 8, 10])
    offsets = np.concatenate([[0], np.cumsum(widths)]).astype(np.uint32)
    # secrets: distinct words within a cell, so drawn from first occurrences only
    firsts = np.array(sorted({bytes(w): i for i, w in reversed(list(enumerate(catalog)))}.values()))
    ids = np.concatenate([rng.choice(firsts, size=w, replace=False) for w in widths]).astype(np.uint32)
    rt = _runtime(pkg, side, A, catalog, offsets, ids, seed=21)
    try:
        _check_evaluation(pkg, orc, rt, catalog, offsets, ids, A, side * side, min_same=0.98)
    finally:
        rt.close()


def test_summary_matches_oracle(pkg, orc, grid16):
    wl, rt = grid16
    rt.evaluate(pkg.runtime.word_counts(512, 512))
    _, fit, _, _ = rt.generation(0)
    s = rt.summary()
    b, m, i = orc.summary(fit)
    assert s.best_fitness == b and s.average_fitness == m and s.best_index == i


def test_plan_bit_exact(pkg, orc, grid16):
    wl, rt = grid16
    counts = pkg.runtime.word_counts(512, 512)
    rt.evaluate(counts)
    _, fit, _, has = rt.generation(0)
    # plant a varied fitness field through the trace-free path: advance plans from the GPU fitness
    cfg = pkg.runtime.default_assembly_config(mutation_probability=0.0)
    seed = 7 + 2
    pre_slots, _, _, _ = rt.generation(0)
    rt.advance(seed, counts, cfg)
    gpu_pairs = rt.plan()
    w = orc.rank_weights(0.5, 256)
    st, ref_pairs = orc.plan(fit, has, 256, 256, 16, 0, 2, w, seed)
    assert st == 0
    assert np.array_equal(gpu_pairs, ref_pairs)
    assert (gpu_pairs[:, 0] != gpu_pairs[:, 1]).all()


def test_plan_varied_fitness_and_radii(pkg, orc):
    """Selection on a planted fitness field with ties, radii 1, 2 and 3 (scan path)."""
    cat = pkg.synth.make_catalog(64)
    sec = pkg.synth.sample_secrets(64, 32)
    off, ids = pkg.synth.block_csr(8, 8, 4, sec)
    for radius in (1, 2, 3):
        rt = _runtime(pkg, 8, 64, cat, off, ids, seed=31, var=2.0, tail=0.5)
        counts = pkg.runtime.word_counts(64, 64)
        cfg = pkg.runtime.default_assembly_config(cellular_breeding_radius=radius, rank_exponent=0.75)
        rt.advance(1234 + radius, counts, cfg)
        gpu_pairs = rt.plan()
        rt.close()
        rt2 = _runtime(pkg, 8, 64, cat, off, ids, seed=31, var=2.0, tail=0.5)
        rt2.evaluate(counts)
        _, fit, _, has = rt2.generation(0)
        rt2.close()
        w = orc.rank_weights(0.75, 64)
        st, ref_pairs = orc.plan(fit, has, 64, 64, 8, 0, radius, w, 1234 + radius)
        assert st == 0 and np.array_equal(gpu_pairs, ref_pairs), radius


def test_assembly_bit_exact(pkg, orc):
    cat = pkg.synth.make_catalog(96)
    sec = pkg.synth.sample_secrets(96, 32)
    off, ids = pkg.synth.block_csr(8, 8, 4, sec)
    A = 96
    rt = _runtime(pkg, 8, A, cat, off, ids, seed=5, var=1.0, tail=0.1)
    counts = pkg.runtime.word_counts(96, A)
    slots, _, _, _ = rt.generation(0)
    parents = [rt.download_slot(s) for s in slots]
    temps = (0.3, 0.2, 0.4)
    cfg = pkg.runtime.default_assembly_config(crossover_temperature_level1=temps[0],
                                              crossover_temperature_level2=temps[1],
                                              crossover_temperature_level3=temps[2], mutation_probability=0.05,
                                              mutation_sigma=0.02, output_tail_row_scale_mutation_probability=0.1)
    seed = 4242
    rt.advance(seed, counts, cfg)
    pairs = rt.plan()
    xrec = rt.download_trace(pkg.runtime.TRACE_XOVER, 64)
    child_slots, _, _, _ = rt.generation(0)
    assert rt.info().current_generation_index == 1
    n_vals = orc.policy_param_count() + 38 * A
    for c in range(64):
        a, b = pairs[c]
        ref, rec = orc.breed_child(parents[a][:n_vals], parents[b][:n_vals], A, A, seed, 0, c, *temps, 0.05, 0.02, 0.1)
        got = rt.download_slot(child_slots[c])
        assert np.array_equal(got[:n_vals].view(np.uint32), ref.view(np.uint32)), c
        assert not got[n_vals:].any()
        assert np.array_equal(xrec[c], rec), c
    rt.close()


def test_copy_first_parent_when_temperatures_and_mutation_are_zero(pkg, orc):
    """Reference KAT (slab_runtime_test.cu:414-481): temperatures 0 and mutation 0 -> child == parent 1."""
    cat = pkg.synth.make_catalog(48)
    sec = pkg.synth.sample_secrets(48, 16)
    off, ids = pkg.synth.block_csr(4, 4, 2, sec)
    rt = _runtime(pkg, 4, 48, cat, off, ids)
    slots, _, _, _ = rt.generation(0)
    parents = [rt.download_slot(s) for s in slots]
    cfg = pkg.runtime.default_assembly_config(crossover_temperature_level1=0.0, crossover_temperature_level2=0.0,
                                              crossover_temperature_level3=0.0, mutation_probability=0.0)
    rt.advance(77, pkg.runtime.word_counts(48, 48), cfg)
    pairs = rt.plan()
    children, _, _, _ = rt.generation(0)
    for c in range(16):
        assert np.array_equal(rt.download_slot(children[c]), parents[pairs[c, 0]])
    rt.close()


def test_masking_end_to_end(pkg):
    """Reference KAT (single_model_device_runtime_test.cu:133-170): a zero genome scores every action 0,
    so it guesses the lowest unmasked index: CRANE, then SLATE (CRANE masked), then TRACE and wins."""
    words = np.stack([np.frombuffer(w.encode(), np.uint8) - 65 for w in ("CRANE", "SLATE", "TRACE")])
    rtm = pkg.runtime
    stride = rtm.slot_stride_bytes(3)
    rt = pkg.SlabRuntime(slab_bytes=4 * stride, generation_bytes=stride, action_count=3, grid_columns=1)
    rt.upload_catalog(words)
    rt.upload_secrets_csr(np.array([0, 1], np.uint32), np.array([2], np.uint32))
    rt.bootstrap(1, seed=1)
    slots, _, _, _ = rt.generation(0)
    rt.upload_slot(slots[0], np.zeros(stride // 4, np.float32))
    rt.set_trace(rtm.TRACE_GAMES)
    # W = 1: the variant-1 prefill equals the solution, so the reference fails with kGuessAppendFailed
    with pytest.raises(RuntimeError) as err:
        rt.evaluate(rtm.word_counts(3, 3))
    assert err.value.status == 6
    rt.close()
    rt = pkg.SlabRuntime(slab_bytes=4 * stride, generation_bytes=stride, action_count=3, grid_columns=1)
    rt.upload_catalog(words)
    rt.upload_secrets_csr(np.array([0, 2], np.uint32), np.array([2, 0], np.uint32))  # W = 2
    rt.bootstrap(1, seed=1)
    slots, _, _, _ = rt.generation(0)
    rt.upload_slot(slots[0], np.zeros(stride // 4, np.float32))
    rt.set_trace(rtm.TRACE_GAMES)
    rt.evaluate(rtm.word_counts(3, 3))
    tr = rt.download_trace(rtm.TRACE_GAMES, 6)
    # fresh game for secret TRACE: CRANE, SLATE, TRACE -> won at turn 3, score 13
    assert list(tr[0]["guesses"][:3]) == [0, 1, 2] and tr[0]["score"] == 13.0
    # W = 2, variant 1 of entry 0 prefills local[1], local[0] = CRANE, TRACE: won by the prefill, score 14
    assert tr[2]["decisions"] == 0 and tr[2]["score"] == 14.0
    rt.close()


def test_slot_bookkeeping_tight_slab(pkg, orc):
    """A slab with room for only 20 extra genomes: children are assembled in several batches as parents
    are released; slot assignment equals the oracle's bookkeeping on the same plan."""
    cat = pkg.synth.make_catalog(48)
    sec = pkg.synth.sample_secrets(48, 16)
    off, ids = pkg.synth.block_csr(8, 8, 4, sec)
    rtm = pkg.runtime
    stride = rtm.slot_stride_bytes(48)
    rt = pkg.SlabRuntime(slab_bytes=(64 + 20) * stride, generation_bytes=64 * stride, action_count=48,
                         grid_columns=8)
    rt.upload_catalog(cat)
    rt.upload_secrets_csr(off, ids)
    rt.bootstrap(64, seed=3, variance_scale=2.0, tail_std=0.5)
    rt.advance(99, rtm.word_counts(48, 48), rtm.default_assembly_config(mutation_probability=0.0))
    pairs = rt.plan()
    children, _, _, _ = rt.generation(0)
    info = rt.info()
    stack, fc, state, cur = orc.slab_after_bootstrap(84, 64)
    r = orc.assemble_slots(84, stack, fc, state, cur, pairs)
    assert r["status"] == 0
    assert children.tolist() == r["child_slots"].tolist()
    assert info.free_slot_count == r["free_count"]
    rt.close()


def test_genuinely_full_slab_reports_slab_full(pkg):
    cat = pkg.synth.make_catalog(8)
    rtm = pkg.runtime
    stride = rtm.slot_stride_bytes(8)
    rt = pkg.SlabRuntime(slab_bytes=2 * stride, generation_bytes=2 * stride, action_count=8, grid_columns=1)
    rt.upload_catalog(cat)
    rt.upload_secrets_csr(np.array([0, 4, 8], np.uint32), np.arange(8, dtype=np.uint32))  # W = 4
    rt.bootstrap(2, seed=3)
    with pytest.raises(RuntimeError) as err:
        rt.advance(5, rtm.word_counts(8, 8))
    assert err.value.status == 14
    rt.close()


def test_cpp_driver_runs_generations(pkg, tmp_path):
    """The C++ host API (host/slab_runtime.hpp) end to end: bootstrap, two TryAdvanceGenerationOnDevice
    calls with a post-fitness callback, a final evaluation (cf. the reference's CLI smoke test,
    CMakeLists.txt:708-714)."""
    import os
    import subprocess
    cat = pkg.synth.make_catalog(64)
    sec = pkg.synth.sample_secrets(64, 32)
    off, ids = pkg.synth.block_csr(8, 8, 4, sec)
    cat.astype(np.uint8).tofile(tmp_path / "catalog.u8")
    off.astype(np.uint32).tofile(tmp_path / "offsets.u32")
    ids.astype(np.uint32).tofile(tmp_path / "secrets.u32")
    driver = os.path.join(os.path.dirname(pkg.runtime.LIB_PATH), "run_ga")
    res = subprocess.run([driver, str(tmp_path), "8", "2", "7"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    lines = res.stdout.strip().splitlines()
    # each advance reports the evaluation of the generation it bred from
    assert lines[0].startswith("generation 0:") and lines[1].startswith("generation 1:")
    assert lines[-1].startswith("final generation 2:") and "2 post-fitness callbacks" in lines[-1]


def _catalog_with_injectable_tail(pkg, orc, n, first):
    """A synthetic catalog whose words from `first` on all have a hint-grid group."""
    cat = pkg.synth.make_catalog(4 * n, seed=0x1234)
    ok = np.array([orc.hint_group(w)[0] == 1 for w in cat])
    head = cat[:first]
    rest = [w for w, good in zip(cat[first:], ok[first:]) if good]
    return np.concatenate([head, np.stack(rest[:n - first])]).astype(np.uint8)


def test_widening_repack_and_injection(pkg, orc):
    """A growth generation: repack to the wider stride, breed with the old action count, inject the new
    rows from hint grids (rows 32..47).  Old part bit-exact, new rows within fp32 tolerance, slots as
    the oracle's repack + assembly bookkeeping."""
    rtm = pkg.runtime
    A0, inj, side = 32, 16, 8
    n = side * side
    cat = _catalog_with_injectable_tail(pkg, orc, 64, A0)
    sec = pkg.synth.sample_secrets(A0, 32, seed=77)
    off, ids = pkg.synth.block_csr(side, side, 4, sec)
    s0, s1 = rtm.slot_stride_bytes(A0), rtm.slot_stride_bytes(A0 + inj)
    slab_bytes = 2 * n * s1 + 3 * s1
    rt = pkg.SlabRuntime(slab_bytes=slab_bytes, generation_bytes=n * s1, action_count=A0, grid_columns=side)
    rt.upload_catalog(cat)
    rt.upload_secrets_csr(off, ids)
    rt.bootstrap(n, seed=21, variance_scale=1.0, tail_std=0.1)
    slots0, _, _, _ = rt.generation(0)
    n0 = orc.policy_param_count() + 38 * A0
    parents = [rt.download_slot(s)[:n0] for s in slots0]
    temps = (0.1, 0.1, 0.2)
    cfg = rtm.default_assembly_config(crossover_temperature_level1=temps[0], crossover_temperature_level2=temps[1],
                                      crossover_temperature_level3=temps[2], mutation_probability=0.02)
    injection = rtm.PendingInjection(enabled=1, reserved0=0, first_catalog_word_index=A0, injection_count=inj)
    rt.advance(606, rtm.word_counts(A0, A0), cfg, injection=injection)
    info = rt.info()
    assert info.action_count == A0 + inj and info.slot_stride_bytes == s1
    # the slab layout's byte size is slot count x stride (device_runtime.cu:1047-1050) and the widened
    # layout keeps it (TryCreateExpandedSlabLayout, layout.hpp:55-62)
    slots1 = (slab_bytes // s0 * s0) // s1
    assert info.slot_count == slots1 and info.slab_bytes == slab_bytes // s0 * s0
    pairs = rt.plan()
    children, _, _, _ = rt.generation(0)
    # slot bookkeeping: repack then assembly
    rep = orc.repack_slots(slots1, slots0, pairs)
    assert rep["status"] == 0
    asm = orc.assemble_slots(slots1, rep["stack"], rep["free_count"], rep["state"], rep["parent_slots"], pairs)
    assert asm["status"] == 0 and children.tolist() == asm["child_slots"].tolist()
    # the generation byte budget was fixed at the old stride, so the population shrinks to whole
    # rows that fit at the new stride (TryPlanNextGenerationShape, slab_runtime.cu:919-939)
    expect_children = ((n * s1 // s0) // side * side * s0 // s1) // side * side
    assert len(children) == len(pairs) == min(n, expect_children)
    n1 = orc.policy_param_count() + 38 * (A0 + inj)
    worst = 0.0
    for c in range(len(children)):
        a, b = pairs[c]
        ref, _ = orc.breed_child(parents[a], parents[b], A0, A0 + inj, 606, 0, c, *temps, 0.02, 0.02, 0.0)
        st, ref_inj = orc.inject(ref, A0, cat, A0, inj, mode=0)
        assert st == 0
        got = rt.download_slot(children[c])
        assert np.array_equal(got[:n0].view(np.uint32), ref[:n0].view(np.uint32)), c
        assert not got[n1:].any()
        target = orc.median_tail_norm(ref[orc.policy_param_count():n0])
        new_got = got[n0:n1].reshape(inj, 38).astype(np.float64)
        new_ref = ref_inj[n0:n1].reshape(inj, 38).astype(np.float64)
        err = np.abs(new_got - new_ref).max() / target
        worst = max(worst, err)
        norms = np.sqrt((new_got ** 2).sum(axis=1))
        assert np.all(np.abs(norms - target) <= 1e-5 * target), c
    assert worst <= 1e-4, worst
    rt.close()


def test_injection_of_a_word_without_hint_grids_fails(pkg, orc):
    rtm = pkg.runtime
    cat = pkg.synth.make_catalog(16, seed=5)
    cat[8] = 0  # AAAAA: no two differing positions to swap
    sec = pkg.synth.sample_secrets(8, 8, seed=3)
    off, ids = pkg.synth.block_csr(4, 4, 4, sec)
    s1 = rtm.slot_stride_bytes(16)
    rt = pkg.SlabRuntime(slab_bytes=3 * 16 * s1, generation_bytes=16 * s1, action_count=8, grid_columns=4)
    rt.upload_catalog(cat)
    rt.upload_secrets_csr(off, ids)
    rt.bootstrap(16, seed=2)
    inj = rtm.PendingInjection(enabled=1, reserved0=0, first_catalog_word_index=8, injection_count=8)
    with pytest.raises(RuntimeError) as err:
        rt.advance(3, rtm.word_counts(8, 8), injection=inj)
    assert err.value.status == 15
    rt.close()


def test_autotuned_shapes_give_identical_results(pkg, grid16):
    """Every launch shape of the persistent evaluation kernel -- waves per workgroup, rounds of tiles per work
    item, workgroups per CU, and the evaluation kernel (accumulators in registers, streamed through scratch,
    two tiles per wave) -- yields the same fitness and counters; the autotuner keeps the fastest and later
    evaluations use it."""
    wl, rt = grid16
    counts = pkg.runtime.word_counts(512, 512)
    ref = None
    for kern in (0, 1, 2):
        for w in (1, 2, 3, 4):
            for r in (1, 2, 4):
                for g in (0, 1, 2):
                    rt.set_eval_shape_ex(w, r, g, kern)
                    rt.evaluate(counts)
                    _, fit, evals, _ = rt.generation(0)
                    st = rt.stage_times()
                    got = (fit.view(np.uint32).copy(), evals.copy(), st.decisions, st.encoded_turns)
                    if ref is None:
                        ref = got
                    assert all(np.array_equal(a, b) for a, b in zip(got, ref)), (kern, w, r, g)
    rt.set_eval_shape_ex(0, 1, 0)
    waves, times, shape = rt.autotune_evaluation(counts)
    assert waves in (1, 2, 3, 4) and all(t > 0 for t in times)
    assert shape.waves == waves and shape.rounds_per_item in (1, 2, 4) and shape.workgroups_per_cu in range(0, 64)
    assert shape.kernel in (0, 1, 2)
    if shape.kernel == 0 and shape.rounds_per_item == 1 and shape.workgroups_per_cu == 0:
        assert times[waves - 1] == min(times)
    _, fit, _, _ = rt.generation(0)
    assert np.array_equal(fit.view(np.uint32), ref[0])
    rt.set_eval_shape_ex(0, 1, 0)


def test_evaluation_kernels_agree_on_ragged_cells(pkg):
    """The streamed and paired evaluation kernels against the register kernel on cells of 4..70 secrets
    (partial tiles, odd tile counts, fresh and prefilled tiles in one wave, waves whose second tile does not
    exist), over several generations so the slab holds bred genomes: identical fitness, evaluation counts,
    decisions and encoded turns for every shape.  The BF16-screened form ignores the knob and gives the same
    fitness as well."""
    rtm, synth = pkg.runtime, pkg.synth
    A = 700
    catalog = synth.make_catalog(A, seed=synth.CATALOG_SEED, distinct_letters=True)
    pool = synth.sample_secrets(A, 200)
    offsets, ids = synth.ragged_csr(256, pool, 4, 70)
    stride = rtm.slot_stride_bytes(A)
    rt = pkg.SlabRuntime(slab_bytes=3 * 256 * stride, generation_bytes=256 * stride, action_count=A, grid_columns=16)
    rt.upload_catalog(catalog)
    rt.upload_secrets_csr(offsets, ids)
    rt.bootstrap(256, seed=5, variance_scale=1.0, tail_std=0.1)
    counts = rtm.word_counts(A, A)
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    for gen in range(3):
        ref = None
        for kern, w, r in ((0, 0, 1), (2, 0, 1), (2, 1, 1), (2, 2, 2), (2, 3, 1), (2, 4, 4), (1, 0, 1), (1, 1, 1),
                           (1, 3, 2), (1, 4, 4)):
            rt.set_eval_precision(rtm.EVAL_FP32)
            rt.set_eval_shape_ex(w, r, 0, kern)
            rt.evaluate(counts)
            _, fit, evals, _ = rt.generation(0)
            st = rt.stage_times()
            got = (fit.view(np.uint32).copy(), evals.copy(), st.decisions, st.encoded_turns)
            if ref is None:
                ref = got
            assert all(np.array_equal(a, b) for a, b in zip(got, ref)), (gen, kern, w, r)
        rt.set_eval_precision(rtm.EVAL_BF16_RESCORED)
        rt.evaluate(counts)
        assert np.array_equal(rt.generation(0)[1].view(np.uint32), ref[0])
        rt.set_eval_precision(rtm.EVAL_FP32)
        rt.set_eval_shape_ex(0, 1, 0, 2)
        rt.advance(11 + gen, counts, cfg)
    rt.close()
