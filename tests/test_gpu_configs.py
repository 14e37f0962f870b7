"""The larger configurations on the HIP path, each checked against the oracle at its real shape.

* Ragged growth (64x64 grid, per-cell shards of U{32..1,024} secrets, actions 512 -> 1,024 -> 1,536 through two
  widening steps with output-embedding injection): population shape as TryPlanNextGenerationShape
  (slab_runtime.cu:919-939), slot bookkeeping as the oracle's repack + assembly, sampled children bit-exact
  with the oracle's breeding of the parents as they were before the repack (so the repack moved them byte
  for byte), injected rows within 1e-4 of the oracle's and at the median norm; ragged-cell fitness equal
  to the oracle's on sampled genomes without near-ties.
* Surrogate-screened evolution at 64x64 / 4,739 actions (brood 4x): the rank-correlation gate at 0.5
  decides the next step as specified; screened steps keep the generation's invariants, and each sampled
  cell's selected child is the oracle's one-pass breeding of one of that cell's four brood members.
* 128x128 / 4,739 actions (16,384 genomes, 8,192 secrets in 256 shards of 32, ~41 GB slab): one evaluation
  and one step; the summary, the plan's locality, the slot bookkeeping, sampled children and sampled
  genomes' fitness against the oracle; the BF16-screened form's fitness equal to the FP32 form's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MARGIN_TOL = 1e-4
THREADS = 16


def _clean_fitness_check(orc, genomes, A, catalog, offsets, ids, cells, fitness):
    """Oracle FP32 evaluation of `cells`; returns (checked, clean) after asserting bit-equal fitness on every
    cell whose oracle decisions all have a top-2 margin above 1e-4 x max(1, |best|)."""
    widths = np.array([offsets[c + 1] - offsets[c] for c in cells])
    sub_off = np.concatenate([[0], np.cumsum(widths)]).astype(np.uint32)
    sub_ids = np.concatenate([ids[offsets[c]:offsets[c + 1]] for c in cells]).astype(np.uint32)
    ref = orc.evaluate(genomes, 0, A, A, catalog, sub_off, sub_ids, mode=0, threads=THREADS, trace=True)
    assert ref["status"] == 0
    tr = ref["traces"]
    played = np.arange(6)[None, :] < tr["decisions"][:, None]
    near = (played & ~(tr["best"] - tr["second"] > MARGIN_TOL * np.maximum(1.0, np.abs(tr["best"])))).any(axis=1)
    clean = 0
    for i, c in enumerate(cells):
        if not near[3 * int(sub_off[i]):3 * int(sub_off[i + 1])].any():
            assert fitness[c] == ref["fitness"][i], (c, fitness[c], ref["fitness"][i])
            clean += 1
    return len(cells), clean


def _moore(rows, cols, focal, r=2):
    fr, fc = divmod(focal, cols)
    return {((fr + dr) % rows) * cols + (fc + dc) % cols for dr in range(-r, r + 1) for dc in range(-r, r + 1)}


# ---- ragged growth ------------------------------------------------------------------------------------
def test_ragged_growth_through_two_widenings(pkg, orc):
    rtm, synth = pkg.runtime, pkg.synth
    side, A_final = 64, 4739
    n = side * side
    catalog = synth.make_catalog(A_final, distinct_letters=True)
    pool = synth.sample_secrets(A_final, 1030)
    offsets, ids = synth.ragged_csr(n, pool, 32, 1024)
    widths = np.diff(offsets.astype(np.int64))
    assert widths.min() >= 32 and widths.max() <= 1024 and widths.max() - widths.min() > 900
    schedule = [512, 512, 1024, 1024, 1536]          # A before step g; widening at steps 1 and 3
    s_final = rtm.slot_stride_bytes(1536)
    slab_bytes = 2 * n * s_final + 4 * s_final
    rt = pkg.SlabRuntime(slab_bytes=slab_bytes, generation_bytes=n * s_final, action_count=512, grid_columns=side)
    rt.upload_catalog(catalog)
    rt.upload_secrets_csr(offsets, ids)
    rt.bootstrap(n, seed=7, variance_scale=1.0, tail_std=0.1)
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    temps = (cfg.crossover_temperature_level1, cfg.crossover_temperature_level2, cfg.crossover_temperature_level3)
    # ragged-cell fitness against the oracle: the 16 narrowest cells and the 2 widest (a cell of ~1,000
    # secrets plays ~3,000 games and almost always holds a near-tie decision, so the bit-exact fitness check
    # mostly lands on the narrow ones; the wide ones still run through the oracle)
    counts0 = rtm.word_counts(512, 512)
    rt.evaluate(counts0)
    slots, fit, _, has = rt.generation(0)
    order = np.argsort(widths, kind="stable")
    cells = sorted(set(order[:16].tolist()) | set(order[-2:].tolist()))
    checked, clean = _clean_fitness_check(orc, [rt.download_slot(slots[c]) for c in cells], 512, catalog, offsets,
                                          ids, cells, fit)
    assert clean >= 2, (checked, clean)
    rng = np.random.default_rng(3)
    widened = 0
    # the generation budget is capacity x initial stride (TryComputeGenerationByteBudgetBytes,
    # slab_runtime.cu:685-696), capacity = whole rows of the configured budget at the initial stride
    s_init = rtm.slot_stride_bytes(512)
    budget = (n * s_final // s_init) // side * side * s_init
    sizes = []
    for g in range(4):
        A0, A1 = schedule[g], schedule[g + 1]
        info0 = rt.info()
        slots0, _, _, _ = rt.generation(0)
        cur = len(slots0)
        inj = None
        if A1 != A0:
            inj = rtm.PendingInjection(enabled=1, reserved0=0, first_catalog_word_index=A0, injection_count=A1 - A0)
            n0 = orc.policy_param_count() + 38 * A0
            parents = {int(k): rt.download_slot(slots0[k])[:n0] for k in range(cur)}
        rt.advance(9 + g, rtm.word_counts(A0, A0), cfg, injection=inj)
        info = rt.info()
        pairs = rt.plan()
        children, _, _, _ = rt.generation(0)
        assert len(set(children.tolist())) == len(children)
        # TryPlanNextGenerationShape (slab_runtime.cu:918-938): whole rows of min(N, budget / new stride)
        expect = min(cur, (budget // rtm.slot_stride_bytes(A1)) // side * side)
        assert len(children) == len(pairs) == info.current_generation_size == expect
        sizes.append(expect)
        if inj is None:
            assert info.action_count == A0
            continue
        widened += 1
        s0, s1 = rtm.slot_stride_bytes(A0), rtm.slot_stride_bytes(A1)
        assert info.action_count == A1 and info.slot_stride_bytes == s1
        # the widened layout keeps slot count x stride bytes (TryCreateExpandedSlabLayout, layout.hpp:55-62)
        assert info.slot_count == (info0.slot_count * s0) // s1
        rep = orc.repack_slots(info.slot_count, slots0, pairs)
        assert rep["status"] == 0
        asm = orc.assemble_slots(info.slot_count, rep["stack"], rep["free_count"], rep["state"], rep["parent_slots"],
                                 pairs)
        assert asm["status"] == 0 and children.tolist() == asm["child_slots"].tolist()
        n1 = orc.policy_param_count() + 38 * A1
        worst = 0.0
        for c in sorted(rng.choice(len(children), 12, replace=False).tolist()):
            a, b = pairs[c]
            ref, _ = orc.breed_child(parents[int(a)], parents[int(b)], A0, A1, 9 + g, g, c, *temps, 0.01, 0.02, 0.0)
            st, ref_inj = orc.inject(ref, A0, catalog, A0, A1 - A0, mode=0)
            assert st == 0
            got = rt.download_slot(children[c])
            assert np.array_equal(got[:n0].view(np.uint32), ref[:n0].view(np.uint32)), (g, c)
            assert not got[n1:].any()
            target = orc.median_tail_norm(ref[orc.policy_param_count():n0])
            new_got = got[n0:n1].reshape(A1 - A0, 38).astype(np.float64)
            worst = max(worst, np.abs(new_got - ref_inj[n0:n1].reshape(A1 - A0, 38)).max() / target)
            norms = np.sqrt((new_got ** 2).sum(axis=1))
            assert np.all(np.abs(norms - target) <= 1e-5 * target), (g, c)
        assert worst <= 1e-4, worst
        del parents
    assert widened == 2
    assert sizes == [4096, 4096, 4096, 4032]  # the last widening no longer fits 64 rows in the budget
    rt.evaluate(rtm.word_counts(1536, 1536))
    _, fit, evals, has = rt.generation(0)
    assert has.all() and (evals == 1).all() and (fit > 0).all() and (fit <= 1).all()
    rt.close()


# ---- surrogate-screened evolution ---------------------------------------------------------------------
@pytest.fixture(scope="module")
def grid64(pkg):
    return pkg.synth.workload("grid64_a4739")


def _screen_runtime(pkg, wl, **kw):
    rtm = pkg.runtime
    A, side = wl["actions"], wl["side"]
    n = side * side
    stride = rtm.slot_stride_bytes(A)
    rt = pkg.SlabRuntime(slab_bytes=6 * n * stride, generation_bytes=4 * n * stride, action_count=A, grid_columns=side)
    rt.upload_catalog(wl["catalog"])
    rt.upload_secrets_csr(wl["offsets"], wl["ids"])
    rt.bootstrap(n, seed=7, variance_scale=1.0, tail_std=0.1)
    rt.set_eval_precision(rtm.EVAL_BF16_RESCORED)
    rt.enable_screening(rtm.default_screen_config(**kw))
    return rt


def test_screened_gate_at_half_follows_the_audit(pkg, grid64):
    """C4's gate: screening runs only while the audit's Spearman rank correlation is at least 0.5 and the
    replay buffer holds min_replay pairs."""
    rtm = pkg.runtime
    rt = _screen_runtime(pkg, grid64, spearman_threshold=0.5)
    counts = rtm.word_counts(4739, 4739)
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    n = 4096
    prev_next = None
    for g in range(4):
        rep = rt.advance_screened(9 + g, counts, cfg)
        if prev_next is not None:
            assert rep.screened == prev_next
        rho = rep.audit_spearman
        if np.isfinite(rho) and abs(rho - 0.5) < 1e-6:
            continue  # the runtime compares in double precision
        assert rep.next_screened == int(rep.replay_count >= 4096 and np.isfinite(rho) and rho >= 0.5), (g, rho)
        assert rep.brood == (4 * n if rep.screened else n)
        slots, _, _, has = rt.generation(0)
        info = rt.info()
        assert has.all() and len(set(slots.tolist())) == n and info.free_slot_count == info.slot_count - n
        prev_next = rep.next_screened
    rt.close()


def test_screened_step_brood_against_the_oracle(pkg, orc, grid64):
    rtm = pkg.runtime
    wl = grid64
    A, side = wl["actions"], wl["side"]
    n = side * side
    rt = _screen_runtime(pkg, wl, spearman_threshold=-1.0, min_replay=1)
    counts = rtm.word_counts(A, A)
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    temps = (cfg.crossover_temperature_level1, cfg.crossover_temperature_level2, cfg.crossover_temperature_level3)
    rep = rt.advance_screened(101, counts, cfg)
    assert rep.screened == 0 and rep.next_screened == 1
    slots, fit, _, has = rt.generation(0)
    seed, gen = 202, rt.info().current_generation_index
    wtable = orc.rank_weights(cfg.rank_exponent, max(n, 25))
    st, pairs = orc.plan_brood(fit, has, n, 4, side, gen, 2, wtable, seed)
    assert st == 0
    cells = sorted(np.random.default_rng(5).choice(n, 6, replace=False).tolist())
    nv = orc.policy_param_count() + 38 * A
    need = {int(p) for c in cells for b in range(4) for p in pairs[b * n + c]}
    parents = {p: rt.download_slot(slots[p])[:nv] for p in need}
    rep = rt.advance_screened(seed, counts, cfg)
    assert rep.screened == 1 and rep.brood == 4 * n and n <= rep.evaluated <= n + rep.audited
    children, cfit, _, chas = rt.generation(0)
    info = rt.info()
    assert chas.all() and len(set(children.tolist())) == n and info.free_slot_count == info.slot_count - n
    for c in cells:
        got = rt.download_slot(children[c])[:nv]
        match = [b for b in range(4) if np.array_equal(got.view(np.uint32), orc.breed_child(
            parents[int(pairs[b * n + c][0])], parents[int(pairs[b * n + c][1])], A, A, seed, gen, b * n + c, *temps,
            cfg.mutation_probability, cfg.mutation_sigma, 0.0)[0].view(np.uint32))]
        assert match, c
        for p in pairs[[b * n + c for b in range(4)]].ravel():
            assert int(p) in _moore(side, side, c)
    # the reported fitness is a full evaluation of the selected children
    rt.evaluate(counts)
    assert np.array_equal(rt.generation(0)[1].view(np.uint32), cfit.view(np.uint32))
    rt.close()


# ---- 128x128 -----------------------------------------------------------------------------------------
def test_grid128_evaluation_and_step(pkg, orc):
    rtm = pkg.runtime
    wl = pkg.synth.workload("grid128_a4739")
    A, side = wl["actions"], wl["side"]
    n = side * side
    assert n == 16384 and int(wl["offsets"][-1]) == 16384 * 32
    stride = rtm.slot_stride_bytes(A)
    rt = pkg.SlabRuntime(slab_bytes=2 * n * stride + stride, generation_bytes=n * stride, action_count=A,
                         grid_columns=side)
    rt.upload_catalog(wl["catalog"])
    rt.upload_secrets_csr(wl["offsets"], wl["ids"])
    rt.bootstrap(n, seed=7, variance_scale=1.0, tail_std=0.1)
    counts = rtm.word_counts(A, A)
    rt.set_eval_precision(rtm.EVAL_BF16_RESCORED)
    rt.evaluate(counts)
    fit16 = rt.generation(0)[1].copy()
    rt.set_eval_precision(rtm.EVAL_FP32)
    rt.evaluate(counts)
    slots, fit, evals, has = rt.generation(0)
    assert np.array_equal(fit16.view(np.uint32), fit.view(np.uint32))
    assert has.all() and (evals == 1).all() and (fit > 0).all() and (fit <= 1).all()
    s = rt.summary()
    best, mean, best_i = orc.summary(fit)
    assert (s.best_fitness, s.average_fitness, s.best_index) == (best, mean, best_i)
    assert s.best_slot_index == slots[best_i]
    cells = np.linspace(0, n - 1, 16).round().astype(int).tolist()
    _, clean = _clean_fitness_check(orc, [rt.download_slot(slots[c]) for c in cells], A, wl["catalog"], wl["offsets"],
                                    wl["ids"], cells, fit)
    assert clean >= 2
    nv = orc.policy_param_count() + 38 * A
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    temps = (cfg.crossover_temperature_level1, cfg.crossover_temperature_level2, cfg.crossover_temperature_level3)
    st, pairs_ref = orc.plan(fit, has, n, n, side, 0, 2, orc.rank_weights(cfg.rank_exponent, n), 31)
    assert st == 0
    sample = sorted(np.random.default_rng(9).choice(n, 6, replace=False).tolist())
    parents = {int(p): rt.download_slot(slots[p])[:nv] for c in sample for p in pairs_ref[c]}
    info0 = rt.info()
    rt.advance(31, counts, cfg)
    pairs = rt.plan()
    assert np.array_equal(pairs, pairs_ref)
    for c in range(n):
        hood = _moore(side, side, c)
        assert pairs[c, 0] != pairs[c, 1] and int(pairs[c, 0]) in hood and int(pairs[c, 1]) in hood
    children, _, _, chas = rt.generation(0)
    stack, fc, state, cur = orc.slab_after_bootstrap(info0.slot_count, n)
    asm = orc.assemble_slots(info0.slot_count, stack, fc, state, cur, pairs)
    assert asm["status"] == 0 and children.tolist() == asm["child_slots"].tolist()
    assert not chas.any()
    for c in sample:
        a, b = pairs[c]
        ref, _ = orc.breed_child(parents[int(a)], parents[int(b)], A, A, 31, 0, c, *temps, 0.01, 0.02, 0.0)
        assert np.array_equal(rt.download_slot(children[c])[:nv].view(np.uint32), ref[:nv].view(np.uint32)), c
    rt.close()
