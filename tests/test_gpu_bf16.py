"""The BF16 screening form against the FP32 form on the same genomes.

Bar: BF16 screening + FP32 rescoring (ne_set_eval_precision(NE_EVAL_BF16_RESCORED)) takes exactly the
FP32 form's decisions: identical guesses, chosen scores equal bit for bit (the rescoring runs the FP32
contraction itself), identical fitness.  The window bound that guarantees this is proven in the kernel
(|S_bf16 - S_fp32| <= 1.03 * 2^-8 * sum_d max_a|E_ad| |p_d|); these tests check the outcome.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _runtime(pkg, side, A, catalog, offsets, ids, seed=9, var=1.0, tail=0.1):
    rtm = pkg.runtime
    stride = rtm.slot_stride_bytes(A)
    n = side * side
    rt = pkg.SlabRuntime(slab_bytes=2 * n * stride, generation_bytes=n * stride, action_count=A, grid_columns=side)
    rt.upload_catalog(catalog)
    rt.upload_secrets_csr(offsets, ids)
    rt.bootstrap(n, seed=seed, variance_scale=var, tail_std=tail)
    return rt


def _evaluate(pkg, rt, A, games, mode):
    rtm = pkg.runtime
    rt.set_eval_precision(mode)
    rt.set_trace(rtm.TRACE_GAMES)
    rt.evaluate(rtm.word_counts(A, A))
    _, fit, _, _ = rt.generation(0)
    tr = rt.download_trace(rtm.TRACE_GAMES, games)
    st = rt.stage_times()
    return fit.copy(), tr, st


def _same_decisions(pkg, rt, A, games, mode=None):
    rtm = pkg.runtime
    f32, t32, s32 = _evaluate(pkg, rt, A, games, rtm.EVAL_FP32)
    f16, t16, s16 = _evaluate(pkg, rt, A, games, rtm.EVAL_BF16_RESCORED if mode is None else mode)
    assert np.array_equal(t16["decisions"], t32["decisions"])
    assert np.array_equal(t16["guesses"], t32["guesses"])
    played = t32["guesses"] != 0xFFFFFFFF
    assert np.array_equal(t16["best"][played].view(np.uint32), t32["best"][played].view(np.uint32))
    assert np.array_equal(f16.view(np.uint32), f32.view(np.uint32))
    assert s16.decisions == s32.decisions
    return s16


def test_bf16_form_takes_the_fp32_decisions_grid16(pkg):
    wl = pkg.synth.workload("grid16_a512")
    rt = _runtime(pkg, 16, 512, wl["catalog"], wl["offsets"], wl["ids"])
    try:
        st = _same_decisions(pkg, rt, 512, 3 * int(wl["offsets"][-1]))
        assert st.rescored >= st.decisions - st.fallbacks
        assert st.fallbacks <= st.decisions // 100, (st.fallbacks, st.decisions)
        print(f"decisions {st.decisions}, candidates rescored {st.rescored} "
              f"({st.rescored / max(1, st.decisions):.2f}/decision), fallbacks {st.fallbacks}")
    finally:
        rt.close()


def test_bf16_form_takes_the_fp32_decisions_4739_actions(pkg):
    A, side = 4739, 4
    catalog = pkg.synth.make_catalog(A)
    secrets = pkg.synth.sample_secrets(A, 64)
    offsets, ids = pkg.synth.block_csr(side, side, 2, secrets)
    rt = _runtime(pkg, side, A, catalog, offsets, ids, seed=33)
    try:
        st = _same_decisions(pkg, rt, A, 3 * int(offsets[-1]))
        assert st.fallbacks <= st.decisions // 50, (st.fallbacks, st.decisions)
        print(f"decisions {st.decisions}, rescored {st.rescored}, fallbacks {st.fallbacks}")
    finally:
        rt.close()


def test_bf16_form_duplicate_words_and_ragged_cells(pkg):
    A, side = 365, 4
    rng = np.random.default_rng(0xD0B1E)
    catalog = pkg.synth.make_catalog(A, seed=0xD0B1E).copy()
    for k in range(40, A, 9):
        catalog[k] = catalog[int(rng.integers(0, k))]
    widths = np.array([2, 4, 5, 16, 17, 33, 70, 6, 9, 48, 2, 64, 7, 31, 8, 10])
    offsets = np.concatenate([[0], np.cumsum(widths)]).astype(np.uint32)
    firsts = np.array(sorted({bytes(w): i for i, w in reversed(list(enumerate(catalog)))}.values()))
    ids = np.concatenate([rng.choice(firsts, size=w, replace=False) for w in widths]).astype(np.uint32)
    rt = _runtime(pkg, side, A, catalog, offsets, ids, seed=21)
    try:
        _same_decisions(pkg, rt, A, 3 * int(offsets[-1]))
    finally:
        rt.close()


def test_bf16_forced_fallback_and_all_tie_genomes(pkg):
    """Every decision through the FP32 fallback scan; then zero genomes (every score 0, all actions
    tie: the window holds them all and the decision must fall back) in the BF16 form."""
    wl = pkg.synth.workload("grid16_a512")
    rtm = pkg.runtime
    rt = _runtime(pkg, 16, 512, wl["catalog"], wl["offsets"], wl["ids"])
    try:
        st = _same_decisions(pkg, rt, 512, 3 * int(wl["offsets"][-1]),
                             mode=rtm.EVAL_BF16_RESCORED | rtm.EVAL_FORCE_FALLBACK)
        assert st.fallbacks == st.decisions and st.rescored == 0
        rt.bootstrap(256, seed=1, variance_scale=0.0, tail_std=0.0)
        st = _same_decisions(pkg, rt, 512, 3 * int(wl["offsets"][-1]))
        assert st.fallbacks == st.decisions
    finally:
        rt.close()


def test_eval_precision_rejects_unknown_modes(pkg):
    wl = pkg.synth.workload("grid16_a512")
    rt = _runtime(pkg, 16, 512, wl["catalog"], wl["offsets"], wl["ids"])
    try:
        for bad in (3, 0x200, 0x101 | 0x1000):
            with pytest.raises(pkg.runtime.RuntimeError_):
                rt.set_eval_precision(bad)
    finally:
        rt.close()


@pytest.mark.parametrize("case", ["plain", "dups", "zero"])
def test_untraced_evaluations_equal_traced_ones(pkg, case):
    """The untraced kernels skip the per-row argmax work when no lane of a wave can change, gated by the
    best score another lane of the game already holds: their fitness and decision counts must equal the
    traced kernels' (which do every row), in both forms.  "zero": all-zero genomes, so every score is
    exactly 0 and every decision is a tie that the lowest unmasked index must win."""
    rtm = pkg.runtime
    dups = case == "dups"
    if dups:
        A, side = 365, 4
        rng = np.random.default_rng(5)
        catalog = pkg.synth.make_catalog(A, seed=5).copy()
        for k in range(30, A, 7):
            catalog[k] = catalog[int(rng.integers(0, k))]
        widths = np.array([2, 4, 5, 16, 17, 33, 70, 6, 9, 48, 2, 64, 7, 31, 8, 10])
        offsets = np.concatenate([[0], np.cumsum(widths)]).astype(np.uint32)
        firsts = np.array(sorted({bytes(w): i for i, w in reversed(list(enumerate(catalog)))}.values()))
        ids = np.concatenate([rng.choice(firsts, size=w, replace=False) for w in widths]).astype(np.uint32)
    else:
        wl = pkg.synth.workload("grid16_a512")
        A, side, catalog, offsets, ids = 512, 16, wl["catalog"], wl["offsets"], wl["ids"]
    rt = _runtime(pkg, side, A, catalog, offsets, ids, seed=41, var=0.0 if case == "zero" else 1.0,
                  tail=0.0 if case == "zero" else 0.1)
    try:
        results = []
        for mode in (rtm.EVAL_FP32, rtm.EVAL_BF16_RESCORED):
            for traced in (True, False):
                rt.set_eval_precision(mode)
                rt.set_trace(rtm.TRACE_GAMES if traced else 0)
                rt.evaluate(rtm.word_counts(A, A))
                _, fit, _, _ = rt.generation(0)
                results.append((fit.copy(), rt.stage_times().decisions))
        f0, d0 = results[0]
        for f, d in results[1:]:
            assert np.array_equal(f.view(np.uint32), f0.view(np.uint32))
            assert d == d0
    finally:
        rt.close()


@pytest.mark.parametrize("case", ["grid16", "a4739", "ragged", "fallback", "zero"])
def test_bf16_kernels_screen_and_decide_alike(pkg, case):
    """The BF16 form's untraced kernels -- one tile per wave with the accumulators in registers or streamed
    through scratch, two tiles per wave -- at several launch shapes: the screening, windows and rescoring
    are per game, so every shape gives the register kernel's fitness bit for bit, its decision count, the
    same number of candidates rescored and the same fallback decisions; and the FP32 form's fitness.
    "ragged": cells of 2 to 70 secrets with duplicate catalog words (odd tile counts, half-empty pairs);
    "fallback": every decision forced through the FP32 scan; "zero": all-zero genomes (every score ties)."""
    rtm = pkg.runtime
    mode = rtm.EVAL_BF16_RESCORED | (rtm.EVAL_FORCE_FALLBACK if case == "fallback" else 0)
    if case == "a4739":
        A, side = 4739, 4
        catalog = pkg.synth.make_catalog(A)
        offsets, ids = pkg.synth.block_csr(side, side, 2, pkg.synth.sample_secrets(A, 64))
    elif case == "ragged":
        A, side = 365, 4
        rng = np.random.default_rng(0xD0B1E)
        catalog = pkg.synth.make_catalog(A, seed=0xD0B1E).copy()
        for k in range(40, A, 9):
            catalog[k] = catalog[int(rng.integers(0, k))]
        widths = np.array([2, 4, 5, 16, 17, 33, 70, 6, 9, 48, 2, 64, 7, 31, 8, 10])
        offsets = np.concatenate([[0], np.cumsum(widths)]).astype(np.uint32)
        firsts = np.array(sorted({bytes(w): i for i, w in reversed(list(enumerate(catalog)))}.values()))
        ids = np.concatenate([rng.choice(firsts, size=w, replace=False) for w in widths]).astype(np.uint32)
    else:
        wl = pkg.synth.workload("grid16_a512")
        A, side, catalog, offsets, ids = 512, 16, wl["catalog"], wl["offsets"], wl["ids"]
    zero = case == "zero"
    rt = _runtime(pkg, side, A, catalog, offsets, ids, seed=23, var=0.0 if zero else 1.0, tail=0.0 if zero else 0.1)
    counts = rtm.word_counts(A, A)
    try:
        rt.set_trace(0)

        def run(m, waves, rounds, kernel):
            rt.set_eval_precision(m)
            rt.set_eval_shape_ex(waves, rounds, 0, kernel)
            rt.evaluate(counts)
            st = rt.stage_times()
            return rt.generation(0)[1].copy(), (int(st.decisions), int(st.rescored), int(st.fallbacks))

        f32, (d32, _, _) = run(rtm.EVAL_FP32, 3, 1, rtm.EVAL_KERNEL_REGISTERS)
        f0, c0 = run(mode, 3, 1, rtm.EVAL_KERNEL_REGISTERS)
        assert np.array_equal(f0.view(np.uint32), f32.view(np.uint32)) and c0[0] == d32
        if case == "fallback":
            assert c0[2] == c0[0] and c0[1] == 0
        # streamed: the 3-wave ring kernel (4 runs at 3; 1 and 2 run the register kernel); paired: 1-3 waves (4
        # runs at 3, its LDS lists' limit), 6-entry lists, so it may send a decision to the FP32 scan where the
        # 8-entry lists of the others keep every window member (never the other way round)
        shapes = [(w, 1, rtm.EVAL_KERNEL_STREAMED) for w in (1, 2, 3, 4)] + [(3, 4, rtm.EVAL_KERNEL_STREAMED)]
        shapes += [(w, 1, rtm.EVAL_KERNEL_PAIRED) for w in (1, 2, 3, 4)] + [(3, 2, rtm.EVAL_KERNEL_PAIRED)]
        for shape in shapes:
            f, c = run(mode, *shape)
            f2, c2 = run(mode, *shape)  # and the same again: no run-to-run differences
            assert c2 == c and np.array_equal(f2.view(np.uint32), f.view(np.uint32)), (shape, c, c2)
            assert np.array_equal(f.view(np.uint32), f0.view(np.uint32)), shape
            assert c[0] == c0[0], (shape, c, c0)
            if shape[2] == rtm.EVAL_KERNEL_PAIRED:
                assert c[2] >= c0[2] and c[1] <= c0[1], (shape, c, c0)
            else:
                assert c == c0, (shape, c, c0)
        print(f"{case}: {c0[0]} decisions, {c0[1]} rescored, {c0[2]} fallbacks, {len(shapes)} shapes identical")
    finally:
        rt.set_eval_shape_ex(0, 1, 0, 0)
        rt.close()
