"""The BF16-trunk form (ne_set_eval_precision(NE_EVAL_BF16_TRUNK)): the BF16 screening form with the trunk's
layers 1 and 2 on BF16 MFMA (split operands, hi.hi + hi.lo + lo.hi, FP32 accumulate).

Bar: against the oracle's BF16-trunk evaluation (mode 2, pinned on the CPU in test_oracle_bf16_trunk.py) the
GPU takes >= 99.9% of the decisions on common trajectories, and every genome whose oracle decisions all have
a top-2 margin above 1e-4 x max(1, |best|) has the oracle's fitness bit for bit; every kernel of the form
(accumulators in registers, the streamed ring kernel, two tiles per wave) and every launch shape gives the
same fitness and decision counts as the traced register kernel.  Agreement with the FP32 form's decisions
at full size is checked in test_gpu_fullsize.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MARGIN_TOL = 1e-4


def _runtime(pkg, side, A, catalog, offsets, ids, seed=9, var=1.0, tail=0.1):
    stride = pkg.runtime.slot_stride_bytes(A)
    n = side * side
    rt = pkg.SlabRuntime(slab_bytes=2 * n * stride, generation_bytes=n * stride, action_count=A, grid_columns=side)
    rt.upload_catalog(catalog)
    rt.upload_secrets_csr(offsets, ids)
    rt.bootstrap(n, seed=seed, variance_scale=var, tail_std=tail)
    return rt


def _path_agreement(a, b):
    """Decisions compared while two games' trajectories coincide (after a differing guess the boards differ)."""
    compared = agreed = 0
    on_path = None
    for k in range(6):
        live = (a["decisions"] > k) & (b["decisions"] > k)
        if k > 0:
            live &= on_path
        eq = a["guesses"][:, k] == b["guesses"][:, k]
        compared += int(live.sum())
        agreed += int((live & eq).sum())
        on_path = live & eq if k == 0 else on_path & eq
    return agreed, compared


@pytest.mark.parametrize("workload", ["grid16_a512", "grid4_a4739"])
def test_bf16_trunk_against_the_oracle(pkg, orc, workload):
    rtm = pkg.runtime
    if workload == "grid16_a512":
        wl = pkg.synth.workload("grid16_a512")
        A, side, catalog, offsets, ids = 512, 16, wl["catalog"], wl["offsets"], wl["ids"]
    else:
        A, side = 4739, 4
        catalog = pkg.synth.make_catalog(A)
        offsets, ids = pkg.synth.block_csr(side, side, 2, pkg.synth.sample_secrets(A, 64))
    n = side * side
    games = 3 * int(offsets[-1])
    rt = _runtime(pkg, side, A, catalog, offsets, ids, seed=17)
    try:
        rt.set_eval_precision(rtm.EVAL_BF16_TRUNK)
        rt.set_trace(rtm.TRACE_GAMES)
        rt.evaluate(rtm.word_counts(A, A))
        slots, fit, _, _ = rt.generation(0)
        fit = fit.copy()
        tr = rt.download_trace(rtm.TRACE_GAMES, games)
        genomes = [rt.download_slot(s) for s in slots]
    finally:
        rt.close()
    ref = orc.evaluate(genomes, 0, A, A, catalog, offsets, ids, mode=2, threads=16, trace=True)
    assert ref["status"] == 0
    g_orc = ref["traces"]
    agreed, compared = _path_agreement(tr, g_orc)
    assert agreed / compared >= 0.999, (agreed, compared)
    margin = g_orc["best"] - g_orc["second"]
    played = np.arange(6)[None, :] < g_orc["decisions"][:, None]
    near = (played & ~(margin > MARGIN_TOL * np.maximum(1.0, np.abs(g_orc["best"])))).any(axis=1)
    clean = np.array([not near[3 * int(offsets[k]):3 * int(offsets[k + 1])].any() for k in range(n)])
    assert clean.sum() >= n // 8, int(clean.sum())
    assert np.array_equal(fit[clean].view(np.uint32), ref["fitness"][clean].view(np.uint32))
    print(f"{workload}: {agreed}/{compared} decisions agree with the oracle's BF16 trunk, "
          f"{int(clean.sum())}/{n} genomes without near-ties all with the oracle's fitness")


@pytest.mark.parametrize("case", ["grid16", "a4739", "zero"])
def test_bf16_trunk_kernels_and_shapes_agree(pkg, case):
    """Every kernel and launch shape of the form plays the traced register kernel's games: same fitness bit for
    bit, same decision count; "zero": all-zero genomes (every score ties, the lowest unmasked index wins)."""
    rtm = pkg.runtime
    if case == "a4739":
        A, side = 4739, 4
        catalog = pkg.synth.make_catalog(A)
        offsets, ids = pkg.synth.block_csr(side, side, 2, pkg.synth.sample_secrets(A, 64))
    else:
        wl = pkg.synth.workload("grid16_a512")
        A, side, catalog, offsets, ids = 512, 16, wl["catalog"], wl["offsets"], wl["ids"]
    zero = case == "zero"
    rt = _runtime(pkg, side, A, catalog, offsets, ids, seed=29, var=0.0 if zero else 1.0, tail=0.0 if zero else 0.1)
    counts = rtm.word_counts(A, A)
    try:
        rt.set_eval_precision(rtm.EVAL_BF16_TRUNK)
        rt.set_trace(rtm.TRACE_GAMES)
        rt.evaluate(counts)
        f0 = rt.generation(0)[1].copy()
        d0 = int(rt.stage_times().decisions)
        rt.set_trace(0)
        shapes = [(w, 1, rtm.EVAL_KERNEL_REGISTERS) for w in (1, 3, 4)]
        shapes += [(3, 1, rtm.EVAL_KERNEL_STREAMED), (3, 4, rtm.EVAL_KERNEL_STREAMED)]
        shapes += [(w, 1, rtm.EVAL_KERNEL_PAIRED) for w in (1, 3)]
        for shape in shapes:
            for _ in range(2):  # and the same again: no run-to-run differences
                rt.set_eval_shape_ex(*shape[:2], 0, shape[2])
                rt.evaluate(counts)
                f = rt.generation(0)[1]
                assert np.array_equal(f.view(np.uint32), f0.view(np.uint32)), shape
                assert int(rt.stage_times().decisions) == d0, shape
        if zero:  # every genome plays the FP32 form's games
            rt.set_eval_precision(rtm.EVAL_FP32)
            rt.evaluate(counts)
            assert np.array_equal(rt.generation(0)[1].view(np.uint32), f0.view(np.uint32))
        print(f"{case}: {d0} decisions, {len(shapes)} shapes identical")
    finally:
        rt.set_eval_shape_ex(0, 1, 0, 0)
        rt.close()
