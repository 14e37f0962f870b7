"""Full-size evaluation report: the 64x64 grid (4,096 genomes), 4,739 actions, 2,048 secrets in 64 block
shards of 32 -> 393,216 games, one evaluation of generation 0 in the FP32 form and in the BF16-screened
form.

Bar: over every game of the grid, the BF16 form takes the FP32 form's guesses with chosen scores equal
bit for bit and identical fitness, and the untraced kernels (the ones timed: three kernels in each form)
give the traced kernels' fitness bit for bit; on the first 1,024 genomes (98,304 games) the FP32 form agrees with the oracle's
FP32 evaluation on >= 99.9% of the decisions taken on a common trajectory, and every genome without a
near-tie decision has the oracle's fitness bit for bit; on 128 genomes spread over the grid (12,288
games), every FP32 policy vector lies within 1e-5 (normwise) of the host fp64 forward on the state the GPU
visited and every decision whose fp64 top-2 margin exceeds 1e-4 x max(1, |best|) is the fp64 argmax.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
MARGIN_TOL = 1e-4
SAMPLE = 128
ORACLE_GENOMES = 1024
THREADS = 16  # the host's CPU share for one GPU


@pytest.fixture(scope="module")
def fullsize(pkg):
    rtm = pkg.runtime
    wl = pkg.synth.workload("grid64_a4739")
    A, side = wl["actions"], wl["side"]
    n = side * side
    stride = rtm.slot_stride_bytes(A)
    rt = pkg.SlabRuntime(slab_bytes=2 * n * stride + stride, generation_bytes=n * stride, action_count=A,
                         grid_columns=side)
    rt.upload_catalog(wl["catalog"])
    rt.upload_secrets_csr(wl["offsets"], wl["ids"])
    rt.bootstrap(n, seed=7, variance_scale=1.0, tail_std=0.1)
    games = 3 * int(wl["offsets"][-1])
    counts = rtm.word_counts(A, A)
    out = {}
    # untraced: the BF16 form and the three FP32 kernels (the streamed one is what the autotuner picks and
    # the bench times here)
    for form, mode, kernel in (("fp32_untraced", rtm.EVAL_FP32, rtm.EVAL_KERNEL_REGISTERS),
                               ("fp32_streamed_untraced", rtm.EVAL_FP32, rtm.EVAL_KERNEL_STREAMED),
                               ("fp32_paired_untraced", rtm.EVAL_FP32, rtm.EVAL_KERNEL_PAIRED),
                               ("bf16_untraced", rtm.EVAL_BF16_RESCORED, rtm.EVAL_KERNEL_REGISTERS),
                               ("bf16_streamed_untraced", rtm.EVAL_BF16_RESCORED, rtm.EVAL_KERNEL_STREAMED),
                               ("bf16_paired_untraced", rtm.EVAL_BF16_RESCORED, rtm.EVAL_KERNEL_PAIRED),
                               ("t16_streamed_untraced", rtm.EVAL_BF16_TRUNK, rtm.EVAL_KERNEL_STREAMED)):
        rt.set_eval_precision(mode)
        rt.set_trace(0)
        rt.set_eval_shape_ex(3, 1, 0, kernel)
        rt.evaluate(counts)
        st = rt.stage_times()
        out[form] = dict(fitness=rt.generation(0)[1].copy(), decisions=int(st.decisions), rescored=int(st.rescored),
                         fallbacks=int(st.fallbacks))
    rt.set_eval_shape_ex(0, 1, 0, rtm.EVAL_KERNEL_REGISTERS)
    for form, mode, trace in (("fp32", rtm.EVAL_FP32, rtm.TRACE_GAMES | rtm.TRACE_POLICY),
                              ("bf16", rtm.EVAL_BF16_RESCORED, rtm.TRACE_GAMES),
                              ("t16", rtm.EVAL_BF16_TRUNK, rtm.TRACE_GAMES)):
        rt.set_eval_precision(mode)
        rt.set_trace(trace)
        rt.evaluate(counts)
        _, fit, _, _ = rt.generation(0)
        st = rt.stage_times()
        out[form] = dict(fitness=fit.copy(), games=rt.download_trace(rtm.TRACE_GAMES, games),
                         decisions=int(st.decisions), rescored=int(st.rescored), fallbacks=int(st.fallbacks))
        if form == "fp32":
            out[form]["policy"] = rt.download_trace(rtm.TRACE_POLICY, games)
            slots, _, _, _ = rt.generation(0)
            keep = set(range(ORACLE_GENOMES)) | set(np.linspace(0, n - 1, SAMPLE).round().astype(int).tolist())
            out["genomes"] = {int(k): rt.download_slot(slots[k]) for k in sorted(keep)}
    rt.set_trace(0)
    rt.close()
    yield wl, games, out


def test_bf16_form_takes_every_fp32_decision_at_full_size(fullsize):
    wl, games, out = fullsize
    g32, g16 = out["fp32"]["games"], out["bf16"]["games"]
    assert games == 393216 and g32.shape[0] == games
    assert np.array_equal(g16["decisions"], g32["decisions"])
    assert np.array_equal(g16["guesses"], g32["guesses"])
    played = g32["guesses"] != 0xFFFFFFFF
    assert np.array_equal(g16["best"][played].view(np.uint32), g32["best"][played].view(np.uint32))
    assert np.array_equal(out["bf16"]["fitness"].view(np.uint32), out["fp32"]["fitness"].view(np.uint32))
    assert out["bf16"]["decisions"] == out["fp32"]["decisions"] == int(g32["decisions"].sum())
    d = out["bf16"]
    print(f"[full size] {games} games, {d['decisions']} decisions: BF16 guesses identical, chosen scores "
          f"bit-identical, fitness identical; {d['rescored'] / d['decisions']:.3f} candidates rescored per "
          f"decision, {d['fallbacks']} fallback decisions")


def test_untraced_kernels_equal_traced_ones_at_full_size(fullsize):
    """Every untraced kernel (the timed ones: the register, streamed and paired kernels of both forms) gives
    the traced kernel's fitness bit for bit and its decision count; the BF16 form's kernels screen alike
    (the same candidates rescored and the same fallback decisions as the traced kernel)."""
    _, _, out = fullsize
    for form, traced in (("fp32", "fp32"), ("fp32_streamed", "fp32"), ("fp32_paired", "fp32"), ("bf16", "bf16"),
                         ("bf16_streamed", "bf16"), ("bf16_paired", "bf16")):
        assert np.array_equal(out[form + "_untraced"]["fitness"].view(np.uint32),
                              out[traced]["fitness"].view(np.uint32)), form
        assert out[form + "_untraced"]["decisions"] == out[traced]["decisions"], form
        if form == "bf16_paired":  # 6-entry lists: maybe more decisions through the FP32 scan
            assert out[form + "_untraced"]["fallbacks"] >= out["bf16"]["fallbacks"], form
            assert out[form + "_untraced"]["rescored"] <= out["bf16"]["rescored"], form
        elif traced == "bf16":
            assert out[form + "_untraced"]["rescored"] == out["bf16"]["rescored"], form
            assert out[form + "_untraced"]["fallbacks"] == out["bf16"]["fallbacks"], form


def test_bf16_trunk_form_agrees_with_the_fp32_form_at_full_size(fullsize):
    """The BF16-trunk form (trunk layers 1 and 2 on split-BF16 MFMA) against the FP32 form over all 393,216
    games: >= 99.9% of the decisions on common trajectories are the FP32 form's (the bar for the BF16 trunk);
    its timed kernel (the streamed ring) has the traced register kernel's fitness and decision count."""
    wl, games, out = fullsize
    a, b = out["fp32"]["games"], out["t16"]["games"]
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
    same_games = int(np.all(a["guesses"] == b["guesses"], axis=1).sum())
    same_fit = float(np.mean(out["t16"]["fitness"] == out["fp32"]["fitness"]))
    assert compared > 1_500_000 and agreed / compared >= 0.999, (agreed, compared)
    tu, tt = out["t16_streamed_untraced"], out["t16"]
    assert np.array_equal(tu["fitness"].view(np.uint32), tt["fitness"].view(np.uint32))
    assert tu["decisions"] == tt["decisions"]
    print(f"[full size] BF16 trunk vs FP32: {agreed}/{compared} = {agreed / compared:.6f} decisions agree on common "
          f"trajectories, {same_games}/{len(a)} games identical, fitness equal on {same_fit:.4f} of the genomes")


def test_fp32_form_agrees_with_the_oracle_on_1024_genomes(fullsize, orc):
    """The oracle plays the first 1,024 genomes' 98,304 games in the reference's sequential FP32 order.  Along
    each game, decisions are compared while the two trajectories coincide (after a differing guess the
    boards differ); a genome whose oracle decisions all have a top-2 margin above 1e-4 x max(1, |best|)
    must have the oracle's fitness bit for bit."""
    wl, games, out = fullsize
    A = wl["actions"]
    offsets, ids = wl["offsets"], wl["ids"]
    K = ORACLE_GENOMES
    ref = orc.evaluate([out["genomes"][n] for n in range(K)], 0, A, A, wl["catalog"], offsets[:K + 1],
                       ids[:int(offsets[K])], mode=0, threads=THREADS, trace=True)
    assert ref["status"] == 0
    g_gpu = out["fp32"]["games"][:3 * int(offsets[K])]
    g_orc = ref["traces"]
    assert g_orc.shape == g_gpu.shape
    compared = agreed = same_games = 0
    for k in range(6):
        live = (g_gpu["decisions"] > k) & (g_orc["decisions"] > k)
        if k > 0:
            live &= on_path
        eq = g_gpu["guesses"][:, k] == g_orc["guesses"][:, k]
        compared += int(live.sum())
        agreed += int((live & eq).sum())
        on_path = live & eq if k == 0 else on_path & eq
    same_games = int(np.all(g_gpu["guesses"] == g_orc["guesses"], axis=1).sum())
    assert compared > 400000 and agreed / compared >= 0.999, (agreed, compared)
    fit = out["fp32"]["fitness"][:K]
    margin = g_orc["best"] - g_orc["second"]
    played = np.arange(6)[None, :] < g_orc["decisions"][:, None]
    near = played & ~(margin > MARGIN_TOL * np.maximum(1.0, np.abs(g_orc["best"])))
    game_near = near.any(axis=1)
    clean = np.array([not game_near[3 * int(offsets[n]):3 * int(offsets[n + 1])].any() for n in range(K)])
    assert clean.sum() >= K // 4
    assert np.array_equal(fit[clean].view(np.uint32), ref["fitness"][clean].view(np.uint32))
    print(f"[full size] oracle on {K} genomes: {agreed}/{compared} decisions agree on common trajectories, "
          f"{same_games}/{len(g_gpu)} games identical, {int(clean.sum())} genomes without near-ties all with the "
          f"oracle's fitness")


def test_fp32_forward_against_fp64_on_a_sample(fullsize, orc):
    wl, games, out = fullsize
    A = wl["actions"]
    catalog, offsets, ids = wl["catalog"], wl["offsets"], wl["ids"]
    gtr, ptr_ = out["fp32"]["games"], out["fp32"]["policy"]
    sample = np.linspace(0, len(offsets) - 2, SAMPLE).round().astype(int).tolist()

    def check_genome(n):
        """(max normwise policy error, max chosen-score error, decisions, near-ties, wrong argmaxes)."""
        genome = out["genomes"][n]
        local = ids[offsets[n]:offsets[n + 1]]
        base = 3 * int(offsets[n])
        m_err = c_err = 0.0
        dec = nr = wrong = 0
        for slot in range(3 * len(local)):
            g = gtr[base + slot]
            d = int(g["decisions"])
            rep = orc.replay_game(genome, A, catalog, local, slot, g["guesses"][:d])
            assert rep["status"] == 0
            for k in range(d):
                p64 = rep["policy"][k]
                scale = max(np.max(np.abs(p64)), 1e-30)
                m_err = max(m_err, float(np.max(np.abs(ptr_[base + slot, k].astype(np.float64) - p64)) / scale))
                c_err = max(c_err, abs(float(g["best"][k]) - rep["chosen"][k]) / max(1.0, scale))
                dec += 1
                if rep["best"][k] - rep["second"][k] > MARGIN_TOL * max(1.0, abs(rep["best"][k])):
                    wrong += int(rep["argmax"][k] != g["guesses"][k])
                else:
                    nr += 1
        return m_err, c_err, dec, nr, wrong

    with ThreadPoolExecutor(THREADS) as pool:  # the oracle's ctypes calls release the GIL
        res = list(pool.map(check_genome, sample))
    max_err = max(r[0] for r in res)
    chosen_err = max(r[1] for r in res)
    checked = sum(r[2] for r in res)
    near = sum(r[3] for r in res)
    assert sum(r[4] for r in res) == 0
    clean = {n: r[3] == 0 for n, r in zip(sample, res)}
    assert max_err <= FWD_TOL, max_err
    # the oracle's FP32 evaluation of the sampled genomes (their own cells, in sample order)
    widths = np.array([offsets[n + 1] - offsets[n] for n in sample])
    sub_off = np.concatenate([[0], np.cumsum(widths)]).astype(np.uint32)
    sub_ids = np.concatenate([ids[offsets[n]:offsets[n + 1]] for n in sample]).astype(np.uint32)
    ref = orc.evaluate([out["genomes"][n] for n in sample], 0, A, A, catalog, sub_off, sub_ids, mode=0, threads=16)
    assert ref["status"] == 0
    fit = out["fp32"]["fitness"]
    same = [fit[n] == ref["fitness"][i] for i, n in enumerate(sample) if clean[n]]
    assert same and all(same)
    print(f"[full size] fp64 check on {len(sample)} genomes: {checked} decisions, {near} near-ties, max normwise "
          f"policy error {max_err:.2e}, max chosen-score error {chosen_err:.2e}, {len(same)} clean genomes with "
          f"the oracle's fitness")
