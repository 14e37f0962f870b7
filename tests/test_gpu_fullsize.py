"""Full-size evaluation report: the 64x64 grid (4,096 genomes), 4,739 actions, 2,048 secrets in 64 block
shards of 32 -> 393,216 games, one evaluation of generation 0 in the FP32 form and in the BF16-screened
form.

Bar: over every game of the grid, the BF16 form takes the FP32 form's guesses with chosen scores equal
bit for bit and identical fitness; on 16 genomes spread over the grid (1,536 games), every FP32 policy
vector lies within 1e-5 (normwise) of the host fp64 forward on the state the GPU visited, every decision
whose fp64 top-2 margin exceeds 1e-4 x max(1, |best|) is the fp64 argmax, and the fitness of every genome
without a near-tie decision equals the oracle's FP32 evaluation.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
MARGIN_TOL = 1e-4
SAMPLE = 16


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
    for form, mode, trace in (("fp32", rtm.EVAL_FP32, rtm.TRACE_GAMES | rtm.TRACE_POLICY),
                              ("bf16", rtm.EVAL_BF16_RESCORED, rtm.TRACE_GAMES)):
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
            out["genomes"] = {int(k): rt.download_slot(slots[k]) for k in
                              np.linspace(0, n - 1, SAMPLE).round().astype(int)}
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


def test_fp32_forward_against_fp64_on_a_sample(fullsize, orc):
    wl, games, out = fullsize
    A = wl["actions"]
    catalog, offsets, ids = wl["catalog"], wl["offsets"], wl["ids"]
    gtr, ptr_ = out["fp32"]["games"], out["fp32"]["policy"]
    sample = sorted(out["genomes"])
    max_err = chosen_err = 0.0
    checked = near = 0
    clean = {}
    for n in sample:
        genome = out["genomes"][n]
        local = ids[offsets[n]:offsets[n + 1]]
        base = 3 * int(offsets[n])
        clean[n] = True
        for slot in range(3 * len(local)):
            g = gtr[base + slot]
            d = int(g["decisions"])
            rep = orc.replay_game(genome, A, catalog, local, slot, g["guesses"][:d])
            assert rep["status"] == 0
            for k in range(d):
                p64 = rep["policy"][k]
                scale = max(np.max(np.abs(p64)), 1e-30)
                max_err = max(max_err, float(np.max(np.abs(ptr_[base + slot, k].astype(np.float64) - p64)) / scale))
                chosen_err = max(chosen_err, abs(float(g["best"][k]) - rep["chosen"][k]) / max(1.0, scale))
                checked += 1
                if rep["best"][k] - rep["second"][k] > MARGIN_TOL * max(1.0, abs(rep["best"][k])):
                    assert rep["argmax"][k] == g["guesses"][k], (n, slot, k)
                else:
                    near += 1
                    clean[n] = False
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
