"""Single-model inference and the winner artifact on the GPU.

Bar: the next guess, its score and the runner-up score equal the oracle's fp32 restatement of the
reference's policy and masked argmax bit for bit (src/inference/dynamic_policy.hpp:62-150), one board per
call and many per launch alike; a closed-form genome picks the closed-form word; invalid, finished and
exhausted boards get the reference's statuses; a run's winner, written as an artifact and read back, plays
its fresh games as the evaluation kernel played them.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TILE = {"G": 0, "Y": 1, "-": 2}


def tiles(s):
    return np.array([TILE[c] for c in s], np.uint8)


def _trunk_golden_genome(pkg, A):
    """A sparse genome in the 145->64 layout whose policy on one turn ABCDE/GY-GY follows by hand: hidden0 =
    (1.5, 2.5, 0), hidden1 = (2, 0), policy = (-1 + 3*2, 0.5 - 2*2, 1.25 + 7*0) = (5, -3.5, 1.25, 0, ...);
    zero tails."""
    g = np.zeros(pkg.runtime.slot_stride_bytes(A) // 4, np.float32)
    W0, B0 = 9344, 9344 + 256 * 321
    W1, B1 = B0 + 256, B0 + 256 + 128 * 256
    W2, B2 = B1 + 128, B1 + 128 + 64 * 128
    g[B0 + 0] = -1.0 + 0.5 * 2.0
    g[W0 + 0 * 321 + 1] = 1.5
    g[B0 + 1] = -0.5
    g[W0 + 1 * 321 + 64] = 1.0
    g[W0 + 1 * 321 + 129] = 2.0
    g[B0 + 2] = -2.0
    g[W0 + 2 * 321 + 5] = 1.0
    g[B1 + 0] = 0.25
    g[W1 + 0], g[W1 + 1] = 2.0, -0.5
    g[B1 + 1] = -3.0
    g[W1 + 256 + 1] = 0.5
    g[B2 + 0] = -1.0
    g[W2 + 0] = 3.0
    g[B2 + 1] = 0.5
    g[W2 + 128 + 0], g[W2 + 128 + 1] = -2.0, 4.0
    g[B2 + 2] = 1.25
    g[W2 + 256 + 1] = 7.0
    g[9280 + 0], g[9280 + 4], g[9280 + 63] = 1.0, 1.0, 3.0
    return g


def test_closed_form_choice(pkg, orc):
    rtm = pkg.runtime
    words = np.stack([orc.word(w) for w in ("CABBY", "CACAO", "FUZZY")])
    g = _trunk_golden_genome(pkg, 3)
    turn = (orc.word("ABCDE"), tiles("GY-GY"))
    pol, _ = orc.policy_forward(g[:orc.policy_param_count()], [turn])
    assert pol[:3].tolist() == [5.0, -3.5, 1.25] and not pol[3:].any()
    m = rtm.Model(g, words)
    try:
        # fixed features A, B, C (count, or -1 when absent): CABBY (1, 2, 1) -> 5 - 7 + 1.25 = -0.75,
        # CACAO (2, -1, 2) -> 10 + 3.5 + 2.5 = 16, FUZZY (-1, -1, -1) -> -5 + 3.5 - 1.25 = -2.75
        r = m.select([turn])
        assert (r.status, r.action_index, r.score, r.second_score, bytes(r.word)) == (
            0, 1, 16.0, -0.75, bytes(words[1]))
        r = m.select([turn, (words[1], tiles("-Y---"))])  # CACAO guessed: masked
        assert (r.status, r.action_index, r.score, r.second_score) == (0, 0, -0.75, -2.75)
    finally:
        m.close()


def _random_genome(pkg, A, seed):
    rng = np.random.default_rng(seed)
    g = np.zeros(pkg.runtime.slot_stride_bytes(A) // 4, np.float32)
    fan = [(0, 9280, 145), (9344, 9344 + 256 * 321, 321), (9344 + 256 * 321 + 256, 9344 + 256 * 321 + 256 + 128 * 256,
                                                          256),
           (9344 + 256 * 321 + 256 + 128 * 256 + 128, 9344 + 256 * 321 + 256 + 128 * 256 + 128 + 64 * 128, 128)]
    for lo, hi, fi in fan:
        g[lo:hi] = rng.standard_normal(hi - lo) * np.sqrt(2.0 / fi)
    P = 132928
    g[P:P + 38 * A] = rng.standard_normal(38 * A) * 0.1
    return g


def _random_boards(orc, catalog, rng, count, max_turns=5):
    A = len(catalog)
    boards = []
    while len(boards) < count:
        secret = catalog[rng.integers(A)]
        turns = []
        for _ in range(rng.integers(0, max_turns + 1)):
            w = catalog[rng.integers(A)]
            fb = orc.feedback(w, secret)
            if (fb == 0).all():
                break
            turns.append((w, fb))
        boards.append(turns)
    return boards


@pytest.mark.parametrize("A,dups", [(512, False), (365, True), (4739, False)])
def test_guesses_equal_the_oracle_bit_for_bit(pkg, orc, A, dups):
    rtm = pkg.runtime
    rng = np.random.default_rng(A)
    catalog = pkg.synth.make_catalog(A, seed=A).copy()
    if dups:  # a guessed word masks every copy of it
        for k in range(40, A, 9):
            catalog[k] = catalog[int(rng.integers(0, k))]
    g = _random_genome(pkg, A, A)
    boards = _random_boards(orc, catalog, rng, 96 if A < 4739 else 32)
    m = rtm.Model(g, catalog)
    try:
        many = m.select_many(boards)
        for i, turns in enumerate(boards):
            ref = orc.select_guess(g, A, catalog, turns)
            assert ref["status"] == 0
            one = m.select(turns)
            for r in (one, many[i]):
                assert r.status == 0 and r.action_index == ref["index"], (i, r.action_index, ref["index"])
                assert np.float32(r.score).view(np.uint32) == ref["best"].view(np.uint32), i
                assert np.float32(r.second_score).view(np.uint32) == ref["second"].view(np.uint32), i
                assert bytes(r.word) == bytes(catalog[ref["index"]])
            # the fp64 choice agrees wherever its top-2 margin is clear
            if ref["best64"] - ref["second64"] > 1e-4 * max(1.0, abs(ref["best64"])):
                assert ref["index64"] == ref["index"]
    finally:
        m.close()


def test_board_statuses(pkg, orc):
    rtm = pkg.runtime
    words = np.stack([orc.word(w) for w in ("CABBY", "CACAO", "FUZZY")])
    m = rtm.Model(_random_genome(pkg, 3, 1), words)
    try:
        bad_letter = [(np.array([0, 1, 2, 3, 26], np.uint8), tiles("-----"))]
        bad_tile = [(orc.word("CRANE"), np.array([0, 1, 2, 3, 3], np.uint8))]
        won = [(orc.word("CRANE"), tiles("GGGGG"))]
        six = [(orc.word(w), tiles("-----")) for w in ("AAAAA", "BBBBB", "DDDDD", "EEEEE", "GGGGG", "HHHHH")]
        exhausted = [(words[i], tiles("-Y---")) for i in range(3)]
        got = [r.status for r in m.select_many([bad_letter, bad_tile, won, six, exhausted, []])]
        assert got == [7, 7, 7, 7, 8, 0]
        for turns, status in ((won, 7), (exhausted, 8)):
            ref = orc.select_guess(_random_genome(pkg, 3, 1), 3, words, turns)
            assert ref["status"] == status
    finally:
        m.close()
    with pytest.raises(RuntimeError) as err:  # genome bytes of another action count
        rtm.Model(_random_genome(pkg, 4, 1), words)
    assert err.value.status == 17


def test_winner_artifact_plays_like_the_evaluation(pkg, orc, tmp_path):
    rtm = pkg.runtime
    wl = pkg.synth.workload("grid16_a512")
    A = 512
    stride = rtm.slot_stride_bytes(A)
    rt = pkg.SlabRuntime(slab_bytes=3 * 256 * stride, generation_bytes=256 * stride, action_count=A, grid_columns=16)
    try:
        rt.upload_catalog(wl["catalog"])
        rt.upload_secrets_csr(wl["offsets"], wl["ids"])
        rt.bootstrap(256, seed=9, variance_scale=1.0, tail_std=0.1)
        counts = rtm.word_counts(A, A)
        cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
        for g in range(2):
            rt.advance(9 + g, counts, cfg)
        rt.set_trace(rtm.TRACE_GAMES)
        rt.evaluate(counts)
        s = rt.summary()
        trace = rt.download_trace(rtm.TRACE_GAMES, 3 * int(wl["offsets"][-1]))
        genome = rt.download_slot(s.best_slot_index)
    finally:
        rt.close()
    meta = rtm.winner_metadata(s.best_fitness, s.generation_index, s.best_index, s.best_slot_index, A, 7)
    b, mpath = rtm.write_winner_artifact_in_directory(str(tmp_path), genome, wl["catalog"], meta)
    meta2, g2, letters = rtm.read_winner_artifact(b, mpath)
    assert g2.tobytes() == genome.tobytes() and meta2.best_index == s.best_index
    n = int(s.best_index)
    off, ids = wl["offsets"], wl["ids"]
    secrets = [letters[i] for i in ids[off[n]:off[n + 1]]]
    m = rtm.Model(g2, letters)
    try:
        boards = [[] for _ in secrets]
        chosen = [[] for _ in secrets]
        live = list(range(len(secrets)))
        while live:
            picks = m.select_many([boards[i] for i in live])
            nxt = []
            for i, r in zip(live, picks):
                assert r.status == 0
                fb = orc.feedback(letters[r.action_index], secrets[i])
                boards[i].append((letters[r.action_index], fb))
                chosen[i].append(int(r.action_index))
                if not (fb == 0).all() and len(boards[i]) < 6:
                    nxt.append(i)
            live = nxt
    finally:
        m.close()
    # the evaluation's fresh games of the winner come first, one per secret of its cell
    base = 3 * int(off[n])
    same = 0
    for i in range(len(boards)):
        t = trace[base + i]
        same += [int(x) for x in t["guesses"][:int(t["decisions"])]] == chosen[i]
    assert same >= 0.95 * len(boards), (same, len(boards))
    print(f"[winner] generation {s.generation_index} fitness {s.best_fitness:.4f}: {same}/{len(boards)} fresh games "
          f"replayed identically by the served model")


def test_cpp_driver_writes_a_servable_winner(pkg, tmp_path):
    """run_ga --winner DIR: the final generation's best genome as an artifact that the model serves."""
    import os
    import subprocess
    rtm = pkg.runtime
    cat = pkg.synth.make_catalog(64)
    sec = pkg.synth.sample_secrets(64, 32)
    off, ids = pkg.synth.block_csr(8, 8, 4, sec)
    cat.astype(np.uint8).tofile(tmp_path / "catalog.u8")
    off.astype(np.uint32).tofile(tmp_path / "offsets.u32")
    ids.astype(np.uint32).tofile(tmp_path / "secrets.u32")
    driver = os.path.join(os.path.dirname(rtm.LIB_PATH), "run_ga")
    res = subprocess.run([driver, str(tmp_path), "8", "2", "7", "0.01", "--winner", str(tmp_path / "win")],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    line = [x for x in res.stdout.splitlines() if x.startswith("winner artifact ")][0]
    binary = line.split(" ", 2)[2]
    meta, genome, letters = rtm.read_winner_artifact(binary, binary[:-4] + ".json")
    assert meta.generation_index == 2 and meta.action_count == 64 and meta.seed == 7
    assert np.array_equal(letters, cat)
    final = [x for x in res.stdout.splitlines() if x.startswith("final generation 2:")][0]
    assert f"best {meta.best_fitness:.6f}" in final
    m = rtm.Model(genome, letters)
    try:
        assert m.select([]).status == 0
    finally:
        m.close()

