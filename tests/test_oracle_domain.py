"""Pins the oracle's integer and model semantics: the reference's own known-answer tests restated,
brute force against vectors written from the reference headers (tests/golden), and closed forms."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_integer_vectors.npz")
TILE = {"G": 0, "Y": 1, "-": 2}


def tiles(s):
    return np.array([TILE[c] for c in s], np.uint8)


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


# ---- feedback (reference tests/wordle/wordle_grid_test.cpp:107-193) ---------------------------
@pytest.mark.parametrize("guess,sol,expect", [("CRANE", "SOLAR", "-YY--"), ("EERIE", "STEEL", "YY---"),
                                              ("HELLO", "LEVEL", "-GYY-"), ("LEVEL", "HELLO", "YG--Y"),
                                              ("ABCDE", "AEBDF", "GY-GY")])
def test_feedback_known_answers(orc, guess, sol, expect):
    assert orc.feedback(orc.word(guess), orc.word(sol)).tolist() == tiles(expect).tolist()


def _feedback_python(guess, sol):
    # an independent statement: greens first, then yellows limited by unmatched solution letters
    fb = [2] * 5
    pool = {}
    for p in range(5):
        if guess[p] == sol[p]:
            fb[p] = 0
        else:
            pool[sol[p]] = pool.get(sol[p], 0) + 1
    for p in range(5):
        if fb[p] == 0:
            continue
        if pool.get(guess[p], 0) > 0:
            fb[p] = 1
            pool[guess[p]] -= 1
    return fb


def test_feedback_all_pairs_against_reference_vectors(orc, golden):
    words, ref_fb = golden["words"], golden["feedback"]
    n = len(words)
    for i in range(n):
        for j in range(n):
            got = orc.feedback(words[i], words[j])
            assert got.tolist() == ref_fb[i, j].tolist(), (i, j)
            assert _feedback_python(list(words[i]), list(words[j])) == ref_fb[i, j].tolist()


def test_feedback_against_reference_probe_when_built(orc):
    ref = orc.ref()
    if ref is None:
        pytest.skip("reference probe not built (no /root/reference)")
    rng = np.random.default_rng(11)
    for _ in range(3000):
        g = rng.integers(0, 6, 5).astype(np.uint8)  # small alphabet: many repeated letters
        s = rng.integers(0, 6, 5).astype(np.uint8)
        out = np.zeros(5, np.uint8)
        assert ref.ref_feedback(orc.ptr(g), orc.ptr(s), orc.ptr(out)) == 1
        assert orc.feedback(g, s).tolist() == out.tolist()


# ---- toroidal grid (reference tests/genetic_algorithm/spatial_grid_test.cpp:79-117) -------------
def test_moore_neighbourhoods_against_reference_vectors(orc, golden):
    for key in golden.files:
        if not key.startswith("moore_") or key.endswith("_count"):
            continue
        rows, cols = map(int, key.split("_")[1].split("x"))
        radius = int(key.split("_r")[1])
        ref_lists, ref_counts = golden[key], golden[key + "_count"]
        for c in range(rows * cols):
            out = np.zeros(25, np.uint64)
            k = orc.lib().oro_moore(rows, cols, c, radius, orc.ptr(out))
            assert k == ref_counts[c]
            assert out[:k].astype(np.int64).tolist() == ref_lists[c, :k].tolist()


def test_moore_closed_forms(orc):
    out = np.zeros(25, np.uint64)
    assert orc.lib().oro_moore(5, 5, 12, 2, orc.ptr(out)) == 25
    assert orc.lib().oro_moore(16, 16, 0, 1, orc.ptr(out)) == 9
    assert orc.lib().oro_moore(3, 3, 4, 2, orc.ptr(out)) == 9   # every cell once on a 3x3 torus
    assert orc.lib().oro_chebyshev(16, 16, 0, 255) == 1          # wraps in both axes


# ---- shard schedule -> per-cell CSR (training_data.hpp:150-316), the product's host builder -------
def test_product_shard_csr_against_reference_vectors(pkg, golden):
    rtm = pkg.runtime
    for key in golden.files:
        if not key.startswith("shard_case"):
            continue
        k = key[len("shard_case"):]
        initial, step, period, introduced, gen, rows, cols, epi_rows, r0, growth = map(int, golden[key])
        counts = rtm.WordCounts(training_word_count=introduced, action_space_word_count=introduced,
                                training_word_schedule=rtm.WordCountSchedule(initial, step, period),
                                shard_initial_radius=r0, shard_radius_growth_period_generations=growth)
        offsets, ids = rtm.build_shard_csr(counts, gen, rows, cols, epi_rows)
        assert offsets.tolist() == golden["shard_offsets" + k].tolist(), key
        assert ids.tolist() == golden["shard_ids" + k].tolist(), key


def test_shard_exposure_closed_form(pkg):
    """Reference KAT (slab_runtime_test.cu:533-575): foundation 20 words everywhere, a second shard of
    10 words only at its golden-ratio centre while its radius is 0."""
    rtm = pkg.runtime
    counts = rtm.word_counts(30, 30, initial=20, step=10, period=1, radius0=0, growth=2)
    offsets, ids = rtm.build_shard_csr(counts, 1, 5, 5)
    w = np.diff(offsets)
    centre = (1 * 11400714819323198485) % (2 ** 64) % 25
    assert w[centre] == 30 and (np.delete(w, centre) == 20).all()
    assert ids[offsets[centre]:offsets[centre + 1]].tolist() == list(range(30))


# ---- model (reference tests/model/*) ----------------------------------------------------------
def _enc128_genome(orc, A=3):
    """Reference fixtures for the 145->128->64 encoder model (shared_encoder_fixture.hpp:43-101,
    policy_model_fixture.hpp:49-107), laid out in this build's fp32 genome order."""
    P = orc.policy_param_count(128)
    g = np.zeros(P + 38 * A, np.float32)
    w1, b1, w2, b2 = 0, 128 * 145, 128 * 145 + 128, 128 * 145 + 128 + 64 * 128
    t = b2 + 64
    W0, B0 = t, t + 256 * 321
    W1, B1 = B0 + 256, B0 + 256 + 128 * 256
    W2, B2 = B1 + 128, B1 + 128 + 64 * 128
    assert B2 + 64 == P
    g[b1 + 0], g[b1 + 1], g[b1 + 2] = -2.0, 1.0, -1.0
    g[w1 + 0 * 145 + 0] = 1.5
    g[w1 + 0 * 145 + 27] = 0.5
    g[w1 + 0 * 145 + 130] = 4.0
    g[w1 + 1 * 145 + 54] = 2.0
    g[w1 + 1 * 145 + 138] = -5.0
    g[w1 + 1 * 145 + 143] = 0.5
    g[w1 + 2 * 145 + 81] = 1.0
    g[w1 + 2 * 145 + 108] = 2.0
    g[w1 + 2 * 145 + 134] = 0.25
    g[w1 + 2 * 145 + 139] = 0.75
    g[b2 + 0], g[b2 + 1], g[b2 + 2] = 0.5, -2.0, 1.25
    g[w2 + 0 * 128 + 0], g[w2 + 0 * 128 + 2] = 2.0, -1.0
    g[w2 + 1 * 128 + 0], g[w2 + 1 * 128 + 2] = -0.5, 4.0
    g[w2 + 2 * 128 + 1] = 7.0
    g[B0 + 0] = -1.0
    g[W0 + 0 * 321 + 1], g[W0 + 0 * 321 + 2] = 1.0, 0.5
    g[W0 + 1 * 321 + 2], g[W0 + 1 * 321 + 3] = -0.25, 2.0
    g[B0 + 2] = -1.5
    g[W0 + 2 * 321 + 0] = 2.0
    g[B1 + 0] = 0.25
    g[W1 + 0], g[W1 + 1] = 0.5, 2.0
    g[B1 + 1] = -1.0
    g[W1 + 256 + 0], g[W1 + 256 + 2] = 0.25, 4.0
    g[B2 + 0] = -2.0
    g[W2 + 0] = 2.0
    g[B2 + 1] = 0.5
    g[W2 + 128 + 0], g[W2 + 128 + 1] = -0.5, 4.0
    g[B2 + 2] = 1.0
    g[W2 + 256 + 1] = 3.0
    return g


def test_encoder_golden(orc):
    g = _enc128_genome(orc)
    o32, o64 = orc.encode_turn(g, orc.word("ABCDE"), tiles("GY-GY"), enc_hidden=128)
    assert o32[:3].tolist() == [5.5, 8.0, 1.25] and not o32[3:].any()
    assert o64[:3].tolist() == [5.5, 8.0, 1.25]


def test_policy_goldens(orc):
    g = _enc128_genome(orc)
    virgin32, virgin64 = orc.policy_forward(g, [], enc_hidden=128)
    assert virgin32[:3].tolist() == [-1.5, 4.375, 4.0] and virgin64[:3].tolist() == [-1.5, 4.375, 4.0]
    one32, _ = orc.policy_forward(g, [(orc.word("ABCDE"), tiles("GY-GY"))], enc_hidden=128)
    # closed form of PopulateExpectedSingleTurnOutput with encoded turn (5.5, 8, 1.25)
    h00, h01 = -1.0 + 5.5 + 0.5 * 8.0, -0.25 * 8.0 + 2.0 * 1.25
    h10, h11 = 0.25 + 0.5 * h00 + 2.0 * h01, -1.0 + 0.25 * h00
    assert one32[:3].tolist() == [-2.0 + 2.0 * h10, 0.5 - 0.5 * h10 + 4.0 * h11, 1.0 + 3.0 * h11]


def test_trunk_golden_through_the_145_64_layout(orc):
    """dense_trunk_test.cpp:37-128: input x = {0:2, 1:1, 5:1, 64:3, 129:1} -> (2.0, -1.5, 1.25).
    The 145->64 encoder is set to the identity-like map that produces exactly that x from one turn."""
    P = orc.policy_param_count(0)
    g = np.zeros(P + 38, np.float32)
    W0, B0 = 9344, 9344 + 256 * 321
    W1, B1 = B0 + 256, B0 + 256 + 128 * 256
    W2, B2 = B1 + 128, B1 + 128 + 64 * 128
    # x[0] is the virgin flag (0 after one turn): fold its weight 0.5*2 into the bias instead
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
    # encoder 145->64 (weights [64][145] at 0, biases at 9280): x[1+e] = enc[e] for the first turn,
    # so enc[0] = 1, enc[4] = 1, enc[63] = 3 produce x[1] = 1, x[5] = 1, x[64] = 3
    g[9280 + 0], g[9280 + 4], g[9280 + 63] = 1.0, 1.0, 3.0
    out32, out64 = orc.policy_forward(g, [(orc.word("ABCDE"), tiles("GY-GY"))])
    # x[129] (turn 1 slot) is empty after one turn, so hidden0_1 = relu(-0.5 + 3) = 2.5 here
    h00 = max(0.0, 0.0 + 1.5 * 1.0)
    h01 = max(0.0, -0.5 + 1.0 * 3.0)
    h02 = max(0.0, -2.0 + 1.0)
    h10 = max(0.0, 0.25 + 2.0 * h00 - 0.5 * h01)
    h11 = max(0.0, -3.0 + 0.5 * h01)
    assert out32[:3].tolist() == [-1.0 + 3.0 * h10, 0.5 - 2.0 * h10 + 4.0 * h11, 1.25 + 7.0 * h11]
    assert out64[:3].tolist() == out32[:3].tolist()


def test_encoder_145_64_closed_form(orc):
    """One linear layer on a one-hot turn: out[o] = b[o] + sum of the 10 active columns."""
    P = orc.policy_param_count(0)
    g = np.zeros(P, np.float32)
    o, k = np.meshgrid(np.arange(64), np.arange(145), indexing="ij")
    g[:64 * 145] = (o * 1000 + k).astype(np.float32).ravel()
    g[9280:9344] = np.arange(64, dtype=np.float32) * 0.5
    w, t = orc.word("CRANE"), tiles("-YY--")
    active = [p * 26 + int(w[p]) for p in range(5)] + [130 + 3 * p + int(t[p]) for p in range(5)]
    o32, _ = orc.encode_turn(g, w, t)
    expect = np.arange(64) * 0.5 + np.array([sum(oo * 1000 + kk for kk in active) for oo in range(64)])
    assert o32.tolist() == expect.astype(np.float32).tolist()


def test_scoring_known_answers(orc):
    """output_embedding_test.cpp:80-137: APPLE 7, BEEFY 11.5, MUMMY -15; CRASS fixed features."""
    pol = np.zeros(64, np.float32)
    pol[0], pol[1], pol[4], pol[12] = 1.0, 2.0, 3.0, -1.0
    pol[26], pol[27] = 4.0, -2.0
    tails = {"APPLE": (0.5, -1.0), "BEEFY": (1.0, 0.25), "MUMMY": (-0.5, 2.0)}
    got = {}
    for wd, (t0, t1) in tails.items():
        tail = np.zeros(38, np.float32)
        tail[0], tail[1] = t0, t1
        got[wd] = orc.score_action(pol, orc.fixed_features(orc.word(wd)), tail)
    assert got == {"APPLE": 7.0, "BEEFY": 11.5, "MUMMY": -15.0}
    f = orc.fixed_features(orc.word("CRASS"))
    assert (f[0], f[1], f[2], f[17], f[18]) == (1.0, -1.0, 1.0, 1.0, 2.0)


def test_end_to_end_single_decision_golden(orc):
    """single_model_device_runtime_test.cu:97-130: fixture policy on grid (AEBDF | ABCDE) over
    {CABBY, CACAO, FUZZY} with zero tails picks CACAO (index 1) with score 24.5."""
    g = _enc128_genome(orc)
    pol, _ = orc.policy_forward(g, [(orc.word("ABCDE"), tiles("GY-GY"))], enc_hidden=128)
    scores = [orc.score_action(pol, orc.fixed_features(orc.word(w)), np.zeros(38, np.float32))
              for w in ("CABBY", "CACAO", "FUZZY")]
    assert int(np.argmax(scores)) == 1 and scores[1] == 24.5


# ---- fitness (evaluation_ops.cuh:87-111) ------------------------------------------------------
def test_fitness_normalisation_closed_forms(orc):
    L = orc.lib()
    assert L.oro_normalize_fitness(0.0, 16) == np.float32(1.40129846e-45)
    assert L.oro_normalize_fitness(45.0 * 16, 16) == 1.0
    assert L.oro_normalize_fitness(14.0, 2) == np.float32(14.0) / np.float32(90.0)
    assert L.oro_normalize_fitness(1e9, 3) == 1.0


# ---- hint grids (reference tests/wordle/hint_grid_test.cpp:138-229) ---------------------------
def _grid_words(letters, tc):
    return ["".join(chr(65 + int(c)) for c in letters[t]) for t in range(tc)]


def test_hint_group_known_answers(orc):
    ok, tc, letters, tl = orc.hint_group(orc.word("SPARE"))
    assert ok == 1
    assert tc[0] == 4 and all((tl[0, t] == 1).all() for t in range(4))  # 4 all-yellow rotations
    assert _grid_words(letters[1], 1) == ["PSARE"] and _grid_words(letters[2], 1) == ["APSRE"]
    ok, tc, letters, tl = orc.hint_group(orc.word("AABCD"))
    assert ok == 1 and _grid_words(letters[0], tc[0]) == ["BCDAA", "CDAAB"]  # rotations 2 and 3 only
    ok, tc, letters, tl = orc.hint_group(orc.word("LEVEL"))
    assert ok == 1 and _grid_words(letters[1], 1) == ["ELVEL"] and tl[1, 0].tolist() == [1, 1, 0, 0, 0]
    assert orc.hint_group(orc.word("AAAAA"))[0] == 0


def test_hint_groups_against_reference_vectors(orc, golden):
    for i, w in enumerate(golden["words"]):
        ok, tc, letters, tl = orc.hint_group(w)
        assert ok == golden["hint_ok"][i], i
        if not ok:
            continue
        assert tc.tolist() == golden["hint_tc"][i].tolist()
        for g in range(3):
            for t in range(tc[g]):
                assert letters[g, t].tolist() == golden["hint_letters"][i, g, t].tolist()
                assert tl[g, t].tolist() == golden["hint_tiles"][i, g, t].tolist()


def test_median_tail_norm_closed_forms(orc):
    rows = np.zeros((5, 38), np.float32)
    for a, v in enumerate([3.0, 1.0, 4.0, 1.5, 9.0]):
        rows[a, 0] = v
    assert orc.median_tail_norm(rows) == 3.0           # odd count: the middle value
    assert orc.median_tail_norm(rows[:4]) == np.float32(0.5) * (np.float32(1.5) + np.float32(3.0))


def test_injected_rows_have_the_median_norm(orc):
    """slab_runtime_test.cu:577-686: every injected tail has the median norm of the existing rows."""
    rng = np.random.default_rng(4)
    P = orc.policy_param_count(0)
    cat = np.stack([orc.word(w) for w in ("CRANE", "SLATE", "TRACE", "SPARE", "LEVEL", "MONEY")])
    g = (rng.standard_normal(P + 38 * 6) * 0.1).astype(np.float32)
    g[P + 38 * 4:] = 0
    st, out = orc.inject(g, 4, cat, 4, 2)
    assert st == 0
    target = orc.median_tail_norm(g[P:P + 38 * 4])
    for k in (4, 5):
        row = out[P + 38 * k:P + 38 * (k + 1)].astype(np.float64)
        assert abs(np.sqrt((row ** 2).sum()) - target) <= 1e-2 * target
    assert np.array_equal(out[:P + 38 * 4], g[:P + 38 * 4])
    st, _ = orc.inject(g, 4, cat, 3, 2)
    assert st == 15  # the first injected word must follow the existing rows
