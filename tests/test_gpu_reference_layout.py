"""The reference's own genome layout on the HIP path (single-model inference): the 145 -> 128 -> 64 encoder
(src/model/input_encoder/encoder_spec.hpp:19-20) and FP16 parameter storage (src/common/float16.hpp:9-13),
genome bytes as ComputeDynamicGenomeStrideBytes lays them out (dynamic_layout.hpp:15-35).

Goldens, with the reference's policy fixture (tests/model/policy_model/policy_model_fixture.hpp:49-107, its
encoder part shared_encoder_fixture.hpp:43-101) over {CABBY, CACAO, FUZZY} with zero tails:
* the single-turn board ABCDE / GY-GY picks CACAO, index 1, score 24.5
  (tests/inference/single_model_device_runtime_test.cu:97-130) -- the encoder golden (5.5, 8.0, 1.25) and the
  single-turn policy closed form lead there;
* the empty board's policy is the virgin golden (-1.5, 4.375, 4.0), so the scores are closed forms of the
  fixed features: CABBY -1.5 + 2 x 4.375 + 4.0 = 11.25, CACAO -3 - 4.375 + 8 = 0.625, FUZZY -6.875;
* a zero genome with CRANE guessed picks SLATE, index 1, score 0 (single_model_device_runtime_test.cu:133-170).
Random genomes of every layout are checked against the oracle's FP32 restatement (values widened from FP16
exactly) bit for bit: index, score and runner-up.
"""
import numpy as np
import pytest

from test_oracle_domain import _enc128_genome, tiles

pytestmark = pytest.mark.gpu


def _words(orc, *ws):
    return np.stack([orc.word(w) for w in ws])


@pytest.mark.parametrize("storage", ["fp16", "fp32"])
def test_reference_fixture_goldens(pkg, orc, storage):
    rtm = pkg.runtime
    words = _words(orc, "CABBY", "CACAO", "FUZZY")
    g = _enc128_genome(orc, A=3)
    st = rtm.STORAGE_FP16 if storage == "fp16" else rtm.STORAGE_FP32
    nbytes = rtm.model_genome_bytes(128, st, 3)
    if storage == "fp16":
        data = g.astype(np.float16)
        assert np.array_equal(data.astype(np.float32), g)  # the fixture is exact in binary16
        assert nbytes == 2 * (150528 + 38 * 3) == data.nbytes
    else:
        data = np.zeros(nbytes // 4, np.float32)
        data[:g.size] = g
    m = rtm.Model(data, words, encoder_hidden=128, storage=st)
    try:
        r = m.select([(orc.word("ABCDE"), tiles("GY-GY"))])
        assert (r.status, r.action_index, r.score, bytes(r.word)) == (0, 1, 24.5, bytes(words[1]))
        r = m.select([])
        assert (r.status, r.action_index, r.score, r.second_score) == (0, 0, 11.25, 0.625)
        r = m.select([(orc.word("CABBY"), tiles("-----"))])  # a guessed word is never chosen again
        assert r.status == 0 and r.action_index != 0
    finally:
        m.close()


def test_zero_genome_masks_the_guessed_word(pkg, orc):
    rtm = pkg.runtime
    words = _words(orc, "CRANE", "SLATE", "TRACE")
    data = np.zeros(rtm.model_genome_bytes(128, rtm.STORAGE_FP16, 3) // 2, np.float16)
    m = rtm.Model(data, words, encoder_hidden=128, storage=rtm.STORAGE_FP16)
    try:
        r = m.select([(words[0], orc.feedback(words[0], words[2]))])
        assert (r.status, r.action_index, r.score, bytes(r.word)) == (0, 1, 0.0, bytes(words[1]))
    finally:
        m.close()


def _random_boards(orc, words, count, seed):
    rng = np.random.default_rng(seed)
    boards = []
    for _ in range(count):
        secret = words[rng.integers(len(words))]
        turns = []
        for _ in range(int(rng.integers(0, 6))):
            w = words[rng.integers(len(words))]
            fb = orc.feedback(w, secret)
            if not fb.any():
                break
            turns.append((w, fb))
        boards.append(turns)
    return boards


@pytest.mark.parametrize("hidden,storage", [(128, "fp16"), (128, "fp32"), (0, "fp16")])
def test_random_genomes_match_the_oracle(pkg, orc, hidden, storage):
    rtm = pkg.runtime
    A = 700
    words = pkg.synth.make_catalog(A, seed=17)
    st = rtm.STORAGE_FP16 if storage == "fp16" else rtm.STORAGE_FP32
    P = orc.policy_param_count(hidden)
    rng = np.random.default_rng(hidden + st)
    g = (rng.standard_normal(P + 38 * A) * 0.08).astype(np.float32)
    if storage == "fp16":
        data = g.astype(np.float16)
        g = data.astype(np.float32)  # the values the model computes with
    else:
        data = np.zeros(rtm.model_genome_bytes(hidden, st, A) // 4, np.float32)
        data[:g.size] = g
    m = rtm.Model(data, words, encoder_hidden=hidden, storage=st)
    try:
        boards = _random_boards(orc, words, 96, 5)
        got = m.select_many(boards)
        for b, r in zip(boards, got):
            ref = orc.select_guess(g, A, words, b, enc_hidden=hidden)
            assert r.status == ref["status"]
            if ref["status"] == 0:
                assert (r.action_index, r.score, r.second_score) == (ref["index"], ref["best"], ref["second"])
    finally:
        m.close()


def test_layout_sizes_and_refusals(pkg):
    rtm = pkg.runtime
    assert rtm.model_genome_bytes(128, rtm.STORAGE_FP16, 4739) == 661220  # the reference's stride at 4,739 actions
    assert rtm.model_genome_bytes(0, rtm.STORAGE_FP32, 4739) == rtm.slot_stride_bytes(4739)
    assert rtm.model_genome_bytes(64, rtm.STORAGE_FP16, 10) == 0 and rtm.model_genome_bytes(128, 7, 10) == 0
    with pytest.raises(RuntimeError) as err:
        rtm.Model(np.zeros(100, np.float16), np.zeros((3, 5), np.uint8), encoder_hidden=128, storage=rtm.STORAGE_FP16)
    assert err.value.status == 17
