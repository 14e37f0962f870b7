"""ne_model_select_guesses_noalloc (the inference entry point that allocates nothing) against
ne_model_select_guesses through the C ABI: equal results byte for byte on two calls in a row, the second
with fewer boards (the shared static staging keeps the first call's bytes beyond them), and a clear
refusal above its maximum.  The boards mix played, fresh, finished and invalid ones."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _boards(pkg, orc, words, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        secret = words[rng.integers(len(words))]
        tc = int(rng.integers(0, 6))
        turns = []
        for _ in range(tc):
            g = words[rng.integers(len(words))]
            turns.append((g, orc.feedback(g, secret)))
        if i % 97 == 0:  # a won board (finished): NE_POLICY_FORWARD_FAILED
            turns = turns[:2] + [(secret, np.zeros(5, np.uint8))]
        elif i % 89 == 0:  # an invalid tile: NE_POLICY_FORWARD_FAILED
            turns = [(words[0], np.array([0, 1, 2, 3, 0], np.uint8))]
        out.append(pkg.runtime.make_board(turns))
    return out


def _random_genome(pkg, A, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal(pkg.runtime.slot_stride_bytes(A) // 4).astype(np.float32) * np.float32(0.05)
    return g


def test_noalloc_entry_equals_the_allocating_one(pkg, orc):
    rtm = pkg.runtime
    A = 1500
    words = pkg.synth.make_catalog(A, seed=11)
    m = rtm.Model(_random_genome(pkg, A, 3), words)
    try:
        big = _boards(pkg, orc, words, 3000, 1)
        small = _boards(pkg, orc, words, 700, 2)
        got_big = bytes(m.select_many_noalloc(big))
        got_small = bytes(m.select_many_noalloc(small))  # right after the larger call
        assert got_big == bytes(m.select_many(big))
        assert got_small == bytes(m.select_many(small))
        res = m.select_many_noalloc(big)
        statuses = {r.status for r in res}
        assert statuses >= {0, 7}, statuses
        assert bytes(res) == got_big
        one = m.select_many_noalloc(small[:1])
        assert bytes(one) == bytes(m.select_many(small[:1]))
        with pytest.raises(RuntimeError) as err:
            m.select_many_noalloc(_boards(pkg, orc, words, rtm.NOALLOC_MAX_BOARDS + 1, 4))
        assert err.value.status == 17
        full = _boards(pkg, orc, words, rtm.NOALLOC_MAX_BOARDS, 5)
        assert bytes(m.select_many_noalloc(full)) == bytes(m.select_many(full))
    finally:
        m.close()


def test_noalloc_entry_serves_two_models(pkg, orc):
    """Two models on one GPU share the staging array; calls in sequence each get their own model's answers."""
    rtm = pkg.runtime
    A = 600
    words = pkg.synth.make_catalog(A, seed=13)
    a = rtm.Model(_random_genome(pkg, A, 7), words)
    b = rtm.Model(_random_genome(pkg, A, 8), words)
    try:
        boards = _boards(pkg, orc, words, 500, 9)
        ra, rb = bytes(a.select_many_noalloc(boards)), bytes(b.select_many_noalloc(boards))
        assert ra == bytes(a.select_many(boards)) and rb == bytes(b.select_many(boards)) and ra != rb
    finally:
        a.close()
        b.close()
