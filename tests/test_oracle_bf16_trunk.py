"""Pins the oracle's BF16-trunk forward (evaluation mode 2: layer 0 in fp32; layers 1 and 2 on split bfloat16
operands -- v = hi + lo with hi = bf16(v), lo = bf16(v - hi) -- taking w.x as hi.hi + hi.lo + lo.hi, exact
products summed in fp64 from the fp32 bias, outputs rounded to fp32).

Checks: the rounding against torch's own float32 -> bfloat16 conversion (round to nearest even) including
exact halfway cases; the reference's trunk golden (dense_trunk_test.cpp:37-128), whose operands are all
bfloat16-exact, so the BF16 trunk must reproduce it exactly; and a random genome whose layer 0 is exact in
any summation order, recomputed independently with torch's rounding and numpy fp64 dot products."""
import numpy as np
import pytest
import torch

ENC_B, W0 = 9280, 9344
B0 = W0 + 256 * 321
W1, B1 = B0 + 256, B0 + 256 + 128 * 256
W2, B2 = B1 + 128, B1 + 128 + 64 * 128


def torch_bf16(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


def test_bf16_rounding_matches_torch(orc):
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 20000),
                        rng.standard_normal(1000).astype(np.float32) * np.float32(1e-39)])  # subnormals too
    # exact halfway cases: 1 + 2^-8 rounds down to 1 (even), 1 + 3*2^-8 up to 1 + 2^-6; negatives mirror
    half = np.array([1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -(1 + 2.0 ** -8), -(1 + 3 * 2.0 ** -8), 3 * 2.0 ** -9],
                    np.float32)
    x = np.concatenate([x, half])
    got = orc.bf16_round(x)
    assert np.array_equal(got.view(np.uint32), torch_bf16(x).view(np.uint32))
    assert got[-5:].tolist() == [1.0, 1 + 2.0 ** -6, -1.0, -(1 + 2.0 ** -6), 3 * 2.0 ** -9]


def test_bf16_trunk_reproduces_the_reference_trunk_golden(orc):
    """Every weight and activation of dense_trunk_test.cpp:37-128 is bfloat16-exact: (2.0, -1.5, 1.25) shape."""
    P = orc.policy_param_count(0)
    g = np.zeros(P + 38, np.float32)
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
    g[ENC_B + 0], g[ENC_B + 4], g[ENC_B + 63] = 1.0, 1.0, 3.0
    turns = [(orc.word("ABCDE"), np.array([0, 1, 2, 0, 1], np.uint8))]
    t16 = orc.policy_forward_t16(g, turns)
    f32, _ = orc.policy_forward(g, turns)
    assert t16.tolist() == f32.tolist()
    h00, h01 = 1.5, 2.5
    h10, h11 = max(0.0, 0.25 + 2.0 * h00 - 0.5 * h01), max(0.0, -3.0 + 0.5 * h01)
    assert t16[:3].tolist() == [-1.0 + 3.0 * h10, 0.5 - 2.0 * h10 + 4.0 * h11, 1.25 + 7.0 * h11]


@pytest.mark.parametrize("seed", [1, 2])
def test_bf16_trunk_against_an_independent_statement(orc, seed):
    """Layer 0 exact in fp32 in any order (dyadic operands of few bits), so h0 is known exactly; layers 1
    and 2 random normal fp32 weights.  Expected: torch's bfloat16 rounding of weights and ReLU'd inputs,
    numpy fp64 dot products from the fp32 bias, each layer's output rounded to fp32."""
    rng = np.random.default_rng(seed)
    P = orc.policy_param_count(0)
    g = np.zeros(P + 38, np.float32)
    enc = rng.integers(-16, 17, 64).astype(np.float32) / 8  # encoder output = its bias (zero weights)
    g[ENC_B:ENC_B + 64] = enc
    g[W0:B0] = (rng.integers(-512, 513, 256 * 321).astype(np.float32) / 256)
    g[B0:B0 + 256] = rng.integers(-4096, 4097, 256).astype(np.float32) / 2048
    g[W1:B1] = (rng.standard_normal(128 * 256) / 16).astype(np.float32)
    g[B1:B1 + 128] = (rng.standard_normal(128) * 0.1).astype(np.float32)
    g[W2:B2] = (rng.standard_normal(64 * 128) / 11).astype(np.float32)
    g[B2:B2 + 64] = (rng.standard_normal(64) * 0.1).astype(np.float32)
    turns = [(orc.word("CRANE"), np.array([2, 1, 1, 2, 2], np.uint8))]
    x = np.zeros(321)
    x[1:65] = enc  # one turn: slot 0 holds the encoder output, the virgin flag is 0
    h0 = np.maximum(g[W0:B0].astype(np.float64).reshape(256, 321) @ x + g[B0:B0 + 256], 0.0)
    assert np.array_equal(h0.astype(np.float32).astype(np.float64), h0), "layer 0 must be exact in fp32"
    h0 = h0.astype(np.float32)
    def split(v):
        v = np.asarray(v, np.float32)
        hi = torch_bf16(v)
        return hi.astype(np.float64), torch_bf16(v - hi).astype(np.float64)

    def dense(w, b, x):
        wh, wl = split(w)
        xh, xl = split(x)
        return (wh @ xh + wh @ xl + wl @ xh + b).astype(np.float32)

    h1 = np.maximum(dense(g[W1:B1].reshape(128, 256), g[B1:B1 + 128], h0), 0.0)
    expect = dense(g[W2:B2].reshape(64, 128), g[B2:B2 + 64], h1)
    got = orc.policy_forward_t16(g, turns)
    scale = np.abs(g[W2:B2].reshape(64, 128).astype(np.float64)) @ h1 + np.abs(g[B2:B2 + 64])
    assert np.all(np.abs(got.astype(np.float64) - expect) <= 2.0 ** -22 * scale)
    # the split is active and is all that differs: the fp32 trunk differs by about 2^-16 relative, more than
    # fp32 rounding and far less than one bfloat16 rounding (2^-9)
    f32, _ = orc.policy_forward(g, turns)
    rel = np.abs(f32.astype(np.float64) - got) / scale
    assert 2e-7 < rel.max() < 2e-4, rel.max()
    one_term = (split(g[W2:B2].reshape(64, 128))[0] @ split(h1)[0] + g[B2:B2 + 64]).astype(np.float32)
    assert np.abs(one_term.astype(np.float64) - got).max() > 10 * np.abs(f32.astype(np.float64) - got).max()


def test_mode2_evaluation_runs_the_bf16_trunk(orc):
    """oro_evaluate mode 2 plays with the BF16-trunk policy: its first decision's best score is the one
    PolicyForwardT16 gives (fp32 sequential scoring), and it differs from mode 0's on some game."""
    import importlib
    synth = importlib.import_module("__graft_entry__").import_package().synth
    catalog = synth.make_catalog(64)
    secrets = synth.sample_secrets(64, 8)
    offsets, ids = synth.block_csr(2, 2, 2, secrets)
    P = orc.policy_param_count(0)
    rng = np.random.default_rng(3)
    genomes = [(rng.standard_normal(P + 38 * 64) * 0.08).astype(np.float32) for _ in range(4)]
    r2 = orc.evaluate(genomes, 0, 64, 64, catalog, offsets, ids, mode=2, trace=True)
    r0 = orc.evaluate(genomes, 0, 64, 64, catalog, offsets, ids, mode=0, trace=True)
    assert r2["status"] == 0 and r0["status"] == 0
    pol = orc.policy_forward_t16(genomes[0], [])
    fixed = np.stack([orc.fixed_features(catalog[a]) for a in range(64)])
    tails = genomes[0][P:].reshape(64, 38)
    s = np.array([orc.score_action(pol, fixed[a], tails[a]) for a in range(64)], np.float32)
    fresh = 0  # game 0 of genome 0 is a fresh game: its first decision is on the empty board
    assert r2["traces"]["guesses"][fresh, 0] == int(np.argmax(s))
    assert r2["traces"]["best"][fresh, 0] == s.max()
    assert not np.array_equal(r2["traces"]["best"], r0["traces"]["best"])
