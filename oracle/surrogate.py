"""CPU restatement of the surrogate-screening pieces (numpy, float64).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.  Surrogate screening has no
counterpart in the reference (which evaluates every child: slab_runtime.cu:1554-1661), so each function
restates this build's own definition, as documented beside ne_advance_generation_screened in
include/ne_runtime.h; the tests pin them against independent routines (torch autograd and
torch.optim.Adam, central finite differences, scipy.stats.spearmanr, brute-force Wordle rules).
"""
import numpy as np

import oracle as orc  # the oracle directory is on sys.path (tests/conftest.py)

IN, H1, H2 = 192, 128, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN * H1
OFF_W2 = OFF_B1 + H1
OFF_B2 = OFF_W2 + H1 * H2
OFF_W3 = OFF_B2 + H2
OFF_B3 = OFF_W3 + H2
N_PARAMS = OFF_B3 + 1  # 33,025


def unpack(params):
    """Flat parameter array -> (W1T [192][128], b1, W2T [128][64], b2, w3 [64], b3)."""
    p = np.asarray(params, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(IN, H1), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H1, H2),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(params, x):
    """y = w3 . relu(W2T^T relu(W1T^T x + b1) + b2) + b3, per row of x [n][192]; returns (y, h1, h2)."""
    w1, b1, w2, b2, w3, b3 = unpack(params)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ w1 + b1, 0.0)
    h2 = np.maximum(h1 @ w2 + b2, 0.0)
    return h2 @ w3 + b3, h1, h2


def loss(params, x, t):
    """1/(2n) sum (y - t)^2."""
    y, _, _ = forward(params, x)
    return 0.5 * np.mean((y - np.asarray(t, np.float64)) ** 2)


def gradient(params, x, t):
    """Gradient of loss() by backpropagation (ReLU derivative 0 at 0), as a flat array."""
    w1, b1, w2, b2, w3, b3 = unpack(params)
    x = np.asarray(x, np.float64)
    n = x.shape[0]
    y, h1, h2 = forward(params, x)
    dy = (y - np.asarray(t, np.float64)) / n
    dh2 = np.outer(dy, w3) * (h2 > 0)
    dh1 = (dh2 @ w2.T) * (h1 > 0)
    g = np.zeros(N_PARAMS)
    g[OFF_W1:OFF_B1] = (x.T @ dh1).ravel()
    g[OFF_B1:OFF_W2] = dh1.sum(0)
    g[OFF_W2:OFF_B2] = (h1.T @ dh2).ravel()
    g[OFF_B2:OFF_W3] = dh2.sum(0)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    return g


def adam_step(params, m, v, step, x, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, with bias correction); `step` counts from 1.  Returns (params, m, v)."""
    g = gradient(params, x, t)
    m = beta1 * np.asarray(m, np.float64) + (1 - beta1) * g
    v = beta2 * np.asarray(v, np.float64) + (1 - beta2) * g * g
    mhat = m / (1 - beta1 ** step)
    vhat = v / (1 - beta2 ** step)
    return np.asarray(params, np.float64) - lr * mhat / (np.sqrt(vhat) + eps), m, v


def average_ranks(a):
    a = np.asarray(a, np.float64)
    order = np.argsort(a, kind="stable")
    r = np.empty(a.size)
    i = 0
    while i < a.size:
        k = i
        while k + 1 < a.size and a[order[k + 1]] == a[order[i]]:
            k += 1
        r[order[i:k + 1]] = 0.5 * (i + k) + 1.0
        i = k + 1
    return r


def rank_correlation(a, b):
    """Spearman's rho: Pearson correlation of average ranks; NaN below 3 points or for a constant side."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.size < 3 or not (np.isfinite(a).all() and np.isfinite(b).all()):
        return float("nan")
    ra, rb = average_ranks(a), average_ranks(b)
    ra -= ra.mean()
    rb -= rb.mean()
    den = np.sqrt((ra * ra).sum() * (rb * rb).sum())
    return float("nan") if den == 0 else float((ra * rb).sum() / den)


def _code(word):
    return int(sum(int(word[q]) << (5 * q) for q in range(5)))


def probe_features(genome, catalog, A, probes, parent_fitness, margin=1e-4):
    """The 192 fingerprint inputs of one genome and, per probe, whether its guess is certain.

    probes: [16][4] catalog indices (secret, then the 3 guesses of board k); parent_fitness: (f1, f2).
    Probe p = 16 d + k plays board k truncated to d turns: the fp64 policy on that board, FP64 scores of
    the first A actions (fixed letter features + the genome's tail rows), guessed words masked by word
    identity, argmax (lowest index on ties).  certain[p] is False when the fp64 top-2 gap is below
    margin * max(1, |best|) -- there the FP32 GPU argmax may legitimately differ."""
    P = orc.policy_param_count()
    tails = np.asarray(genome[P:P + 38 * A], np.float64).reshape(A, 38)
    fixed = np.stack([orc.fixed_features(catalog[a]) for a in range(A)]).astype(np.float64)
    emb = np.concatenate([fixed, tails], 1)
    codes = np.array([_code(catalog[a]) for a in range(A)])
    out = np.zeros(192)
    certain = np.ones(64, bool)
    for d in range(4):
        for k in range(16):
            p = 16 * d + k
            secret = catalog[probes[k][0]]
            turns = [(catalog[probes[k][1 + t]], orc.feedback(catalog[probes[k][1 + t]], secret)) for t in range(d)]
            _, pol64 = orc.policy_forward(genome, turns)
            s = emb @ pol64
            masked = np.zeros(A, bool)
            for t in range(d):
                masked |= codes == _code(catalog[probes[k][1 + t]])
            s = np.where(masked, -np.inf, s)
            g = int(np.argmax(s))
            srt = np.sort(s[np.isfinite(s)])
            if srt.size > 1 and srt[-1] - srt[-2] < margin * max(1.0, abs(srt[-1])):
                certain[p] = False
            fb = orc.feedback(catalog[g], secret)
            out[3 * p] = (2 * np.sum(fb == 0) + np.sum(fb == 1)) / 10.0
            if d == 0:
                out[3 * p + 1] = 10.0 * parent_fitness[k % 2]
            else:
                out[3 * p + 1] = float(all(np.array_equal(orc.feedback(w, catalog[g]), f) for w, f in turns))
            out[3 * p + 2] = float(_code(catalog[g]) == _code(secret))
    return out, certain
