"""Pins the oracle's draw-driven GA stages -- generation-0 bootstrap, cellular rank-roulette selection,
crossover and mutation -- against the reference's semantics without retyping the oracle's formulas:

* the reference's own selection known-answer test (distinct parents, both inside the radius-2
  neighbourhood; tests/genetic_algorithm/device/slab_runtime_test.cu:483-531);
* special cases that reduce to a copy, a flip or an always-mutate (temperatures and probabilities 0 and 1,
  src/genetic_algorithm/device/genome_ops.cuh:35-44,169-221,259-376);
* exact boundaries (the roulette threshold `<=`, the Bernoulli `<`, the tie keys of SplitMix64 pinned by
  its published first output);
* frequencies against closed forms (rank^e / sum, uniform neuron / row / splice-point indices, the
  mutated fraction, per-layer standard deviations with fan-in), with 5-sigma binomial bounds.

Each check is chosen so that a plausible mistake fails it: an inverted rank order, a wrong exponent, an
L2 flip applied to the wrong base, a neuron taken from the donor instead of the other parent, a swapped
splice prefix, `<=` for `<`, or a standard deviation from fan-out.
"""
import numpy as np
import pytest
from scipy import stats

# Reference layer shapes, row-major [out][in] then biases (shared_encoder.hpp:16-30 with the
# 145->64 encoder, dense_trunk.hpp:25-40): (fan_in, fan_out).
LAYERS = [(145, 64), (321, 256), (256, 128), (128, 64)]
TAIL = 38
MASK64 = (1 << 64) - 1


def segments():
    """[(fan_in, fan_out, (w0, w1), (b0, b1))] in genome order, and the policy parameter count."""
    off, out = 0, []
    for fin, fout in LAYERS:
        w = (off, off + fin * fout)
        off += fin * fout
        b = (off, off + fout)
        off += fout
        out.append((fin, fout, w, b))
    return out, off


SEGS, P = segments()


def test_layout_matches_reference_shapes(orc):
    assert orc.policy_param_count(0) == P == 132928


# ---- selection --------------------------------------------------------------------------------------
def moore_set(rows, cols, focal, r):
    fr, fc = divmod(focal, cols)
    return {((fr + dr) % rows) * cols + (fc + dc) % cols for dr in range(-r, r + 1) for dc in range(-r, r + 1)}


def chebyshev_set(rows, cols, focal, r):
    fr, fc = divmod(focal, cols)
    out = set()
    for i in range(rows * cols):
        ir, ic = divmod(i, cols)
        dr = min(abs(ir - fr), rows - abs(ir - fr))
        dc = min(abs(ic - fc), cols - abs(ic - fc))
        if max(dr, dc) <= r:
            out.add(i)
    return out


@pytest.mark.parametrize("rows,cols", [(5, 5), (16, 16), (64, 64)])
def test_selection_known_answer_parents_distinct_and_local(orc, rows, cols):
    """slab_runtime_test.cu:483-531: every child has two different parents, both in the radius-2 Moore
    neighbourhood of its own cell."""
    n = rows * cols
    w = orc.rank_weights(0.5, 25)
    rng = np.random.default_rng(rows)
    for gen in range(3):
        fit = rng.random(n).astype(np.float32)
        fit[rng.integers(0, n, n // 4)] = 0.25  # ties
        st, pairs = orc.plan(fit, np.ones(n, np.uint8), n, n, cols, gen, 2, w, 1000 + gen)
        assert st == 0
        for c in range(n):
            a, b = int(pairs[c, 0]), int(pairs[c, 1])
            hood = moore_set(rows, cols, c, 2)
            assert a != b and a in hood and b in hood, (c, a, b)


@pytest.mark.parametrize("radius", [1, 3])
def test_selection_other_radii_stay_inside(orc, radius):
    """Radius 1 (Moore list) and radius 3 (the Chebyshev scan of selection_ops.cuh:163-216)."""
    rows = cols = 12
    n = rows * cols
    fit = np.random.default_rng(radius).random(n).astype(np.float32)
    st, pairs = orc.plan(fit, np.ones(n, np.uint8), n, n, cols, 4, radius, orc.rank_weights(0.5, n), 77)
    assert st == 0
    for c in range(n):
        hood = chebyshev_set(rows, cols, c, radius)
        assert pairs[c, 0] != pairs[c, 1] and pairs[c, 0] in hood and pairs[c, 1] in hood


def test_selection_unfitted_cells_never_chosen_and_failure(orc):
    rows = cols = 5
    n = 25
    fit = np.linspace(0, 1, n).astype(np.float32)
    has = np.ones(n, np.uint8)
    has[[3, 7, 20]] = 0
    st, pairs = orc.plan(fit, has, n, n, cols, 0, 2, orc.rank_weights(0.5, 25), 5)
    assert st == 0 and not set(pairs.ravel().tolist()) & {3, 7, 20}
    has[:] = 0
    has[4] = 1  # one fitted cell: no distinct second parent -> kParentSelectionFailed (11)
    st, _ = orc.plan(fit, has, n, n, cols, 0, 2, orc.rank_weights(0.5, 25), 5)
    assert st == 11


def expected_frequencies(fit, has, exponent):
    """Closed form of the rank roulette over one candidate set: rank 1 = lowest fitness, ties in a random
    order (each tie group shares its ranks evenly), weight r^e; second parent = the same over the set
    without the first."""
    elig = [i for i in range(len(fit)) if has[i]]

    def probs(cells):
        order = sorted(cells, key=lambda i: fit[i])
        wts = np.array([float(np.float32(r ** exponent)) for r in range(1, len(cells) + 1)])
        p, k = {}, 0
        while k < len(order):
            j = k
            while j < len(order) and fit[order[j]] == fit[order[k]]:
                j += 1
            for i in order[k:j]:
                p[i] = wts[k:j].mean() / wts.sum()
            k = j
        return p

    p1 = probs(elig)
    p2 = {i: 0.0 for i in elig}
    for j in elig:
        rest = probs([i for i in elig if i != j])
        for i, q in rest.items():
            p2[i] += p1[j] * q
    return p1, p2


@pytest.mark.parametrize("exponent", [0.0, 0.5, 1.0, 2.0])
def test_rank_roulette_frequencies(orc, exponent):
    """5x5 torus, radius 2: every cell's candidate list is the whole grid, so each of the 25 children of a
    generation is an independent draw over the same field.  1e5 draws per exponent."""
    n, cols = 25, 5
    fit = np.array([0.1, 0.5, 0.5, 0.9, 0.3, 0.3, 0.3, 0.05, 0.7, 0.2, 0.6, 0.45, 0.8, 0.15, 0.95,
                    0.35, 0.55, 0.65, 0.75, 0.85, 0.25, 0.4, 0.0, 0.12, 0.33], np.float32)
    has = np.ones(n, np.uint8)
    has[[2, 13, 21]] = 0
    w = orc.rank_weights(exponent, n)
    c1 = np.zeros(n, np.int64)
    c2 = np.zeros(n, np.int64)
    gens = 4000
    for g in range(gens):
        st, pairs = orc.plan(fit, has, n, n, cols, g, 2, w, 0xBEEF)
        assert st == 0
        c1 += np.bincount(pairs[:, 0], minlength=n)
        c2 += np.bincount(pairs[:, 1], minlength=n)
    N = gens * n
    p1, p2 = expected_frequencies(fit, has, exponent)
    for i in range(n):
        if not has[i]:
            assert c1[i] == 0 and c2[i] == 0
            continue
        for count, p in ((c1[i], p1[i]), (c2[i], p2[i])):
            assert abs(count - N * p) <= 5 * np.sqrt(N * p * (1 - p)) + 1, (exponent, i, count, N * p)
    if exponent == 0.0:  # uniform over the 22 fitted cells
        assert np.allclose(p1[0], 1 / 22)


def test_rank_weight_table_closed_form(orc):
    for e in (0.0, 0.5, 1.0, 2.0):
        w = orc.rank_weights(e, 30)
        assert w[0] == 0
        assert all(abs(w[r] - r ** e) <= 1e-6 * r ** e for r in range(1, 31))


def test_roulette_threshold_is_inclusive_and_walks_list_order(orc):
    """selection_ops.cuh:142-160: threshold = u * total; the first candidate in list order whose cumulative
    weight reaches the threshold (<=) wins."""
    fit = np.array([0.2, 0.7, 0.4, 0.4], np.float32)
    has = np.ones(4, np.uint8)
    w0 = orc.rank_weights(0.0, 4)  # every weight 1
    sel = lambda cand, u, **kw: orc.select_one(fit, has, cand, w0, 1234, np.float32(u), **kw)
    assert sel([0, 1], 0.5) == 0                   # threshold 1.0 == cumulative 1.0
    assert sel([0, 1], 0.5 + 2 ** -24) == 1
    assert sel([0, 1], 2 ** -24) == 0 and sel([0, 1], 1.0) == 1
    assert sel([3, 1, 2], 2 ** -24) == 3          # list order, not index order
    assert sel([3, 1, 2], 2 ** -24, excluded=3) == 1
    h2 = has.copy()
    h2[3] = 0
    assert orc.select_one(fit, h2, [3, 1, 2], w0, 1234, np.float32(2 ** -24)) == 1
    assert orc.select_one(fit, np.zeros(4, np.uint8), [0, 1], w0, 1, np.float32(0.5)) is None
    # e = 1: the fitter cell ranks 2 (weight 2), the other 1; u = 0.5 -> threshold 1.5
    w1 = orc.rank_weights(1.0, 4)
    assert orc.select_one(fit, has, [0, 1], w1, 7, np.float32(0.5)) == 1   # 1 < 1.5 <= 3
    assert orc.select_one(fit, has, [1, 0], w1, 7, np.float32(0.5)) == 1   # 2 >= 1.5
    # a zero total weight selects nothing (selection_ops.cuh:138-140)
    assert orc.select_one(fit, has, [0, 1], np.zeros(5, np.float32), 7, np.float32(0.5)) is None


def test_excluded_parent_leaves_the_ranking(orc):
    """The second draw ranks only the candidates other than the first parent (selection_ops.cuh:77-79):
    fitness 0.1 < 0.5 < 0.9 with the worst excluded gives ranks 1, 2 (total 3, threshold 0.4 * 3 = 1.2 ->
    the fitter one); counting the excluded cell would give ranks 2, 3 (threshold 2.0 -> the other)."""
    fit = np.array([0.1, 0.5, 0.9], np.float32)
    has = np.ones(3, np.uint8)
    w1 = orc.rank_weights(1.0, 3)
    assert orc.select_one(fit, has, [0, 1, 2], w1, 3, np.float32(0.4), excluded=0) == 2
    assert orc.select_one(fit, has, [0, 1, 2], w1, 3, np.float32(0.4)) == 1  # ranks 1,2,3: 2.4 -> cum 3 at cell 1


def splitmix64_mix(v):
    """SplitMix64 (Vigna) output function applied to state v + golden gamma."""
    z = (v + 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def test_tie_break_keys_are_splitmix64(orc):
    """Equal fitness: rank order by SplitMix64(seed ^ index) (selection_ops.cuh:32-57).  With two tied
    candidates, e = 1 and u = 0.5 the first-listed one wins exactly when its key is the larger."""
    assert splitmix64_mix(0) == 0xE220A8397B1DCDAF  # the generator's published first output from seed 0
    fit = np.full(8, 0.5, np.float32)
    has = np.ones(8, np.uint8)
    w1 = orc.rank_weights(1.0, 8)
    rng = np.random.default_rng(9)
    wins = 0
    for _ in range(300):
        seed = int(rng.integers(0, 2 ** 63)) * 2 + int(rng.integers(0, 2))
        a, b = (int(x) for x in rng.choice(8, 2, replace=False))
        got = orc.select_one(fit, has, [a, b], w1, seed, np.float32(0.5))
        expect = a if splitmix64_mix(seed ^ a) > splitmix64_mix(seed ^ b) else b
        assert got == expect
        wins += got == a
    assert 100 < wins < 200


def test_brood_plan_extends_the_plain_plan(orc):
    n, cols = 64, 8
    fit = np.random.default_rng(4).random(n).astype(np.float32)
    w = orc.rank_weights(0.5, 25)
    st, plain = orc.plan(fit, np.ones(n, np.uint8), n, n, cols, 3, 2, w, 99)
    st2, brood = orc.plan_brood(fit, np.ones(n, np.uint8), n, 4, cols, 3, 2, w, 99)
    assert st == st2 == 0
    assert (brood[:n] == plain).all()
    for c in range(4 * n):
        hood = moore_set(8, 8, c % n, 2)
        assert brood[c, 0] != brood[c, 1] and brood[c, 0] in hood and brood[c, 1] in hood


# ---- crossover --------------------------------------------------------------------------------------
A = 40


def breed_sources(orc, t, child, A_parent=A, seed=5, gen=2):
    """Child bred from parent 1 (all 1.0) and parent 2 (all 2.0) without mutation: 1 = first, 2 = second."""
    n = P + TAIL * A_parent
    child_vec, rec = orc.breed_child(np.ones(n, np.float32), np.full(n, 2.0, np.float32), A_parent, A_parent,
                                     seed, gen, child, t[0], t[1], t[2], 0.0, 0.0, 0.0)
    src = np.where(child_vec == 1.0, 1, np.where(child_vec == 2.0, 2, 0))
    assert (src > 0).all()
    return src, rec


def layer_rows(src, seg):
    fin, fout, (w0, w1), (b0, b1) = seg
    return src[w0:w1].reshape(fout, fin), src[b0:b1]


def test_zero_temperatures_copy_the_first_parent(orc):
    """slab_runtime_test.cu:414-481: temperatures 0, mutation 0 -> the child is its first parent."""
    for c in range(20):
        src, _ = breed_sources(orc, (0.0, 0.0, 0.0), c)
        assert (src == 1).all()


def test_level1_flips_every_subsystem(orc):
    """genome_ops.cuh:269-283,350-351: at t1 = 1 each subsystem's base is the second parent; nothing else
    fires at t2 = t3 = 0."""
    for c in range(20):
        src, _ = breed_sources(orc, (1.0, 0.0, 0.0), c)
        assert (src == 2).all()


def test_level2_flips_layers_and_swaps_one_tail_row(orc):
    """t2 = 1: MaybeUseAlternateParent flips every dense layer off its subsystem's base (first -> second);
    the tail keeps its base (first) except one row taken from the other parent (genome_ops.cuh:357-365)."""
    rows_seen = set()
    for c in range(200):
        src, _ = breed_sources(orc, (0.0, 1.0, 0.0), c)
        assert (src[:P] == 2).all()
        tail = src[P:].reshape(A, TAIL)
        second_rows = np.flatnonzero((tail == 2).all(axis=1))
        assert len(second_rows) == 1 and ((tail == 1).all(axis=1)).sum() == A - 1
        rows_seen.add(int(second_rows[0]))
    assert len(rows_seen) > A // 2


def check_neuron_copies(src, base, counts):
    """Each dense layer: every weight row and bias from `base`, except one neuron (its whole weight row
    and its bias) from the other parent."""
    other = 3 - base
    for li, seg in enumerate(SEGS):
        W, b = layer_rows(src, seg)
        flipped = np.flatnonzero((W == other).all(axis=1))
        assert len(flipped) == 1, li
        k = int(flipped[0])
        assert (np.delete(W, k, axis=0) == base).all() and b[k] == other and (np.delete(b, k) == base).all()
        counts[li][k] += 1


def check_splice(tail, base, row2, points):
    """The splice row: features below the point from the row's own source, the rest from the other."""
    rows = [r for r in range(A) if not ((tail[r] == tail[r][0]).all())]
    if not rows:
        return None  # spliced row 0 < point... never: point >= 1 and < 38 always mixes
    assert len(rows) == 1
    r = rows[0]
    row_src = (3 - base) if r == row2 else base
    pt = int(np.argmax(tail[r] != row_src))
    assert pt >= 1 and (tail[r][:pt] == row_src).all() and (tail[r][pt:] == 3 - row_src).all()
    points[pt] += 1
    return r


def test_level3_neuron_copy_and_splice(orc):
    """t3 = 1 (genome_ops.cuh:216-220,367-375): one random neuron per layer from the other parent; one tail
    row spliced at a point in [1, 37].  Indices uniform (chi-square, 2000 children)."""
    counts = [np.zeros(fout, np.int64) for _, fout in LAYERS]
    points = np.zeros(TAIL, np.int64)
    for c in range(2000):
        src, _ = breed_sources(orc, (0.0, 0.0, 1.0), c)
        check_neuron_copies(src, 1, counts)
        tail = src[P:].reshape(A, TAIL)
        r = check_splice(tail, 1, -1, points)
        assert r is not None
        assert ((tail == 1).all(axis=1)).sum() == A - 1
    assert points[0] == 0 and (points[1:] > 0).all()
    assert stats.chisquare(points[1:]).pvalue > 1e-4
    for cnt in counts:
        assert stats.chisquare(cnt).pvalue > 1e-4


def test_all_levels_at_temperature_one(orc):
    """t1 = t2 = t3 = 1: subsystem bases flip to the second parent, the layer flips bring every dense layer
    back to the first, the neuron copy comes from the second; the tail's base is the second parent, the
    level-2 row comes from the first, and the spliced row's prefix comes from that row's source."""
    counts = [np.zeros(fout, np.int64) for _, fout in LAYERS]
    points = np.zeros(TAIL, np.int64)
    same_row = 0
    for c in range(600):
        src, rec = breed_sources(orc, (1.0, 1.0, 1.0), c)
        check_neuron_copies(src, 1, counts)
        tail = src[P:].reshape(A, TAIL)
        first_rows = np.flatnonzero((tail == 1).all(axis=1))
        row2 = int(rec[13])
        spliced = check_splice(tail, 2, row2, points)
        if spliced == row2:
            same_row += 1
            assert len(first_rows) == 0
        else:
            assert first_rows.tolist() == [row2]
        assert (np.delete(tail, list({row2, spliced}), axis=0) == 2).all()
    assert same_row > 0
    assert (points[1:] > 0).all()


def test_fractional_temperatures_flip_at_their_rates(orc):
    """Bernoulli(t) decisions fire with frequency t: t1 = 0.3 on the subsystem bases, t2 = 0.25 on layers."""
    n = 3000
    enc2 = trunk2 = tail2 = differ = 0
    for c in range(n):
        rec = orc.xover_decisions(A, 17, 1, c, 0.3, 0.0, 0.0)
        enc2 += rec[0] == 0
        trunk2 += rec[1] == 0
        tail2 += rec[12] == 0
        differ += rec[0] != rec[1]
        assert rec[1] == rec[2] == rec[3]  # the trunk layers share their subsystem's base when t2 = 0
    # the encoder's and the trunk's level-1 draws are separate (genome_ops.cuh:269-270,282-283)
    pd = 2 * 0.3 * 0.7
    assert abs(differ - n * pd) <= 5 * np.sqrt(n * pd * (1 - pd)), differ
    flips = 0
    for c in range(n):
        rec = orc.xover_decisions(A, 18, 1, c, 0.0, 0.25, 0.0)
        flips += int((rec[:4] == 0).sum())
    for count, trials, p in ((enc2, n, 0.3), (trunk2, n, 0.3), (tail2, n, 0.3), (flips, 4 * n, 0.25)):
        assert abs(count - trials * p) <= 5 * np.sqrt(trials * p * (1 - p)), (count, trials, p)


# ---- mutation ---------------------------------------------------------------------------------------
def test_bernoulli_boundaries(orc):
    """SampleBernoulli (genome_ops.cuh:35-44): p <= 0 never, p >= 1 always, else u < p with u in (0, 1]."""
    B = orc.lib().oro_bernoulli
    u_half = (2 ** 23 - 1) << 8          # U01 = 2^23 / 2^24 = 0.5 exactly
    assert B(np.float32(0.5), u_half) == 0
    assert B(np.float32(0.5), u_half - 256) == 1
    assert B(np.float32(0.0), 0) == 0 and B(np.float32(-1.0), 0) == 0
    assert B(np.float32(1.0), 0xFFFFFFFF) == 1 and B(np.float32(2.0), 0xFFFFFFFF) == 1
    assert B(np.float32(0.999), 0xFFFFFFFF) == 0  # u = 1.0


def random_parents(seed, n):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 0.1, n).astype(np.float32), rng.normal(0, 0.1, n).astype(np.float32)


@pytest.mark.parametrize("p,sigma", [(0.0, 0.02), (0.5, 0.0), (-1.0, 0.02)])
def test_no_mutation_is_a_copy(orc, p, sigma):
    n = P + TAIL * A
    p1, p2 = random_parents(1, n)
    child, _ = orc.breed_child(p1, p2, A, A, 3, 1, 0, 0.0, 0.0, 0.0, p, sigma, 0.0)
    assert (child == p1).all()


def test_mutation_always_applies_sigma_times_a_standard_normal(orc):
    """p = 1: every value moves by sigma * z.  z (sigma = 1 on zero parents) is N(0, 1) (KS test) and is
    the same draw at sigma = 0.02 on random parents: child = fma(0.02, z, x) up to one rounding."""
    n = P + TAIL * A
    zeros = np.zeros(n, np.float32)
    z, _ = orc.breed_child(zeros, zeros, A, A, 8, 3, 11, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0)
    assert (z != 0).all()
    assert stats.kstest(z.astype(np.float64), "norm").pvalue > 1e-4
    p1, p2 = random_parents(2, n)
    child, _ = orc.breed_child(p1, p2, A, A, 8, 3, 11, 0.0, 0.0, 0.0, 1.0, 0.02, 0.0)
    expect = (np.float64(np.float32(0.02)) * z.astype(np.float64) + p1.astype(np.float64)).astype(np.float32)
    ulp = np.spacing(np.abs(expect))
    assert (np.abs(child - expect) <= ulp).all()
    assert (child == expect).mean() > 0.999
    # a different child index draws different normals
    z2, _ = orc.breed_child(zeros, zeros, A, A, 8, 3, 12, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0)
    assert (z2 != z).mean() > 0.999


@pytest.mark.parametrize("p", [0.01, 0.3])
def test_mutated_fraction_matches_probability(orc, p):
    n = P + TAIL * 512
    p1, p2 = random_parents(3, n)
    child, _ = orc.breed_child(p1, p2, 512, 512, 21, 5, 2, 0.0, 0.0, 0.0, p, 0.02, 0.0)
    k = int((child != p1).sum())
    assert abs(k - n * p) <= 5 * np.sqrt(n * p * (1 - p)), (k, n * p)


def test_tail_row_scale(orc):
    """MutateOutputTailRow's row scale (genome_ops.cuh:250-256, output_tail_ops.hpp:7-12): with probability
    1 every parent row is multiplied by 1.02 or 0.98, each about half the time; the policy is untouched."""
    n = P + TAIL * 400
    p1, p2 = random_parents(4, n)
    child, _ = orc.breed_child(p1, p2, 400, 400, 6, 1, 0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    assert (child[:P] == p1[:P]).all()
    rows, src = child[P:].reshape(400, TAIL), p1[P:].reshape(400, TAIL)
    up = (rows == src * np.float32(1.02)).all(axis=1)
    down = (rows == src * np.float32(0.98)).all(axis=1)
    assert (up ^ down).all()
    assert abs(int(up.sum()) - 200) <= 5 * 10
    child0, _ = orc.breed_child(p1, p2, 400, 400, 6, 1, 0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    assert (child0 == p1).all()


def test_new_tail_rows_are_zero(orc):
    p1, p2 = random_parents(5, P + TAIL * 30)
    child, _ = orc.breed_child(p1, p2, 30, 45, 6, 1, 0, 0.0, 0.0, 0.0, 0.01, 0.02, 0.0)
    assert child.size == P + TAIL * 45 and (child[P + TAIL * 30:] == 0).all()


# ---- bootstrap --------------------------------------------------------------------------------------
def test_bootstrap_layer_statistics(orc):
    """InitializeRandomGenome (genome_ops.cuh:58-143): weights N(0, std_l^2) with std_l from the layer's
    fan-in, biases exactly 0, tails N(0, tail_std^2).  Here std_l = 1/sqrt(fan_in) (variance 1/fan_in)."""
    Aa = 512
    stds = np.array([1.0 / np.sqrt(fin) for fin, _ in LAYERS], np.float32)
    g = orc.bootstrap_genome(Aa, 41, 0, 3, stds, 0.1)
    for (fin, fout, (w0, w1), (b0, b1)), s in zip(SEGS, stds):
        w = g[w0:w1].astype(np.float64)
        m = w.size
        assert (g[b0:b1] == 0).all()
        assert abs(w.mean()) <= 5 * s / np.sqrt(m)
        assert abs(w.std() / s - 1) <= 5 / np.sqrt(2 * m), (fin, w.std(), s)
        assert stats.kstest(w / s, "norm").pvalue > 1e-4
    tail = g[P:].astype(np.float64)
    assert tail.size == TAIL * Aa
    assert abs(tail.std() / 0.1 - 1) <= 5 / np.sqrt(2 * tail.size)
    assert abs(tail.mean()) <= 5 * 0.1 / np.sqrt(tail.size)


def test_bootstrap_he_defaults_and_determinism(orc):
    """The reference default He-normal std sqrt(2 / fan_in) (genome_ops.cuh:61-63,100-119); same seed and
    organism -> identical genome (slab_runtime_test.cu:220-290), another organism -> another genome."""
    he = orc.layer_stds(2.0)
    assert np.allclose(he, [np.sqrt(2.0 / fin) for fin, _ in LAYERS], rtol=1e-6)
    a = orc.bootstrap_genome(64, 9, 0, 5, he, 0.05)
    b = orc.bootstrap_genome(64, 9, 0, 5, he, 0.05)
    c = orc.bootstrap_genome(64, 9, 0, 6, he, 0.05)
    assert (a == b).all() and (a != c).mean() > 0.9
    w0 = SEGS[1][2]
    assert abs(a[w0[0]:w0[1]].std() / np.sqrt(2.0 / 321) - 1) < 0.03
