"""Pins the telemetry oracle (oracle/telemetry.py) and checks the library's host-side sample plan against it.
CPU only.

Pins: the reference's own convergence vectors (tests/genetic_algorithm/genetic_convergence_test.cpp), the
published SplitMix64 outputs, the C++ standard's mt19937_64 check value, numpy's inverted-CDF percentile
(the nearest-rank definition), exact sums, and closed forms of the distances.
"""
import math

import numpy as np
import pytest

import telemetry as tlo


def test_reference_convergence_vectors():
    # TestIdenticalSyntheticGenotypesHaveZeroDistance: three identical organisms
    same = np.array([1, 2, 3, 4] * 3, np.float32).reshape(3, 4)
    m = tlo.convergence_metrics(same, [(0, 1), (1, 2), (2, 0)])
    assert all(abs(float(v)) <= 1e-6 for v in m.values())
    # TestDifferentSyntheticGenotypesHavePositiveDistance: (0,0,0) and (3,4,0)
    diff = np.array([0, 0, 0, 3, 4, 0], np.float32).reshape(2, 3)
    m = tlo.convergence_metrics(diff, [(0, 1), (1, 0)])
    for k in ("centroid_distance_mean", "centroid_distance_min", "centroid_distance_max"):
        assert abs(float(m[k]) - 2.5) <= 1e-6
    for k in ("pairwise_distance_mean", "pairwise_distance_min", "pairwise_distance_max"):
        assert abs(float(m[k]) - 5.0) <= 1e-6


def test_convergence_closed_forms():
    rng = np.random.default_rng(5)
    v = rng.standard_normal((16, 64)).astype(np.float32)
    pairs = [(i, (i + 3) % 16) for i in range(16)]
    m = tlo.convergence_metrics(v, pairs)
    c = v.astype(np.float64).mean(0)
    cd = np.sqrt(((v - c) ** 2).sum(1))
    pd = np.array([np.sqrt(((v[a].astype(np.float64) - v[b]) ** 2).sum()) for a, b in pairs])
    for key, ref in (("centroid_distance_mean", cd.mean()), ("centroid_distance_min", cd.min()),
                     ("centroid_distance_max", cd.max()), ("pairwise_distance_mean", pd.mean()),
                     ("pairwise_distance_min", pd.min()), ("pairwise_distance_max", pd.max())):
        assert abs(float(m[key]) - ref) <= 1e-5 * ref, key
    # two organisms: each lies half the pair's distance from the centroid
    two = tlo.convergence_metrics(v[:2], [(0, 1)])
    assert abs(float(two["centroid_distance_mean"]) * 2 - float(two["pairwise_distance_mean"])) <= 1e-5


def test_seed_mixing_and_engine_vectors():
    # SplitMix64: the finaliser of state 0 + golden gamma, and the stream from seed 1234567
    assert tlo.splitmix64_finalize(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF
    state, out = 1234567, []
    for _ in range(5):
        state = (state + 0x9E3779B97F4A7C15) & tlo.MASK
        out.append(tlo.splitmix64_finalize(state))
    assert out == [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                   16408922859458223821]
    # [rand.predef]: the 10000th output of a default-constructed mt19937_64 (seed 5489)
    assert tlo.mt19937_64_nth(5489, 10000) == 9981545732273789042
    # the generation enters the seed: distinct generations, distinct engines
    assert len({tlo.mix_sampling_seed(7, g) for g in range(64)}) == 64


@pytest.mark.parametrize("population,scalars", [(4096, 132928 + 38 * 4739), (256, 20000), (100, 5000), (2, 8192)])
def test_sample_plan_properties(population, scalars):
    org, w, pairs = tlo.sample_plan(population, scalars, 7, 11)
    k = min(256, population)
    assert org.size == k and len(set(org.tolist())) == k and org.max() < population
    if k == population:
        assert org.tolist() == list(range(population))
    kw = min(8192, scalars)
    assert w.size == kw and len(set(w.tolist())) == kw and w.max() < scalars
    if kw == scalars:
        assert w.tolist() == list(range(scalars))
    assert pairs.shape == (512, 2) and (pairs < k).all() and (pairs[:, 0] != pairs[:, 1]).all()
    again = tlo.sample_plan(population, scalars, 7, 11)
    assert all(np.array_equal(a, b) for a, b in zip((org, w, pairs), again))
    other = tlo.sample_plan(population, scalars, 7, 12)
    assert not np.array_equal(other[2], pairs)


@pytest.mark.parametrize("population,scalars,seed,gen", [(4096, 132928 + 38 * 512, 7, 3), (192, 9000, 11, 0),
                                                         (16384, 132928 + 38 * 4739, 0xFFFFFFFF, 123456)])
def test_library_sample_plan_equals_oracle(pkg, population, scalars, seed, gen):
    lib_plan = pkg.runtime.convergence_sample_plan(population, scalars, seed, gen)
    ora_plan = tlo.sample_plan(population, scalars, seed, gen)
    for a, b in zip(lib_plan, ora_plan):
        assert np.array_equal(a, b)


def test_fitness_telemetry_against_independent_routines():
    rng = np.random.default_rng(9)
    for n in (1, 2, 7, 100, 256, 4096, 1000):
        f = np.round(rng.random(n), 3).astype(np.float32)  # ties
        row = tlo.fitness_telemetry(f)
        s = np.sort(f)
        for p, key in ((90, "fitness_p90"), (99, "fitness_p99")):
            assert row[key] == np.percentile(f, p, method="inverted_cdf"), (n, p)
        exact = math.fsum(float(x) for x in f) / n
        assert row["fitness_mean"] == np.float32(exact) or abs(float(row["fitness_mean"]) - exact) <= 1e-7 * exact
        assert abs(float(row["fitness_stddev"]) - np.std(f.astype(np.float64))) <= 1e-6
        assert row["fitness_min"] == s[0] and row["fitness_max"] == s[-1]
        assert row["distinct_fitness_count"] == np.unique(f).size
        med = np.median(f.astype(np.float64))
        assert abs(float(row["fitness_median"]) - med) <= 1e-7
    # constant population
    row = tlo.fitness_telemetry(np.full(64, 0.25, np.float32), pairs=[(0, 0), (1, 2)])
    assert row["fitness_stddev"] == 0 and row["distinct_fitness_count"] == 1
    assert (row["parent_childless_count"], row["parent_one_child_count"], row["parent_multiple_children_count"]) == (
        61, 2, 1)


def test_release_gates():
    pol = dict(min_gap_generations=3, first_new_release_generation=10, centroid_distance_threshold=6.0,
               fitness_p99_threshold=0.2)
    d = tlo.release_decision
    assert d(4739, 16, 4739, 50, [0], pol) == "none"          # catalog exhausted
    assert d(32, 0, 96, 50, [0], pol) == "none"               # no step
    assert d(32, 16, 96, 8, [0], pol) == "gated"              # foundation only: not before generation 10
    assert d(32, 16, 96, 9, [0], pol) == "measure"            # releases into generation 10
    assert d(48, 16, 96, 11, [0, 10], pol) == "gated"         # 2 generations since the last release
    assert d(48, 16, 96, 12, [0, 10], pol) == "measure"
    assert d(48, 16, 96, 12, [0, 10], pol, 0.21, 6.0) == ("release", 48, 16)
    assert d(48, 16, 96, 12, [0, 10], pol, 0.20, 6.0) == "hold"  # p99 must exceed its threshold
    assert d(48, 16, 96, 12, [0, 10], pol, 0.9, 6.01) == "hold"  # centroid distance must not exceed its own
    assert d(88, 16, 96, 40, [0, 10], pol, 0.9, 1.0) == ("release", 88, 8)  # the catalog's last words
