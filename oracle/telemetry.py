"""CPU restatement of the reference driver's generation telemetry and adaptive shard release (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.  Follows src/cli/run_genetic_algorithm.cu
(TryCollectGenerationFitnessTelemetry :1080-1186, TryMaybeReleaseTrainingDataShard :1200-1286) and
src/genetic_algorithm/genetic_convergence.hpp (MixSamplingSeed :48-57, MakeSamplePlan :87-117,
TryComputeMetrics :119-192).  Sums the reference accumulates one value at a time are accumulated one value at a
time here too (np.cumsum), in the same precision.
"""
import math

import numpy as np

import oracle as orc  # the oracle directory is on sys.path (tests/conftest.py)

MASK = (1 << 64) - 1
SAMPLE_ORGANISMS, SAMPLE_WEIGHTS, SAMPLE_PAIRS = 256, 8192, 512


def fitness_telemetry(fitness, has=None, pairs=None):
    """The fitness row: order statistics, sequential double sums, distinct count, breeding metrics."""
    f = np.asarray(fitness, np.float32)
    n = f.size
    if n == 0 or (has is not None and not np.all(np.asarray(has))):
        raise ValueError("every organism needs a fitness")
    s = np.sort(f, kind="stable")
    total = np.cumsum(s.astype(np.float64))[-1]
    mean = total / n
    squares = np.cumsum((s.astype(np.float64) - mean) ** 2)[-1]
    m = n // 2
    median = np.float32((s[m - 1] + s[m]) * np.float32(0.5)) if n % 2 == 0 else s[m]

    def rank(p):
        r = math.ceil((p / 100.0) * n)
        i = 0 if r == 0 else r - 1
        return s[min(i, n - 1)]

    row = dict(population_size=n, fitness_min=s[0], fitness_mean=np.float32(mean), fitness_median=median,
               fitness_p90=rank(90.0), fitness_p99=rank(99.0), fitness_max=s[-1],
               fitness_stddev=np.float32(math.sqrt(squares / n)),
               distinct_fitness_count=1 + int(np.count_nonzero(s[1:] != s[:-1])))
    if pairs is not None:
        counts = np.zeros(n, np.int64)
        for a, b in np.asarray(pairs).reshape(-1, 2):
            counts[a] += 1
            counts[b] += 1
        row.update(parent_childless_count=int(np.sum(counts == 0)), parent_one_child_count=int(np.sum(counts == 1)),
                   parent_multiple_children_count=int(np.sum(counts >= 2)))
    return row


def splitmix64_finalize(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def mix_sampling_seed(run_seed, generation):
    v = 0x9E3779B97F4A7C15 ^ run_seed
    v ^= (generation + 0xBF58476D1CE4E5B9 + ((v << 6) & MASK) + (v >> 2)) & MASK
    return splitmix64_finalize(v)


def mt19937_64_nth(seed, n):
    return int(orc.lib().oro_mt19937_64_nth(seed, n))


def sample_plan(population, scalars, run_seed, generation, organisms=SAMPLE_ORGANISMS, weights=SAMPLE_WEIGHTS,
                pairs=SAMPLE_PAIRS):
    """(organism indices, scalar indices, [pairs][2]) of MakeSamplePlan."""
    org = np.zeros(min(organisms, population), np.uint32)
    w = np.zeros(min(weights, scalars), np.uint32)
    pr = np.zeros((pairs, 2), np.uint32)
    counts = np.zeros(3, np.uint64)
    orc.lib().oro_sample_plan(population, scalars, mix_sampling_seed(run_seed, generation), organisms, weights, pairs,
                              orc.ptr(org), orc.ptr(w), orc.ptr(pr), orc.ptr(counts))
    return org, w, pr[:int(counts[2])]


def convergence_metrics(values, pairs):
    """TryComputeMetrics on values [organisms][weights] (float32): centroid and pairwise L2 distances."""
    v = np.asarray(values, np.float32)
    n_org = v.shape[0]
    centroid = np.zeros(v.shape[1], np.float32)
    for o in range(n_org):  # organism order, float adds
        centroid = centroid + v[o]
    inverse = np.float32(1.0) / np.float32(n_org)
    centroid = centroid * inverse

    def norm(d):  # weight order, float adds, then a correctly rounded float square root
        return np.sqrt(np.cumsum(d * d, dtype=np.float32)[-1])

    dist = [norm(v[o] - centroid) for o in range(n_org)]
    out = {}
    total = np.float32(0.0)
    for d in dist:
        total = np.float32(total + d)
    out["centroid_distance_mean"] = np.float32(total * inverse)
    out["centroid_distance_min"] = min(dist)
    out["centroid_distance_max"] = max(max(dist), np.float32(0.0))
    pd = [norm(v[a] - v[b]) for a, b in np.asarray(pairs).reshape(-1, 2)]
    if not pd:
        out.update(pairwise_distance_mean=np.float32(0), pairwise_distance_min=np.float32(0),
                   pairwise_distance_max=np.float32(0))
        return out
    total = np.float32(0.0)
    for d in pd:
        total = np.float32(total + d)
    out["pairwise_distance_mean"] = np.float32(total / np.float32(len(pd)))
    out["pairwise_distance_min"] = min(pd)
    out["pairwise_distance_max"] = max(max(pd), np.float32(0.0))
    return out


def release_decision(action_words, step, catalog_words, generation, history, policy, fitness_p99=None,
                     centroid_mean=None):
    """TryMaybeReleaseTrainingDataShard's gates.  Returns "none" (no release possible: catalog exhausted or no
    step), "gated" (too early), "measure" (telemetry needed: when fitness_p99 / centroid_mean are None) or
    ("release", first, count) / "hold"."""
    if not history:
        raise ValueError("a foundation release record is required")
    if action_words >= catalog_words or step == 0:
        return "none"
    release_generation = generation + 1
    previous = history[-1]
    since = release_generation - previous if release_generation >= previous else 0
    if len(history) == 1 and release_generation < policy["first_new_release_generation"]:
        return "gated"
    if since < policy["min_gap_generations"]:
        return "gated"
    if fitness_p99 is None or centroid_mean is None:
        return "measure"
    if centroid_mean <= policy["centroid_distance_threshold"] and fitness_p99 > policy["fitness_p99_threshold"]:
        return ("release", action_words, min(catalog_words - action_words, step))
    return "hold"
