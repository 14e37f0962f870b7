"""Synthetic, seeded inputs of the GA workloads: Zipf(1.1) word catalogs, sampled secrets and the
per-cell CSR of local training secrets (block shards or ragged per-cell lists)."""
import numpy as np

CATALOG_SEED = 0xC0FFEE01
SECRETS_SEED = 0xC0FFEE02
RAGGED_SEED = 0xC0FFEE03


def zipf_letter_probabilities(s=1.1, letters=26):
    ranks = np.arange(1, letters + 1, dtype=np.float64)
    w = ranks ** (-s)
    return w / w.sum()


def make_catalog(count, seed=CATALOG_SEED, s=1.1, distinct_letters=False):
    """`count` unique 5-letter words, letters i.i.d. Zipf(s) over ranks 1..26 (rank k -> letter k-1),
    duplicates rejected in draw order (and, with `distinct_letters`, words repeating a letter, which
    always have the hint grids output-embedding injection needs).  Returns uint8 [count, 5]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = zipf_letter_probabilities(s)
    seen = set()
    out = []
    while len(out) < count:
        batch = rng.choice(26, size=(max(256, 2 * (count - len(out))), 5), p=p).astype(np.uint8)
        for w in batch:
            key = w.tobytes()
            if key in seen or (distinct_letters and len(set(key)) < 5):
                continue
            seen.add(key)
            out.append(w)
            if len(out) == count:
                break
    return np.stack(out).astype(np.uint8)


def sample_secrets(catalog_size, count, seed=SECRETS_SEED):
    """`count` distinct catalog indices sampled without replacement."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.choice(catalog_size, size=count, replace=False).astype(np.uint32)


def block_csr(rows, cols, block, secrets):
    """One shard per block x block square of cells: block b (row-major over blocks) owns the next
    len(secrets)/blocks secrets; every cell of the block plays them in block order."""
    br, bc = rows // block, cols // block
    nblocks = br * bc
    per = len(secrets) // nblocks
    offsets = np.zeros(rows * cols + 1, np.uint32)
    ids = []
    for cell in range(rows * cols):
        r, c = divmod(cell, cols)
        b = (r // block) * bc + (c // block)
        ids.append(secrets[b * per:(b + 1) * per])
        offsets[cell + 1] = offsets[cell] + per
    return offsets, np.concatenate(ids).astype(np.uint32)


def ragged_csr(cells, secret_pool, low=32, high=1024, seed=RAGGED_SEED):
    """Per-cell shard sizes W_c ~ U{low..high}; cell c plays a contiguous (cyclic) window of the pool."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = rng.integers(low, high + 1, size=cells)
    starts = rng.integers(0, len(secret_pool), size=cells)
    offsets = np.zeros(cells + 1, np.uint32)
    ids = []
    for c in range(cells):
        idx = (starts[c] + np.arange(sizes[c])) % len(secret_pool)
        ids.append(secret_pool[idx])
        offsets[c + 1] = offsets[c] + sizes[c]
    return offsets, np.concatenate(ids).astype(np.uint32)


WORKLOADS = {
    # name: grid side, actions, secrets, block side
    "grid16_a512": dict(side=16, actions=512, secrets=256, block=4),
    "grid64_a4739": dict(side=64, actions=4739, secrets=2048, block=8),
    "grid128_a4739": dict(side=128, actions=4739, secrets=8192, block=8),
}


def workload(name):
    w = WORKLOADS[name]
    catalog = make_catalog(w["actions"])
    if w["secrets"] <= w["actions"]:
        secrets = sample_secrets(w["actions"], w["secrets"])
    else:
        # more secrets than words: each shard draws its own distinct secrets, shards may share words
        blocks = (w["side"] // w["block"]) ** 2
        per = w["secrets"] // blocks
        rng = np.random.Generator(np.random.PCG64(SECRETS_SEED))
        secrets = np.concatenate([rng.choice(w["actions"], size=per, replace=False) for _ in range(blocks)])
        secrets = secrets.astype(np.uint32)
    offsets, ids = block_csr(w["side"], w["side"], w["block"], secrets)
    return dict(w, catalog=catalog, secrets=secrets, offsets=offsets, ids=ids)
