"""Writes tests/golden/reference_integer_vectors.npz from the reference's own integer-only headers
(wordle feedback, hint grids, toroidal Moore neighbourhoods, shard schedule), compiled under
oracle/_ref by oracle/build_oracle.py.  Run in a container where /root/reference exists:

    python oracle/build_oracle.py && python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def words_for_feedback():
    import __graft_entry__ as ge
    pkg = ge.import_package()
    cat = pkg.synth.make_catalog(160, seed=0x5EED01)
    extra = [oracle.word(w) for w in ("CRANE", "SOLAR", "EERIE", "STEEL", "HELLO", "LEVEL", "AAAAA", "ABABA",
                                      "SPARE", "AABCD", "LLAMA", "ALLAY")]
    return np.concatenate([cat, np.stack(extra)]).astype(np.uint8)


def main():
    ref = oracle.ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libneref.so is missing (needs /root/reference)")
    words = words_for_feedback()
    n = len(words)
    fb = np.zeros((n, n, 5), np.uint8)
    for i in range(n):
        for j in range(n):
            ok = ref.ref_feedback(oracle.ptr(words[i]), oracle.ptr(words[j]), oracle.ptr(fb[i, j]))
            assert ok == 1
    hint_ok = np.zeros(n, np.uint8)
    hint_tc = np.zeros((n, 3), np.uint32)
    hint_letters = np.zeros((n, 3, 5, 5), np.uint8)
    hint_tiles = np.zeros((n, 3, 5, 5), np.uint8)
    for i in range(n):
        hint_ok[i] = ref.ref_hint_group(oracle.ptr(words[i]), oracle.ptr(hint_tc[i]), oracle.ptr(hint_letters[i]),
                                        oracle.ptr(hint_tiles[i]))
    moore = {}
    for rows, cols in ((3, 3), (5, 5), (4, 7), (16, 16), (1, 6)):
        for radius in (1, 2):
            out = np.zeros((rows * cols, 25), np.int64) - 1
            counts = np.zeros(rows * cols, np.int32)
            for c in range(rows * cols):
                buf = np.zeros(25, np.uint64)
                k = ref.ref_moore(rows, cols, c, radius, oracle.ptr(buf))
                counts[c] = k
                out[c, :max(k, 0)] = buf[:max(k, 0)]
            moore[f"moore_{rows}x{cols}_r{radius}"] = out
            moore[f"moore_{rows}x{cols}_r{radius}_count"] = counts
    shard_cases = [
        # initial, step, period, introduced, generation, rows, cols, epi_rows, radius0, growth
        (20, 0, 1, 20, 0, 8, 8, 8, 0, 2),
        (20, 10, 2, 50, 0, 8, 8, 8, 0, 2),
        (20, 10, 2, 50, 5, 8, 8, 8, 0, 2),
        (20, 10, 2, 50, 9, 6, 8, 8, 1, 3),
        (16, 8, 1, 64, 4, 16, 16, 16, 0, 2),
        (20, 10, 1, 30, 25, 50, 50, 50, 0, 2),
    ]
    shards = {}
    for k, case in enumerate(shard_cases):
        cells = case[5] * case[6]
        offsets = np.zeros(cells + 1, np.uint32)
        ids = np.zeros(cells * 64, np.uint32)
        st = ref.ref_shard_csr(*case, None, 0, oracle.ptr(offsets), oracle.ptr(ids), ids.size)
        assert st > 0, (case, st)
        shards[f"shard_case{k}"] = np.array(case, np.uint64)
        shards[f"shard_offsets{k}"] = offsets
        shards[f"shard_ids{k}"] = ids[:offsets[-1]]
    np.savez_compressed(os.path.join(HERE, "reference_integer_vectors.npz"), words=words, feedback=fb,
                        hint_ok=hint_ok, hint_tc=hint_tc, hint_letters=hint_letters, hint_tiles=hint_tiles,
                        **moore, **shards)
    print("wrote", os.path.join(HERE, "reference_integer_vectors.npz"))


if __name__ == "__main__":
    main()
