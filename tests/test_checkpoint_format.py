"""The runtime-checkpoint file format (CPU only; no device calls).

The oracle's restatement (oracle/checkpoint.py) is pinned by the published FNV-1a 64 vectors and by byte
images of the configuration structs built from the reference's own headers (oracle/_ref); the library's
reader, writer, checksum and identity hash are then checked against the oracle on synthetic files,
including files whose struct padding holds arbitrary bytes, and on corrupt files both must reject.
"""
import os
import struct

import numpy as np
import pytest

import checkpoint as cko

# FNV-1a 64 test vectors published with the algorithm (offset basis 0xcbf29ce484222325)
FNV_VECTORS = [(b"", 0xCBF29CE484222325), (b"a", 0xAF63DC4C8601EC8C), (b"foobar", 0x85944171F73967E8)]


def test_fnv1a_published_vectors():
    for data, expect in FNV_VECTORS:
        assert cko.fnv1a64(data, cko.FNV_OFFSET_BASIS) == expect
    # the reference's start value is the published basis with its last decimal digit dropped
    assert cko.REFERENCE_FNV_START == cko.FNV_OFFSET_BASIS // 10
    assert cko.fnv1a64(b"") == cko.REFERENCE_FNV_START


def test_struct_images_match_reference_headers(orc):
    R = orc.ref()
    if R is None:
        pytest.skip("reference headers not present (oracle/_ref not built)")
    sizes = np.zeros(5, np.uint64)
    R.ref_checkpoint_sizes(orc.ptr(sizes))
    assert sizes.tolist() == [5, 8, 4, 8, cko.RELEASE_CAPACITY]
    v = np.array([40, 48, 32, 8, 5, 1, 2], np.uint64)
    img = np.zeros(64, np.uint8)
    assert R.ref_word_counts_image(orc.ptr(v), orc.ptr(img)) == cko.WORD_COUNTS_BYTES
    assert img[:56].tobytes() == cko.pack_word_counts(*v.tolist())
    f = np.array([0.75, 0.125, 0.0625, 0.03125, 3e-4, 0.05, 0.25], np.float32)
    assert R.ref_assembly_image(3, orc.ptr(f), orc.ptr(img)) == cko.ASSEMBLY_BYTES
    assert img[:40].tobytes() == cko.pack_assembly(3, *f.tolist())
    radius = np.zeros(1, np.uint64)
    d = np.zeros(7, np.float32)
    R.ref_assembly_defaults(orc.ptr(radius), orc.ptr(d))
    assert int(radius[0]) == 2 and d.tolist() == np.float32([0.5, 0.02, 0.01, 0.005, 1e-4, 0.02, 0.0]).tolist()
    assert R.ref_grid_shape_image(6, 7, orc.ptr(img)) == cko.GRID_SHAPE_BYTES
    assert img[:24].tobytes() == cko.pack_grid_shape(6, 7)


A, STRIDE, COLS = 48, 512, 4


def _synthetic(seed=1, n=12, children=8, releases=3, holes=(3,)):
    """A checkpoint dict with arbitrary padding bytes in the two padded struct images."""
    rng = np.random.default_rng(seed)
    slots_total = 40
    slots = rng.permutation(slots_total)[:n].astype(np.uint32)
    for h in holes:
        slots[h] = 0xFFFFFFFF  # an organism without a slot has no genome record
    c = {
        "schema_version": 3, "genome_layout_version": cko.GENOME_LAYOUT_VERSION, "resume_phase": 1,
        "generation_seed": 0xDEADBEEF,
        "word_counts": cko.pack_word_counts(40, A, 32, 8, 5, 1, 2),
        "release_generations": list(range(5, 5 + 5 * releases, 5)),
        "release_p99": np.float32(rng.random(releases)).tolist(),
        "assembly": cko.pack_assembly(2, 0.5, 0.02, 0.01, 0.005, 1e-4, 0.02, 0.0, padding=b"\xab\xcd\xef\x01"),
        "injection": cko.pack_injection(True, A - 16, 16, padding=bytes(range(1, 8))),
        "runtime_config": struct.pack("<6Q", slots_total * STRIDE, n * STRIDE, 0, A, n, COLS),
        "slab_layout": struct.pack("<4Q", A, STRIDE, slots_total, slots_total * STRIDE),
        "shapes": [cko.pack_grid_shape(n // COLS, COLS), cko.pack_grid_shape(children // COLS, COLS),
                   cko.pack_grid_shape(n // COLS, COLS)],
        "generation_index": 7,
        "slots": slots.tolist(),
        "fitness": np.float32(rng.random(n)).tolist(),
        "evals": rng.integers(0, 3, n).tolist(),
        "has": rng.integers(0, 2, n).tolist(),
        "pairs": [tuple(int(x) for x in rng.integers(0, n, 2)) for _ in range(children)],
        "records": [(i, rng.integers(0, 256, STRIDE, dtype=np.uint8).tobytes()) for i in range(n) if i not in holes],
    }
    c["identity_hash"] = cko.identity_hash(c["word_counts"], (c["release_generations"], c["release_p99"]))
    c["checksum"] = cko.checksum(c)
    return c


def test_oracle_round_trip():
    c = _synthetic()
    assert cko.read(cko.write(c)) == c


def test_library_reads_the_oracle_file_and_writes_it_back(pkg, tmp_path):
    rtm = pkg.runtime
    c = _synthetic()
    data = cko.write(c)
    (tmp_path / "in.ckpt").write_bytes(data)
    ck = rtm.Checkpoint.read(str(tmp_path / "in.ckpt"))
    i = ck.info()
    assert (i.schema_version, i.genome_layout_version, i.resume_phase, i.generation_seed) == (
        3, cko.GENOME_LAYOUT_VERSION, 1, 0xDEADBEEF)
    assert i.checksum == c["checksum"] and i.training_data_identity_hash == c["identity_hash"]
    assert (i.generation_index, i.population, i.child_count, i.live_genotypes) == (7, 12, 8, 11)
    assert (i.action_count, i.slot_stride_bytes, i.slot_count, i.slab_bytes) == (A, STRIDE, 40, 40 * STRIDE)
    assert (i.generation_byte_budget, i.grid_column_count, i.release_count) == (12 * STRIDE, COLS, 3)
    w = i.word_counts
    assert (w.training_word_count, w.action_space_word_count, w.training_word_schedule.initial_word_count,
            w.training_word_schedule.word_count_step, w.training_word_schedule.word_count_step_period_generations,
            w.shard_initial_radius, w.shard_radius_growth_period_generations) == (40, A, 32, 8, 5, 1, 2)
    a = i.assembly_config
    got = (a.cellular_breeding_radius, a.rank_exponent, a.crossover_temperature_level1,
           a.crossover_temperature_level2, a.crossover_temperature_level3, a.mutation_probability,
           a.mutation_sigma, a.output_tail_row_scale_mutation_probability)
    assert got == cko.unpack_assembly(c["assembly"])
    assert (i.injection.enabled, i.injection.first_catalog_word_index, i.injection.injection_count) == (1, A - 16, 16)
    # CheckpointPayloadByteCount: plan pairs, history, per-organism rows, genome bytes
    assert i.payload_bytes == 8 * 8 + 8 + 3 * 12 + 12 * 13 + 11 * STRIDE
    slots, fit, evals, has = ck.generation()
    assert slots.tolist() == c["slots"] and fit.tolist() == c["fitness"]
    assert evals.tolist() == c["evals"] and has.tolist() == c["has"]
    assert [tuple(p) for p in ck.plan().tolist()] == c["pairs"]
    for r, (org, genome) in enumerate(c["records"]):
        o, g = ck.genome(r)
        assert o == org and g.tobytes() == genome
    ck.write(str(tmp_path / "sub" / "dir" / "out.ckpt"))
    assert (tmp_path / "sub" / "dir" / "out.ckpt").read_bytes() == data
    assert not (tmp_path / "sub" / "dir" / "out.ckpt.tmp").exists()
    # a clone writes the same bytes; writing over an existing file replaces it
    cl = ck.clone()
    ck.close()
    cl.write(str(tmp_path / "sub" / "dir" / "out.ckpt"))
    assert (tmp_path / "sub" / "dir" / "out.ckpt").read_bytes() == data
    cl.close()


def _corruptions():
    def field(c, **kw):
        c = dict(c)
        c.update(kw)
        c["checksum"] = cko.checksum(c)
        return cko.write(c)

    base = _synthetic()
    good = cko.write(base)
    flip = bytearray(good)
    flip[-5] ^= 0x40  # a genome byte
    magic = bytearray(good)
    magic[0] ^= 1
    too_many = _synthetic(releases=cko.RELEASE_CAPACITY + 1)
    bad_size = dict(base)
    bad_size["records"] = [(o, g[:-4]) for o, g in base["records"]]
    bad_size["checksum"] = cko.checksum(bad_size)
    return {
        "flipped genome byte": bytes(flip),
        "magic": bytes(magic),
        "truncated": good[:-100],
        "trailing header only": good[:40],
        "schema 2": field(base, schema_version=2),
        "the reference's fp16 genome layout": field(base, genome_layout_version=cko.REFERENCE_GENOME_LAYOUT_VERSION),
        "release history over capacity": cko.write(too_many),
        "empty generation": field(base, slots=[], fitness=[], evals=[], has=[], records=[]),
        "empty plan": field(base, pairs=[]),
        "genome size differs from the stride": cko.write(bad_size),
        "more records than organisms": field(base, records=base["records"] * 2),
    }


@pytest.mark.parametrize("case", sorted(_corruptions()))
def test_corrupt_files_are_rejected_by_both(pkg, tmp_path, case):
    data = _corruptions()[case]
    with pytest.raises(cko.Invalid):
        cko.read(data)
    path = tmp_path / "bad.ckpt"
    path.write_bytes(data)
    with pytest.raises(RuntimeError) as err:
        pkg.runtime.Checkpoint.read(str(path))
    assert err.value.status == 20


def test_missing_file_is_an_io_failure(pkg, tmp_path):
    with pytest.raises(RuntimeError) as err:
        pkg.runtime.Checkpoint.read(str(tmp_path / "absent.ckpt"))
    assert err.value.status == 21


def test_identity_hash_matches_oracle(pkg):
    rtm = pkg.runtime
    wc = rtm.word_counts(40, 48, initial=32, step=8, period=5, radius0=1, growth=2)
    image = cko.pack_word_counts(40, 48, 32, 8, 5, 1, 2)
    assert rtm.training_data_identity_hash(wc) == cko.identity_hash(image)
    gens, p99 = [5, 10, 15], np.float32([0.25, 0.5, 0.75])
    assert rtm.training_data_identity_hash(wc, gens, p99) == cko.identity_hash(image, (gens, p99.tolist()))
    assert rtm.training_data_identity_hash(wc, gens) == cko.identity_hash(image, (gens, [0.0] * 3))
    # an empty history is mixed (its count), an absent one is not
    assert rtm.training_data_identity_hash(wc, []) == cko.identity_hash(image, ([], []))
    assert rtm.training_data_identity_hash(wc, []) != rtm.training_data_identity_hash(wc)
    cat = pkg.synth.make_catalog(48)
    assert rtm.training_data_identity_hash(wc, gens, p99, catalog=cat) == cko.identity_hash(
        image, (gens, p99.tolist()), cat.reshape(-1, 5))


def test_status_names_cover_the_checkpoint_codes(pkg):
    names = pkg.runtime.STATUS_NAMES
    assert names[20] == "checkpoint invalid" and names[21] == "checkpoint I/O failure"
    assert os.path.exists(pkg.runtime.LIB_PATH)


@pytest.mark.parametrize("stride", [(1 << 64) - 8, (1 << 64) - 12, 0, 1 << 40])
def test_corrupt_stride_is_rejected_without_sizing_from_it(pkg, tmp_path, stride):
    """The slot stride is read from the file; the library must refuse one that cannot describe the genome
    records that follow before sizing anything from it (a stride near 2^64 would wrap `stride + 12` and
    `live * stride`, zero would divide by zero, 2^40 is larger than the file)."""
    c = _synthetic()
    A_, _, slots_total, _ = cko.slab_layout(c)
    c["slab_layout"] = struct.pack("<4Q", A_, stride, slots_total, (slots_total * stride) % (1 << 64))
    c["checksum"] = cko.checksum(c)
    path = tmp_path / "stride.ckpt"
    path.write_bytes(cko.write(c))
    with pytest.raises(RuntimeError) as err:
        pkg.runtime.Checkpoint.read(str(path))
    assert err.value.status == 20
