"""The winner artifact (CPU only: the library's host-side writer and reader).

The JSON carries the keys of the reference's writer (src/model_artifact/winner_artifact.cpp:149-182) with
format_version 2, plus "genome_layout": "FP32"; the reader follows winner_artifact_reader.cpp:203-247 and
refuses artifacts whose bytes are not this build's fp32 slot.
"""
import json
import os
import re

import numpy as np
import pytest

KEYS = ["format_version", "genome_layout", "timestamp_local", "generation_index", "best_fitness", "best_index",
        "best_slot_index", "action_count", "genome_byte_count", "seed", "action_space_path", "action_space_words"]


def _inputs(pkg, A=7, seed=3):
    rtm = pkg.runtime
    rng = np.random.default_rng(seed)
    genome = rng.standard_normal(rtm.slot_stride_bytes(A) // 4).astype(np.float32)
    letters = pkg.synth.make_catalog(A + 2, seed=seed)
    meta = rtm.winner_metadata(best_fitness=np.float32(0.123456789), generation_index=41, best_index=17,
                               best_slot_index=903, action_count=A, seed=0xBEEF, action_space_path='data/"words".txt')
    return genome, letters, meta


def test_written_artifact(pkg, tmp_path):
    rtm = pkg.runtime
    genome, letters, meta = _inputs(pkg)
    b, m = tmp_path / "w" / "win.bin", tmp_path / "w" / "win.json"
    rtm.write_winner_artifact(str(b), str(m), genome, letters, meta)
    assert b.read_bytes() == genome.tobytes()
    text = m.read_text()
    doc = json.loads(text)
    assert list(doc) == KEYS
    assert doc["format_version"] == 2 and doc["genome_layout"] == "FP32"
    assert re.fullmatch(r"\d{4}-\d\d-\d\dT\d\d:\d\d:\d\d", doc["timestamp_local"])
    assert doc["timestamp_local"] == meta.timestamp_local.decode()
    assert np.float32(doc["best_fitness"]) == np.float32(0.123456789)
    assert (doc["generation_index"], doc["best_index"], doc["best_slot_index"], doc["seed"]) == (41, 17, 903, 0xBEEF)
    assert doc["action_count"] == 7 and doc["genome_byte_count"] == rtm.slot_stride_bytes(7)
    assert doc["action_space_path"] == 'data/"words".txt'
    assert doc["action_space_words"] == ["".join(chr(65 + int(c)) for c in w) for w in letters[:7]]
    assert not list((tmp_path / "w").glob("*.tmp"))
    meta2, g2, l2 = rtm.read_winner_artifact(str(b), str(m))
    assert g2.tobytes() == genome.tobytes() and np.array_equal(l2, letters[:7])
    assert (meta2.generation_index, meta2.best_index, meta2.best_slot_index, meta2.action_count, meta2.seed) == (
        41, 17, 903, 7, 0xBEEF)
    assert meta2.best_fitness == np.float32(0.123456789)
    assert meta2.action_space_path == b'data/"words".txt' and meta2.timestamp_local == meta.timestamp_local


def test_directory_names_do_not_collide(pkg, tmp_path):
    rtm = pkg.runtime
    genome, letters, meta = _inputs(pkg)
    first = rtm.write_winner_artifact_in_directory(str(tmp_path), genome, letters, meta)
    second = rtm.write_winner_artifact_in_directory(str(tmp_path), genome, letters, meta)
    stem = os.path.basename(first[0])[:-4]
    assert re.fullmatch(r"winner-\d{4}-\d\d-\d\d_\d\d-\d\d-\d\d-g41-seed48879", stem)
    assert first[1] == first[0][:-4] + ".json" and second[0] != first[0]
    if os.path.basename(second[0]).startswith(stem):  # same second: the -2 suffix
        assert os.path.basename(second[0]) == stem + "-2.bin"
    for b, m in (first, second):
        assert rtm.read_winner_artifact(b, m)[1].tobytes() == genome.tobytes()


def test_control_bytes_in_the_path_stay_valid_json(pkg, tmp_path):
    """Control bytes below 0x20 other than \\n \\r \\t are written as \\u00XX (RFC 8259 section 7)."""
    rtm = pkg.runtime
    genome, letters, meta = _inputs(pkg)
    path = "dir\x01name\x1f/words\x07.txt\ttab"
    meta.action_space_path = path.encode()
    b, m = tmp_path / "c.bin", tmp_path / "c.json"
    rtm.write_winner_artifact(str(b), str(m), genome, letters, meta)
    text = m.read_text()
    assert "\\u0001" in text and "\\u001f" in text and "\\u0007" in text
    assert json.loads(text)["action_space_path"] == path
    assert rtm.read_winner_artifact(str(b), str(m))[0].action_space_path == path.encode()


def _rewrite(path, **changes):
    doc = json.loads(path.read_text())
    for k, v in changes.items():
        if v is None:
            doc.pop(k)
        else:
            doc[k] = v
    path.write_text(json.dumps(doc, indent=2))


@pytest.mark.parametrize("case", ["format 1", "no layout", "fp16 layout", "short binary", "count mismatch",
                                  "lower-case word", "stride of another count", "missing file"])
def test_reader_refuses(pkg, tmp_path, case):
    rtm = pkg.runtime
    genome, letters, meta = _inputs(pkg)
    b, m = tmp_path / "a.bin", tmp_path / "a.json"
    rtm.write_winner_artifact(str(b), str(m), genome, letters, meta)
    words = json.loads(m.read_text())["action_space_words"]
    status = 20
    if case == "format 1":
        _rewrite(m, format_version=1)
    elif case == "no layout":
        _rewrite(m, genome_layout=None)
    elif case == "fp16 layout":
        _rewrite(m, genome_layout="FP16")
    elif case == "short binary":
        b.write_bytes(genome.tobytes()[:-4])
    elif case == "count mismatch":
        _rewrite(m, action_space_words=words[:-1])
    elif case == "lower-case word":
        _rewrite(m, action_space_words=[words[0].lower()] + words[1:])
    elif case == "stride of another count":
        _rewrite(m, action_count=6, action_space_words=words[:6])
    else:
        os.remove(b)
        status = 21
    with pytest.raises(RuntimeError) as err:
        rtm.read_winner_artifact(str(b), str(m))
    assert err.value.status == status


def test_writer_refuses_a_genome_of_another_stride(pkg, tmp_path):
    rtm = pkg.runtime
    genome, letters, meta = _inputs(pkg)
    meta.genome_byte_count += 256
    with pytest.raises(RuntimeError) as err:
        rtm.write_winner_artifact(str(tmp_path / "x.bin"), str(tmp_path / "x.json"), genome, letters, meta)
    assert err.value.status == 17
    assert not os.path.exists(tmp_path / "x.bin")
