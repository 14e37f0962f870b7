"""Checkpoint and resume on the GPU.

Bar: a run checkpointed at the pre-recombination boundary, written, read back, restored into a new
runtime and resumed is bit-identical to the uninterrupted run (every genome of every later generation,
compared by organism, and every fitness); so is an in-process create + resume.  The file holds exactly
the runtime's state (checked with the oracle's reader, oracle/checkpoint.py); the restored runtime
places the k-th genome record in slot k (TryAllocateSlabSlot on a fresh slab).
"""
import struct

import numpy as np
import pytest

import checkpoint as cko

pytestmark = pytest.mark.gpu


def _runtime(pkg, wl, side, A, seed=9, slack=3):
    rtm = pkg.runtime
    stride = rtm.slot_stride_bytes(A)
    n = side * side
    rt = pkg.SlabRuntime(slab_bytes=slack * n * stride, generation_bytes=n * stride, action_count=A,
                         grid_columns=side)
    rt.upload_catalog(wl["catalog"])
    rt.upload_secrets_csr(wl["offsets"], wl["ids"])
    rt.bootstrap(n, seed=seed, variance_scale=1.0, tail_std=0.1)
    return rt


def _state(rt):
    """Genomes (uint32 views) and fitness of the current generation, by organism."""
    slots, fit, _, _ = rt.generation(0)
    return [rt.download_slot(s).view(np.uint32).copy() for s in slots], fit.copy()


def _same_state(a, b):
    ga, fa = a
    gb, fb = b
    assert len(ga) == len(gb)
    for i, (x, y) in enumerate(zip(ga, gb)):
        assert np.array_equal(x, y), f"organism {i}"
    assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))


@pytest.fixture(scope="module")
def grid16(pkg):
    return pkg.synth.workload("grid16_a512")


CFG = dict(mutation_probability=0.01, mutation_sigma=0.02)
SEED = lambda g: 7 + 2 + g  # noqa: E731  (the reference CLI's generation seeds)


def _run(pkg, wl, gens, stop=None, mode=None, tmp_path=None):
    """Advances `gens` generations; at generation `stop` checkpoints instead: mode "inplace" resumes the
    same runtime, mode "file" writes, reads, restores into a new runtime and resumes there.  Returns the
    final state (after one more evaluation) and the summaries of the generations after the checkpoint."""
    rtm = pkg.runtime
    counts = rtm.word_counts(512, 512)
    cfg = rtm.default_assembly_config(**CFG)
    rt = _runtime(pkg, wl, 16, 512)
    sums = []
    for g in range(gens):
        if g == stop:
            ck = rt.create_checkpoint(SEED(g), counts, cfg, catalog=wl["catalog"])
            if mode == "inplace":
                rt.resume(ck)
            else:
                ck.write(str(tmp_path / "gen.ckpt"))
                ck.close()
                rt.close()
                ck = rtm.Checkpoint.read(str(tmp_path / "gen.ckpt"))
                rt = pkg.SlabRuntime.restore(ck)
                slots, _, _, _ = rt.generation(0)
                assert slots.tolist() == list(range(ck.info().live_genotypes))
                rt.upload_catalog(wl["catalog"])
                rt.upload_secrets_csr(wl["offsets"], wl["ids"])
                rt.resume(ck)
            ck.close()
        else:
            rt.advance(SEED(g), counts, cfg)
            if stop is not None and g > stop:
                s = rt.summary()
                sums.append((s.best_fitness, s.average_fitness, s.best_index, s.generation_index))
    rt.evaluate(counts)
    out = _state(rt), sums
    rt.close()
    return out


def test_resumed_runs_equal_the_uninterrupted_run(pkg, grid16, tmp_path):
    ref, ref_sums = _run(pkg, grid16, 5, stop=None)
    ref_stopped, sums_ref = _run(pkg, grid16, 5, stop=2, mode="inplace")
    _same_state(ref, ref_stopped)
    restored, sums_restored = _run(pkg, grid16, 5, stop=2, mode="file", tmp_path=tmp_path)
    _same_state(ref, restored)
    assert sums_ref == sums_restored and len(sums_ref) == 2


def test_checkpoint_holds_the_runtime_state(pkg, grid16, tmp_path):
    rtm = pkg.runtime
    counts = rtm.word_counts(512, 512)
    cfg = rtm.default_assembly_config(**CFG)
    rt = _runtime(pkg, grid16, 16, 512)
    try:
        rt.advance(SEED(0), counts, cfg)
        ck = rt.create_checkpoint(SEED(1), counts, cfg, catalog=grid16["catalog"])
        ck.write(str(tmp_path / "a.ckpt"))
        c = cko.read((tmp_path / "a.ckpt").read_bytes())  # validates magic, versions and checksum
        slots, fit, evals, has = rt.generation(0)
        assert c["slots"] == slots.tolist() and c["evals"] == evals.tolist() and c["has"] == has.tolist()
        assert np.array_equal(np.float32(c["fitness"]).view(np.uint32), fit.view(np.uint32))
        assert has.all() and (evals == 1).all()
        assert [tuple(p) for p in rt.plan().tolist()] == c["pairs"]
        assert [o for o, _ in c["records"]] == list(range(256))
        for org, genome in c["records"]:
            assert genome == rt.download_slot(int(slots[org])).tobytes(), org
        info = rt.info()
        assert cko.slab_layout(c) == (512, info.slot_stride_bytes, info.slot_count, info.slab_bytes)
        assert info.slab_bytes == info.slot_count * info.slot_stride_bytes
        assert cko.runtime_config(c) == (info.slab_bytes, 256 * info.slot_stride_bytes, 0, 512, 256, 16)
        assert cko.grid_shapes(c) == [(16, 16, 256)] * 3
        assert c["generation_index"] == 1 and c["generation_seed"] == SEED(1)
        assert c["assembly"] == cko.pack_assembly(2, *np.float32([0.5, 0.02, 0.01, 0.005, 0.01, 0.02, 0.0]).tolist())
        assert c["injection"] == cko.pack_injection(False, 0, 0)
        wc = cko.pack_word_counts(512, 512, 512, 0, 1, 0, 2)
        assert c["word_counts"] == wc
        assert c["identity_hash"] == cko.identity_hash(wc, None, grid16["catalog"].reshape(-1, 5))
        i = ck.info()
        assert i.checksum == c["checksum"] and i.payload_bytes == (256 * 8 + 8 + 256 * 13 +
                                                                    256 * info.slot_stride_bytes)
        # the runtime stands at the boundary: a second resume of another generation's checkpoint fails
        rt.resume(ck)
        with pytest.raises(RuntimeError) as err:
            rt.resume(ck)
        assert err.value.status == 4
        ck.close()
    finally:
        rt.close()


def _catalog_with_injectable_tail(pkg, orc, n, first):
    cat = pkg.synth.make_catalog(4 * n, seed=0x1234)
    ok = np.array([orc.hint_group(w)[0] == 1 for w in cat])
    rest = [w for w, good in zip(cat[first:], ok[first:]) if good]
    return np.concatenate([cat[:first], np.stack(rest[:n - first])]).astype(np.uint8)


def test_checkpoint_across_a_widening_step(pkg, orc, tmp_path):
    """The boundary of a growth generation lies after the widening repack: the checkpoint carries the
    wider layout and the pending injection, and the resumed step injects the new rows."""
    rtm = pkg.runtime
    A0, inj_n, side = 32, 16, 8
    n = side * side
    cat = _catalog_with_injectable_tail(pkg, orc, 64, A0)
    sec = pkg.synth.sample_secrets(A0, 32, seed=77)
    off, ids = pkg.synth.block_csr(side, side, 4, sec)
    s1 = rtm.slot_stride_bytes(A0 + inj_n)
    wl = {"catalog": cat, "offsets": off, "ids": ids}
    cfg = rtm.default_assembly_config(crossover_temperature_level1=0.1, crossover_temperature_level2=0.1,
                                      crossover_temperature_level3=0.2, mutation_probability=0.02)
    inj = rtm.PendingInjection(enabled=1, reserved0=0, first_catalog_word_index=A0, injection_count=inj_n)
    counts0, counts1 = rtm.word_counts(A0, A0), rtm.word_counts(A0, A0 + inj_n)

    def make():
        rt = pkg.SlabRuntime(slab_bytes=2 * n * s1 + 3 * s1, generation_bytes=n * s1, action_count=A0,
                             grid_columns=side)
        rt.upload_catalog(cat)
        rt.upload_secrets_csr(off, ids)
        rt.bootstrap(n, seed=21, variance_scale=1.0, tail_std=0.1)
        return rt

    rt = make()
    rt.advance(606, counts0, cfg, injection=inj)
    rt.advance(607, counts1, cfg)
    expect = _state(rt)
    rt.close()

    rt = make()
    ck = rt.create_checkpoint(606, counts0, cfg, injection=inj, catalog=cat)
    i = ck.info()
    assert i.action_count == A0 + inj_n and i.slot_stride_bytes == s1
    assert i.injection.enabled == 1 and i.injection.first_catalog_word_index == A0
    ck.write(str(tmp_path / "grow.ckpt"))
    ck.close()
    rt.close()
    ck = rtm.Checkpoint.read(str(tmp_path / "grow.ckpt"))
    rt = pkg.SlabRuntime.restore(ck)
    try:
        rt.upload_catalog(wl["catalog"])
        rt.upload_secrets_csr(off, ids)
        rt.resume(ck)
        assert rt.info().action_count == A0 + inj_n
        rt.advance(607, counts1, cfg)
        _same_state(expect, _state(rt))
    finally:
        ck.close()
        rt.close()


def test_restore_rejects_inconsistent_checkpoints(pkg, tmp_path):
    """Files the reader accepts but the restore must refuse (slab_runtime.cu:1734-1742, 1790-1813)."""
    rtm = pkg.runtime
    A = 8
    stride = rtm.slot_stride_bytes(A)

    def make(**kw):
        n, cols = 4, 2
        c = {"schema_version": 3, "genome_layout_version": cko.GENOME_LAYOUT_VERSION, "resume_phase": 1,
             "generation_seed": 1,
             "identity_hash": 0, "word_counts": cko.pack_word_counts(A, A, A, 0, 1, 0, 2),
             "release_generations": [], "release_p99": [],
             "assembly": cko.pack_assembly(2, 0.5, 0.02, 0.01, 0.005, 1e-4, 0.02, 0.0),
             "injection": cko.pack_injection(False, 0, 0),
             "runtime_config": struct.pack("<6Q", 8 * stride, n * stride, 0, A, n, cols),
             "slab_layout": struct.pack("<4Q", A, stride, 8, 8 * stride),
             "shapes": [cko.pack_grid_shape(2, 2)] * 3, "generation_index": 3,
             "slots": [0, 1, 2, 3], "fitness": [0.5] * 4, "evals": [1] * 4, "has": [1] * 4,
             "pairs": [(0, 1), (1, 2), (2, 3), (3, 0)],
             "records": [(i, bytes(stride)) for i in range(4)]}
        c.update(kw)
        c["checksum"] = cko.checksum(c)
        p = tmp_path / "c.ckpt"
        p.write_bytes(cko.write(c))
        return rtm.Checkpoint.read(str(p))

    ok = make()
    rt = pkg.SlabRuntime.restore(ok)  # the consistent one restores
    assert rt.generation(0)[0].tolist() == [0, 1, 2, 3] and rt.info().slot_count == 8
    rt.close()
    ok.close()
    bad = {
        "stride of another action count": dict(slab_layout=struct.pack("<4Q", A + 1, stride, 8, 8 * stride)),
        "slot count off the slab bytes": dict(slab_layout=struct.pack("<4Q", A, stride, 9, 8 * stride)),
        "duplicate organism": dict(records=[(0, bytes(stride)), (0, bytes(stride))]),
        "record of a slotless organism": dict(slots=[0, 1, 2, 0xFFFFFFFF]),
        "parent without a genome": dict(records=[(i, bytes(stride)) for i in range(3)]),
        "parent index out of range": dict(pairs=[(0, 1), (1, 2), (2, 3), (3, 4)]),
        "pre-recombination phase only": dict(resume_phase=2),
    }
    for name, kw in bad.items():
        ck = make(**kw)
        with pytest.raises(RuntimeError) as err:
            pkg.SlabRuntime.restore(ck)
        assert err.value.status == 20, name
        ck.close()


def test_cpp_driver_checkpoints_and_resumes(pkg, tmp_path):
    """The C++ host API end to end (host/run_ga.cpp): an uninterrupted run, a run whose generation-2 step
    goes through create -> background write -> resume, and a new process resumed from that file print
    the same summaries."""
    import os
    import subprocess
    cat = pkg.synth.make_catalog(64)
    sec = pkg.synth.sample_secrets(64, 32)
    off, ids = pkg.synth.block_csr(8, 8, 4, sec)
    cat.astype(np.uint8).tofile(tmp_path / "catalog.u8")
    off.astype(np.uint32).tofile(tmp_path / "offsets.u32")
    ids.astype(np.uint32).tofile(tmp_path / "secrets.u32")
    driver = os.path.join(os.path.dirname(pkg.runtime.LIB_PATH), "run_ga")
    ckpt = str(tmp_path / "ck" / "run.ckpt")

    def run(*extra):
        res = subprocess.run([driver, str(tmp_path), "8", "4", "7", "0.01", *extra], capture_output=True, text=True,
                             timeout=120)
        assert res.returncode == 0, res.stdout + res.stderr
        return res.stdout.strip().splitlines()

    plain = run()
    checkpointed = run("--checkpoint", ckpt, "--checkpoint-every", "2")
    resumed = run("--resume", ckpt)
    strip = lambda line: line.split(" (")[0]  # noqa: E731  (drop the callback count and the time)
    assert [strip(x) for x in checkpointed] == [strip(x) for x in plain]
    assert resumed[0].startswith("resumed generation 3")
    assert [strip(x) for x in resumed[1:]] == [strip(x) for x in plain[3:]]
    assert cko.read(open(ckpt, "rb").read())["generation_index"] == 2
