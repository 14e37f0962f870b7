"""Boundary calls of the device runtime on the GPU: the whole-slab download (TryDownloadSlabFromDevice),
the generation-buffer swap (SwapDeviceSlabGenerationBuffers) and the verbose progress lines of a step.

The slab download restates the reference's one-generation test (tests/genetic_algorithm/device/
slab_runtime_test.cu:414-481): 8 slots, 4 genomes, temperatures and mutation 0, so each child copies its first
parent; afterwards 4 slots are free and every child slot holds one generation reference."""
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_runtime(pkg, A=24, side=2, slots=8, n=4):
    rtm = pkg.runtime
    stride = rtm.slot_stride_bytes(A)
    cat = pkg.synth.make_catalog(A, seed=4)
    sec = pkg.synth.sample_secrets(A, 8, seed=5)
    off, ids = pkg.synth.block_csr(side, side, 1, sec[:side * side * 2].reshape(-1))
    rt = pkg.SlabRuntime(slab_bytes=slots * stride, generation_bytes=n * stride, action_count=A, grid_columns=side)
    rt.upload_catalog(cat)
    rt.upload_secrets_csr(off, ids)
    rt.bootstrap(n, seed=77, generation_index=2)
    return rt, rtm.word_counts(A, A)


def test_whole_slab_download_after_one_generation(pkg):
    rtm = pkg.runtime
    rt, counts = _small_runtime(pkg)
    try:
        before, states0, stack0, fc0 = rt.download_slab()
        info = rt.info()
        assert before.shape == (8, info.slot_stride_bytes) and fc0 == 4
        assert states0.tolist() == [1, 1, 1, 1, 0, 0, 0, 0]
        assert stack0[:fc0].tolist() == [7, 6, 5, 4]
        parents, _, _, _ = rt.generation(0)
        for s in range(8):
            if states0[s]:
                assert rt.download_slot(s).tobytes() == before[s].tobytes()
        cfg = rtm.default_assembly_config(crossover_temperature_level1=0.0, crossover_temperature_level2=0.0,
                                          crossover_temperature_level3=0.0, mutation_probability=0.0)
        rt.advance(19, counts, cfg)
        after, states, stack, fc = rt.download_slab()
        children, _, evals, has = rt.generation(0)
        pairs = rt.plan()
        assert rt.info().current_generation_index == 3 and fc == 4
        assert not has.any() and not evals.any()
        nbytes = (132928 + 38 * 24) * 4
        for c, s in enumerate(children):
            assert states[s] == 1
            assert after[s, :nbytes].tobytes() == before[parents[pairs[c][0]], :nbytes].tobytes()
        assert sorted(stack[:fc].tolist() + children.tolist()) == list(range(8))
    finally:
        rt.close()


def test_generation_swap_exchanges_and_restores(pkg):
    rt, counts = _small_runtime(pkg)
    try:
        rt.advance(5, counts)
        g0 = rt.generation(0)
        rt.swap_generations()
        assert rt.info().current_generation_size == 0  # the step left no next generation behind
        rt.swap_generations()
        g1 = rt.generation(0)
        assert all(np.array_equal(a, b) for a, b in zip(g0, g1))
        assert rt.info().current_generation_index == 3
        rt.evaluate(counts)  # still a working runtime
        rt.advance(6, counts)
    finally:
        rt.close()


def test_verbose_step_prints_timestamped_stage_lines(pkg, capfd):
    rt, counts = _small_runtime(pkg)
    try:
        rt.advance(5, counts)
        quiet = capfd.readouterr().out
        assert "Generation" not in quiet
        rt.advance(6, counts, verbose=True)
        out = capfd.readouterr().out.strip().splitlines()
    finally:
        rt.close()
    assert out and all(re.match(r"^\[\d{4}-\d\d-\d\d \d\d:\d\d:\d\d\] Generation 4: ", line) for line in out), out
    text = "\n".join(out)
    for stage in ("starting advancement from generation 3", "evaluating current generation fitness",
                  "fitness evaluated", "next generation shape planned", "parent plan built", "assembly batch",
                  "advancement finished"):
        assert stage in text, stage
    assert re.search(r" in \d+\.\d ms", text)
