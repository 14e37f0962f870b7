"""Captured generation steps (ne_set_step_graph) on the GPU.

Bar: a runtime that replays its plain steps as hipGraphs ends every step in the state of a runtime that
runs them eagerly, bit for bit -- summary, generation slots, fitness and evaluation counts, the plan, the
free-slot count and every genome of the new generation -- in the FP32 and the BF16-screened forms; the
first step of each buffer parity captures, later ones replay; a changed configuration captures anew; an
eager evaluation between steps (which rewrites the work-list buffers), a step that does not qualify (a slab
without room for a whole generation), a widening step with injection (new stride: both parities capture
anew) and surrogate-screened steps in between leave the results unchanged.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _make(pkg, wl, mode, slots_extra=None, graph=False):
    rtm = pkg.runtime
    A, side = wl["actions"], wl["side"]
    n = side * side
    stride = rtm.slot_stride_bytes(A)
    slots = 2 * n + 1 if slots_extra is None else n + slots_extra
    rt = pkg.SlabRuntime(slab_bytes=slots * stride, generation_bytes=n * stride, action_count=A, grid_columns=side)
    rt.upload_catalog(wl["catalog"])
    rt.upload_secrets_csr(wl["offsets"], wl["ids"])
    rt.set_eval_precision(mode)
    rt.bootstrap(n, seed=5, variance_scale=1.0, tail_std=0.1)
    if graph:
        rt.set_step_graph(True)
    return rt


def _state(rt, genomes=True):
    s = rt.summary()
    slots, fit, evals, has = rt.generation(0)
    info = rt.info()
    out = {
        "summary": (np.float32(s.best_fitness).view(np.uint32), np.float32(s.average_fitness).view(np.uint32),
                    s.best_index, s.best_slot_index, s.generation_index, s.population_size),
        "slots": slots.tolist(), "fit": fit.view(np.uint32).tolist(), "evals": evals.tolist(),
        "has": has.tolist(), "free": info.free_slot_count, "plan": rt.plan().tolist(),
    }
    if genomes:
        out["genomes"] = [rt.download_slot(x).view(np.uint32).tobytes() for x in slots]
    return out


def _compare(a, b, where):
    for k in a:
        assert a[k] == b[k], (where, k)


@pytest.mark.parametrize("mode_name,kernel", [("EVAL_FP32", 0), ("EVAL_BF16_RESCORED", 0), ("EVAL_FP32", 1),
                                              ("EVAL_FP32", 2), ("EVAL_BF16_RESCORED", 1),
                                              ("EVAL_BF16_RESCORED", 2)])
def test_graph_steps_equal_eager_steps(pkg, mode_name, kernel):
    """The captured runtime uses the given evaluation kernel (registers, streamed, paired), the eager one the
    register kernel."""
    rtm = pkg.runtime
    mode = getattr(rtm, mode_name)
    wl = pkg.synth.workload("grid16_a512")
    counts = rtm.word_counts(wl["actions"], wl["actions"])
    cfg = rtm.default_assembly_config(mutation_probability=0.02, mutation_sigma=0.03)
    cfg2 = rtm.default_assembly_config(mutation_probability=0.05, mutation_sigma=0.01)
    eager = _make(pkg, wl, mode)
    graph = _make(pkg, wl, mode, graph=True)
    graph.set_eval_shape_ex(0, 1, 0, kernel)
    try:
        kinds = []
        for g in range(6):
            c = cfg if g < 4 else cfg2
            eager.advance(40 + g, counts, c)
            graph.advance(40 + g, counts, c)
            st = graph.stage_times()
            kinds.append(int(st.step_graph))
            assert st.step_graph_ms > 0.0
            assert eager.stage_times().step_graph == 0
            assert st.decisions == eager.stage_times().decisions and st.games == eager.stage_times().games
            _compare(_state(eager), _state(graph), f"step {g}")
            if g == 2:  # an eager evaluation in between rewrites the CSR and work-list buffers
                eager.evaluate(counts)
                graph.evaluate(counts)
                _compare(_state(eager, False), _state(graph, False), "evaluate")
        # each buffer parity captures once; a new configuration captures both parities again
        assert kinds == [2, 2, 1, 1, 2, 2], kinds
        print(f"[step-graph] {mode_name}: step kinds {kinds}, replay {graph.stage_times().step_graph_ms:.3f} ms")
    finally:
        eager.close()
        graph.close()


def test_step_without_room_runs_eagerly(pkg):
    rtm = pkg.runtime
    wl = pkg.synth.workload("grid16_a512")
    counts = rtm.word_counts(wl["actions"], wl["actions"])
    cfg = rtm.default_assembly_config(mutation_probability=0.02, mutation_sigma=0.03)
    n = wl["side"] ** 2
    # a slab of 2n - 1 slots: the free count at the start of a step is one short of a generation, so the
    # step may need a second assembly batch, which only the eager step takes
    eager = _make(pkg, wl, rtm.EVAL_FP32, slots_extra=n - 1)
    graph = _make(pkg, wl, rtm.EVAL_FP32, slots_extra=n - 1, graph=True)
    try:
        for g in range(2):
            eager.advance(70 + g, counts, cfg)
            graph.advance(70 + g, counts, cfg)
            assert graph.stage_times().step_graph == 0
            _compare(_state(eager), _state(graph), f"step {g}")
        graph.set_step_graph(False)
    finally:
        eager.close()
        graph.close()


def test_graph_step_at_full_size(pkg):
    """64x64 genomes, 4,739 actions: graph steps with the streamed kernel (the autotuner's choice there) and
    eager steps with the register kernel agree, and the replay is timed."""
    rtm = pkg.runtime
    wl = pkg.synth.workload("grid64_a4739")
    counts = rtm.word_counts(wl["actions"], wl["actions"])
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    eager = _make(pkg, wl, rtm.EVAL_FP32)
    graph = _make(pkg, wl, rtm.EVAL_FP32, graph=True)
    graph.set_eval_shape_ex(3, 1, 0, rtm.EVAL_KERNEL_STREAMED)
    try:
        for g in range(3):
            eager.advance(9 + g, counts, cfg)
            graph.advance(9 + g, counts, cfg)
            a, b = _state(eager, False), _state(graph, False)
            _compare(a, b, f"step {g}")
            for o in (0, 1, 2047, 4095):  # a sample of the new generation's genomes
                assert np.array_equal(eager.download_slot(a["slots"][o]).view(np.uint32),
                                      graph.download_slot(b["slots"][o]).view(np.uint32)), (g, o)
        st = graph.stage_times()
        assert st.step_graph == 1
        print(f"[step-graph] 64x64 / 4739: replay {st.step_graph_ms:.3f} ms, eager evaluate "
              f"{eager.stage_times().evaluate_ms:.3f} ms + assemble {eager.stage_times().assemble_ms:.3f} ms")
    finally:
        eager.close()
        graph.close()


def test_cpp_driver_with_captured_steps(pkg, tmp_path):
    """run_ga --graph: the same per-generation summaries as the eager driver."""
    import os
    import subprocess
    rtm = pkg.runtime
    cat = pkg.synth.make_catalog(64)
    sec = pkg.synth.sample_secrets(64, 32)
    off, ids = pkg.synth.block_csr(8, 8, 4, sec)
    cat.astype(np.uint8).tofile(tmp_path / "catalog.u8")
    off.astype(np.uint32).tofile(tmp_path / "offsets.u32")
    ids.astype(np.uint32).tofile(tmp_path / "secrets.u32")
    driver = os.path.join(os.path.dirname(rtm.LIB_PATH), "run_ga")
    out = []
    for extra in ([], ["--graph"]):
        res = subprocess.run([driver, str(tmp_path), "8", "6", "11", "0.02"] + extra, capture_output=True, text=True,
                             timeout=120)
        assert res.returncode == 0, res.stdout + res.stderr
        out.append([x for x in res.stdout.splitlines() if x.startswith("generation ")])
    assert len(out[0]) == 6 and out[0] == out[1]


def test_graph_steps_across_a_widening(pkg):
    """A widening step (injection, eager) between captured steps: the new stride captures anew."""
    rtm = pkg.runtime
    side, a0, a1 = 16, 512, 1024
    n = side * side
    catalog = pkg.synth.make_catalog(a1, seed=pkg.synth.CATALOG_SEED, distinct_letters=True)
    off, ids = pkg.synth.block_csr(side, side, 4, pkg.synth.sample_secrets(a0, 256))
    s1 = rtm.slot_stride_bytes(a1)
    cfg = rtm.default_assembly_config(mutation_probability=0.02, mutation_sigma=0.03)
    inj = rtm.PendingInjection(enabled=1, reserved0=0, first_catalog_word_index=a0, injection_count=a1 - a0)

    def make(graph):
        rt = pkg.SlabRuntime(slab_bytes=2 * n * s1 + 4 * s1, generation_bytes=n * s1, action_count=a0,
                             grid_columns=side)
        rt.upload_catalog(catalog)
        rt.upload_secrets_csr(off, ids)
        rt.bootstrap(n, seed=21, variance_scale=1.0, tail_std=0.1)
        rt.set_step_graph(graph)
        return rt

    eager, graph = make(False), make(True)
    try:
        kinds = []
        for g in range(6):
            a = a0 if g <= 2 else a1
            counts = rtm.word_counts(a, a)
            for rt in (eager, graph):
                rt.advance(60 + g, counts, cfg, injection=inj if g == 2 else None)
            kinds.append(int(graph.stage_times().step_graph))
            _compare(_state(eager), _state(graph), f"step {g}")
        assert graph.info().action_count == a1
        assert kinds == [2, 2, 0, 2, 2, 1], kinds
    finally:
        eager.close()
        graph.close()


def test_graph_steps_between_screened_steps(pkg):
    """Surrogate-screened steps (eager) interleaved with captured plain steps."""
    rtm = pkg.runtime
    wl = pkg.synth.workload("grid16_a512")
    A, side = wl["actions"], wl["side"]
    n = side * side
    stride = rtm.slot_stride_bytes(A)
    counts = rtm.word_counts(A, A)
    cfg = rtm.default_assembly_config(mutation_probability=0.02, mutation_sigma=0.03)
    scfg = rtm.default_screen_config(brood_factor=2, min_replay=256, spearman_threshold=0.0)

    def make(graph):
        rt = pkg.SlabRuntime(slab_bytes=4 * n * stride, generation_bytes=2 * n * stride, action_count=A,
                             grid_columns=side)
        rt.upload_catalog(wl["catalog"])
        rt.upload_secrets_csr(wl["offsets"], wl["ids"])
        rt.set_eval_precision(rtm.EVAL_BF16_RESCORED)
        rt.bootstrap(n, seed=33, variance_scale=1.0, tail_std=0.1)
        rt.enable_screening(scfg)
        rt.set_step_graph(graph)
        return rt

    eager, graph = make(False), make(True)
    try:
        kinds = []
        for g, screened in enumerate([False, True, False, True, True, False, False]):
            for rt in (eager, graph):
                if screened:
                    rt.advance_screened(80 + g, counts, cfg)
                else:
                    rt.advance(80 + g, counts, cfg)
            kinds.append(None if screened else int(graph.stage_times().step_graph))
            _compare(_state(eager), _state(graph), f"step {g}")
        assert all(k in (1, 2) for k in kinds if k is not None), kinds
    finally:
        eager.close()
        graph.close()
