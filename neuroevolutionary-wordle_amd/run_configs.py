"""Runs the measured configurations of the GA generation step on one GPU; one JSON record per run.

  grid16    16x16 grid (256 genomes), 512 actions, 256 secrets in 16 block shards of 16, 20 generations
  grid128   128x128 grid (16,384 genomes, ~41 GB slab), 4,739 actions, 8,192 secrets in 256 shards of 32
  ragged    64x64 grid, per-cell shards of U{32..1024} secrets, actions growing 512 -> 4,739 in steps of
            512 every 5 generations (9 widening repacks + output-embedding injections), 50 generations
  published the reference's two published evaluation workloads (1,024 genomes / 20 actions / 20 words per cell;
            672 genomes / 2,000 actions / 2,000 words per cell), evaluation only, against its published times
  target    64x64 grid, 4,739 actions: wall-clock to the plain FP32 run's best fitness (over G generations,
            default 200) for the BF16-screened form (eager and captured) and the surrogate-screened form
            (brood 4x, best-predicted 25% + 5% audit, BF16-screened evaluation; rank-correlation gate 0.5
            and 0), per seed (default 10 seeds), then the geometric mean, median and total-time ratio of the
            speedups

Every run is timed in the plain FP32 form and in the BF16-screened form (which takes the same decisions);
`--forms fp32` skips the latter; grid16, grid128 and ragged also take the forms fp32-graph and bf16-graph
(every step that needs no host decision replayed as one captured hipGraph -- in the ragged run every step
but the widening ones -- same results), and target times the BF16-screened form both ways.  The ragged run
autotunes the evaluation's launch shape on generation 0 and reports each widening's repack throughput.
Usage: python run_configs.py grid16|grid128|ragged|target [--forms fp32,bf16,fp32-graph,bf16-graph]
[--generations G] [--seeds S]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import __graft_entry__ as ge  # noqa: E402

pkg = ge.import_package()
rtm, synth = pkg.runtime, pkg.synth


def _run_generations(rt, gens, counts_for, injection_for, cfg, seed=7, repacks=None):
    """Advances `gens` generations; returns per-generation (seconds, games, decisions, best, mean, rescored,
    fallbacks, step kind: 0 eager / 1 replayed / 2 captured).  Widening steps append (survivors, old stride,
    copy ms, whole repack ms) to `repacks`."""
    out = []
    for g in range(gens):
        counts = counts_for(g)
        inj = injection_for(g)
        t0 = time.perf_counter()
        rt.advance(seed + 2 + g, counts, cfg, injection=inj)
        dt = time.perf_counter() - t0
        st = rt.stage_times()
        s = rt.summary()
        if inj is not None and repacks is not None:
            repacks.append((int(st.repack_survivors), int(st.repack_old_stride), float(st.repack_copy_ms),
                            float(st.repack_ms)))
        out.append((dt, int(st.games), int(st.decisions), float(s.best_fitness), float(s.average_fitness),
                    int(st.rescored), int(st.fallbacks), int(st.step_graph)))
    return out


def _record(name, form, rows, extra):
    secs = sum(r[0] for r in rows)
    games = sum(r[1] for r in rows)
    rec = {"config": name, "form": form, "generations": len(rows), "seconds": secs,
           "game_evals_per_s": games / secs, "generations_per_s": len(rows) / secs,
           "decisions": sum(r[2] for r in rows), "best_fitness": [r[3] for r in rows],
           "mean_fitness_last": rows[-1][4], "rescored": sum(r[5] for r in rows),
           "fallbacks": sum(r[6] for r in rows),
           "steps_eager_replayed_captured": [sum(r[7] == k for r in rows) for k in (0, 1, 2)]}
    rec.update(extra)
    print(json.dumps(rec), flush=True)
    return rec


def plain_config(name, gens, forms):
    wl = synth.workload(name)
    A, side = wl["actions"], wl["side"]
    n = side * side
    stride = rtm.slot_stride_bytes(A)
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    recs = []
    for form in forms:
        rt = pkg.SlabRuntime(slab_bytes=2 * n * stride + stride, generation_bytes=n * stride, action_count=A,
                             grid_columns=side)
        try:
            rt.upload_catalog(wl["catalog"])
            rt.upload_secrets_csr(wl["offsets"], wl["ids"])
            rt.set_eval_precision(rtm.EVAL_FP32 if form.startswith("fp32") else rtm.EVAL_BF16_RESCORED)
            rt.set_step_graph(form.endswith("-graph"))
            rt.bootstrap(n, seed=7, variance_scale=1.0, tail_std=0.1)
            rows = _run_generations(rt, gens, lambda g: rtm.word_counts(A, A), lambda g: None, cfg)
            recs.append(_record(name, form, rows, {"genomes": n, "actions": A,
                                                   "slab_bytes": 2 * n * stride + stride}))
        finally:
            rt.close()
    if len(recs) >= 2:
        print(json.dumps({"config": name, "same_best_fitness_trajectory":
                          all(r["best_fitness"] == recs[0]["best_fitness"] for r in recs)}), flush=True)


def ragged_config(forms, gens=50, side=64, a0=512, a_final=4739, step=512, period=5, autotune=True):
    n = side * side
    catalog = synth.make_catalog(a_final, seed=synth.CATALOG_SEED, distinct_letters=True)
    pool = synth.sample_secrets(a_final, 1030)
    offsets, ids = synth.ragged_csr(n, pool, 32, 1024)
    schedule = [min(a0 + step * (g // period), a_final) for g in range(gens + 1)]
    s_final = rtm.slot_stride_bytes(a_final)
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)

    def injection_for(g):
        if schedule[g + 1] == schedule[g]:
            return None
        return rtm.PendingInjection(enabled=1, reserved0=0, first_catalog_word_index=schedule[g],
                                    injection_count=schedule[g + 1] - schedule[g])

    recs = []
    for form in forms:
        # the generation budget becomes capacity x initial stride (slab_runtime.cu:685-696) and every
        # widening takes whole rows of it at the new stride (:918-938); one row beyond n at the final stride
        # keeps all 64 rows, so the population stays 4,096 while the action count grows (n x final stride
        # alone would give 63 rows at 4,739 actions); the slab holds two generations at the final stride
        # plus repack staging
        rt = pkg.SlabRuntime(slab_bytes=2 * n * s_final + 4 * s_final, generation_bytes=(n + side) * s_final,
                             action_count=a0, grid_columns=side)
        try:
            rt.upload_catalog(catalog)
            rt.upload_secrets_csr(offsets, ids)
            rt.set_eval_precision(rtm.EVAL_FP32 if form.startswith("fp32") else rtm.EVAL_BF16_RESCORED)
            rt.bootstrap(n, seed=7, variance_scale=1.0, tail_std=0.1)
            # the launch shape for this workload (waves, rounds per item, workgroups per CU), chosen on
            # generation 0 by the autotuner, or the host's default shape
            if autotune:
                _, wave_ms, shape = rt.autotune_evaluation(rtm.word_counts(a0, a0))
            else:
                wave_ms, shape = [], rt.eval_shape()
            rt.set_step_graph(form.endswith("-graph"))
            repacks = []
            rows = _run_generations(rt, gens, lambda g: rtm.word_counts(schedule[g], schedule[g]), injection_for,
                                    cfg, repacks=repacks)
            # repack throughput: algorithmic bytes 2 x survivors x old stride (each surviving parent read and
            # written once) over the gather + scatter kernels
            gbs = [2 * sv * stv / (ms * 1e-3) / 1e9 for sv, stv, ms, _ in repacks if ms > 0]
            recs.append(_record("ragged", form, rows, {
                "autotuned": autotune,
                "autotuned_shape": {"waves": shape.waves, "rounds_per_item": shape.rounds_per_item,
                                    "workgroups_per_cu": shape.workgroups_per_cu,
                                    "waves_sweep_ms": [round(x, 3) for x in wave_ms]},
                "repacks": [{"survivors": sv, "old_stride": stv, "copy_ms": round(ms, 3), "repack_ms": round(rm, 3),
                             "algorithmic_GBps": round(2 * sv * stv / (ms * 1e-3) / 1e9, 1) if ms > 0 else None}
                            for sv, stv, ms, rm in repacks],
                "repack_algorithmic_GBps_mean": float(np.mean(gbs)) if gbs else None,
                "genomes": n, "actions": [schedule[0], schedule[gens]], "widenings": sum(
                    schedule[g + 1] != schedule[g] for g in range(gens)),
                "secrets_per_cell": [int(np.min(np.diff(offsets))), int(np.max(np.diff(offsets))),
                                     float(np.mean(np.diff(offsets)))],
                "final_action_count": int(rt.info().action_count),
                "final_population": int(rt.info().current_generation_size)}))
        finally:
            rt.close()
    if len(recs) >= 2:
        print(json.dumps({"config": "ragged", "same_best_fitness_trajectory":
                          all(r["best_fitness"] == recs[0]["best_fitness"] for r in recs)}), flush=True)


def published_config(forms, reps=5):
    """The reference's published evaluation workloads (docs/genetic-algorithm/fitness-evaluation-concurrency-
    results.md:8-35): generation 0 with 1,024 genomes on a 32-column grid, 20 actions and every cell playing
    all 20 words (61,440 games), and generation 10 with 672 genomes (21 x 32), 2,000 actions and all 2,000
    words per cell (4,032,000 games).  Synthetic Zipf catalogs, random-init genomes; evaluation only,
    timed with HIP events (median of `reps` evaluations of the same generation)."""
    published_ms = {"gen0_a20": 676.9, "gen10_a2000": 98116.9}
    for name, rows, cols, A in (("gen0_a20", 32, 32, 20), ("gen10_a2000", 21, 32, 2000)):
        n = rows * cols
        stride = rtm.slot_stride_bytes(A)
        catalog = synth.make_catalog(A)
        offsets = (np.arange(n + 1, dtype=np.uint64) * A).astype(np.uint32)
        ids = np.tile(np.arange(A, dtype=np.uint32), n)
        for form in forms:
            rt = pkg.SlabRuntime(slab_bytes=2 * n * stride + stride, generation_bytes=n * stride, action_count=A,
                                 grid_columns=cols)
            try:
                rt.upload_catalog(catalog)
                rt.upload_secrets_csr(offsets, ids)
                rt.set_eval_precision(rtm.EVAL_FP32 if form == "fp32" else rtm.EVAL_BF16_RESCORED)
                rt.bootstrap(n, seed=7, variance_scale=1.0, tail_std=0.1)
                counts = rtm.word_counts(A, A)
                ms, wall = [], []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    rt.evaluate(counts)
                    wall.append(time.perf_counter() - t0)
                    st = rt.stage_times()
                    ms.append(st.prepare_ms + st.evaluate_ms + st.finalize_ms)
                games = 3 * n * A
                eval_ms = float(np.median(ms))
                print(json.dumps({"config": "published", "workload": name, "form": form, "genomes": n, "actions": A,
                                  "words_per_cell": A, "games": games, "evaluation_ms": eval_ms,
                                  "wall_ms": 1e3 * float(np.median(wall)), "game_evals_per_s": games / eval_ms * 1e3,
                                  "decisions": int(st.decisions), "best_fitness": float(rt.summary().best_fitness),
                                  "reference_rtx5050_laptop_ms": published_ms[name],
                                  "speedup_vs_published": published_ms[name] / eval_ms}), flush=True)
            finally:
                rt.close()


TARGET_FORMS = ("bf16", "bf16-graph", "t16", "surrogate", "surrogate-rho0", "surrogate-t16")


def wallclock_to_target(seeds, gens, cap_factor=2, forms=TARGET_FORMS, log=sys.stderr, workload="grid64_a4739"):
    """Wall-clock to target (64x64 / 4,739 actions unless `workload` names another): per seed, the plain FP32
    run's best fitness over `gens` generations is the target; every other form runs until its best-so-far
    reaches it (at most cap_factor * gens generations).  Forms: "bf16" (BF16-screened evaluation, the FP32
    form's decisions), "bf16-graph" (the same, every step one captured hipGraph), "surrogate"
    (surrogate-screened evolution, brood 4x, best-predicted 25% + 5% audit, BF16-screened evaluation in the
    persistent kernel, plain steps while the audit's rank correlation is below 0.5), "surrogate-rho0" (the
    same gated at 0), "t16" (the BF16-trunk form: BF16 screening with the trunk's layers 1 and 2 on split-BF16
    MFMA; its decisions are the FP32 form's on >= 99.9% of the decisions, not all, so its trajectory may
    differ) and "surrogate-t16" (surrogate screening, gate 0.5, on the BF16-trunk form).  Every form runs at the launch shape the autotuner picks for it on generation 0.  Host
    wall-clock per step, steps only.  Returns (per-seed records, aggregate)."""
    wl = synth.workload(workload)
    A, side = wl["actions"], wl["side"]
    n = side * side
    stride = rtm.slot_stride_bytes(A)
    counts = rtm.word_counts(A, A)
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    gates = {"surrogate": 0.5, "surrogate-rho0": 0.0, "surrogate-t16": 0.5}
    out = []
    for seed in seeds:
        def make(mode, brood):
            rt = pkg.SlabRuntime(slab_bytes=(brood + 2) * n * stride, generation_bytes=brood * n * stride,
                                 action_count=A, grid_columns=side)
            rt.upload_catalog(wl["catalog"])
            rt.upload_secrets_csr(wl["offsets"], wl["ids"])
            rt.set_eval_precision(mode)
            rt.bootstrap(n, seed=seed, variance_scale=1.0, tail_std=0.1)
            # each form at its own autotuned launch shape (chosen on generation 0, before the timed steps;
            # results do not depend on it)
            rt.autotune_evaluation(counts)
            return rt

        def run(rt, step, limit, label, target=None):
            t, best, trace, screened, rhos = 0.0, -1.0, [], 0, []
            for g in range(limit):
                t0 = time.perf_counter()
                rep = step(rt, g)
                t += time.perf_counter() - t0
                best = max(best, float(rt.summary().best_fitness))
                trace.append((round(t, 4), best))
                if rep is not None:
                    screened += rep.screened
                    if np.isfinite(rep.audit_spearman):
                        rhos.append(float(rep.audit_spearman))
                if (g + 1) % 100 == 0 and log is not None:
                    print(f"[target] seed {seed} {label} generation {g + 1} best {best:.5f} {t:.2f} s", file=log,
                          flush=True)
                if target is not None and best >= target:
                    break
            return trace, screened, rhos

        plain = lambda rt, g: rt.advance(seed + 2 + g, counts, cfg)  # noqa: E731
        screened_step = lambda rt, g: rt.advance_screened(seed + 2 + g, counts, cfg)  # noqa: E731
        rt = make(rtm.EVAL_FP32, 1)
        try:
            trace_p, _, _ = run(rt, plain, gens, "fp32")
        finally:
            rt.close()
        target = trace_p[-1][1]
        t_plain = next(tt for tt, b in trace_p if b >= target)
        g_plain = next(i + 1 for i, (tt, b) in enumerate(trace_p) if b >= target)
        rec = {"config": "target", "workload": workload, "seed": seed, "genomes": n, "actions": A,
               "plain_generations": gens,
               "target_fitness": target, "forms": {}}
        rec["forms"]["fp32"] = {"seconds_to_target": t_plain, "generations_to_target": g_plain, "reached": True,
                                "seconds_per_generation": trace_p[-1][0] / len(trace_p)}
        for form in forms:
            surrogate = form in gates
            rt = make(rtm.EVAL_BF16_TRUNK if form.endswith("t16") else rtm.EVAL_BF16_RESCORED, 4 if surrogate else 1)
            try:
                if surrogate:
                    rt.enable_screening(rtm.default_screen_config(spearman_threshold=gates[form]))
                rt.set_step_graph(form == "bf16-graph")
                tr, scr, rhos = run(rt, screened_step if surrogate else plain, cap_factor * gens, form, target)
            finally:
                rt.close()
            reached = tr[-1][1] >= target
            rec["forms"][form] = {"seconds_to_target": tr[-1][0] if reached else None,
                                  "generations_to_target": len(tr) if reached else None, "reached": reached,
                                  "seconds_per_generation": tr[-1][0] / len(tr),
                                  "speedup_vs_fp32": t_plain / tr[-1][0] if reached else None}
            if surrogate:
                rec["forms"][form].update({"spearman_threshold": gates[form], "screened_generations": scr,
                                           "generations": len(tr),
                                           "audit_spearman_mean": float(np.mean(rhos)) if rhos else None})
        out.append(rec)
    agg = {"seeds": [r["seed"] for r in out], "plain_generations": gens, "cap_generations": cap_factor * gens}
    for form in forms:
        sp = [r["forms"][form]["speedup_vs_fp32"] for r in out if r["forms"][form]["reached"]]
        t_form = sum(r["forms"][form]["seconds_to_target"] for r in out if r["forms"][form]["reached"])
        t_fp32 = sum(r["forms"]["fp32"]["seconds_to_target"] for r in out if r["forms"][form]["reached"])
        agg[form] = {"reached": len(sp), "geometric_mean_speedup": float(np.exp(np.mean(np.log(sp)))) if sp else None,
                     "median_speedup": float(np.median(sp)) if sp else None,
                     "total_time_ratio": t_fp32 / t_form if sp else None,
                     "per_seed": [r["forms"][form]["speedup_vs_fp32"] for r in out]}
    return out, agg


def agreement_config(seeds, checkpoints=(0, 25, 100), workload="grid64_a4739"):
    """Decision agreement of the BF16-trunk form with the FP32 form on evolving populations: per seed, plain FP32
    generations run, and at each checkpoint generation the current population is evaluated traced in both forms;
    decisions are compared while a game's two trajectories coincide."""
    wl = synth.workload(workload)
    A, side = wl["actions"], wl["side"]
    n = side * side
    stride = rtm.slot_stride_bytes(A)
    counts = rtm.word_counts(A, A)
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    games = 3 * int(wl["offsets"][-1])
    for seed in seeds:
        rt = pkg.SlabRuntime(slab_bytes=2 * n * stride + stride, generation_bytes=n * stride, action_count=A,
                             grid_columns=side)
        rt.upload_catalog(wl["catalog"])
        rt.upload_secrets_csr(wl["offsets"], wl["ids"])
        rt.bootstrap(n, seed=seed, variance_scale=1.0, tail_std=0.1)
        try:
            for g in range(max(checkpoints) + 1):
                if g in checkpoints:
                    out = {}
                    for name, mode in (("fp32", rtm.EVAL_FP32), ("t16", rtm.EVAL_BF16_TRUNK)):
                        rt.set_eval_precision(mode)
                        rt.set_trace(rtm.TRACE_GAMES)
                        rt.evaluate(counts)
                        out[name] = (rt.download_trace(rtm.TRACE_GAMES, games), rt.generation(0)[1].copy())
                    rt.set_trace(0)
                    rt.set_eval_precision(rtm.EVAL_FP32)
                    a, b = out["fp32"][0], out["t16"][0]
                    compared = agreed = 0
                    on_path = None
                    for k in range(6):
                        live = (a["decisions"] > k) & (b["decisions"] > k)
                        if k > 0:
                            live &= on_path
                        eq = a["guesses"][:, k] == b["guesses"][:, k]
                        compared += int(live.sum())
                        agreed += int((live & eq).sum())
                        on_path = live & eq if k == 0 else on_path & eq
                    print(json.dumps({"config": "agreement", "workload": workload, "seed": seed, "generation": g,
                                      "decisions_compared": compared, "decisions_agreeing": agreed,
                                      "agreement": agreed / compared,
                                      "games_identical": int(np.all(a["guesses"] == b["guesses"], axis=1).sum()),
                                      "games": games,
                                      "fitness_equal_fraction": float(np.mean(out["fp32"][1] == out["t16"][1])),
                                      "best_fitness_fp32": float(out["fp32"][1].max())}), flush=True)
                rt.advance(seed + 2 + g, counts, cfg)
        finally:
            rt.close()


def target_config(seeds, gens, cap_factor=2, forms=TARGET_FORMS, workload="grid64_a4739"):
    recs, agg = wallclock_to_target(seeds, gens, cap_factor, forms=forms, workload=workload)
    for r in recs:
        print(json.dumps(r), flush=True)
    print(json.dumps(dict(agg, config="target", workload=workload)), flush=True)
    return recs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("config", choices=["grid16", "grid128", "ragged", "target", "target128", "published", "agreement"])
    ap.add_argument("--forms", default="")
    ap.add_argument("--generations", type=int, default=0)
    ap.add_argument("--seeds", type=int, default=10)
    ap.add_argument("--first-seed", type=int, default=7)
    ap.add_argument("--no-autotune", action="store_true", help="ragged: the default launch shape")
    args = ap.parse_args()
    forms = [f for f in args.forms.split(",") if f]
    seeds = [args.first_seed + 10 * i for i in range(args.seeds)]
    if args.config == "agreement":  # BF16-trunk vs FP32 decisions on evolving populations (64x64 / 4,739)
        agreement_config(seeds)
        return
    if args.config in ("target", "target128"):
        target_forms = tuple(forms) if forms else TARGET_FORMS
        if args.config == "target":
            target_config(seeds, args.generations or 200, forms=target_forms)
        else:  # the combined form at 128x128: BF16 evaluation + persistent kernel + surrogate screening
            target_config(seeds, args.generations or 100, forms=tuple(forms) if forms else ("bf16", "surrogate"),
                          workload="grid128_a4739")
        return
    forms = forms or ["fp32", "bf16"]
    if args.config == "grid16":
        plain_config("grid16_a512", args.generations or 20, forms)
    elif args.config == "grid128":
        plain_config("grid128_a4739", args.generations or 3, forms)
    elif args.config == "published":
        published_config(forms)
    else:
        ragged_config(forms, gens=args.generations or 50, autotune=not args.no_autotune)


if __name__ == "__main__":
    main()
