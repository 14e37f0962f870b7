"""Benchmark of the GA generation step on one MI355X.

A step = one full generation (ne_advance_generation): evaluate every genome's games, plan parents,
assemble children (crossover + mutation), swap.  Workload: 64x64 toroidal grid (4,096 genomes),
4,739 synthetic actions, 2,048 secrets in 64 block shards of 32 -> 393,216 games per generation.
The plain FP32 form is `value`; the BF16-screened form with exact FP32 rescoring (same decisions,
same fitness) is timed over the same generations and reported beside it, and so are both forms with every
step replayed as one captured hipGraph (ne_set_step_graph: same results; only the whole replay can be
timed, so `value` and the roofline stay on the eager steps, whose kernels are timed with HIP events).
Prints one JSON line.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "neuroevolutionary-wordle_amd"))

PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
WORKLOAD = "grid64_a4739"


def algorithmic_eval_flops(decisions, encoded_turns, A):
    # per decision: 2*(321*256 + 256*128 + 128*64) + 128*A ; per encoded turn (145->64): 2*145*64
    return decisions * (246272 + 128 * A) + encoded_turns * 18560


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "evaluate_grid64_a4739_traffic.json")
KERNEL_SOURCES = [os.path.join(ROOT, "neuroevolutionary-wordle_amd", "csrc", f) for f in ("ne_kernels.hip", "ne_device.hpp")]


def kernel_source_hash():
    """sha256 over the evaluation kernel's sources (what a traffic measurement is tied to)."""
    h = hashlib.sha256()
    for path in KERNEL_SOURCES:
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


KERNEL_NAMES = ("EvaluateKernel<false,false,false>", "EvaluateStreamedKernel", "EvaluatePairKernel")
# the BF16-screened form's kernels (its streamed kernel is the 3-wave ring; at other sizes the register kernel runs)
KERNEL_NAMES_BF16 = ("EvaluateKernel<false,true,false>", "EvaluateStreamedBf16RingKernel<false>",
                     "EvaluatePairBf16Kernel<false>")
# the BF16-trunk form's kernels (the BF16 kernels with the split-BF16 trunk)
KERNEL_NAMES_T16 = ("EvaluateKernel<false,true,true>", "EvaluateStreamedBf16RingKernel<true>",
                    "EvaluatePairBf16Kernel<true>")
PEAK_BF16_MFMA_TFLOPS = 2500.0


def measured_traffic(kernel=0, bf16=False, t16=False):
    """HBM bytes per evaluation launch of one form's `kernel` (0 registers, 1 streamed, 2 paired) from the
    committed FETCH_SIZE / WRITE_SIZE measurement of this workload (scaled by the plain-copy calibration
    recorded with it) and whether it was measured on the kernel sources of this tree: (bytes or None, stale)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, True
    with open(TRAFFIC_FILE) as f:
        t = json.load(f)
    stale = t.get("kernel_source_sha256") != kernel_source_hash()
    k = t.get("kernels", {}).get((KERNEL_NAMES_T16 if t16 else KERNEL_NAMES_BF16 if bf16 else KERNEL_NAMES)[kernel])
    if k is None:
        return None, True
    return k["fetch_bytes"] * t["fetch_scale"] + k["write_bytes"] * t["write_scale"], stale


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(pkg, rt, wl, A, seconds=3.0, reps=5):
    """The oracle (fp32, scalar C++) on bounded samples of the same workload: the host's cores for one GPU
    (at most 16, the box's share) and 1 thread, each the median of `reps` timed runs of about `seconds`
    after a warm-up run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    share = max(1, min(share, 16))
    slots, _, _, _ = rt.generation(0)
    offsets, ids = wl["offsets"], wl["ids"]
    pool = []

    def timed(n, threads):
        while len(pool) < n:
            pool.append(rt.download_slot(slots[len(pool)]))
        t0 = time.perf_counter()
        r = orc.evaluate(pool[:n], 0, A, A, wl["catalog"], offsets[:n + 1], ids, mode=0, threads=threads)
        dt = time.perf_counter() - t0
        assert r["status"] == 0
        return 3 * int(offsets[n]) / dt, r["decisions"], dt

    def measure(threads):
        probe = max(threads, 2)
        rate, _, _ = timed(probe, threads)  # warm-up and sizing
        per_genome = 3 * int(offsets[probe]) / probe
        n = int(min(len(slots), max(probe, rate * seconds / per_genome)))
        n = max(probe, (n // threads) * threads)
        runs = [timed(n, threads) for _ in range(reps)]
        rates = sorted(r[0] for r in runs)
        return {"value": float(np.median(rates)), "threads": threads, "genomes": n, "games": 3 * int(offsets[n]),
                "decisions": int(runs[0][1]), "seconds_per_run": float(np.median([r[2] for r in runs])),
                "runs": [float(x) for x in rates]}

    many = measure(share)
    one = measure(1)
    return {"value": many["value"], "unit": "game-evals/s", "cores": share, "kind": "port",
            "sample": f"fitness evaluation of the first {many['genomes']} genomes of generation 0 ({many['games']} "
                      f"games, {many['decisions']} decisions) at A={A}: median of {reps} runs of "
                      f"{many['seconds_per_run']:.1f} s on {share} threads",
            "one_thread": {"value": one["value"], "unit": "game-evals/s",
                           "sample": f"first {one['genomes']} genomes ({one['games']} games): median of {reps} runs "
                                     f"of {one['seconds_per_run']:.1f} s"},
            "runs": many["runs"], "one_thread_runs": one["runs"],
            "hardware_concurrency": os.cpu_count(), "cpu_model": cpu_model(),
            "build": "g++ -O3 -march=x86-64-v3 -ffp-contract=off, no fast-math (oracle/build_oracle.py)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-wallclock", action="store_true", help="skip the wall-clock-to-target leg")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="time every form at the built-in launch shape instead of the autotuner's")
    ap.add_argument("--eval-shape", default="", metavar="W,R,G,K",
                    help="time every form at this launch shape (waves, rounds per item, workgroups per CU, "
                         "evaluation kernel: 0 registers, 1 streamed, 2 paired) instead of the autotuner's")
    ap.add_argument("--wallclock-seeds", default="207,217,227,237,247",
                    help="seeds of the wall-clock-to-target leg (none of them was used to choose a setting)")
    ap.add_argument("--wallclock-generations", type=int, default=200,
                    help="plain FP32 generations that define the target (other forms stop at twice as many)")
    ap.add_argument("--fp32-only", action="store_true", help="skip the BF16-screened form")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="first run one plain device-to-device copy of BYTES (a reference kernel of known "
                         "traffic for hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU)",
              file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        import torch
        x = torch.empty(args.calibrate_copy // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        x.fill_(1.0)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy of {args.calibrate_copy} bytes done", file=sys.stderr, flush=True)
        del x, y

    import __graft_entry__ as ge
    pkg = ge.import_package()
    rtm = pkg.runtime
    wl = pkg.synth.workload(WORKLOAD)
    A, side = wl["actions"], wl["side"]
    n = side * side
    stride = rtm.slot_stride_bytes(A)
    rt = pkg.SlabRuntime(slab_bytes=2 * n * stride + stride, generation_bytes=n * stride, action_count=A,
                         grid_columns=side)
    rt.upload_catalog(wl["catalog"])
    rt.upload_secrets_csr(wl["offsets"], wl["ids"])
    rt.bootstrap(n, seed=7, variance_scale=1.0, tail_std=0.1)
    counts = rtm.word_counts(A, A)
    cfg = rtm.default_assembly_config(mutation_probability=0.01, mutation_sigma=0.02)
    games_per_gen = 3 * int(wl["offsets"][-1])

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(pkg, rt, wl, A)
        print(f"[bench] cpu baseline {cpu['value']:.1f} game-evals/s on {cpu['cores']} threads", file=sys.stderr,
              flush=True)

    shapes = {}

    def run_form(mode, graph=False):
        """Bootstraps the same generation 0 and times `steps` generation steps in one form (eagerly, or as
        captured step graphs: then only the whole replay is timed)."""
        rt.set_eval_precision(mode)
        rt.set_step_graph(False)
        rt.bootstrap(n, seed=7, variance_scale=1.0, tail_std=0.1)
        if args.eval_shape:
            w, r, g, k = (int(x) for x in args.eval_shape.split(","))
            rt.set_eval_shape_ex(w, r, g, k)
            shapes[(mode, graph)] = {"waves": w, "rounds_per_item": r, "workgroups_per_cu": g, "kernel_id": k,
                                     "kernel": ["registers", "streamed", "paired"][k], "fixed": True}
        elif args.autotune:
            # the evaluation's launch shape for this form, chosen on generation 0 (results do not depend on it)
            _, wave_ms, shape = rt.autotune_evaluation(counts)
            shapes[(mode, graph)] = {"waves": shape.waves, "rounds_per_item": shape.rounds_per_item,
                                     "kernel_id": shape.kernel,
                                     "workgroups_per_cu": shape.workgroups_per_cu,
                                     "kernel": ["registers", "streamed", "paired"][shape.kernel],
                                     "waves_sweep_ms": [round(x, 3) for x in wave_ms]}
        rt.set_step_graph(graph)
        gen = 0
        for _ in range(args.warmup):
            rt.advance(7 + 2 + gen, counts, cfg)
            gen += 1
        eval_ms, flops, decisions, rescored, fallbacks = [], [], 0, 0, 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rt.advance(7 + 2 + gen, counts, cfg)
            gen += 1
            st = rt.stage_times()
            eval_ms.append(st.evaluate_ms)
            flops.append(algorithmic_eval_flops(st.decisions, st.encoded_turns, A))
            decisions += st.decisions
            rescored += st.rescored
            fallbacks += st.fallbacks
        elapsed = time.perf_counter() - t0
        _, fit, _, _ = rt.generation(0)
        rt.set_step_graph(False)
        return dict(elapsed=elapsed, eval_ms=float(np.mean(eval_ms)), flops=float(np.mean(flops)),
                    decisions=decisions, rescored=rescored, fallbacks=fallbacks, stages=rt.stage_times(),
                    summary=rt.summary(), fitness=fit.copy())

    fp32 = run_form(rtm.EVAL_FP32)
    bf16 = run_form(rtm.EVAL_BF16_RESCORED) if not args.fp32_only else None
    trunk16 = run_form(rtm.EVAL_BF16_TRUNK) if not args.fp32_only else None
    fp32_graph = run_form(rtm.EVAL_FP32, graph=True)
    bf16_graph = run_form(rtm.EVAL_BF16_RESCORED, graph=True) if not args.fp32_only else None
    rt.close()
    st = fp32["stages"]
    value = games_per_gen * args.steps / fp32["elapsed"]
    eval_avg_ms = fp32["eval_ms"]
    achieved = fp32["flops"] / (eval_avg_ms * 1e-3) / 1e12
    summary = fp32["summary"]
    # the evaluation kernel the FP32 form was timed with (the built-in one unless a shape was chosen)
    fp32_kernel = shapes.get((rtm.EVAL_FP32, False), {}).get("kernel_id", 0)
    traffic, traffic_stale = measured_traffic(fp32_kernel)
    bf16_out = None
    if bf16 is not None:
        sh = shapes.get((rtm.EVAL_BF16_RESCORED, False), {})
        k16 = sh.get("kernel_id", 0)
        if k16 == 1 and sh.get("waves") != 3:
            k16 = 0  # the BF16 form's streamed kernel exists at 3 waves only
        t16, t16_stale = measured_traffic(k16, bf16=True)
        f16 = bf16["flops"] / (bf16["eval_ms"] * 1e-3) / 1e12
        bf16_out = {
            "value": games_per_gen * args.steps / bf16["elapsed"], "unit": "game-evals/s",
            "ms_per_step": bf16["elapsed"] / args.steps * 1e3, "generations_per_s": args.steps / bf16["elapsed"],
            "evaluate_ms": bf16["eval_ms"],
            "kernel": KERNEL_NAMES_BF16[k16],
            # the same algorithmic FLOP as the FP32 form's roofline, over the BF16 dense MFMA peak (most of the
            # work -- the scoring -- runs on BF16 MFMA; the trunk and the rescoring stay FP32), and the measured
            # HBM bytes of the kernel per launch over its time against the 8 TB/s peak
            "roofline": {"achieved_tflops": f16, "frac_of_bf16_peak": f16 / PEAK_BF16_MFMA_TFLOPS,
                         "traffic": None if t16_stale else t16,
                         "traffic_GBps": None if (t16_stale or t16 is None) else t16 / (bf16["eval_ms"] * 1e-3) / 1e9,
                         "traffic_frac_of_hbm_peak": None if (t16_stale or t16 is None)
                         else t16 / (bf16["eval_ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS},
            "candidates_rescored_per_decision": bf16["rescored"] / max(1, bf16["decisions"]),
            "fp32_fallback_decisions": bf16["fallbacks"],
            "same_fitness_as_fp32": bool(np.array_equal(bf16["fitness"].view(np.uint32),
                                                        fp32["fitness"].view(np.uint32))),
        }
    t16_out = None
    if trunk16 is not None:
        sh = shapes.get((rtm.EVAL_BF16_TRUNK, False), {})
        kt = sh.get("kernel_id", 0)
        if kt == 1 and sh.get("waves") != 3:
            kt = 0
        ft = trunk16["flops"] / (trunk16["eval_ms"] * 1e-3) / 1e12
        tt, tt_stale = measured_traffic(kt, t16=True)
        # not bit-identical to the FP32 form: the trunk's split-BF16 products differ by ~2^-16 relative, so the
        # steps' trajectories part after the first differing decision (decision agreement at full size is a GPU
        # test); fitness agreement is reported on the last generation of the same steps
        t16_out = {
            "value": games_per_gen * args.steps / trunk16["elapsed"], "unit": "game-evals/s",
            "ms_per_step": trunk16["elapsed"] / args.steps * 1e3, "generations_per_s": args.steps / trunk16["elapsed"],
            "evaluate_ms": trunk16["eval_ms"], "kernel": KERNEL_NAMES_T16[kt],
            "roofline": {"achieved_tflops": ft, "frac_of_bf16_peak": ft / PEAK_BF16_MFMA_TFLOPS,
                         "traffic": None if tt_stale else tt,
                         "traffic_GBps": None if (tt_stale or tt is None) else tt / (trunk16["eval_ms"] * 1e-3) / 1e9},
            "candidates_rescored_per_decision": trunk16["rescored"] / max(1, trunk16["decisions"]),
            "fp32_fallback_decisions": trunk16["fallbacks"],
            "fitness_equal_to_fp32_fraction": float(np.mean(trunk16["fitness"] == fp32["fitness"])),
            "best_fitness": float(trunk16["summary"].best_fitness),
        }

    def graph_record(rec, eager):
        return {"value": games_per_gen * args.steps / rec["elapsed"], "unit": "game-evals/s",
                "ms_per_step": rec["elapsed"] / args.steps * 1e3, "generations_per_s": args.steps / rec["elapsed"],
                "replay_ms": float(rec["stages"].step_graph_ms),
                "same_fitness_as_eager": bool(np.array_equal(rec["fitness"].view(np.uint32),
                                                             eager["fitness"].view(np.uint32)))}
    graph_out = {"fp32": graph_record(fp32_graph, fp32)}
    if bf16 is not None:
        graph_out["bf16_screened_fp32_rescored"] = graph_record(bf16_graph, bf16)
    out = {
        "metric": "game-evals/sec (plain FP32 generation step)",
        "value": value,
        "unit": "game-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp32["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "64x64 toroidal grid (4096 genomes), 4739 Zipf(1.1) synthetic actions, 2048 "
                               "secrets in 64 block shards of 32, 3 episodes per secret",
                   "genomes": n, "actions": A, "games_per_generation": games_per_gen,
                   "mutation_probability": 0.01, "mutation_sigma": 0.02},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_MFMA_TFLOPS, "traffic": None if traffic_stale else traffic,
                     "kernel": KERNEL_NAMES[fp32_kernel], "kernel_ms": eval_avg_ms},
        "traffic_measurement": {"bytes_per_launch": traffic, "stale": traffic_stale, "kernel": KERNEL_NAMES[fp32_kernel],
                                "file": os.path.relpath(TRAFFIC_FILE, ROOT),
                                "note": "FETCH_SIZE + WRITE_SIZE passes of this kernel; stale when the kernel sources "
                                        "changed since (then roofline.traffic is null)"},
        "stages_ms": {"evaluate": st.evaluate_ms, "finalize": st.finalize_ms, "plan": st.plan_ms,
                      "assemble": st.assemble_ms, "bookkeeping": st.bookkeeping_ms},
        "generations_per_s": args.steps / fp32["elapsed"],
        "decisions_per_generation": fp32["decisions"] / args.steps,
        "bf16_screened_fp32_rescored": bf16_out,
        "bf16_trunk_bf16_screened": t16_out,
        "captured_step_graph": graph_out,
        "best_fitness": summary.best_fitness,
        "launch_shapes": {{rtm.EVAL_FP32: "fp32", rtm.EVAL_BF16_RESCORED: "bf16", rtm.EVAL_BF16_TRUNK: "bf16_trunk"}[m]
                          + ("_graph" if g else ""):
                          {k: x for k, x in v.items() if k != "kernel_id"} for (m, g), v in shapes.items()},
        "cpu_baseline": cpu,
    }
    if not args.no_wallclock:
        import run_configs  # the configuration runner beside the package (imports the same package)
        seeds = [int(x) for x in args.wallclock_seeds.split(",") if x]
        print(f"[bench] wall-clock to target: seeds {seeds}, {args.wallclock_generations} plain FP32 generations",
              file=sys.stderr, flush=True)
        recs, agg = run_configs.wallclock_to_target(seeds, args.wallclock_generations, cap_factor=2)
        out["wallclock_to_target"] = dict(agg, workload="64x64 grid, 4739 actions (the bench workload)",
                                          metric="speedup = plain FP32 seconds to its own best fitness after "
                                                 "the plain generations / the form's seconds to reach it",
                                          target_fitness=[r["target_fitness"] for r in recs],
                                          fp32_seconds_to_target=[r["forms"]["fp32"]["seconds_to_target"]
                                                                  for r in recs])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
