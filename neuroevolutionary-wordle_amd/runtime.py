"""ctypes mirror of include/ne_runtime.h and a thin Python wrapper named after the reference's
slab_device runtime calls (src/genetic_algorithm/device/slab_runtime.hpp:119-186).

The wrapper only marshals arguments; every call goes through the HIP library libne_wordle.so and
raises RuntimeError with the runtime's status when a call fails.  Loading fails loudly when the
library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libne_wordle.so")

STATUS_NAMES = {
    0: "ok", 1: "device failure", 2: "invalid runtime config", 3: "invalid slab", 4: "invalid generation",
    5: "invalid training shard", 6: "guess append failed", 7: "policy forward failed",
    8: "action selection failed", 9: "population not evaluated", 10: "invalid assembly config",
    11: "parent selection failed", 12: "invalid assembly plan", 13: "invalid parent index", 14: "slab full",
    15: "output embedding injection failed", 16: "slab repack failed", 17: "invalid argument",
    18: "unsupported configuration", 19: "callback failed", 20: "checkpoint invalid", 21: "checkpoint I/O failure",
}

EVAL_FP32, EVAL_BF16_RESCORED, EVAL_BF16_TRUNK, EVAL_FORCE_FALLBACK = 0, 1, 2, 0x100
# evaluation kernels of the untraced FP32 form (identical results; ne_eval_shape.kernel)
EVAL_KERNEL_REGISTERS, EVAL_KERNEL_STREAMED, EVAL_KERNEL_PAIRED = 0, 1, 2
TRACE_GAMES = 1
TRACE_POLICY = 2
TRACE_XOVER = 4

POLICY_PARAMS = 132928
TAIL = 38


class RuntimeConfig(C.Structure):
    _fields_ = [("genotype_slab_byte_budget", C.c_uint64), ("generation_byte_budget", C.c_uint64),
                ("host_spillover_byte_budget", C.c_uint64), ("action_count", C.c_uint64),
                ("population_size_ceiling", C.c_uint64), ("grid_column_count", C.c_uint64),
                ("encoder_hidden", C.c_uint32), ("reserved0", C.c_uint32), ("stream", C.c_void_p)]


class WordCountSchedule(C.Structure):
    _fields_ = [("initial_word_count", C.c_uint64), ("word_count_step", C.c_uint64),
                ("word_count_step_period_generations", C.c_uint64)]


class WordCounts(C.Structure):
    _fields_ = [("training_word_count", C.c_uint64), ("action_space_word_count", C.c_uint64),
                ("training_word_schedule", WordCountSchedule), ("shard_initial_radius", C.c_uint64),
                ("shard_radius_growth_period_generations", C.c_uint64)]


class ShardReleaseHistory(C.Structure):
    _fields_ = [("release_generations", C.POINTER(C.c_uint64)), ("release_count", C.c_uint64),
                ("release_fitness_p99", C.POINTER(C.c_float))]


class AssemblyConfig(C.Structure):
    _fields_ = [("cellular_breeding_radius", C.c_uint64), ("rank_exponent", C.c_float),
                ("crossover_temperature_level1", C.c_float), ("crossover_temperature_level2", C.c_float),
                ("crossover_temperature_level3", C.c_float), ("mutation_probability", C.c_float),
                ("mutation_sigma", C.c_float), ("output_tail_row_scale_mutation_probability", C.c_float),
                ("reserved0", C.c_float)]


class PendingInjection(C.Structure):
    _fields_ = [("enabled", C.c_uint32), ("reserved0", C.c_uint32), ("first_catalog_word_index", C.c_uint64),
                ("injection_count", C.c_uint64)]


class InitConfig(C.Structure):
    _fields_ = [("weight_variance_scale", C.c_float), ("weight_gain", C.c_float), ("tail_stddev", C.c_float),
                ("reserved0", C.c_float)]


class FitnessSummary(C.Structure):
    _fields_ = [("best_fitness", C.c_float), ("average_fitness", C.c_float), ("best_index", C.c_uint64),
                ("best_slot_index", C.c_uint32), ("reserved0", C.c_uint32), ("generation_index", C.c_uint64),
                ("action_count", C.c_uint64), ("population_size", C.c_uint64)]


class SlabInfo(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("action_count", "slot_stride_bytes", "slot_count", "slab_bytes",
                                          "policy_parameter_count", "max_generation_size",
                                          "current_generation_index", "current_generation_size",
                                          "next_generation_size", "free_slot_count")]


class StageTimes(C.Structure):
    _fields_ = [("evaluate_ms", C.c_float), ("finalize_ms", C.c_float), ("plan_ms", C.c_float),
                ("assemble_ms", C.c_float), ("repack_ms", C.c_float), ("inject_ms", C.c_float),
                ("bookkeeping_ms", C.c_float), ("prepare_ms", C.c_float), ("decisions", C.c_uint64),
                ("encoded_turns", C.c_uint64), ("games", C.c_uint64), ("children", C.c_uint64),
                ("rescored", C.c_uint64), ("fallbacks", C.c_uint64), ("spilled", C.c_uint64),
                ("step_graph_ms", C.c_float), ("step_graph", C.c_uint32), ("repack_survivors", C.c_uint64),
                ("repack_old_stride", C.c_uint64), ("repack_copy_ms", C.c_float), ("reserved0", C.c_uint32)]


class EvalShape(C.Structure):
    _fields_ = [("waves", C.c_uint32), ("rounds_per_item", C.c_uint32), ("workgroups_per_cu", C.c_uint32),
                ("kernel", C.c_uint32)]


GAME_TRACE_DTYPE = np.dtype([("guesses", "<u4", 6), ("best", "<f4", 6), ("second", "<f4", 6),
                             ("decisions", "<u4"), ("turn_count", "<u4"), ("score", "<f4"), ("won", "<u4")])

POST_FITNESS_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(WordCounts), C.c_void_p)
SHARD_RELEASE_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(WordCounts), C.POINTER(PendingInjection), C.c_void_p)


class ScreenConfig(C.Structure):
    _fields_ = [("brood_factor", C.c_uint32), ("replay_capacity", C.c_uint32), ("audit_fraction", C.c_float),
                ("spearman_threshold", C.c_float), ("learning_rate", C.c_float), ("train_steps", C.c_uint32),
                ("batch_size", C.c_uint32), ("min_replay", C.c_uint32), ("seed", C.c_uint32),
                ("flags", C.c_uint32)]


class ScreenReport(C.Structure):
    _fields_ = [("screened", C.c_uint32), ("next_screened", C.c_uint32), ("brood", C.c_uint32),
                ("evaluated", C.c_uint32), ("audited", C.c_uint32), ("audit_spearman", C.c_float),
                ("train_loss", C.c_float), ("fingerprint_ms", C.c_float), ("surrogate_ms", C.c_float),
                ("best_fitness", C.c_float), ("average_fitness", C.c_float), ("reserved0", C.c_uint32),
                ("replay_count", C.c_uint64)]


SURROGATE_PARAMS = 33025
SCREEN_EAGER_TRAINING = 1
FINGERPRINT_DIMS = 192


def default_screen_config(**kw):
    """The surrogate-screened configuration: brood 4x, top 25% (each cell's best-predicted child) plus a
    5% random audit, Adam lr 1e-3 on a replay buffer of 65,536 pairs, plain steps while the audit Spearman
    rank correlation is below 0.5."""
    c = ScreenConfig(brood_factor=4, replay_capacity=65536, audit_fraction=0.05, spearman_threshold=0.5,
                     learning_rate=1e-3, train_steps=32, batch_size=256, min_replay=4096, seed=0x5C2EE7, flags=0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class CheckpointInfo(C.Structure):
    _fields_ = ([(n, C.c_uint32) for n in ("schema_version", "genome_layout_version", "resume_phase",
                                          "generation_seed")] +
                [(n, C.c_uint64) for n in ("checksum", "training_data_identity_hash", "generation_index",
                                          "population", "child_count", "live_genotypes", "payload_bytes",
                                          "action_count", "slot_stride_bytes", "slot_count", "slab_bytes",
                                          "generation_byte_budget", "grid_column_count", "release_count")] +
                [("word_counts", WordCounts), ("assembly_config", AssemblyConfig), ("injection", PendingInjection)])


class FitnessTelemetry(C.Structure):
    _fields_ = ([(n, C.c_uint64) for n in ("generation", "population_size", "training_word_count")] +
                [(n, C.c_float) for n in ("fitness_min", "fitness_mean", "fitness_median", "fitness_p90",
                                          "fitness_p99", "fitness_max", "fitness_stddev")] +
                [("has_breeding_metrics", C.c_uint32)] +
                [(n, C.c_uint64) for n in ("distinct_fitness_count", "parent_childless_count",
                                          "parent_one_child_count", "parent_multiple_children_count")])


class ConvergenceTelemetry(C.Structure):
    _fields_ = ([(n, C.c_uint64) for n in ("generation", "sample_organisms", "sample_weights", "pair_count")] +
                [(n, C.c_float) for n in ("centroid_distance_mean", "centroid_distance_min", "centroid_distance_max",
                                          "pairwise_distance_mean", "pairwise_distance_min", "pairwise_distance_max",
                                          "elapsed_ms", "reserved0")])


class ShardReleasePolicy(C.Structure):
    _fields_ = [("min_gap_generations", C.c_uint64), ("first_new_release_generation", C.c_uint64),
                ("centroid_distance_threshold", C.c_float), ("fitness_p99_threshold", C.c_float),
                ("run_seed", C.c_uint32), ("reserved0", C.c_uint32)]


def default_release_policy(**kw):
    """The reference CLI's defaults (training_data.hpp:17-20)."""
    p = ShardReleasePolicy(min_gap_generations=3, first_new_release_generation=10, centroid_distance_threshold=6.0,
                           fitness_p99_threshold=0.20, run_seed=7, reserved0=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Board(C.Structure):
    _fields_ = [("turn_count", C.c_uint32), ("guesses", (C.c_uint8 * 5) * 6), ("tiles", (C.c_uint8 * 5) * 6)]


class SelectedAction(C.Structure):
    _fields_ = [("status", C.c_uint32), ("action_index", C.c_uint32), ("score", C.c_float),
                ("second_score", C.c_float), ("word", C.c_uint8 * 5), ("reserved", C.c_uint8 * 3)]


class WinnerMetadata(C.Structure):
    _fields_ = [("best_fitness", C.c_float), ("best_slot_index", C.c_uint32), ("generation_index", C.c_uint64),
                ("best_index", C.c_uint64), ("action_count", C.c_uint64), ("genome_byte_count", C.c_uint64),
                ("seed", C.c_uint32), ("reserved0", C.c_uint32), ("action_space_path", C.c_char * 1024),
                ("timestamp_local", C.c_char * 32)]


def make_board(turns):
    """An ne_board from [(word letters[5], tiles[5]), ...]."""
    b = Board()
    b.turn_count = len(turns)
    for t, (w, tl) in enumerate(turns):
        for p in range(5):
            b.guesses[t][p] = int(w[p])
            b.tiles[t][p] = int(tl[p])
    return b


class Callbacks(C.Structure):
    _fields_ = [("post_fitness_evaluation", POST_FITNESS_CB), ("training_shard_release", SHARD_RELEASE_CB),
                ("user", C.c_void_p)]


EXPORTED_SYMBOLS = [
    "ne_create", "ne_destroy", "ne_upload_catalog", "ne_upload_secrets_csr", "ne_clear_secrets", "ne_bootstrap",
    "ne_evaluate", "ne_read_summary", "ne_advance_generation", "ne_get_slab_info", "ne_download_slot",
    "ne_download_slab", "ne_swap_generations", "ne_upload_slot", "ne_download_generation", "ne_download_plan", "ne_download_local_word_counts",
    "ne_read_status", "ne_status_string", "ne_get_stage_times", "ne_set_trace", "ne_download_trace",
    "ne_build_shard_csr", "ne_set_eval_precision", "ne_enable_screening", "ne_advance_generation_screened",
    "ne_fingerprint", "ne_get_probe_states", "ne_surrogate_predict", "ne_surrogate_train_step",
    "ne_surrogate_get_params", "ne_surrogate_set_params", "ne_rank_correlation", "ne_autotune_evaluation",
    "ne_set_eval_shape", "ne_set_eval_shape_ex", "ne_get_eval_shape", "ne_set_step_graph", "ne_create_prebreeding_checkpoint", "ne_restore_prebreeding_checkpoint",
    "ne_resume_generation_from_checkpoint", "ne_write_checkpoint", "ne_read_checkpoint", "ne_clone_checkpoint",
    "ne_get_checkpoint_info", "ne_get_checkpoint_generation", "ne_get_checkpoint_plan", "ne_get_checkpoint_genome",
    "ne_free_checkpoint", "ne_training_data_identity_hash", "ne_collect_fitness_telemetry",
    "ne_collect_convergence_telemetry", "ne_convergence_sample_plan", "ne_maybe_release_training_shard",
    "ne_model_create", "ne_model_create_with_layout", "ne_model_genome_bytes", "ne_model_destroy", "ne_model_select_guess", "ne_model_select_guesses",
    "ne_model_select_guesses_noalloc",
    "ne_write_winner_artifact", "ne_write_winner_artifact_in_directory", "ne_read_winner_artifact",
]

_lib = None


def load_library(path=LIB_PATH):
    """Loads libne_wordle.so (raises if it was not built: there is no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP runtime library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    p, u64, u32, i32 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
    sig = {
        "ne_create": (i32, [C.POINTER(RuntimeConfig), C.POINTER(p)]),
        "ne_destroy": (None, [p]),
        "ne_upload_catalog": (i32, [p, p, u64]),
        "ne_upload_secrets_csr": (i32, [p, p, u64, p, u64]),
        "ne_clear_secrets": (i32, [p]),
        "ne_bootstrap": (i32, [p, u64, u32, u64, C.POINTER(InitConfig)]),
        "ne_evaluate": (i32, [p, C.POINTER(WordCounts), C.POINTER(ShardReleaseHistory)]),
        "ne_read_summary": (i32, [p, C.POINTER(FitnessSummary)]),
        "ne_advance_generation": (i32, [p, u32, C.POINTER(WordCounts), C.POINTER(AssemblyConfig),
                                        C.POINTER(PendingInjection), u32, C.POINTER(Callbacks),
                                        C.POINTER(ShardReleaseHistory)]),
        "ne_get_slab_info": (i32, [p, C.POINTER(SlabInfo)]),
        "ne_download_slot": (i32, [p, u32, p, u64, C.POINTER(u64)]),
        "ne_download_slab": (i32, [p, p, u64, p, p, u64, p]),
        "ne_swap_generations": (i32, [p]),
        "ne_upload_slot": (i32, [p, u32, p, u64]),
        "ne_download_generation": (i32, [p, i32, p, p, p, p, u64, C.POINTER(u64)]),
        "ne_download_plan": (i32, [p, p, u64, C.POINTER(u64)]),
        "ne_download_local_word_counts": (i32, [p, p, u64, C.POINTER(u64)]),
        "ne_read_status": (i32, [p]),
        "ne_status_string": (C.c_char_p, [i32]),
        "ne_get_stage_times": (i32, [p, C.POINTER(StageTimes)]),
        "ne_set_trace": (i32, [p, u32]),
        "ne_set_eval_precision": (i32, [p, u32]),
        "ne_download_trace": (i32, [p, u32, p, u64, C.POINTER(u64)]),
        "ne_build_shard_csr": (i32, [C.POINTER(WordCounts), C.POINTER(ShardReleaseHistory), u64, u64, u64, u64,
                                     p, p, u64, C.POINTER(u64)]),
        "ne_enable_screening": (i32, [p, C.POINTER(ScreenConfig)]),
        "ne_advance_generation_screened": (i32, [p, u32, C.POINTER(WordCounts), C.POINTER(AssemblyConfig),
                                                 C.POINTER(ShardReleaseHistory), C.POINTER(ScreenReport)]),
        "ne_fingerprint": (i32, [p, p, p, u64, p]),
        "ne_get_probe_states": (i32, [p, p, u64, C.POINTER(u64)]),
        "ne_surrogate_predict": (i32, [p, p, u64, p]),
        "ne_surrogate_train_step": (i32, [p, p, p, u64]),
        "ne_surrogate_get_params": (i32, [p, p, u64, C.POINTER(u64)]),
        "ne_surrogate_set_params": (i32, [p, p, u64]),
        "ne_rank_correlation": (C.c_double, [p, p, u64]),
        "ne_autotune_evaluation": (i32, [p, C.POINTER(WordCounts), C.POINTER(ShardReleaseHistory), C.POINTER(u32),
                                         p]),
        "ne_set_eval_shape": (i32, [p, u32]),
        "ne_set_eval_shape_ex": (i32, [p, C.POINTER(EvalShape)]),
        "ne_get_eval_shape": (i32, [p, C.POINTER(EvalShape)]),
        "ne_set_step_graph": (i32, [p, u32]),
        "ne_create_prebreeding_checkpoint": (i32, [p, u32, C.POINTER(WordCounts), C.POINTER(AssemblyConfig),
                                                   C.POINTER(PendingInjection), p, u64, u32, C.POINTER(Callbacks),
                                                   C.POINTER(ShardReleaseHistory), C.POINTER(p)]),
        "ne_restore_prebreeding_checkpoint": (i32, [p, p, C.POINTER(p)]),
        "ne_resume_generation_from_checkpoint": (i32, [p, p, u32]),
        "ne_write_checkpoint": (i32, [p, C.c_char_p]),
        "ne_read_checkpoint": (i32, [C.c_char_p, C.POINTER(p)]),
        "ne_clone_checkpoint": (i32, [p, C.POINTER(p)]),
        "ne_get_checkpoint_info": (i32, [p, C.POINTER(CheckpointInfo)]),
        "ne_get_checkpoint_generation": (i32, [p, p, p, p, p, u64]),
        "ne_get_checkpoint_plan": (i32, [p, p, u64]),
        "ne_get_checkpoint_genome": (i32, [p, u64, C.POINTER(u32), p, u64]),
        "ne_free_checkpoint": (None, [p]),
        "ne_training_data_identity_hash": (u64, [C.POINTER(WordCounts), C.POINTER(ShardReleaseHistory), p, u64]),
        "ne_collect_fitness_telemetry": (i32, [p, C.POINTER(WordCounts), u32, C.POINTER(FitnessTelemetry)]),
        "ne_collect_convergence_telemetry": (i32, [p, u32, C.POINTER(ConvergenceTelemetry)]),
        "ne_convergence_sample_plan": (i32, [u64, u64, u32, u64, p, p, p, p]),
        "ne_maybe_release_training_shard": (i32, [p, C.POINTER(WordCounts), C.POINTER(ShardReleasePolicy), u64, p, p,
                                                  C.POINTER(u64), u64, C.POINTER(PendingInjection),
                                                  C.POINTER(u64), C.POINTER(u32)]),
        "ne_model_create": (i32, [p, u64, p, u64, p, C.POINTER(p)]),
        "ne_model_destroy": (None, [p]),
        "ne_model_select_guess": (i32, [p, C.POINTER(Board), C.POINTER(SelectedAction)]),
        "ne_model_create_with_layout": (i32, [p, u64, C.POINTER(ModelLayout), p, u64, p, C.POINTER(p)]),
        "ne_model_genome_bytes": (u64, [C.POINTER(ModelLayout), u64]),
        "ne_model_select_guesses": (i32, [p, p, u64, p]),
        "ne_model_select_guesses_noalloc": (i32, [p, p, u64, p]),
        "ne_write_winner_artifact": (i32, [C.c_char_p, C.c_char_p, p, p, u64, C.POINTER(WinnerMetadata)]),
        "ne_write_winner_artifact_in_directory": (i32, [C.c_char_p, p, p, u64, C.POINTER(WinnerMetadata), p, p,
                                                        u64]),
        "ne_read_winner_artifact": (i32, [C.c_char_p, C.c_char_p, C.POINTER(WinnerMetadata), p, u64, p, u64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class RuntimeError_(RuntimeError):
    def __init__(self, call, status):
        super().__init__(f"{call} failed: status {status} ({STATUS_NAMES.get(status, 'unknown')})")
        self.status = status


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _history(history, fitness_p99=None):
    """(pointer to an ne_shard_release_history or None, arrays to keep alive) for release generations and,
    optionally, the fitness p99 at each release."""
    if history is None:
        return None, None
    gens = np.ascontiguousarray(history, dtype=np.uint64)
    p99 = None if fitness_p99 is None else np.ascontiguousarray(fitness_p99, dtype=np.float32)
    if p99 is not None and p99.size != gens.size:
        raise ValueError("one fitness p99 per release generation")
    h = ShardReleaseHistory(gens.ctypes.data_as(C.POINTER(C.c_uint64)), gens.size,
                            None if p99 is None else p99.ctypes.data_as(C.POINTER(C.c_float)))
    return C.byref(h), (h, gens, p99)


def default_assembly_config(**kw):
    c = AssemblyConfig(cellular_breeding_radius=2, rank_exponent=0.5, crossover_temperature_level1=0.02,
                       crossover_temperature_level2=0.01, crossover_temperature_level3=0.005,
                       mutation_probability=0.0001, mutation_sigma=0.02,
                       output_tail_row_scale_mutation_probability=0.0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def word_counts(training, actions, initial=None, step=0, period=1, radius0=0, growth=2):
    return WordCounts(training_word_count=training, action_space_word_count=actions,
                      training_word_schedule=WordCountSchedule(initial if initial is not None else training, step,
                                                               period),
                      shard_initial_radius=radius0, shard_radius_growth_period_generations=growth)


def slot_stride_bytes(action_count):
    return ((POLICY_PARAMS + TAIL * action_count) * 4 + 255) // 256 * 256


class SlabRuntime:
    """One runtime handle: TryCreateDeviceSlabGARuntimeBuffers ... DestroyDeviceSlabGARuntimeBuffers."""

    def __init__(self, slab_bytes, generation_bytes, action_count, grid_columns, population_ceiling=0,
                 host_spill_bytes=0):
        self.lib = load_library()
        cfg = RuntimeConfig(genotype_slab_byte_budget=slab_bytes, generation_byte_budget=generation_bytes,
                            host_spillover_byte_budget=host_spill_bytes, action_count=action_count,
                            population_size_ceiling=population_ceiling, grid_column_count=grid_columns,
                            encoder_hidden=0, reserved0=0, stream=None)
        h = C.c_void_p()
        st = self.lib.ne_create(C.byref(cfg), C.byref(h))
        if st != 0:
            raise RuntimeError_("ne_create", st)
        self.h = h
        self.grid_columns = grid_columns

    def close(self):
        if self.h:
            self.lib.ne_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, call, st):
        if st != 0:
            raise RuntimeError_(call, st)

    # --- inputs
    def upload_catalog(self, letters):
        letters = np.ascontiguousarray(letters, dtype=np.uint8)
        self._check("ne_upload_catalog", self.lib.ne_upload_catalog(self.h, _ptr(letters), letters.shape[0]))

    def upload_secrets_csr(self, offsets, secrets):
        offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
        secrets = np.ascontiguousarray(secrets, dtype=np.uint32)
        self._check("ne_upload_secrets_csr", self.lib.ne_upload_secrets_csr(
            self.h, _ptr(offsets), offsets.size - 1, _ptr(secrets), secrets.size))

    def clear_secrets(self):
        self._check("ne_clear_secrets", self.lib.ne_clear_secrets(self.h))

    # --- generation step
    def bootstrap(self, generation_size, seed, generation_index=0, variance_scale=2.0, gain=1.0, tail_std=0.05):
        ic = InitConfig(variance_scale, gain, tail_std, 0.0)
        self._check("ne_bootstrap", self.lib.ne_bootstrap(self.h, generation_size, seed, generation_index,
                                                          C.byref(ic)))

    def evaluate(self, counts, history=None):
        hist, self._hist_keep = _history(history)
        self._check("ne_evaluate", self.lib.ne_evaluate(self.h, C.byref(counts), hist))

    def advance(self, seed, counts, config=None, injection=None, callbacks=None, history=None, verbose=False):
        cfg = C.byref(config) if config is not None else None
        inj = C.byref(injection) if injection is not None else None
        cbs = C.byref(callbacks) if callbacks is not None else None
        hist, self._hist_keep = _history(history)
        self._check("ne_advance_generation",
                    self.lib.ne_advance_generation(self.h, seed, C.byref(counts), cfg, inj, 1 if verbose else 0, cbs,
                                                   hist))

    # --- telemetry and the curriculum controller
    def fitness_telemetry(self, counts, breeding=False):
        out = FitnessTelemetry()
        self._check("ne_collect_fitness_telemetry",
                    self.lib.ne_collect_fitness_telemetry(self.h, C.byref(counts), 1 if breeding else 0,
                                                          C.byref(out)))
        return out

    def convergence_telemetry(self, run_seed):
        out = ConvergenceTelemetry()
        self._check("ne_collect_convergence_telemetry",
                    self.lib.ne_collect_convergence_telemetry(self.h, run_seed, C.byref(out)))
        return out

    def maybe_release_shard(self, counts, policy, catalog_words, history):
        """TryMaybeReleaseTrainingDataShard: (released, injection, next_word_count); appends to `history`."""
        inj = PendingInjection()
        nxt = C.c_uint64()
        rel = C.c_uint32()
        self._check("ne_maybe_release_training_shard", self.lib.ne_maybe_release_training_shard(
            self.h, C.byref(counts), C.byref(policy), catalog_words, _ptr(history.generations), _ptr(history.p99),
            C.byref(history.count), history.generations.size, C.byref(inj), C.byref(nxt), C.byref(rel)))
        return bool(rel.value), inj, nxt.value

    # --- checkpoint and resume
    def create_checkpoint(self, seed, counts, config=None, injection=None, catalog=None, callbacks=None,
                          history=None, fitness_p99=None):
        """TryCreatePrebreedingCheckpointOnDevice: the first half of a generation step (evaluation, plan,
        widening) and the snapshot; `resume` finishes the step."""
        cfg = C.byref(config) if config is not None else None
        inj = C.byref(injection) if injection is not None else None
        cbs = C.byref(callbacks) if callbacks is not None else None
        hist, self._hist_keep = _history(history, fitness_p99)
        cat = None
        n = 0
        if catalog is not None:
            cat = np.ascontiguousarray(catalog, dtype=np.uint8).reshape(-1, 5)
            n = cat.shape[0]
        h = C.c_void_p()
        self._check("ne_create_prebreeding_checkpoint", self.lib.ne_create_prebreeding_checkpoint(
            self.h, seed, C.byref(counts), cfg, inj, None if cat is None else _ptr(cat), n, 0, cbs, hist,
            C.byref(h)))
        return Checkpoint(h)

    def resume(self, checkpoint):
        """TryResumeGenerationFromCheckpointOnDevice: assembly, injection and swap from the boundary."""
        self._check("ne_resume_generation_from_checkpoint",
                    self.lib.ne_resume_generation_from_checkpoint(self.h, checkpoint.h, 0))

    @classmethod
    def restore(cls, checkpoint):
        """TryRestorePrebreedingCheckpointToDevice: a new runtime at the checkpoint's boundary."""
        self = cls.__new__(cls)
        self.lib = load_library()
        self.h = None
        h = C.c_void_p()
        st = self.lib.ne_restore_prebreeding_checkpoint(checkpoint.h, None, C.byref(h))
        if st != 0:
            raise RuntimeError_("ne_restore_prebreeding_checkpoint", st)
        self.h = h
        self.grid_columns = checkpoint.info().grid_column_count
        return self

    def summary(self):
        s = FitnessSummary()
        self._check("ne_read_summary", self.lib.ne_read_summary(self.h, C.byref(s)))
        return s

    # --- transfers
    def info(self):
        s = SlabInfo()
        self._check("ne_get_slab_info", self.lib.ne_get_slab_info(self.h, C.byref(s)))
        return s

    def generation(self, which=0):
        info = self.info()
        cap = info.max_generation_size
        slots = np.zeros(cap, np.uint32)
        fit = np.zeros(cap, np.float32)
        evals = np.zeros(cap, np.uint32)
        has = np.zeros(cap, np.uint8)
        n = C.c_uint64()
        self._check("ne_download_generation", self.lib.ne_download_generation(
            self.h, which, _ptr(slots), _ptr(fit), _ptr(evals), _ptr(has), cap, C.byref(n)))
        k = n.value
        return slots[:k].copy(), fit[:k].copy(), evals[:k].copy(), has[:k].copy()

    def download_slot(self, slot):
        info = self.info()
        buf = np.zeros(info.slot_stride_bytes // 4, np.float32)
        w = C.c_uint64()
        self._check("ne_download_slot", self.lib.ne_download_slot(self.h, int(slot), _ptr(buf), buf.nbytes,
                                                                  C.byref(w)))
        return buf

    def download_slab(self, storage=True):
        """TryDownloadSlabFromDevice: (slot bytes [slot_count, stride] as uint8 or None, slot liveness counts,
        free-slot stack, free count)."""
        info = self.info()
        S = info.slot_count
        data = np.zeros((S, info.slot_stride_bytes), np.uint8) if storage else None
        states = np.zeros(S, np.uint32)
        stack = np.zeros(S, np.uint32)
        fc = C.c_uint32()
        self._check("ne_download_slab", self.lib.ne_download_slab(
            self.h, _ptr(data) if storage else None, data.nbytes if storage else 0, _ptr(states), _ptr(stack), S,
            C.byref(fc)))
        return data, states, stack, int(fc.value)

    def swap_generations(self):
        """SwapDeviceSlabGenerationBuffers."""
        self._check("ne_swap_generations", self.lib.ne_swap_generations(self.h))

    def upload_slot(self, slot, values):
        values = np.ascontiguousarray(values, dtype=np.float32)
        self._check("ne_upload_slot", self.lib.ne_upload_slot(self.h, int(slot), _ptr(values), values.nbytes))

    def plan(self):
        info = self.info()
        cap = info.max_generation_size
        pairs = np.zeros((cap, 2), np.uint32)
        n = C.c_uint64()
        self._check("ne_download_plan", self.lib.ne_download_plan(self.h, _ptr(pairs), cap, C.byref(n)))
        return pairs[:n.value].copy()

    def local_word_counts(self):
        info = self.info()
        out = np.zeros(info.current_generation_size, np.uint32)
        n = C.c_uint64()
        self._check("ne_download_local_word_counts",
                    self.lib.ne_download_local_word_counts(self.h, _ptr(out), out.size, C.byref(n)))
        return out

    def stage_times(self):
        s = StageTimes()
        self._check("ne_get_stage_times", self.lib.ne_get_stage_times(self.h, C.byref(s)))
        return s

    def set_trace(self, flags):
        self._check("ne_set_trace", self.lib.ne_set_trace(self.h, flags))

    def set_eval_precision(self, mode):
        """EVAL_FP32, EVAL_BF16_RESCORED or EVAL_BF16_TRUNK (the BF16 ones optionally | EVAL_FORCE_FALLBACK)."""
        self._check("ne_set_eval_precision", self.lib.ne_set_eval_precision(self.h, mode))

    def download_trace(self, kind, count):
        if kind == TRACE_GAMES:
            out = np.zeros(count, GAME_TRACE_DTYPE)
        elif kind == TRACE_POLICY:
            out = np.zeros((count, 6, 64), np.float32)
        else:
            out = np.zeros((count, 16), np.uint32)
        w = C.c_uint64()
        self._check("ne_download_trace", self.lib.ne_download_trace(self.h, kind, _ptr(out), out.nbytes,
                                                                    C.byref(w)))
        return out

    def status(self):
        return self.lib.ne_read_status(self.h)

    # --- evaluation shape
    def autotune_evaluation(self, counts):
        """Coordinate search over the evaluation's launch shape (waves 1..4, rounds per item, workgroups per
        CU); keeps the fastest; returns (waves, [ms for 1..4 waves], the chosen EvalShape)."""
        w = C.c_uint32()
        t = np.zeros(4, np.float32)
        self._check("ne_autotune_evaluation",
                    self.lib.ne_autotune_evaluation(self.h, C.byref(counts), None, C.byref(w), _ptr(t)))
        return w.value, t.tolist(), self.eval_shape()

    def eval_shape(self):
        s = EvalShape()
        self._check("ne_get_eval_shape", self.lib.ne_get_eval_shape(self.h, C.byref(s)))
        return s

    def set_eval_shape_ex(self, waves=0, rounds_per_item=1, workgroups_per_cu=0, kernel=0):
        s = EvalShape(waves=waves, rounds_per_item=rounds_per_item, workgroups_per_cu=workgroups_per_cu,
                      kernel=kernel)
        self._check("ne_set_eval_shape_ex", self.lib.ne_set_eval_shape_ex(self.h, C.byref(s)))

    def set_eval_shape(self, waves):
        self._check("ne_set_eval_shape", self.lib.ne_set_eval_shape(self.h, waves))

    def set_step_graph(self, enabled=True):
        """Plain steps as one captured hipGraph each (bit-identical to the eager step)."""
        self._check("ne_set_step_graph", self.lib.ne_set_step_graph(self.h, 1 if enabled else 0))

    # --- surrogate screening
    def enable_screening(self, config=None):
        cfg = config if config is not None else default_screen_config()
        self._check("ne_enable_screening", self.lib.ne_enable_screening(self.h, C.byref(cfg)))

    def advance_screened(self, seed, counts, config=None, history=None):
        cfg = C.byref(config) if config is not None else None
        hist, self._hist_keep = _history(history)
        rep = ScreenReport()
        self._check("ne_advance_generation_screened",
                    self.lib.ne_advance_generation_screened(self.h, seed, C.byref(counts), cfg, hist, C.byref(rep)))
        return rep

    def fingerprint(self, slots, parent_fitness):
        slots = np.ascontiguousarray(slots, dtype=np.uint32)
        pf = np.ascontiguousarray(parent_fitness, dtype=np.float32).reshape(-1, 2)
        out = np.zeros((slots.size, FINGERPRINT_DIMS), np.float32)
        self._check("ne_fingerprint", self.lib.ne_fingerprint(self.h, _ptr(slots), _ptr(pf), slots.size, _ptr(out)))
        return out

    def probe_states(self):
        """[16][4]: per probe board, the secret's catalog index then its 3 turns' guess indices."""
        out = np.zeros((16, 4), np.uint32)
        n = C.c_uint64()
        self._check("ne_get_probe_states", self.lib.ne_get_probe_states(self.h, _ptr(out), out.size, C.byref(n)))
        return out

    def surrogate_predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FINGERPRINT_DIMS)
        y = np.zeros(x.shape[0], np.float32)
        self._check("ne_surrogate_predict", self.lib.ne_surrogate_predict(self.h, _ptr(x), x.shape[0], _ptr(y)))
        return y

    def surrogate_train_step(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FINGERPRINT_DIMS)
        t = np.ascontiguousarray(t, dtype=np.float32)
        self._check("ne_surrogate_train_step",
                    self.lib.ne_surrogate_train_step(self.h, _ptr(x), _ptr(t), x.shape[0]))

    def surrogate_params(self):
        out = np.zeros(SURROGATE_PARAMS, np.float32)
        n = C.c_uint64()
        self._check("ne_surrogate_get_params",
                    self.lib.ne_surrogate_get_params(self.h, _ptr(out), out.size, C.byref(n)))
        return out

    def set_surrogate_params(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        self._check("ne_surrogate_set_params", self.lib.ne_surrogate_set_params(self.h, _ptr(params), params.size))


class Checkpoint:
    """A pre-recombination checkpoint (RuntimeCheckpoint, slab_runtime.hpp:98-117), host memory only.

    Files are those of TryWriteRuntimeCheckpointAtomically / TryReadRuntimeCheckpoint
    (slab_runtime.cu:1934-2081)."""

    def __init__(self, handle):
        self.lib = load_library()
        self.h = handle

    @classmethod
    def read(cls, path):
        lib = load_library()
        h = C.c_void_p()
        st = lib.ne_read_checkpoint(os.fsencode(path), C.byref(h))
        if st != 0:
            raise RuntimeError_("ne_read_checkpoint", st)
        return cls(h)

    def write(self, path):
        st = self.lib.ne_write_checkpoint(self.h, os.fsencode(path))
        if st != 0:
            raise RuntimeError_("ne_write_checkpoint", st)

    def clone(self):
        h = C.c_void_p()
        st = self.lib.ne_clone_checkpoint(self.h, C.byref(h))
        if st != 0:
            raise RuntimeError_("ne_clone_checkpoint", st)
        return Checkpoint(h)

    def info(self):
        i = CheckpointInfo()
        st = self.lib.ne_get_checkpoint_info(self.h, C.byref(i))
        if st != 0:
            raise RuntimeError_("ne_get_checkpoint_info", st)
        return i

    def generation(self):
        """(slot_indices, fitness, evaluation_counts, has_fitness) of the current generation."""
        n = self.info().population
        slots, fit = np.zeros(n, np.uint32), np.zeros(n, np.float32)
        evals, has = np.zeros(n, np.uint32), np.zeros(n, np.uint8)
        st = self.lib.ne_get_checkpoint_generation(self.h, _ptr(slots), _ptr(fit), _ptr(evals), _ptr(has), n)
        if st != 0:
            raise RuntimeError_("ne_get_checkpoint_generation", st)
        return slots, fit, evals, has

    def plan(self):
        n = self.info().child_count
        pairs = np.zeros((n, 2), np.uint32)
        st = self.lib.ne_get_checkpoint_plan(self.h, _ptr(pairs), n)
        if st != 0:
            raise RuntimeError_("ne_get_checkpoint_plan", st)
        return pairs

    def genome(self, record):
        """(organism index, fp32 view of the record's slot bytes)."""
        stride = self.info().slot_stride_bytes
        out = np.zeros(stride // 4, np.float32)
        org = C.c_uint32()
        st = self.lib.ne_get_checkpoint_genome(self.h, record, C.byref(org), _ptr(out), stride)
        if st != 0:
            raise RuntimeError_("ne_get_checkpoint_genome", st)
        return org.value, out

    def close(self):
        if self.h:
            self.lib.ne_free_checkpoint(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def training_data_identity_hash(counts, history=None, fitness_p99=None, catalog=None):
    """The identity hash a checkpoint records for its training data (host only)."""
    lib = load_library()
    hist, keep = _history(history, fitness_p99)
    cat = None
    n = 0
    if catalog is not None:
        cat = np.ascontiguousarray(catalog, dtype=np.uint8).reshape(-1, 5)
        n = cat.shape[0]
    return int(lib.ne_training_data_identity_hash(C.byref(counts), hist, None if cat is None else _ptr(cat), n))


def convergence_sample_plan(population, trainable_scalars, run_seed, generation):
    """MakeSamplePlan through the library (host only): (organisms, scalar indices, [pairs][2])."""
    lib = load_library()
    org = np.zeros(256, np.uint32)
    w = np.zeros(8192, np.uint32)
    pr = np.zeros((512, 2), np.uint32)
    counts = np.zeros(3, np.uint64)
    st = lib.ne_convergence_sample_plan(population, trainable_scalars, run_seed, generation, _ptr(org), _ptr(w),
                                        _ptr(pr), _ptr(counts))
    if st != 0:
        raise RuntimeError_("ne_convergence_sample_plan", st)
    return org[:int(counts[0])], w[:int(counts[1])], pr[:int(counts[2])]


class ReleaseHistory:
    """A caller-owned shard release history (generations and fitness p99), as the reference CLI keeps it."""

    def __init__(self, capacity=4739):
        self.generations = np.zeros(capacity, np.uint64)
        self.p99 = np.zeros(capacity, np.float32)
        self.count = C.c_uint64(0)

    def record(self, generation, p99):
        self.generations[self.count.value] = generation
        self.p99[self.count.value] = p99
        self.count.value += 1

    def releases(self):
        n = self.count.value
        return self.generations[:n].tolist(), self.p99[:n].tolist()


NOALLOC_MAX_BOARDS = 8192  # NE_NOALLOC_MAX_BOARDS
STORAGE_FP32, STORAGE_FP16 = 0, 1


class ModelLayout(C.Structure):
    _fields_ = [("encoder_hidden", C.c_uint32), ("storage", C.c_uint32), ("reserved", C.c_uint32 * 2)]


def model_genome_bytes(encoder_hidden, storage, action_count):
    lay = ModelLayout(encoder_hidden=encoder_hidden, storage=storage)
    return int(load_library().ne_model_genome_bytes(C.byref(lay), action_count))


class Model:
    """TryCreateSingleModelDeviceRuntime ... DestroySingleModelDeviceRuntime: one genome serving guesses."""

    def __init__(self, genome, action_letters, encoder_hidden=None, storage=None):
        """genome: this build's fp32 slot bytes (145 -> 64 encoder), or -- with encoder_hidden and storage
        given -- the bytes of that layout (e.g. 128, STORAGE_FP16: the reference's own genome)."""
        self.lib = load_library()
        self.h = None
        g = np.ascontiguousarray(genome).view(np.uint8)
        w = np.ascontiguousarray(action_letters, dtype=np.uint8).reshape(-1, 5)
        h = C.c_void_p()
        if encoder_hidden is None and storage is None:
            st = self.lib.ne_model_create(_ptr(g), g.size, _ptr(w), w.shape[0], None, C.byref(h))
        else:
            lay = ModelLayout(encoder_hidden=encoder_hidden or 0, storage=storage or STORAGE_FP32)
            st = self.lib.ne_model_create_with_layout(_ptr(g), g.size, C.byref(lay), _ptr(w), w.shape[0], None,
                                                      C.byref(h))
        if st != 0:
            raise RuntimeError_("ne_model_create", st)
        self.h = h
        self.action_count = w.shape[0]

    def select(self, turns):
        """TrySelectNextGuessWithSingleModelDeviceRuntime on one board: an ne_selected_action."""
        out = SelectedAction()
        self.lib.ne_model_select_guess(self.h, C.byref(make_board(turns)), C.byref(out))
        return out

    def select_many(self, boards):
        """Many boards in one launch (a list of ne_board or of turn lists): a SelectedAction array."""
        arr = (Board * len(boards))(*[b if isinstance(b, Board) else make_board(b) for b in boards])
        out = (SelectedAction * len(boards))()
        st = self.lib.ne_model_select_guesses(self.h, C.cast(arr, C.c_void_p), len(boards), C.cast(out, C.c_void_p))
        if st != 0:
            raise RuntimeError_("ne_model_select_guesses", st)
        return out

    def select_many_noalloc(self, boards):
        """select_many through ne_model_select_guesses_noalloc (no device allocation; at most
        NOALLOC_MAX_BOARDS boards per call)."""
        arr = (Board * len(boards))(*[b if isinstance(b, Board) else make_board(b) for b in boards])
        out = (SelectedAction * len(boards))()
        st = self.lib.ne_model_select_guesses_noalloc(self.h, C.cast(arr, C.c_void_p), len(boards),
                                                      C.cast(out, C.c_void_p))
        if st != 0:
            raise RuntimeError_("ne_model_select_guesses_noalloc", st)
        return out

    def close(self):
        if self.h:
            self.lib.ne_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def winner_metadata(best_fitness, generation_index, best_index, best_slot_index, action_count, seed,
                    action_space_path=""):
    m = WinnerMetadata(best_fitness=best_fitness, best_slot_index=best_slot_index, generation_index=generation_index,
                       best_index=best_index, action_count=action_count,
                       genome_byte_count=slot_stride_bytes(action_count), seed=seed)
    m.action_space_path = action_space_path.encode()
    return m


def write_winner_artifact(binary_path, metadata_path, genome, action_letters, metadata):
    """TryWriteWinnerArtifactToPaths; returns the metadata with its timestamp."""
    lib = load_library()
    g = np.ascontiguousarray(genome).view(np.uint8)
    w = np.ascontiguousarray(action_letters, dtype=np.uint8).reshape(-1, 5)
    st = lib.ne_write_winner_artifact(os.fsencode(binary_path), os.fsencode(metadata_path), _ptr(g), _ptr(w),
                                      w.shape[0], C.byref(metadata))
    if st != 0:
        raise RuntimeError_("ne_write_winner_artifact", st)
    return metadata


def write_winner_artifact_in_directory(directory, genome, action_letters, metadata):
    """TryWriteWinnerArtifact: (binary path, metadata path) of a new winner-... pair in directory."""
    lib = load_library()
    g = np.ascontiguousarray(genome).view(np.uint8)
    w = np.ascontiguousarray(action_letters, dtype=np.uint8).reshape(-1, 5)
    b, m = C.create_string_buffer(4096), C.create_string_buffer(4096)
    st = lib.ne_write_winner_artifact_in_directory(os.fsencode(directory), _ptr(g), _ptr(w), w.shape[0],
                                                   C.byref(metadata), b, m, 4096)
    if st != 0:
        raise RuntimeError_("ne_write_winner_artifact_in_directory", st)
    return b.value.decode(), m.value.decode()


def read_winner_artifact(binary_path, metadata_path):
    """TryReadWinnerArtifact: (metadata, fp32 genome, action letters [n][5])."""
    lib = load_library()
    meta = WinnerMetadata()
    b, m = os.fsencode(binary_path), os.fsencode(metadata_path)
    st = lib.ne_read_winner_artifact(b, m, C.byref(meta), None, 0, None, 0)
    if st != 0:
        raise RuntimeError_("ne_read_winner_artifact", st)
    genome = np.zeros(meta.genome_byte_count // 4, np.float32)
    letters = np.zeros((meta.action_count, 5), np.uint8)
    st = lib.ne_read_winner_artifact(b, m, C.byref(meta), _ptr(genome), genome.nbytes, _ptr(letters),
                                     meta.action_count)
    if st != 0:
        raise RuntimeError_("ne_read_winner_artifact", st)
    return meta, genome, letters


def rank_correlation(a, b):
    """Spearman rank correlation through the library's host routine (average ranks for ties)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return float(load_library().ne_rank_correlation(_ptr(a), _ptr(b), a.size))


def build_shard_csr(counts, generation_index, rows, cols, epicenter_rows=None, history=None):
    """Host-only shard schedule -> (offsets, ids) through ne_build_shard_csr."""
    lib = load_library()
    epi = rows if epicenter_rows is None else epicenter_rows
    offsets = np.zeros(rows * cols + 1, np.uint32)
    n = C.c_uint64()
    hist, keep = _history(history)
    st = lib.ne_build_shard_csr(C.byref(counts), hist, generation_index, rows, cols, epi, _ptr(offsets), None, 0,
                                C.byref(n))
    if st not in (0, 17):
        raise RuntimeError_("ne_build_shard_csr", st)
    ids = np.zeros(max(n.value, 1), np.uint32)
    st = lib.ne_build_shard_csr(C.byref(counts), hist, generation_index, rows, cols, epi, _ptr(offsets), _ptr(ids),
                                ids.size, C.byref(n))
    if st != 0:
        raise RuntimeError_("ne_build_shard_csr", st)
    return offsets, ids[:n.value]
