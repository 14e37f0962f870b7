/* ne_runtime.h -- C ABI of the MI355X GA generation-step runtime.
 *
 * Each entry point replaces one call of the reference's device runtime
 * (src/genetic_algorithm/device/slab_runtime.hpp and
 * src/genetic_algorithm/genotype_slab/device_runtime.hpp); the replaced call is cited beside it.
 * Plain pointers and sizes only.  Every call returns an ne_status (0 = ok) and has finished all of
 * its device work when it returns.  A handle is not thread-safe; different handles are independent.
 * All host buffers are owned by the caller.
 */
#ifndef NE_RUNTIME_H
#define NE_RUNTIME_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  0..16 are the reference's DeviceSlabGARuntimeStatusCode values
 * (slab_runtime.hpp:21-39); codes above 16 are new. */
typedef enum ne_status {
    NE_OK = 0,
    NE_DEVICE_FAILURE = 1,            /* kCudaFailure: a HIP call failed */
    NE_INVALID_RUNTIME_CONFIG = 2,
    NE_INVALID_SLAB = 3,
    NE_INVALID_GENERATION = 4,
    NE_INVALID_TRAINING_SHARD = 5,
    NE_GUESS_APPEND_FAILED = 6,
    NE_POLICY_FORWARD_FAILED = 7,
    NE_ACTION_SELECTION_FAILED = 8,
    NE_POPULATION_NOT_EVALUATED = 9,
    NE_INVALID_ASSEMBLY_CONFIG = 10,
    NE_PARENT_SELECTION_FAILED = 11,
    NE_INVALID_ASSEMBLY_PLAN = 12,
    NE_INVALID_PARENT_INDEX = 13,
    NE_SLAB_FULL = 14,
    NE_OUTPUT_EMBEDDING_INJECTION_FAILED = 15,
    NE_SLAB_REPACK_FAILED = 16,
    NE_INVALID_ARGUMENT = 17,         /* null pointer, short buffer, size out of range */
    NE_UNSUPPORTED = 18,              /* a configuration this build does not implement */
    NE_CALLBACK_FAILED = 19,          /* a caller callback returned non-zero */
    NE_CHECKPOINT_INVALID = 20,       /* a checkpoint or winner artifact failed validation */
    NE_IO_FAILURE = 21                /* a checkpoint file could not be opened, read or written */
} ne_status;

typedef struct ne_runtime ne_runtime;

/* DeviceSlabGARuntimeConfig (slab_runtime.hpp:41-48) plus layout parameters. */
typedef struct ne_runtime_config {
    uint64_t genotype_slab_byte_budget;   /* slab bytes; slot count = budget / slot stride */
    uint64_t generation_byte_budget;      /* bytes one generation may occupy */
    uint64_t host_spillover_byte_budget;  /* host bytes for assembled children when the slab fills (0: none) */
    uint64_t action_count;                /* A: output-embedding rows per genome */
    uint64_t population_size_ceiling;     /* 0 = none; else a multiple of grid_column_count */
    uint64_t grid_column_count;           /* columns of the toroidal grid */
    uint32_t encoder_hidden;              /* 0: 145->64 turn encoder (only value built) */
    uint32_t reserved0;
    void *stream;                         /* hipStream_t to run on, or NULL for a private stream */
} ne_runtime_config;

/* WordCountSchedule (src/training_folder/training_data.hpp:24-28). */
typedef struct ne_word_count_schedule {
    uint64_t initial_word_count;
    uint64_t word_count_step;
    uint64_t word_count_step_period_generations;
} ne_word_count_schedule;

/* RuntimeWordCounts (src/genetic_algorithm/device/runtime_common.hpp:10-16). */
typedef struct ne_word_counts {
    uint64_t training_word_count;
    uint64_t action_space_word_count;
    ne_word_count_schedule training_word_schedule;
    uint64_t shard_initial_radius;
    uint64_t shard_radius_growth_period_generations;
} ne_word_counts;

/* TrainingDataShardReleaseHistory (training_data.hpp:49-53): release generation per shard and the
 * fitness p99 at its release (kept in checkpoints only; NULL reads as zeros). */
typedef struct ne_shard_release_history {
    const uint64_t *release_generations;
    uint64_t release_count;
    const float *release_fitness_p99;
} ne_shard_release_history;

/* ParentSelectionConfig / BreedingConfig / MutationConfig
 * (selection.hpp:14-17, breeding.hpp:10-14, mutation.hpp:11-15). */
typedef struct ne_assembly_config {
    uint64_t cellular_breeding_radius;    /* reference default 2 */
    float rank_exponent;                  /* reference default 0.5 */
    float crossover_temperature_level1;   /* 0.02 */
    float crossover_temperature_level2;   /* 0.01 */
    float crossover_temperature_level3;   /* 0.005 */
    float mutation_probability;           /* 0.0001 */
    float mutation_sigma;                 /* 0.02 */
    float output_tail_row_scale_mutation_probability; /* 0 */
    float reserved0;
} ne_assembly_config;

/* PendingOutputEmbeddingInjection (genotype_slab/device_runtime.hpp:26-32). */
typedef struct ne_pending_injection {
    uint32_t enabled;
    uint32_t reserved0;
    uint64_t first_catalog_word_index;    /* must equal the current action count */
    uint64_t injection_count;
} ne_pending_injection;

/* Generation-0 initialisation: weights ~ N(0, s_l^2) with s_l = gain*sqrt(scale/fan_in),
 * biases 0, tails ~ N(0, tail_stddev^2).  The reference uses scale 2, gain 1, tail 0.05
 * (genome_ops.cuh:58-133). */
typedef struct ne_init_config {
    float weight_variance_scale;
    float weight_gain;
    float tail_stddev;
    float reserved0;
} ne_init_config;

/* PopulationFitnessSummary (runtime_common.hpp:18-26). */
typedef struct ne_fitness_summary {
    float best_fitness;
    float average_fitness;
    uint64_t best_index;
    uint32_t best_slot_index;
    uint32_t reserved0;
    uint64_t generation_index;
    uint64_t action_count;
    uint64_t population_size;
} ne_fitness_summary;

/* GenotypeSlabLayout (genotype_slab/layout.hpp:20-25) and generation shape. */
typedef struct ne_slab_info {
    uint64_t action_count;
    uint64_t slot_stride_bytes;
    uint64_t slot_count;
    uint64_t slab_bytes;
    uint64_t policy_parameter_count;      /* fp32 values before the tail rows */
    uint64_t max_generation_size;
    uint64_t current_generation_index;
    uint64_t current_generation_size;
    uint64_t next_generation_size;
    uint64_t free_slot_count;
} ne_slab_info;

/* Device timings of the last call (milliseconds, HIP events) and on-device work counters. */
typedef struct ne_stage_times {
    float evaluate_ms;                    /* game-play kernel */
    float finalize_ms;                    /* fitness normalisation + summary */
    float plan_ms;                        /* parent selection */
    float assemble_ms;                    /* crossover + mutation (all batches) */
    float repack_ms;                      /* widening repack */
    float inject_ms;                      /* output-embedding injection */
    float bookkeeping_ms;                 /* refcount, GC, releases */
    float prepare_ms;                     /* per-evaluation encoder transpose */
    uint64_t decisions;                   /* policy decisions taken by the last evaluation */
    uint64_t encoded_turns;               /* turns run through the encoder */
    uint64_t games;                       /* episodes played */
    uint64_t children;                    /* children assembled by the last advance */
    uint64_t rescored;                    /* BF16 mode: candidates rescored in FP32 */
    uint64_t fallbacks;                   /* BF16 mode: decisions settled by the FP32 scan */
    uint64_t spilled;                     /* children the last advance held in host memory (slab full) */
    float step_graph_ms;                  /* captured step (ne_set_step_graph): the whole graph's replay;
                                             the per-stage fields above are then 0 */
    uint32_t step_graph;                  /* 0: eager step, 1: a captured step replayed, 2: captured, then replayed */
    uint64_t repack_survivors;            /* widening: parents moved to the wider layout */
    uint64_t repack_old_stride;           /* widening: their slot stride before the repack (bytes) */
    float repack_copy_ms;                 /* widening: the gather and scatter kernels alone */
    uint32_t reserved0;
} ne_stage_times;

/* Evaluation precision (ne_set_eval_precision).  NE_EVAL_FP32 is the reference's arithmetic.
 * NE_EVAL_BF16_RESCORED screens every action on BF16 MFMA (FP32 accumulate), then rescores each action
 * whose BF16 score lies within a proven error bound of the best in FP32 with the very contraction of the
 * FP32 mode: guesses, chosen scores and fitness equal the FP32 mode's bit for bit. */
enum {
    NE_EVAL_FP32 = 0,
    NE_EVAL_BF16_RESCORED = 1,
    NE_EVAL_BF16_TRUNK = 2,          /* BF16_RESCORED with the trunk's layers 1 and 2 (256->128->64) on BF16
                                        MFMA, FP32 accumulate: the candidates are rescored in FP32 against
                                        the policy this trunk produces, so guesses equal the FP32 mode's
                                        wherever the trunk's rounding does not reorder the top candidates
                                        (agreement is measured, not exact) */
    NE_EVAL_FORCE_FALLBACK = 0x100   /* with BF16: settle every decision by the FP32 scan (diagnostic) */
};

/* Callbacks of TryAdvanceGenerationOnDevice (slab_runtime.hpp:81-84).  Non-zero return = failure. */
typedef struct ne_callbacks {
    int (*post_fitness_evaluation)(ne_runtime *rt, const ne_word_counts *counts, void *user);
    int (*training_shard_release)(ne_runtime *rt, const ne_word_counts *counts,
                                  ne_pending_injection *inout_injection, void *user);
    void *user;
} ne_callbacks;

/* Trace kinds for ne_download_trace (parity mode, see ne_set_trace). */
enum {
    NE_TRACE_GAMES = 1,        /* per game: ne_game_trace */
    NE_TRACE_POLICY = 2,       /* per game and decision: 64 fp32 policy values */
    NE_TRACE_XOVER = 4         /* per child: 16 uint32 crossover decision words */
};

typedef struct ne_game_trace {
    uint32_t guesses[6];       /* action index per decision, 0xFFFFFFFF when unused */
    float best[6];             /* fp32 score of the chosen action */
    float second[6];           /* best fp32 score among the other unmasked actions */
    uint32_t decisions;
    uint32_t turn_count;
    float score;
    uint32_t won;
} ne_game_trace;

/* ---- lifecycle ---------------------------------------------------------------------------- */
/* TryCreateDeviceSlabGARuntimeBuffers (slab_runtime.hpp:119). */
int ne_create(const ne_runtime_config *config, ne_runtime **out_runtime);
/* DestroyDeviceSlabGARuntimeBuffers (slab_runtime.hpp:121). */
void ne_destroy(ne_runtime *rt);

/* UploadTrainingWordCatalogToDeviceConstantMemory (training_data.hpp:375-381): letters[5*n], each
 * 0..25.  The catalog belongs to the handle, not to the process. */
int ne_upload_catalog(ne_runtime *rt, const uint8_t *letters, uint64_t word_count);

/* Per-cell training secrets as CSR: cell c plays secret_ids[cell_offsets[c] .. cell_offsets[c+1]).
 * Replaces the shard runtime set upload (slab_runtime.cu:207-236) when set; ne_clear_secrets
 * returns to the reference shard schedule built from ne_word_counts each evaluation. */
int ne_upload_secrets_csr(ne_runtime *rt, const uint32_t *cell_offsets, uint64_t cell_count,
                          const uint32_t *secret_ids, uint64_t secret_count);
int ne_clear_secrets(ne_runtime *rt);

/* TryBootstrapRandomCurrentGenerationOnDevice (slab_runtime.hpp:123-124).  init may be NULL. */
int ne_bootstrap(ne_runtime *rt, uint64_t generation_size, uint32_t generation_seed,
                 uint64_t generation_index, const ne_init_config *init);

/* TryEvaluateCurrentGenerationFitnessOnDevice (slab_runtime.hpp:137-139). */
int ne_evaluate(ne_runtime *rt, const ne_word_counts *counts, const ne_shard_release_history *history);

/* TryReadPopulationFitnessSummaryFromDevice (slab_runtime.hpp:141-142). */
int ne_read_summary(ne_runtime *rt, ne_fitness_summary *out);

/* TryAdvanceGenerationOnDevice (slab_runtime.hpp:144-150): evaluate, plan, widen+inject when
 * injection->enabled, assemble, swap.  config/injection/callbacks/history may be NULL. */
int ne_advance_generation(ne_runtime *rt, uint32_t generation_seed, const ne_word_counts *counts,
                          const ne_assembly_config *config, const ne_pending_injection *injection,
                          uint32_t verbose, const ne_callbacks *callbacks,
                          const ne_shard_release_history *history);

/* ---- transfers ---------------------------------------------------------------------------- */
int ne_get_slab_info(ne_runtime *rt, ne_slab_info *out);
/* TryDownloadSlabSlotBytesFromDevice (slab_runtime.hpp:128-129). */
int ne_download_slot(ne_runtime *rt, uint32_t slot_index, void *dst, uint64_t capacity, uint64_t *written);
/* TryDownloadSlabFromDevice (slab_runtime.hpp:126; genotype_slab/device_runtime.hpp:91): the whole slab,
 * any part of it -- slot_count x slot_stride_bytes of slot bytes into `storage` (storage_capacity bytes),
 * each slot's liveness count (slot_states; 0 = free), the free-slot stack (free_stack; the next slot handed
 * out is free_stack[*free_count - 1]) and its count.  slot_states and free_stack hold slot_capacity
 * entries; any output pointer may be NULL. */
int ne_download_slab(ne_runtime *rt, void *storage, uint64_t storage_capacity, uint32_t *slot_states,
                     uint32_t *free_stack, uint64_t slot_capacity, uint32_t *free_count);
/* SwapDeviceSlabGenerationBuffers (slab_runtime.hpp:181; slab_runtime.cu:2128-2135): the current and next
 * generation exchange slot handles, fitness, evaluation counts, has-fitness flags, indices and sizes (a
 * generation step ends with this swap; calling it again exchanges them back). */
int ne_swap_generations(ne_runtime *rt);
/* TryUploadSlabToDevice for one slot (genotype_slab/device_runtime.hpp:89); the slot must be occupied. */
int ne_upload_slot(ne_runtime *rt, uint32_t slot_index, const void *src, uint64_t byte_count);
/* TryDownloadCurrent/NextGenerationFromDevice (slab_runtime.hpp:131-135); which: 0 current,
 * 1 next.  Any output pointer may be NULL.  Arrays hold capacity entries. */
int ne_download_generation(ne_runtime *rt, int which, uint32_t *slot_indices, float *fitness,
                           uint32_t *evaluation_counts, uint8_t *has_fitness, uint64_t capacity,
                           uint64_t *count);
/* Parent pairs of the last plan, 2 uint32 per child. */
int ne_download_plan(ne_runtime *rt, uint32_t *pairs, uint64_t capacity_pairs, uint64_t *count);
/* Per-cell local training-word counts of the last evaluation. */
int ne_download_local_word_counts(ne_runtime *rt, uint32_t *dst, uint64_t capacity, uint64_t *count);

/* ---- status, timing, parity ----------------------------------------------------------------- */
/* Host only, no device work: the reference shard schedule as per-cell CSR -- a global foundation
 * shard plus golden-ratio-centred shards whose radius grows every growth period
 * (TryBuildTrainingDataShardRuntimeSet, DoesTrainingDataShardCoverCell and the local word order of
 * TryResolveLocalTrainingWord: training_data.hpp:150-316, evaluation_ops.cuh:122-169).
 * offsets holds rows*cols+1 entries; *id_count receives the number of ids (ids may be NULL to size). */
int ne_build_shard_csr(const ne_word_counts *counts, const ne_shard_release_history *history,
                       uint64_t generation_index, uint64_t rows, uint64_t cols, uint64_t epicenter_rows,
                       uint32_t *offsets, uint32_t *ids, uint64_t ids_capacity, uint64_t *id_count);

/* TryReadDeviceSlabGARuntimeStatus (slab_runtime.hpp:183-184): last sticky status. */
int ne_read_status(ne_runtime *rt);
/* DeviceSlabGARuntimeStatusCodeString (slab_runtime.hpp:186). */
const char *ne_status_string(int status);
int ne_get_stage_times(ne_runtime *rt, ne_stage_times *out);
/* Parity mode: flags = OR of NE_TRACE_*; the next evaluate/advance records traces. */
int ne_set_trace(ne_runtime *rt, uint32_t flags);
/* Selects the evaluation precision for later evaluate/advance calls (no reference counterpart: the
 * reference evaluates in its storage precision only, evaluation_ops.cuh:244-508). */
int ne_set_eval_precision(ne_runtime *rt, uint32_t mode);
/* Shape of the persistent evaluation kernel's launch (no reference counterpart).  Three kernels give the same
 * results bit for bit within each precision mode: REGISTERS plays one tile of 16 games per wave and keeps its
 * layer-0 accumulators in registers (3 waves per SIMD); STREAMED plays one tile per wave and keeps them in
 * device scratch between the fold and the trunk (4 waves per SIMD); PAIRED plays two tiles per wave, so one
 * workgroup streams a genome's operands once per decision for twice the games (in the BF16-screened mode at
 * most 3 waves per workgroup: larger requests run at 3). */
#define NE_EVAL_KERNEL_REGISTERS 0u
#define NE_EVAL_KERNEL_STREAMED 1u
#define NE_EVAL_KERNEL_PAIRED 2u
typedef struct ne_eval_shape {
    uint32_t waves;              /* waves (tiles of 16 games) per workgroup, 1..4; 0: the built-in choice
                                    (the size leaving the fewest idle wave slots) */
    uint32_t rounds_per_item;    /* rounds of `waves` tiles one work item plays (1: a genome's tiles spread
                                    over ceil(tiles / waves) items; r: over ceil(tiles / (r x waves))) */
    uint32_t workgroups_per_cu;  /* persistent workgroups per CU; 0: as many as the kernel's registers and
                                    LDS allow */
    uint32_t kernel;             /* NE_EVAL_KERNEL_*: which evaluation kernel (untraced evaluations; traced
                                    ones always use NE_EVAL_KERNEL_REGISTERS) */
} ne_eval_shape;
/* Autotuner of the evaluation's launch shape (no reference counterpart): a coordinate search over the
 * four knobs of ne_eval_shape, each candidate timed with HIP events on one evaluation of the current
 * generation -- waves 1..4 (the REGISTERS kernel, rounds 1, all workgroups), then the STREAMED and PAIRED
 * kernels at every workgroup size they allow (untraced evaluations), then rounds 2 and 4, then one or two
 * workgroups per CU fewer
 * than fit -- keeping the fastest for later evaluations of this handle.  Results do not depend on
 * the shape; the generation is left evaluated (with the chosen shape) as by ne_evaluate.  chosen_waves
 * (optional) receives the chosen waves, timings_ms (optional) the 4 times of the waves sweep;
 * ne_get_eval_shape reports the whole choice. */
int ne_autotune_evaluation(ne_runtime *rt, const ne_word_counts *counts, const ne_shard_release_history *history,
                           uint32_t *chosen_waves, float *timings_ms);
/* Sets the evaluation's waves per workgroup (1..4), or 0 for the built-in choice; the other knobs keep their
 * values. */
int ne_set_eval_shape(ne_runtime *rt, uint32_t waves);
/* Sets or reads every knob of the evaluation's launch shape. */
int ne_set_eval_shape_ex(ne_runtime *rt, const ne_eval_shape *shape);
int ne_get_eval_shape(ne_runtime *rt, ne_eval_shape *out);

/* Captured generation steps (no counterpart in the reference, which launches every stage from the
 * host).  With enabled = 1, ne_advance_generation runs a plain step -- no injection, no callbacks, no
 * trace, the secrets given as a CSR (ne_upload_secrets_csr), the same population size, and a free count
 * at the start of the step that holds every child in one assembly batch -- as one hipGraph: the same
 * kernels with the same arguments as the eager step, captured once per generation-buffer parity and
 * replayed with the generation index and seed in a device block, then one host read of the status,
 * summary and counters.  Results equal the eager step's bit for bit; ne_get_stage_times reports the
 * graph's replay time only (step_graph_ms).  Any other step, or a change of configuration, shape or
 * buffers, runs eagerly or captures anew; a capture that cannot be made runs the step eagerly and turns
 * the mode off.  enabled = 0 (the default) releases the graphs. */
int ne_set_step_graph(ne_runtime *rt, uint32_t enabled);
/* Copies a recorded trace; game traces are ordered by individual, then game slot (W fresh games by
 * entry, then 2W prefilled games: slot W+q is entry q/2, variant 1+q%2). */
int ne_download_trace(ne_runtime *rt, uint32_t kind, void *dst, uint64_t capacity_bytes, uint64_t *written);

/* ---- surrogate-screened evolution ---------------------------------------------------------------
 * No counterpart in the reference (which evaluates every child, slab_runtime.cu:1554-1661): an
 * oversampled brood per cell is assembled, every child is fingerprinted (its guesses on 64 fixed probe
 * states) and scored by a small learned surrogate, and only each cell's best-predicted child plus a
 * random audit slice of the brood is evaluated in full.  The surrogate (MLP 192 -> 128 -> 64 -> 1)
 * trains online with Adam on the (fingerprint, fitness) pairs of every fully evaluated child.  While
 * the audit's rank correlation is below the threshold, or too few pairs have been seen, generations
 * run plain (one child per cell, all evaluated) and keep feeding the surrogate. */
typedef struct ne_screen_config {
    uint32_t brood_factor;       /* children per cell while screening (4: the best-predicted 25% go on) */
    uint32_t replay_capacity;    /* (fingerprint, fitness) pairs kept for training (65,536) */
    float audit_fraction;        /* random slice of the brood always evaluated in full (0.05) */
    float spearman_threshold;    /* screening pauses while the audit's rank correlation is below (0.5;
                                    0 screens whenever the correlation is positive, which raises the
                                    expected fitness of the chosen child; bench.py times both) */
    float learning_rate;         /* Adam step size (1e-3) */
    uint32_t train_steps;        /* minibatch steps per generation (32) */
    uint32_t batch_size;         /* minibatch size, 1..1024 (256) */
    uint32_t min_replay;         /* pairs seen before screening may start (4,096) */
    uint32_t seed;               /* probe states, surrogate initialisation, audit and minibatch draws */
    uint32_t flags;              /* NE_SCREEN_*; 0: a generation's training steps replay as one hipGraph */
} ne_screen_config;

enum {
    NE_SCREEN_EAGER_TRAINING = 1   /* launch the training steps one by one instead of replaying a graph */
};

typedef struct ne_screen_report {
    uint32_t screened;           /* 1: this step screened a brood; 0: plain step (warm-up or fallback) */
    uint32_t next_screened;      /* what the next step will do */
    uint32_t brood;              /* children assembled */
    uint32_t evaluated;          /* children evaluated in full */
    uint32_t audited;            /* brood members in the audit slice */
    float audit_spearman;        /* rank correlation of prediction and fitness over the audit slice
                                    (predictions made before training on them; NaN below 3 members) */
    float train_loss;            /* 1/2 mean squared error of the last minibatch (standardised fitness) of the
                                    most recently completed training: the previous step's (a step's training
                                    runs on its own stream, overlapping the next step's planning, brood
                                    assembly and fingerprints; that step's predictions wait for it) */
    float fingerprint_ms;        /* encoder transpose + fingerprint kernel */
    float surrogate_ms;          /* prediction + the most recently completed training */
    float best_fitness;          /* of the new generation */
    float average_fitness;
    uint32_t reserved0;
    uint64_t replay_count;       /* pairs in the replay buffer */
} ne_screen_report;

/* Allocates the surrogate, its Adam state, the replay buffer and the 64 probe states.  Probe boards use,
 * and are answered from, the probe vocabulary: the first min(action_count, 512) catalog words.  The generation budget must hold brood_factor x population
 * children and the slab one more population.  Calling it again resets the surrogate. */
int ne_enable_screening(ne_runtime *rt, const ne_screen_config *config);
/* One screened generation step (no widening or injection).  When the current generation has no
 * fitness of its own evaluation yet it is evaluated first; on return the new generation is evaluated
 * (its fitness is that of the full evaluation of each selected child). */
int ne_advance_generation_screened(ne_runtime *rt, uint32_t generation_seed, const ne_word_counts *counts,
                                   const ne_assembly_config *config, const ne_shard_release_history *history,
                                   ne_screen_report *report);
/* Fingerprints of `count` genomes (slot indices) with the given parent fitness pairs ([count][2]):
 * out[count][192], per probe p = 16 d + k (board k at depth d): [3p] (2 greens + yellows) / 10 of the
 * genome's guess (its masked argmax over the probe vocabulary) against the probe secret, [3p+1] d >= 1: 1 if the guess fits every hint on the board,
 * d = 0: 10 x the fitness of parent k % 2, [3p+2] 1 if the guess is the secret. */
int ne_fingerprint(ne_runtime *rt, const uint32_t *slot_indices, const float *parent_fitness, uint64_t count,
                   float *out);
/* The probe states: per board k (16), the secret's catalog index then its 3 turns' guess indices. */
int ne_get_probe_states(ne_runtime *rt, uint32_t *dst, uint64_t capacity, uint64_t *count);
/* Surrogate access (host arrays): predictions, one Adam step on (x, t), parameters (33,025 floats:
 * W1T [192][128], b1 [128], W2T [128][64], b2 [64], w3 [64], b3).  Setting parameters zeroes the Adam
 * state. */
int ne_surrogate_predict(ne_runtime *rt, const float *x, uint64_t count, float *y);
int ne_surrogate_train_step(ne_runtime *rt, const float *x, const float *t, uint64_t count);
int ne_surrogate_get_params(ne_runtime *rt, float *dst, uint64_t capacity, uint64_t *count);
int ne_surrogate_set_params(ne_runtime *rt, const float *src, uint64_t count);
/* Host only: Spearman rank correlation (average ranks for ties); NaN when n < 3 or a side is constant. */
double ne_rank_correlation(const float *a, const float *b, uint64_t n);

/* ---- checkpoint and resume ---------------------------------------------------------------
 * A checkpoint is the logical pre-recombination snapshot of RuntimeCheckpoint (slab_runtime.hpp:86-117):
 * the current generation (slots, fitness, evaluation counts), the parent plan of the next one, every
 * live genome's slot bytes, and the configuration and seed that finish the step.  Files are
 * byte-compatible with TryWriteRuntimeCheckpointAtomically / TryReadRuntimeCheckpoint
 * (slab_runtime.cu:1934-2081): little-endian, schema 3, genome layout 1, FNV-1a 64 checksum. */
typedef struct ne_checkpoint ne_checkpoint;

typedef struct ne_checkpoint_info {
    uint32_t schema_version;
    uint32_t genome_layout_version;
    uint32_t resume_phase;               /* 1: pre-recombination, pre-mutation */
    uint32_t generation_seed;            /* seed of the step the checkpoint interrupts */
    uint64_t checksum;
    uint64_t training_data_identity_hash;
    uint64_t generation_index;           /* of the current (parent) generation */
    uint64_t population;                 /* current generation size */
    uint64_t child_count;                /* children planned for the next generation */
    uint64_t live_genotypes;             /* genome records */
    uint64_t payload_bytes;              /* CheckpointPayloadByteCount (run_genetic_algorithm.cu:1360-1372) */
    uint64_t action_count;               /* slab layout after any widening of this step */
    uint64_t slot_stride_bytes;
    uint64_t slot_count;
    uint64_t slab_bytes;
    uint64_t generation_byte_budget;
    uint64_t grid_column_count;
    uint64_t release_count;              /* shard release history length */
    ne_word_counts word_counts;
    ne_assembly_config assembly_config;
    ne_pending_injection injection;
} ne_checkpoint_info;

/* TryCreatePrebreedingCheckpointOnDevice (slab_runtime.hpp:152-158; slab_runtime.cu:1707-1731): the first
 * half of ne_advance_generation (evaluation, callbacks, plan, widening), then the snapshot.  The runtime
 * is left at the boundary: ne_resume_generation_from_checkpoint finishes the step.  catalog_letters
 * (5 letter indices per word) may be NULL; when given, the identity hash covers the catalog. */
int ne_create_prebreeding_checkpoint(ne_runtime *rt, uint32_t generation_seed, const ne_word_counts *counts,
                                     const ne_assembly_config *config, const ne_pending_injection *injection,
                                     const uint8_t *catalog_letters, uint64_t catalog_word_count, uint32_t verbose,
                                     const ne_callbacks *callbacks, const ne_shard_release_history *history,
                                     ne_checkpoint **out_checkpoint);
/* TryRestorePrebreedingCheckpointToDevice (slab_runtime.hpp:160; slab_runtime.cu:1733-1866): a new
 * runtime holding the checkpoint's boundary state.  The k-th genome record takes slot k of a fresh slab
 * of checkpoint slot_count slots.  stream: hipStream_t or NULL (own stream).  The caller uploads the
 * catalog and secrets before resuming, as for a new runtime. */
int ne_restore_prebreeding_checkpoint(const ne_checkpoint *checkpoint, void *stream, ne_runtime **out_runtime);
/* TryResumeGenerationFromCheckpointOnDevice (slab_runtime.hpp:162-163; slab_runtime.cu:1868-1880): the
 * second half of the step (assembly, injection, swap) with the checkpoint's seed and configuration.
 * The runtime must stand at that checkpoint's boundary (same generation index and planned children). */
int ne_resume_generation_from_checkpoint(ne_runtime *rt, const ne_checkpoint *checkpoint, uint32_t verbose);
/* TryWriteRuntimeCheckpointAtomically (slab_runtime.hpp:165-166): writes <path>.tmp, then renames it
 * over path; creates missing parent directories. */
int ne_write_checkpoint(const ne_checkpoint *checkpoint, const char *path);
/* TryReadRuntimeCheckpoint (slab_runtime.hpp:168): reads and validates (magic, versions, checksum). */
int ne_read_checkpoint(const char *path, ne_checkpoint **out_checkpoint);
/* TryCloneRuntimeCheckpoint (run_genetic_algorithm.cu:1374-1415): a deep copy, e.g. for a background
 * write while the original resumes. */
int ne_clone_checkpoint(const ne_checkpoint *checkpoint, ne_checkpoint **out_checkpoint);
int ne_get_checkpoint_info(const ne_checkpoint *checkpoint, ne_checkpoint_info *out);
/* Current generation arrays (each may be NULL; capacity entries) and the plan ([child][2] parent
 * indices) of a checkpoint. */
int ne_get_checkpoint_generation(const ne_checkpoint *checkpoint, uint32_t *slot_indices, float *fitness,
                                 uint32_t *evaluation_counts, uint8_t *has_fitness, uint64_t capacity);
int ne_get_checkpoint_plan(const ne_checkpoint *checkpoint, uint32_t *pairs, uint64_t capacity_pairs);
/* Genome record `record`: its organism index and slot bytes (capacity >= slot_stride_bytes). */
int ne_get_checkpoint_genome(const ne_checkpoint *checkpoint, uint64_t record, uint32_t *organism_index,
                             void *dst, uint64_t capacity);
void ne_free_checkpoint(ne_checkpoint *checkpoint);
/* The training-data identity hash of TryDownloadBoundaryCheckpointPayload (slab_runtime.cu:1282-1296):
 * FNV-1a 64 over the word counts, the release history and, when given, the catalog.  Host only. */
uint64_t ne_training_data_identity_hash(const ne_word_counts *counts, const ne_shard_release_history *history,
                                        const uint8_t *catalog_letters, uint64_t catalog_word_count);

/* ---- generation telemetry and the curriculum controller ------------------------------------
 * The statistics the reference's driver logs and releases training shards by
 * (src/cli/run_genetic_algorithm.cu:594-624, 941-1286), computed where the generation lives. */

/* GenerationFitnessTelemetryRow (run_genetic_algorithm.cu:594-610). */
typedef struct ne_fitness_telemetry {
    uint64_t generation;
    uint64_t population_size;
    uint64_t training_word_count;
    float fitness_min;
    float fitness_mean;
    float fitness_median;
    float fitness_p90;                   /* nearest rank */
    float fitness_p99;                   /* nearest rank */
    float fitness_max;
    float fitness_stddev;                /* population standard deviation */
    uint32_t has_breeding_metrics;
    uint64_t distinct_fitness_count;
    uint64_t parent_childless_count;     /* parents the plan gives no child */
    uint64_t parent_one_child_count;
    uint64_t parent_multiple_children_count;
} ne_fitness_telemetry;

/* TryCollectGenerationFitnessTelemetry (run_genetic_algorithm.cu:1080-1186) on the current generation,
 * which must be evaluated.  Breeding metrics need the next generation's plan (from the post-fitness
 * callback on).  At most 16,384 organisms (NE_UNSUPPORTED above). */
int ne_collect_fitness_telemetry(ne_runtime *rt, const ne_word_counts *counts, uint32_t collect_breeding_metrics,
                                 ne_fitness_telemetry *out);

/* GeneticConvergenceTelemetryRow (run_genetic_algorithm.cu:612-624). */
typedef struct ne_convergence_telemetry {
    uint64_t generation;
    uint64_t sample_organisms;
    uint64_t sample_weights;
    uint64_t pair_count;
    float centroid_distance_mean;
    float centroid_distance_min;
    float centroid_distance_max;
    float pairwise_distance_mean;
    float pairwise_distance_min;
    float pairwise_distance_max;
    float elapsed_ms;
    float reserved0;
} ne_convergence_telemetry;

/* TryCollectGeneticConvergenceTelemetry (run_genetic_algorithm.cu:941-1052): up to 256 organisms x 8,192
 * trainable scalars x 512 organism pairs sampled by MakeSamplePlan (genetic_convergence.hpp:87-117) from
 * the run seed and the generation index; centroid and pairwise L2 distances over the sample. */
int ne_collect_convergence_telemetry(ne_runtime *rt, uint32_t run_seed, ne_convergence_telemetry *out);

/* MakeSamplePlan (genetic_convergence.hpp:87-117), host only: organism indices (<= 256), weight indices
 * (<= 8,192) and organism pairs ([512][2], indices into the organism sample).  Any output may be NULL;
 * counts receive the three lengths. */
int ne_convergence_sample_plan(uint64_t population, uint64_t trainable_scalars, uint32_t run_seed,
                               uint64_t generation, uint32_t *organisms, uint32_t *weights, uint32_t *pairs,
                               uint64_t counts[3]);

/* The adaptive shard-release thresholds of the reference CLI (training_data.hpp:17-20 defaults). */
typedef struct ne_shard_release_policy {
    uint64_t min_gap_generations;        /* 3: generations between releases */
    uint64_t first_new_release_generation; /* 10: the first non-foundation release comes no earlier */
    float centroid_distance_threshold;   /* 6.0: release when the centroid distance mean is <= this ... */
    float fitness_p99_threshold;         /* 0.20: ... and the fitness p99 is > this */
    uint32_t run_seed;                   /* seeds the convergence sample */
    uint32_t reserved0;
} ne_shard_release_policy;

/* TryMaybeReleaseTrainingDataShard (run_genetic_algorithm.cu:1200-1286), for the training-shard release
 * callback: on the evaluated current generation, decides whether the next catalog words are released for
 * the next generation.  On a release *injection is set for them, (next generation, fitness p99) is
 * appended to the caller's history arrays (capacity entries; the foundation record must be there) and
 * *released is 1; otherwise *injection is cleared.  *next_word_count is the action count afterwards. */
int ne_maybe_release_training_shard(ne_runtime *rt, const ne_word_counts *counts,
                                    const ne_shard_release_policy *policy, uint64_t catalog_word_count,
                                    uint64_t *release_generations, float *release_fitness_p99,
                                    uint64_t *release_count, uint64_t release_capacity,
                                    ne_pending_injection *injection, uint64_t *next_word_count,
                                    uint32_t *released);

/* ---- single-model inference and the winner artifact ------------------------------------------
 * One genome serving next-guess requests (src/inference/single_model_device_runtime.hpp:9-49), and the
 * artifact that carries a run's best genome (src/model_artifact/winner_artifact*.hpp). */
typedef struct ne_model ne_model;

/* A played Wordle board (wordle/wordle_grid.hpp:13-36): turn_count turns of a guess and its tiles. */
typedef struct ne_board {
    uint32_t turn_count;                 /* 0..6 */
    uint8_t guesses[6][5];               /* letter indices 0..25 */
    uint8_t tiles[6][5];                 /* 0 green, 1 yellow, 2 grey */
} ne_board;

/* SelectedAction (output_embedding.hpp) plus the runner-up score. */
typedef struct ne_selected_action {
    uint32_t status;                     /* NE_OK; NE_POLICY_FORWARD_FAILED: invalid or finished board;
                                            NE_ACTION_SELECTION_FAILED: every action already guessed */
    uint32_t action_index;
    float score;
    float second_score;                  /* best score among the other unguessed actions; -inf if none */
    uint8_t word[5];
    uint8_t reserved[3];
} ne_selected_action;

/* TryCreateSingleModelDeviceRuntime (single_model_device_runtime.hpp:38-40): genome_byte_count must be
 * the slot stride of action_count actions; action_letters holds action_count words of 5 letter indices.
 * stream: hipStream_t or NULL (own stream). */
int ne_model_create(const void *genome_bytes, uint64_t genome_byte_count, const uint8_t *action_letters,
                    uint64_t action_count, void *stream, ne_model **out_model);
/* Genome layouts a model may serve: the encoder's hidden width (0: one 145 -> 64 layer, this build's
 * evaluation layout; 128: the reference's 145 -> 128 -> 64, encoder_spec.hpp:19-20) and the storage of every
 * parameter (NE_STORAGE_FP32, or NE_STORAGE_FP16: IEEE binary16 as the reference stores its genomes,
 * common/float16.hpp:9-13; computed in fp32 either way). */
enum { NE_STORAGE_FP32 = 0, NE_STORAGE_FP16 = 1 };
typedef struct ne_model_layout {
    uint32_t encoder_hidden;             /* 0 or 128 */
    uint32_t storage;                    /* NE_STORAGE_* */
    uint32_t reserved[2];
} ne_model_layout;
/* The genome bytes of a layout at action_count actions (0 for an invalid layout): FP16, the reference's
 * ComputeDynamicGenomeStrideBytes (dynamic_layout.hpp:15-35): 2 x (parameters + 38 x action_count); FP32,
 * this build's slot stride: 4 x (parameters + 38 x action_count) rounded up to 256 bytes.  Parameters:
 * 132,928 (145 -> 64 encoder) or 150,528 (145 -> 128 -> 64). */
uint64_t ne_model_genome_bytes(const ne_model_layout *layout, uint64_t action_count);
/* TryCreateSingleModelDeviceRuntime (single_model_device_runtime.hpp:38-40) for a genome of any layout
 * above -- with {128, NE_STORAGE_FP16} the reference's own genome bytes; genome_byte_count must be
 * ne_model_genome_bytes(layout, action_count). */
int ne_model_create_with_layout(const void *genome_bytes, uint64_t genome_byte_count, const ne_model_layout *layout,
                                const uint8_t *action_letters, uint64_t action_count, void *stream,
                                ne_model **out_model);
void ne_model_destroy(ne_model *model);
/* TrySelectNextGuessWithSingleModelDeviceRuntime (single_model_device_runtime.hpp:44-47): the policy
 * forward on the board, the masked argmax over the actions (a guessed word, any copy of it, is never
 * guessed again; ties go to the lowest index).  Returns the board's status. */
int ne_model_select_guess(ne_model *model, const ne_board *board, ne_selected_action *out);
/* No counterpart (the reference serves one board per call): `count` boards in one launch, one status each
 * in out[i].status; the call's return is NE_OK unless the call itself failed. */
int ne_model_select_guesses(ne_model *model, const ne_board *boards, uint64_t count, ne_selected_action *out);
/* No counterpart in the reference: ne_model_select_guesses for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It allocates nothing: the
 * boards and results are staged in one static device array per GPU (NE_NOALLOC_MAX_BOARDS boards,
 * 704 KB, held until the process ends; its address is looked up by ne_model_create).  count above
 * NE_NOALLOC_MAX_BOARDS returns NE_INVALID_ARGUMENT.  Calls that overlap on one GPU (other threads or
 * streams, any model) share the array and race. */
#define NE_NOALLOC_MAX_BOARDS 8192
int ne_model_select_guesses_noalloc(ne_model *model, const ne_board *boards, uint64_t count,
                                    ne_selected_action *out);

/* WinnerArtifactMetadata (winner_artifact.hpp:12-21). */
typedef struct ne_winner_metadata {
    float best_fitness;
    uint32_t best_slot_index;
    uint64_t generation_index;
    uint64_t best_index;
    uint64_t action_count;
    uint64_t genome_byte_count;
    uint32_t seed;
    uint32_t reserved0;
    char action_space_path[1024];        /* NUL-terminated */
    char timestamp_local[32];            /* set by the writer and the reader: YYYY-MM-DDTHH:MM:SS */
} ne_winner_metadata;

/* TryWriteWinnerArtifactToPaths (winner_artifact.hpp:32-35): the genome bytes to binary_path and the JSON
 * metadata (format_version 2, the first action_count words) to metadata_path, each through <path>.tmp. */
int ne_write_winner_artifact(const char *binary_path, const char *metadata_path, const void *genome_bytes,
                             const uint8_t *action_letters, uint64_t action_word_count,
                             ne_winner_metadata *metadata);
/* TryWriteWinnerArtifact (winner_artifact.hpp:28-30): a new winner-<timestamp>-g<generation>-seed<seed>
 * pair in directory (a -2, -3, ... suffix when the name exists); the chosen paths are returned. */
int ne_write_winner_artifact_in_directory(const char *directory, const void *genome_bytes,
                                          const uint8_t *action_letters, uint64_t action_word_count,
                                          ne_winner_metadata *metadata, char *binary_path_out,
                                          char *metadata_path_out, uint64_t path_capacity);
/* TryReadWinnerArtifact (winner_artifact_reader.hpp:20-21): fills metadata; genome_dst (capacity bytes)
 * and letters_dst (capacity words) may be NULL to learn the sizes first. */
int ne_read_winner_artifact(const char *binary_path, const char *metadata_path, ne_winner_metadata *metadata,
                            void *genome_dst, uint64_t genome_capacity, uint8_t *letters_dst,
                            uint64_t letters_capacity_words);

#ifdef __cplusplus
}
#endif

#endif /* NE_RUNTIME_H */
