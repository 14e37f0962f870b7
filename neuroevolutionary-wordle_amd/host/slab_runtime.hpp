// slab_runtime.hpp -- C++ host API with the reference's names and argument meaning
// (src/genetic_algorithm/device/slab_runtime.hpp:119-186), implemented over the C ABI of
// include/ne_runtime.h.  Header-only; link against libne_wordle.so.
//
// Differences from the reference, all at the boundary: the catalog and the per-cell secrets belong
// to the runtime handle (not to process-global constant memory), downloads fill std::vector, and
// every call returns false with the runtime status readable through TryReadDeviceSlabGARuntimeStatus.
#pragma once

#include <chrono>
#include <cstddef>
#include <cstdint>
#include <filesystem>
#include <functional>
#include <future>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ne_runtime.h"

namespace ne_amd::slab_device {

enum class DeviceSlabGARuntimeStatusCode : int {
    kOk = 0,
    kCudaFailure = 1,
    kInvalidRuntimeConfig = 2,
    kInvalidSlab = 3,
    kInvalidGeneration = 4,
    kInvalidTrainingShard = 5,
    kGuessAppendFailed = 6,
    kPolicyForwardFailed = 7,
    kActionSelectionFailed = 8,
    kPopulationNotEvaluated = 9,
    kInvalidAssemblyConfig = 10,
    kParentSelectionFailed = 11,
    kInvalidAssemblyPlan = 12,
    kInvalidParentIndex = 13,
    kSlabFull = 14,
    kOutputEmbeddingInjectionFailed = 15,
    kSlabRepackFailed = 16,
    kInvalidArgument = 17,
    kUnsupported = 18,
    kCallbackFailed = 19,
};

struct DeviceSlabGARuntimeConfig {
    std::size_t genotype_slab_byte_budget_bytes = 0;
    std::size_t generation_byte_budget_bytes = 0;
    std::size_t host_spillover_byte_budget_bytes = 0;
    std::size_t action_count = 0;
    std::size_t population_size_ceiling = 0;
    std::size_t grid_column_count = 0;
};

struct WordCountSchedule {
    std::size_t initial_word_count = 20;
    std::size_t word_count_step = 0;
    std::size_t word_count_step_period_generations = 1;
};

struct RuntimeWordCounts {
    std::size_t training_word_count = 20;
    std::size_t action_space_word_count = 20;
    WordCountSchedule training_word_schedule{};
    std::size_t shard_initial_radius = 0;
    std::size_t shard_radius_growth_period_generations = 2;
};

struct ParentSelectionConfig {
    std::size_t cellular_breeding_radius = 2;
    float rank_exponent = 0.5f;
};
struct BreedingConfig {
    float crossover_temperature_level1 = 0.02f;
    float crossover_temperature_level2 = 0.01f;
    float crossover_temperature_level3 = 0.005f;
};
struct MutationConfig {
    float mutation_probability = 0.0001f;
    float mutation_sigma = 0.02f;
    float output_tail_row_scale_mutation_probability = 0.0f;
};
struct GenerationAssemblyConfig {
    ParentSelectionConfig parent_selection{};
    BreedingConfig breeding{};
    MutationConfig mutation{};
};

struct PendingOutputEmbeddingInjection {
    bool enabled = false;
    std::size_t first_catalog_word_index = 0;
    std::size_t injection_count = 0;
};

struct PopulationFitnessSummary {
    float best_fitness = 0.0f;
    float average_fitness = 0.0f;
    std::size_t best_index = 0;
    std::uint32_t best_slot_index = 0xFFFFFFFFu;
    std::size_t generation_index = 0;
    std::size_t action_count = 0;
    std::size_t population_size = 0;
};

struct SlabGeneration {
    std::size_t generation_index = 0;
    std::vector<std::uint32_t> slot_indices;
    std::vector<float> fitness;
    std::vector<std::uint32_t> evaluation_counts;
    std::vector<std::uint8_t> has_fitness;
};

struct TrainingDataShardReleaseHistory {
    std::vector<std::uint64_t> release_generations;
    std::vector<float> release_fitness_p99;  // empty, or one per release generation
};

struct DeviceSlabGARuntimeBuffers {
    ne_runtime* handle = nullptr;
    std::size_t grid_column_count = 0;
};

using PostFitnessEvaluationCallback = std::function<bool(const DeviceSlabGARuntimeBuffers&, const RuntimeWordCounts&)>;
using TrainingShardReleaseCallback = std::function<bool(const DeviceSlabGARuntimeBuffers&, const RuntimeWordCounts&,
                                                        PendingOutputEmbeddingInjection&)>;

namespace detail {

inline ne_word_counts ToC(const RuntimeWordCounts& w) {
    ne_word_counts c{};
    c.training_word_count = w.training_word_count;
    c.action_space_word_count = w.action_space_word_count;
    c.training_word_schedule.initial_word_count = w.training_word_schedule.initial_word_count;
    c.training_word_schedule.word_count_step = w.training_word_schedule.word_count_step;
    c.training_word_schedule.word_count_step_period_generations =
        w.training_word_schedule.word_count_step_period_generations;
    c.shard_initial_radius = w.shard_initial_radius;
    c.shard_radius_growth_period_generations = w.shard_radius_growth_period_generations;
    return c;
}

inline RuntimeWordCounts FromC(const ne_word_counts& c) {
    RuntimeWordCounts w{};
    w.training_word_count = c.training_word_count;
    w.action_space_word_count = c.action_space_word_count;
    w.training_word_schedule = {c.training_word_schedule.initial_word_count, c.training_word_schedule.word_count_step,
                                c.training_word_schedule.word_count_step_period_generations};
    w.shard_initial_radius = c.shard_initial_radius;
    w.shard_radius_growth_period_generations = c.shard_radius_growth_period_generations;
    return w;
}

// The C view of a history (null when there is none).
inline const ne_shard_release_history* HistoryView(const TrainingDataShardReleaseHistory* history,
                                                   ne_shard_release_history& view) {
    if (!history) return nullptr;
    view = {};
    view.release_generations = history->release_generations.data();
    view.release_count = history->release_generations.size();
    if (history->release_fitness_p99.size() == history->release_generations.size())
        view.release_fitness_p99 = history->release_fitness_p99.data();
    return &view;
}

struct CallbackContext {
    const DeviceSlabGARuntimeBuffers* buffers;
    const PostFitnessEvaluationCallback* post;
    const TrainingShardReleaseCallback* release;
};

inline int PostTrampoline(ne_runtime*, const ne_word_counts* counts, void* user) {
    auto* ctx = static_cast<CallbackContext*>(user);
    return (*ctx->post)(*ctx->buffers, FromC(*counts)) ? 0 : 1;
}

inline int ReleaseTrampoline(ne_runtime*, const ne_word_counts* counts, ne_pending_injection* inj, void* user) {
    auto* ctx = static_cast<CallbackContext*>(user);
    PendingOutputEmbeddingInjection p{inj->enabled != 0, inj->first_catalog_word_index, inj->injection_count};
    if (!(*ctx->release)(*ctx->buffers, FromC(*counts), p)) return 1;
    inj->enabled = p.enabled ? 1u : 0u;
    inj->first_catalog_word_index = p.first_catalog_word_index;
    inj->injection_count = p.injection_count;
    return 0;
}

}  // namespace detail

// TryCreateDeviceSlabGARuntimeBuffers (slab_runtime.hpp:119)
inline bool TryCreateDeviceSlabGARuntimeBuffers(DeviceSlabGARuntimeBuffers& buffers,
                                                const DeviceSlabGARuntimeConfig& config) {
    ne_runtime_config c{};
    c.genotype_slab_byte_budget = config.genotype_slab_byte_budget_bytes;
    c.generation_byte_budget = config.generation_byte_budget_bytes;
    c.host_spillover_byte_budget = config.host_spillover_byte_budget_bytes;
    c.action_count = config.action_count;
    c.population_size_ceiling = config.population_size_ceiling;
    c.grid_column_count = config.grid_column_count;
    buffers = {};
    if (ne_create(&c, &buffers.handle) != NE_OK) return false;
    buffers.grid_column_count = config.grid_column_count;
    return true;
}

// DestroyDeviceSlabGARuntimeBuffers (slab_runtime.hpp:121)
inline void DestroyDeviceSlabGARuntimeBuffers(DeviceSlabGARuntimeBuffers& buffers) noexcept {
    ne_destroy(buffers.handle);
    buffers = {};
}

// UploadTrainingWordCatalogToDeviceConstantMemory (training_data.hpp:375-381), per handle.
inline bool TryUploadTrainingWordCatalog(DeviceSlabGARuntimeBuffers& buffers, const std::vector<std::uint8_t>& letters) {
    return ne_upload_catalog(buffers.handle, letters.data(), letters.size() / 5) == NE_OK;
}

// Per-cell secrets as CSR; replaces the shard schedule when set.
inline bool TryUploadTrainingSecrets(DeviceSlabGARuntimeBuffers& buffers, const std::vector<std::uint32_t>& offsets,
                                     const std::vector<std::uint32_t>& secret_ids) {
    return ne_upload_secrets_csr(buffers.handle, offsets.data(), offsets.size() - 1, secret_ids.data(),
                                 secret_ids.size()) == NE_OK;
}

// TryBootstrapRandomCurrentGenerationOnDevice (slab_runtime.hpp:123-124)
inline bool TryBootstrapRandomCurrentGenerationOnDevice(DeviceSlabGARuntimeBuffers& buffers,
                                                        std::size_t generation_size, std::uint32_t generation_seed,
                                                        std::size_t generation_index = 0,
                                                        const ne_init_config* init = nullptr) {
    return ne_bootstrap(buffers.handle, generation_size, generation_seed, generation_index, init) == NE_OK;
}

// TryEvaluateCurrentGenerationFitnessOnDevice (slab_runtime.hpp:137-139)
inline bool TryEvaluateCurrentGenerationFitnessOnDevice(DeviceSlabGARuntimeBuffers& buffers,
                                                        const RuntimeWordCounts& counts,
                                                        const TrainingDataShardReleaseHistory* history = nullptr) {
    const ne_word_counts c = detail::ToC(counts);
    ne_shard_release_history h{};
    return ne_evaluate(buffers.handle, &c, detail::HistoryView(history, h)) == NE_OK;
}

// TryReadPopulationFitnessSummaryFromDevice (slab_runtime.hpp:141-142)
inline bool TryReadPopulationFitnessSummaryFromDevice(const DeviceSlabGARuntimeBuffers& buffers,
                                                      PopulationFitnessSummary& summary) {
    ne_fitness_summary s{};
    if (ne_read_summary(buffers.handle, &s) != NE_OK) return false;
    summary = {s.best_fitness, s.average_fitness, s.best_index, s.best_slot_index, s.generation_index,
               s.action_count, s.population_size};
    return true;
}

namespace detail {
inline ne_assembly_config ToC(const GenerationAssemblyConfig& config) {
    ne_assembly_config a{};
    a.cellular_breeding_radius = config.parent_selection.cellular_breeding_radius;
    a.rank_exponent = config.parent_selection.rank_exponent;
    a.crossover_temperature_level1 = config.breeding.crossover_temperature_level1;
    a.crossover_temperature_level2 = config.breeding.crossover_temperature_level2;
    a.crossover_temperature_level3 = config.breeding.crossover_temperature_level3;
    a.mutation_probability = config.mutation.mutation_probability;
    a.mutation_sigma = config.mutation.mutation_sigma;
    a.output_tail_row_scale_mutation_probability = config.mutation.output_tail_row_scale_mutation_probability;
    return a;
}
}  // namespace detail

// TryAdvanceGenerationOnDevice (slab_runtime.hpp:144-150)
inline bool TryAdvanceGenerationOnDevice(DeviceSlabGARuntimeBuffers& buffers, std::uint32_t generation_seed,
                                         const RuntimeWordCounts& counts, const GenerationAssemblyConfig& config = {},
                                         const PendingOutputEmbeddingInjection& injection = {}, bool verbose = false,
                                         const PostFitnessEvaluationCallback& post_fitness_evaluation_callback = {},
                                         const TrainingShardReleaseCallback& training_shard_release_callback = {},
                                         const TrainingDataShardReleaseHistory* history = nullptr) {
    const ne_word_counts c = detail::ToC(counts);
    const ne_assembly_config a = detail::ToC(config);
    ne_pending_injection inj{};
    inj.enabled = injection.enabled ? 1u : 0u;
    inj.first_catalog_word_index = injection.first_catalog_word_index;
    inj.injection_count = injection.injection_count;
    detail::CallbackContext ctx{&buffers, &post_fitness_evaluation_callback, &training_shard_release_callback};
    ne_callbacks cbs{};
    cbs.post_fitness_evaluation = post_fitness_evaluation_callback ? detail::PostTrampoline : nullptr;
    cbs.training_shard_release = training_shard_release_callback ? detail::ReleaseTrampoline : nullptr;
    cbs.user = &ctx;
    ne_shard_release_history h{};
    return ne_advance_generation(buffers.handle, generation_seed, &c, &a, &inj, verbose ? 1u : 0u, &cbs,
                                 detail::HistoryView(history, h)) == NE_OK;
}

inline bool TryDownloadGenerationFromDevice(const DeviceSlabGARuntimeBuffers& buffers, int which,
                                            SlabGeneration& generation) {
    ne_slab_info info{};
    if (ne_get_slab_info(buffers.handle, &info) != NE_OK) return false;
    const std::size_t cap = info.max_generation_size;
    generation.slot_indices.assign(cap, 0);
    generation.fitness.assign(cap, 0.0f);
    generation.evaluation_counts.assign(cap, 0);
    generation.has_fitness.assign(cap, 0);
    uint64_t n = 0;
    if (ne_download_generation(buffers.handle, which, generation.slot_indices.data(), generation.fitness.data(),
                               generation.evaluation_counts.data(), generation.has_fitness.data(), cap, &n) != NE_OK)
        return false;
    generation.slot_indices.resize(n);
    generation.fitness.resize(n);
    generation.evaluation_counts.resize(n);
    generation.has_fitness.resize(n);
    generation.generation_index = which == 0 ? info.current_generation_index : info.current_generation_index + 1;
    return true;
}

// TryDownloadCurrentGenerationFromDevice / TryDownloadNextGenerationFromDevice (slab_runtime.hpp:131-135)
inline bool TryDownloadCurrentGenerationFromDevice(const DeviceSlabGARuntimeBuffers& b, SlabGeneration& g) {
    return TryDownloadGenerationFromDevice(b, 0, g);
}
inline bool TryDownloadNextGenerationFromDevice(const DeviceSlabGARuntimeBuffers& b, SlabGeneration& g) {
    return TryDownloadGenerationFromDevice(b, 1, g);
}

// Surrogate-screened evolution (no counterpart in the reference, whose step is TryAdvanceGenerationOnDevice
// above): an oversampled brood per cell, ranked by a surrogate trained online, only the best-predicted
// child of each cell plus a random audit slice evaluated in full.
struct SurrogateScreeningConfig {
    std::uint32_t brood_factor = 4;
    std::uint32_t replay_capacity = 65536;
    float audit_fraction = 0.05f;
    float spearman_threshold = 0.5f;
    float learning_rate = 1e-3f;
    std::uint32_t train_steps = 32;
    std::uint32_t batch_size = 256;
    std::uint32_t min_replay = 4096;
    std::uint32_t seed = 0x5C2EE7u;
    bool eager_training = false;  // true: launch the training steps one by one instead of a hipGraph replay
};

// No counterpart in the reference (which launches every stage from the host): qualifying plain steps of
// TryAdvanceGenerationOnDevice run as one captured hipGraph each, with results equal to the eager step's.
inline bool TrySetCapturedGenerationSteps(DeviceSlabGARuntimeBuffers& buffers, bool enabled) {
    return ne_set_step_graph(buffers.handle, enabled ? 1u : 0u) == NE_OK;
}

inline bool TryEnableSurrogateScreening(DeviceSlabGARuntimeBuffers& buffers, const SurrogateScreeningConfig& config = {}) {
    ne_screen_config c{};
    c.brood_factor = config.brood_factor;
    c.replay_capacity = config.replay_capacity;
    c.audit_fraction = config.audit_fraction;
    c.spearman_threshold = config.spearman_threshold;
    c.learning_rate = config.learning_rate;
    c.train_steps = config.train_steps;
    c.batch_size = config.batch_size;
    c.min_replay = config.min_replay;
    c.seed = config.seed;
    c.flags = config.eager_training ? NE_SCREEN_EAGER_TRAINING : 0u;
    return ne_enable_screening(buffers.handle, &c) == NE_OK;
}

inline bool TryAdvanceScreenedGenerationOnDevice(DeviceSlabGARuntimeBuffers& buffers, std::uint32_t generation_seed,
                                                 const RuntimeWordCounts& counts,
                                                 const GenerationAssemblyConfig& config = {},
                                                 const TrainingDataShardReleaseHistory* history = nullptr,
                                                 ne_screen_report* report = nullptr) {
    const ne_word_counts c = detail::ToC(counts);
    const ne_assembly_config a = detail::ToC(config);
    ne_shard_release_history h{};
    return ne_advance_generation_screened(buffers.handle, generation_seed, &c, &a, detail::HistoryView(history, h),
                                          report) == NE_OK;
}

// ---- checkpoint and resume (slab_runtime.hpp:86-178) ------------------------------------------

// RuntimeCheckpoint (slab_runtime.hpp:98-117): owns the library's checkpoint; move-only.  Info()
// carries the header fields, the layout and the decoded configuration.
class RuntimeCheckpoint {
  public:
    RuntimeCheckpoint() = default;
    explicit RuntimeCheckpoint(ne_checkpoint* handle) : handle_(handle) {}
    RuntimeCheckpoint(const RuntimeCheckpoint&) = delete;
    RuntimeCheckpoint& operator=(const RuntimeCheckpoint&) = delete;
    RuntimeCheckpoint(RuntimeCheckpoint&& other) noexcept : handle_(std::exchange(other.handle_, nullptr)) {}
    RuntimeCheckpoint& operator=(RuntimeCheckpoint&& other) noexcept {
        if (this != &other) {
            ne_free_checkpoint(handle_);
            handle_ = std::exchange(other.handle_, nullptr);
        }
        return *this;
    }
    ~RuntimeCheckpoint() { ne_free_checkpoint(handle_); }

    bool Valid() const { return handle_ != nullptr; }
    ne_checkpoint* Handle() const { return handle_; }
    ne_checkpoint_info Info() const {
        ne_checkpoint_info i{};
        if (handle_) (void)ne_get_checkpoint_info(handle_, &i);
        return i;
    }
    void Reset(ne_checkpoint* handle = nullptr) {
        ne_free_checkpoint(handle_);
        handle_ = handle;
    }

  private:
    ne_checkpoint* handle_ = nullptr;
};

// TryCreatePrebreedingCheckpointOnDevice (slab_runtime.hpp:152-158)
inline bool TryCreatePrebreedingCheckpointOnDevice(
    DeviceSlabGARuntimeBuffers& buffers, std::uint32_t generation_seed, const RuntimeWordCounts& counts,
    const GenerationAssemblyConfig& config, const PendingOutputEmbeddingInjection& injection,
    RuntimeCheckpoint& checkpoint_out, const std::vector<std::uint8_t>* host_training_word_catalog = nullptr,
    bool verbose = false, const PostFitnessEvaluationCallback& post_fitness_evaluation_callback = {},
    const TrainingShardReleaseCallback& training_shard_release_callback = {},
    const TrainingDataShardReleaseHistory* history = nullptr) {
    checkpoint_out.Reset();
    const ne_word_counts c = detail::ToC(counts);
    const ne_assembly_config a = detail::ToC(config);
    ne_pending_injection inj{};
    inj.enabled = injection.enabled ? 1u : 0u;
    inj.first_catalog_word_index = injection.first_catalog_word_index;
    inj.injection_count = injection.injection_count;
    detail::CallbackContext ctx{&buffers, &post_fitness_evaluation_callback, &training_shard_release_callback};
    ne_callbacks cbs{};
    cbs.post_fitness_evaluation = post_fitness_evaluation_callback ? detail::PostTrampoline : nullptr;
    cbs.training_shard_release = training_shard_release_callback ? detail::ReleaseTrampoline : nullptr;
    cbs.user = &ctx;
    ne_shard_release_history h{};
    ne_checkpoint* out = nullptr;
    const std::uint8_t* letters = host_training_word_catalog ? host_training_word_catalog->data() : nullptr;
    const std::uint64_t words = host_training_word_catalog ? host_training_word_catalog->size() / 5 : 0;
    if (ne_create_prebreeding_checkpoint(buffers.handle, generation_seed, &c, &a, &inj, letters, words,
                                         verbose ? 1u : 0u, &cbs, detail::HistoryView(history, h), &out) != NE_OK)
        return false;
    checkpoint_out.Reset(out);
    return true;
}

// TryRestorePrebreedingCheckpointToDevice (slab_runtime.hpp:160): `buffers` receives a new runtime.
inline bool TryRestorePrebreedingCheckpointToDevice(const RuntimeCheckpoint& checkpoint,
                                                    DeviceSlabGARuntimeBuffers& buffers) {
    buffers = {};
    if (ne_restore_prebreeding_checkpoint(checkpoint.Handle(), nullptr, &buffers.handle) != NE_OK) return false;
    buffers.grid_column_count = checkpoint.Info().grid_column_count;
    return true;
}

// TryResumeGenerationFromCheckpointOnDevice (slab_runtime.hpp:162-163)
inline bool TryResumeGenerationFromCheckpointOnDevice(DeviceSlabGARuntimeBuffers& buffers,
                                                      const RuntimeCheckpoint& checkpoint, bool verbose = false) {
    return ne_resume_generation_from_checkpoint(buffers.handle, checkpoint.Handle(), verbose ? 1u : 0u) == NE_OK;
}

// TryWriteRuntimeCheckpointAtomically / TryReadRuntimeCheckpoint (slab_runtime.hpp:165-168)
inline bool TryWriteRuntimeCheckpointAtomically(const RuntimeCheckpoint& checkpoint,
                                                const std::filesystem::path& checkpoint_path) {
    return ne_write_checkpoint(checkpoint.Handle(), checkpoint_path.c_str()) == NE_OK;
}

inline bool TryReadRuntimeCheckpoint(const std::filesystem::path& checkpoint_path, RuntimeCheckpoint& checkpoint_out) {
    ne_checkpoint* out = nullptr;
    checkpoint_out.Reset();
    if (ne_read_checkpoint(checkpoint_path.c_str(), &out) != NE_OK) return false;
    checkpoint_out.Reset(out);
    return true;
}

// TryCloneRuntimeCheckpoint (run_genetic_algorithm.cu:1374-1415)
inline bool TryCloneRuntimeCheckpoint(const RuntimeCheckpoint& source, RuntimeCheckpoint& clone_out) {
    ne_checkpoint* out = nullptr;
    clone_out.Reset();
    if (ne_clone_checkpoint(source.Handle(), &out) != NE_OK) return false;
    clone_out.Reset(out);
    return true;
}

// RuntimeCheckpointAsyncWriter (slab_runtime.hpp:170-178; slab_runtime.cu:2083-2126): one background
// write at a time; a finished write's result is collected before the next starts.
class RuntimeCheckpointAsyncWriter {
  public:
    bool TryStartWrite(RuntimeCheckpoint checkpoint, std::filesystem::path checkpoint_path) {
        if (IsWriteInProgress()) return false;
        if (pending_.valid() && !pending_.get()) return false;
        pending_ = std::async(std::launch::async, [c = std::move(checkpoint), p = std::move(checkpoint_path)]() {
            return TryWriteRuntimeCheckpointAtomically(c, p);
        });
        return true;
    }
    bool IsWriteInProgress() {
        return pending_.valid() && pending_.wait_for(std::chrono::seconds(0)) != std::future_status::ready;
    }
    bool TryCollectFinishedWrite(bool& write_finished_out, bool& write_succeeded_out) {
        write_finished_out = false;
        write_succeeded_out = true;
        if (!pending_.valid() || IsWriteInProgress()) return true;
        write_finished_out = true;
        write_succeeded_out = pending_.get();
        return write_succeeded_out;
    }
    bool TryWaitForWrite() { return pending_.valid() ? pending_.get() : true; }

  private:
    std::future<bool> pending_{};
};

// ---- generation telemetry and the adaptive shard release (run_genetic_algorithm.cu:594-624, 941-1286) ----
using GenerationFitnessTelemetryRow = ne_fitness_telemetry;
using GeneticConvergenceTelemetryRow = ne_convergence_telemetry;

// TryCollectGenerationFitnessTelemetry (run_genetic_algorithm.cu:1080-1186)
inline bool TryCollectGenerationFitnessTelemetry(const DeviceSlabGARuntimeBuffers& buffers,
                                                 const RuntimeWordCounts& counts, bool collect_breeding_metrics,
                                                 GenerationFitnessTelemetryRow& row_out) {
    const ne_word_counts c = detail::ToC(counts);
    row_out = {};
    return ne_collect_fitness_telemetry(buffers.handle, &c, collect_breeding_metrics ? 1u : 0u, &row_out) == NE_OK;
}

// TryCollectGeneticConvergenceTelemetry (run_genetic_algorithm.cu:941-1052)
inline bool TryCollectGeneticConvergenceTelemetry(const DeviceSlabGARuntimeBuffers& buffers, std::uint32_t run_seed,
                                                  GeneticConvergenceTelemetryRow& row_out) {
    row_out = {};
    return ne_collect_convergence_telemetry(buffers.handle, run_seed, &row_out) == NE_OK;
}

// The CLI thresholds of TryMaybeReleaseTrainingDataShard (training_data.hpp:17-20 defaults).
struct ShardReleasePolicy {
    std::size_t shard_release_min_gap_generations = 3;
    std::size_t first_new_shard_release_generation = 10;
    float shard_release_centroid_distance_threshold = 6.0f;
    float shard_release_fitness_p99_threshold = 0.20f;
    std::uint32_t seed = 0;
};

// TryMaybeReleaseTrainingDataShard (run_genetic_algorithm.cu:1200-1286), for the training-shard release
// callback; `history` must hold the foundation record and grows by one entry per release.
inline bool TryMaybeReleaseTrainingDataShard(const DeviceSlabGARuntimeBuffers& buffers,
                                             const RuntimeWordCounts& counts, const ShardReleasePolicy& policy,
                                             std::size_t catalog_word_count, TrainingDataShardReleaseHistory& history,
                                             PendingOutputEmbeddingInjection& injection,
                                             std::size_t& next_word_count_out) {
    const ne_word_counts c = detail::ToC(counts);
    ne_shard_release_policy p{};
    p.min_gap_generations = policy.shard_release_min_gap_generations;
    p.first_new_release_generation = policy.first_new_shard_release_generation;
    p.centroid_distance_threshold = policy.shard_release_centroid_distance_threshold;
    p.fitness_p99_threshold = policy.shard_release_fitness_p99_threshold;
    p.run_seed = policy.seed;
    history.release_fitness_p99.resize(history.release_generations.size(), 0.0f);
    std::uint64_t count = history.release_generations.size();
    history.release_generations.resize(count + 1);
    history.release_fitness_p99.resize(count + 1);
    ne_pending_injection inj{};
    std::uint64_t next = 0;
    std::uint32_t released = 0;
    const int st = ne_maybe_release_training_shard(buffers.handle, &c, &p, catalog_word_count,
                                                   history.release_generations.data(),
                                                   history.release_fitness_p99.data(), &count, count + 1, &inj,
                                                   &next, &released);
    history.release_generations.resize(count);
    history.release_fitness_p99.resize(count);
    injection = {inj.enabled != 0, inj.first_catalog_word_index, inj.injection_count};
    next_word_count_out = next;
    return st == NE_OK;
}

// TryDownloadSlabSlotBytesFromDevice (slab_runtime.hpp:128-129)
inline bool TryDownloadSlabSlotBytesFromDevice(const DeviceSlabGARuntimeBuffers& buffers, std::uint32_t slot_index,
                                               std::vector<std::uint8_t>& slot_bytes) {
    ne_slab_info info{};
    if (ne_get_slab_info(buffers.handle, &info) != NE_OK) return false;
    slot_bytes.assign(info.slot_stride_bytes, 0);
    uint64_t written = 0;
    return ne_download_slot(buffers.handle, slot_index, slot_bytes.data(), slot_bytes.size(), &written) == NE_OK;
}

// HostGenotypeSlab (genotype_slab/slab_allocator.hpp:23-30) as this build downloads it: the layout, every
// slot's bytes, each slot's liveness count (0 = free) and the free-slot stack with its count.
struct HostGenotypeSlab {
    ne_slab_info layout{};
    std::vector<std::uint8_t> storage;
    std::vector<std::uint32_t> slot_liveness;
    std::vector<std::uint32_t> free_slot_stack;
    std::uint32_t free_slot_count = 0;
    const std::uint8_t* SlotBytes(std::uint32_t slot) const {
        return storage.data() + std::size_t(slot) * layout.slot_stride_bytes;
    }
};

// TryDownloadSlabFromDevice (slab_runtime.hpp:126)
inline bool TryDownloadSlabFromDevice(const DeviceSlabGARuntimeBuffers& buffers, HostGenotypeSlab& host_buffer) {
    host_buffer = {};
    if (ne_get_slab_info(buffers.handle, &host_buffer.layout) != NE_OK) return false;
    const std::uint64_t slots = host_buffer.layout.slot_count;
    host_buffer.storage.assign(slots * host_buffer.layout.slot_stride_bytes, 0);
    host_buffer.slot_liveness.assign(slots, 0);
    host_buffer.free_slot_stack.assign(slots, 0);
    return ne_download_slab(buffers.handle, host_buffer.storage.data(), host_buffer.storage.size(),
                            host_buffer.slot_liveness.data(), host_buffer.free_slot_stack.data(), slots,
                            &host_buffer.free_slot_count) == NE_OK;
}

// SwapDeviceSlabGenerationBuffers (slab_runtime.hpp:181)
inline void SwapDeviceSlabGenerationBuffers(DeviceSlabGARuntimeBuffers& buffers) noexcept {
    (void)ne_swap_generations(buffers.handle);
}

// TryReadDeviceSlabGARuntimeStatus (slab_runtime.hpp:183-184)
inline bool TryReadDeviceSlabGARuntimeStatus(const DeviceSlabGARuntimeBuffers& buffers,
                                             DeviceSlabGARuntimeStatusCode& status_code) {
    if (!buffers.handle) return false;
    status_code = static_cast<DeviceSlabGARuntimeStatusCode>(ne_read_status(buffers.handle));
    return true;
}

// DeviceSlabGARuntimeStatusCodeString (slab_runtime.hpp:186)
inline const char* DeviceSlabGARuntimeStatusCodeString(DeviceSlabGARuntimeStatusCode status_code) noexcept {
    return ne_status_string(static_cast<int>(status_code));
}

}  // namespace ne_amd::slab_device

// ---- single-model inference (src/inference/single_model_device_runtime.hpp:9-49) -------------------
namespace ne_amd::inference {

enum class SingleModelDeviceRuntimeStatusCode : int {
    kOk = 0,
    kInvalidRuntime = 1,
    kCudaMemcpyFailed = 2,
    kKernelLaunchFailed = 3,
    kKernelExecutionFailed = 4,
    kPolicyForwardFailed = 5,
    kActionSelectionFailed = 6,
};

struct SingleModelDeviceRuntime {
    ne_model* handle = nullptr;
    std::size_t action_count = 0;
};

// TryCreateSingleModelDeviceRuntime (single_model_device_runtime.hpp:38-40); action_letters holds 5 letter
// indices per word.
inline bool TryCreateSingleModelDeviceRuntime(const std::uint8_t* host_genome_bytes, std::size_t genome_byte_count,
                                              const std::vector<std::uint8_t>& action_letters,
                                              SingleModelDeviceRuntime& runtime_out) {
    runtime_out = {};
    ne_model* m = nullptr;
    if (ne_model_create(host_genome_bytes, genome_byte_count, action_letters.data(), action_letters.size() / 5, nullptr,
                        &m) != NE_OK)
        return false;
    runtime_out = {m, action_letters.size() / 5};
    return true;
}

// The same for a genome of another layout, e.g. the reference's own bytes ({128, NE_STORAGE_FP16}: the
// 145 -> 128 -> 64 encoder stored as binary16, ComputeDynamicGenomeStrideBytes(action_count) bytes).
inline bool TryCreateSingleModelDeviceRuntime(const std::uint8_t* host_genome_bytes, std::size_t genome_byte_count,
                                              const ne_model_layout& layout,
                                              const std::vector<std::uint8_t>& action_letters,
                                              SingleModelDeviceRuntime& runtime_out) {
    runtime_out = {};
    ne_model* m = nullptr;
    if (ne_model_create_with_layout(host_genome_bytes, genome_byte_count, &layout, action_letters.data(),
                                    action_letters.size() / 5, nullptr, &m) != NE_OK)
        return false;
    runtime_out = {m, action_letters.size() / 5};
    return true;
}

inline void DestroySingleModelDeviceRuntime(SingleModelDeviceRuntime& runtime) noexcept {
    ne_model_destroy(runtime.handle);
    runtime = {};
}

// TrySelectNextGuessWithSingleModelDeviceRuntime (single_model_device_runtime.hpp:44-47)
inline bool TrySelectNextGuessWithSingleModelDeviceRuntime(const SingleModelDeviceRuntime& runtime,
                                                           const ne_board& grid, ne_selected_action& selected_out,
                                                           SingleModelDeviceRuntimeStatusCode* status_out = nullptr) {
    using S = SingleModelDeviceRuntimeStatusCode;
    S st = S::kInvalidRuntime;
    if (runtime.handle) {
        switch (ne_model_select_guess(runtime.handle, &grid, &selected_out)) {
            case NE_OK: st = S::kOk; break;
            case NE_POLICY_FORWARD_FAILED: st = S::kPolicyForwardFailed; break;
            case NE_ACTION_SELECTION_FAILED: st = S::kActionSelectionFailed; break;
            default: st = S::kKernelExecutionFailed;
        }
    }
    if (status_out) *status_out = st;
    return st == S::kOk;
}

}  // namespace ne_amd::inference

// ---- the winner artifact (src/model_artifact/winner_artifact*.hpp) --------------------------------
namespace ne_amd::model_artifact {

using WinnerArtifactMetadata = ne_winner_metadata;

struct WinnerArtifactPaths {
    std::filesystem::path binary_path{};
    std::filesystem::path metadata_path{};
    std::string timestamp_local{};
};

struct LoadedWinnerArtifact {
    WinnerArtifactMetadata metadata{};
    std::vector<std::uint8_t> action_letters{};
    std::vector<std::uint8_t> genome_bytes{};
};

// TryWriteWinnerArtifact (winner_artifact.hpp:28-30)
inline bool TryWriteWinnerArtifact(const std::filesystem::path& directory, const std::uint8_t* genome_bytes,
                                   const std::vector<std::uint8_t>& action_letters, WinnerArtifactMetadata metadata,
                                   WinnerArtifactPaths& paths_out) {
    paths_out = {};
    char bin[4096], meta[4096];
    if (ne_write_winner_artifact_in_directory(directory.c_str(), genome_bytes, action_letters.data(),
                                              action_letters.size() / 5, &metadata, bin, meta, sizeof(bin)) != NE_OK)
        return false;
    paths_out = {bin, meta, metadata.timestamp_local};
    return true;
}

// TryWriteWinnerArtifactToPaths (winner_artifact.hpp:32-35)
inline bool TryWriteWinnerArtifactToPaths(const std::filesystem::path& binary_path,
                                          const std::filesystem::path& metadata_path, const std::uint8_t* genome_bytes,
                                          const std::vector<std::uint8_t>& action_letters,
                                          WinnerArtifactMetadata metadata, WinnerArtifactPaths& paths_out) {
    paths_out = {};
    if (ne_write_winner_artifact(binary_path.c_str(), metadata_path.c_str(), genome_bytes, action_letters.data(),
                                 action_letters.size() / 5, &metadata) != NE_OK)
        return false;
    paths_out = {binary_path, metadata_path, metadata.timestamp_local};
    return true;
}

// TryReadWinnerArtifact (winner_artifact_reader.hpp:20-21)
inline bool TryReadWinnerArtifact(const std::filesystem::path& binary_path, const std::filesystem::path& metadata_path,
                                  LoadedWinnerArtifact& artifact_out) {
    artifact_out = {};
    WinnerArtifactMetadata m{};
    if (ne_read_winner_artifact(binary_path.c_str(), metadata_path.c_str(), &m, nullptr, 0, nullptr, 0) != NE_OK)
        return false;
    artifact_out.genome_bytes.resize(m.genome_byte_count);
    artifact_out.action_letters.resize(5 * m.action_count);
    if (ne_read_winner_artifact(binary_path.c_str(), metadata_path.c_str(), &m, artifact_out.genome_bytes.data(),
                                m.genome_byte_count, artifact_out.action_letters.data(), m.action_count) != NE_OK) {
        artifact_out = {};
        return false;
    }
    artifact_out.metadata = m;
    return true;
}

}  // namespace ne_amd::model_artifact
