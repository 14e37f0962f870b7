// run_ga.cpp -- thin C++ driver of the GA generation step through the host API of slab_runtime.hpp:
// bootstrap, then `generations` calls of TryAdvanceGenerationOnDevice with the reference CLI's
// generation-seed convention (seed + 2 + g, src/cli/run_genetic_algorithm.cu:1883-1884), then a
// final evaluation, printing one summary line per generation.  With --checkpoint-every K the steps
// from generations K, 2K, ... go through TryCreatePrebreedingCheckpointOnDevice, a background write to
// --checkpoint PATH and TryResumeGenerationFromCheckpointOnDevice; --resume PATH restores a checkpoint
// into a new runtime and continues from it (the reference CLI's flow, run_genetic_algorithm.cu:1605-1660,
// 1900-1967); --winner DIR writes the final generation's best genome as a winner artifact; --graph runs
// the plain steps as captured hipGraphs (TrySetCapturedGenerationSteps; the steps then carry no
// post-fitness callback, which would keep them eager, and evaluations are counted per step instead).
//
// usage: run_ga <workload_dir> <grid_columns> <generations> <seed> [mutation_probability]
//               [--checkpoint PATH --checkpoint-every K] [--resume PATH] [--winner DIR] [--graph]
// workload_dir holds catalog.u8 (5 letter indices per word), offsets.u32 (cells+1) and secrets.u32.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "slab_runtime.hpp"

using namespace ne_amd::slab_device;

template <typename T>
static bool ReadAll(const std::string& path, std::vector<T>& out) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return false;
    const std::streamsize n = f.tellg();
    f.seekg(0);
    out.resize(static_cast<std::size_t>(n) / sizeof(T));
    return static_cast<bool>(f.read(reinterpret_cast<char*>(out.data()), n));
}

static int Fail(const DeviceSlabGARuntimeBuffers& b, const char* what) {
    DeviceSlabGARuntimeStatusCode st = DeviceSlabGARuntimeStatusCode::kOk;
    (void)TryReadDeviceSlabGARuntimeStatus(b, st);
    std::fprintf(stderr, "%s failed: %s\n", what, DeviceSlabGARuntimeStatusCodeString(st));
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr,
                     "usage: %s <workload_dir> <grid_columns> <generations> <seed> [mutation_p] "
                     "[--checkpoint PATH --checkpoint-every K] [--resume PATH] [--winner DIR] [--graph]\n",
                     argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const std::size_t cols = std::strtoull(argv[2], nullptr, 10);
    const std::size_t generations = std::strtoull(argv[3], nullptr, 10);
    const std::uint32_t seed = static_cast<std::uint32_t>(std::strtoul(argv[4], nullptr, 10));
    float mutation_p = 0.01f;
    std::string checkpoint_path, resume_path, winner_dir;
    std::size_t checkpoint_every = 0;
    bool graph = false;
    for (int a = 5; a < argc; ++a) {
        const std::string arg = argv[a];
        if (arg == "--graph") {
            graph = true;
        } else if (arg == "--checkpoint" && a + 1 < argc) {
            checkpoint_path = argv[++a];
        } else if (arg == "--checkpoint-every" && a + 1 < argc) {
            checkpoint_every = std::strtoull(argv[++a], nullptr, 10);
        } else if (arg == "--resume" && a + 1 < argc) {
            resume_path = argv[++a];
        } else if (arg == "--winner" && a + 1 < argc) {
            winner_dir = argv[++a];
        } else if (a == 5) {
            mutation_p = std::strtof(argv[a], nullptr);
        } else {
            std::fprintf(stderr, "unknown argument %s\n", argv[a]);
            return 2;
        }
    }
    if (checkpoint_every != 0 && checkpoint_path.empty()) {
        std::fprintf(stderr, "--checkpoint-every needs --checkpoint PATH\n");
        return 2;
    }
    std::vector<std::uint8_t> catalog;
    std::vector<std::uint32_t> offsets, secrets;
    if (!ReadAll(dir + "/catalog.u8", catalog) || !ReadAll(dir + "/offsets.u32", offsets) ||
        !ReadAll(dir + "/secrets.u32", secrets) || offsets.size() < 2) {
        std::fprintf(stderr, "cannot read the workload in %s\n", dir.c_str());
        return 2;
    }
    const std::size_t actions = catalog.size() / 5;
    const std::size_t cells = offsets.size() - 1;
    const std::size_t stride = ((132928 + 38 * actions) * 4 + 255) / 256 * 256;

    DeviceSlabGARuntimeConfig config{};
    config.genotype_slab_byte_budget_bytes = 2 * cells * stride + stride;
    config.generation_byte_budget_bytes = cells * stride;
    config.action_count = actions;
    config.grid_column_count = cols;
    DeviceSlabGARuntimeBuffers buffers{};
    RuntimeCheckpoint resumed{};
    if (!resume_path.empty()) {
        if (!TryReadRuntimeCheckpoint(resume_path, resumed)) {
            std::fprintf(stderr, "could not read or validate checkpoint %s\n", resume_path.c_str());
            return 1;
        }
        if (!TryRestorePrebreedingCheckpointToDevice(resumed, buffers)) {
            std::fprintf(stderr, "could not restore checkpoint %s\n", resume_path.c_str());
            return 1;
        }
    } else if (!TryCreateDeviceSlabGARuntimeBuffers(buffers, config)) {
        std::fprintf(stderr, "runtime creation failed\n");
        return 1;
    }
    if (!TryUploadTrainingWordCatalog(buffers, catalog) || !TryUploadTrainingSecrets(buffers, offsets, secrets))
        return Fail(buffers, "input upload");
    if (graph && !TrySetCapturedGenerationSteps(buffers, true)) return Fail(buffers, "captured steps");
    std::size_t first_generation = 0;
    if (resumed.Valid()) {
        if (!TryResumeGenerationFromCheckpointOnDevice(buffers, resumed)) return Fail(buffers, "checkpoint resume");
        first_generation = resumed.Info().generation_index + 1;
        std::printf("resumed generation %zu from %s\n", first_generation, resume_path.c_str());
    } else {
        const ne_init_config init{1.0f, 1.0f, 0.1f, 0.0f};
        if (!TryBootstrapRandomCurrentGenerationOnDevice(buffers, cells, seed, 0, &init))
            return Fail(buffers, "bootstrap");
    }

    RuntimeWordCounts counts{};
    counts.training_word_count = actions;
    counts.action_space_word_count = actions;
    counts.training_word_schedule.initial_word_count = actions;
    GenerationAssemblyConfig assembly{};
    assembly.mutation.mutation_probability = mutation_p;
    std::size_t evaluated = 0;
    const PostFitnessEvaluationCallback count_evaluations = [&](const DeviceSlabGARuntimeBuffers&,
                                                                const RuntimeWordCounts&) {
        ++evaluated;
        return true;
    };
    RuntimeCheckpointAsyncWriter writer{};
    const auto start = std::chrono::steady_clock::now();
    for (std::size_t g = first_generation; g < generations; ++g) {
        const std::uint32_t generation_seed = seed + 2u + static_cast<std::uint32_t>(g);
        if (checkpoint_every != 0 && g != 0 && g % checkpoint_every == 0) {
            RuntimeCheckpoint checkpoint{}, for_write{};
            if (!TryCreatePrebreedingCheckpointOnDevice(buffers, generation_seed, counts, assembly, {}, checkpoint,
                                                        &catalog, false, count_evaluations))
                return Fail(buffers, "checkpoint creation");
            if (!writer.TryWaitForWrite() || !TryCloneRuntimeCheckpoint(checkpoint, for_write) ||
                !writer.TryStartWrite(std::move(for_write), checkpoint_path)) {
                std::fprintf(stderr, "checkpoint write failed\n");
                return 1;
            }
            if (!TryResumeGenerationFromCheckpointOnDevice(buffers, checkpoint)) return Fail(buffers, "resume");
        } else if (!TryAdvanceGenerationOnDevice(buffers, generation_seed, counts, assembly, {}, false,
                                                 graph ? PostFitnessEvaluationCallback{} : count_evaluations)) {
            return Fail(buffers, "advance");
        } else if (graph) {
            ++evaluated;
        }
        PopulationFitnessSummary s{};
        if (!TryReadPopulationFitnessSummaryFromDevice(buffers, s)) return Fail(buffers, "summary");
        std::printf("generation %zu: best %.6f mean %.6f best_index %zu population %zu\n", s.generation_index,
                    s.best_fitness, s.average_fitness, s.best_index, s.population_size);
    }
    if (!writer.TryWaitForWrite()) {
        std::fprintf(stderr, "checkpoint write failed\n");
        return 1;
    }
    if (!TryEvaluateCurrentGenerationFitnessOnDevice(buffers, counts)) return Fail(buffers, "final evaluation");
    PopulationFitnessSummary s{};
    (void)TryReadPopulationFitnessSummaryFromDevice(buffers, s);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::printf("final generation %zu: best %.6f mean %.6f (%zu post-fitness callbacks, %.3f s)\n",
                s.generation_index, s.best_fitness, s.average_fitness, evaluated, secs);
    SlabGeneration current{};
    if (!TryDownloadCurrentGenerationFromDevice(buffers, current) || current.slot_indices.size() != cells)
        return Fail(buffers, "download");
    if (!winner_dir.empty()) {  // TryPersistCheckpointWinningGenome's artifact (run_genetic_algorithm.cu:1925-1934)
        std::vector<std::uint8_t> genome;
        if (!TryDownloadSlabSlotBytesFromDevice(buffers, s.best_slot_index, genome)) return Fail(buffers, "winner");
        ne_amd::model_artifact::WinnerArtifactMetadata meta{};
        meta.best_fitness = s.best_fitness;
        meta.generation_index = s.generation_index;
        meta.best_index = s.best_index;
        meta.best_slot_index = s.best_slot_index;
        meta.action_count = actions;
        meta.genome_byte_count = genome.size();
        meta.seed = seed;
        std::snprintf(meta.action_space_path, sizeof(meta.action_space_path), "%s/catalog.u8", dir.c_str());
        ne_amd::model_artifact::WinnerArtifactPaths paths{};
        if (!ne_amd::model_artifact::TryWriteWinnerArtifact(winner_dir, genome.data(), catalog, meta, paths)) {
            std::fprintf(stderr, "could not write the winner artifact into %s\n", winner_dir.c_str());
            return 1;
        }
        std::printf("winner artifact %s\n", paths.binary_path.c_str());
    }
    DestroyDeviceSlabGARuntimeBuffers(buffers);
    return evaluated == generations - first_generation ? 0 : 1;
}
