// ref_probe.cpp -- exposes a few integer-only functions of the reference, compiled directly from
// its headers (src/wordle/*.hpp, src/spatial/grid.hpp, src/training_folder/training_data.hpp), as
// a C ABI for the oracle's tests.  TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ by
// oracle/build_oracle.py when the reference sources are present; never shipped or loaded by the
// product.
#include <cstdint>
#include <cstring>

#include "genetic_algorithm/device/runtime_common.hpp"
#include "genetic_algorithm/generation_assembly.hpp"
#include "spatial/grid.hpp"
#include "training_folder/training_data.hpp"
#include "wordle/feedback.hpp"
#include "wordle/hint_grid.hpp"
#include "wordle/wordle_grid.hpp"

namespace wd = neuroevolution::wordle;
namespace sp = neuroevolution::spatial;
namespace tf = neuroevolution::training_folder;

static wd::Word MakeWord(const uint8_t* letters) {
    wd::Word w{};
    for (int p = 0; p < 5; ++p) w.letter_indices[p] = letters[p];
    return w;
}

extern "C" {

int ref_feedback(const uint8_t* guess, const uint8_t* solution, uint8_t* out) {
    wd::Feedback fb{};
    if (!wd::TryProvideFeedback(MakeWord(guess), MakeWord(solution), fb)) return 0;
    for (int p = 0; p < 5; ++p) out[p] = static_cast<uint8_t>(fb[p]);
    return 1;
}

// Hint-grid group of a word: for each of the 3 grids, turn count then up to 5 guesses (5 letters)
// and their tiles.  Returns 1 on success, 0 when the group cannot be built.
int ref_hint_group(const uint8_t* word, uint32_t* turn_counts, uint8_t* letters, uint8_t* tiles) {
    wd::HintGridGroup g{};
    if (!wd::TryBuildHintGridGroup(MakeWord(word), g)) return 0;
    for (int i = 0; i < 3; ++i) {
        turn_counts[i] = static_cast<uint32_t>(g.grids[i].turn_count);
        for (std::size_t t = 0; t < g.grids[i].turn_count && t < 5; ++t)
            for (int p = 0; p < 5; ++p) {
                letters[(i * 5 + t) * 5 + p] = g.grids[i].turns[t].guess.letter_indices[p];
                tiles[(i * 5 + t) * 5 + p] = static_cast<uint8_t>(g.grids[i].turns[t].feedback[p]);
            }
    }
    return 1;
}

int ref_moore(uint64_t rows, uint64_t cols, uint64_t focal, uint64_t radius, uint64_t* out) {
    sp::CellularGridShape shape{};
    if (!sp::TryMakeRectangularCellularGridShape(rows, cols, shape)) return -1;
    sp::CellularNeighborList list{};
    if (!sp::TryCollectMooreRadiusNeighbors(shape, focal, radius, list)) return -1;
    for (std::size_t i = 0; i < list.count; ++i) out[i] = list.indices[i];
    return static_cast<int>(list.count);
}

uint64_t ref_chebyshev(uint64_t rows, uint64_t cols, uint64_t a, uint64_t b) {
    sp::CellularGridShape shape{};
    if (!sp::TryMakeRectangularCellularGridShape(rows, cols, shape)) return ~uint64_t(0);
    return sp::ToroidalChebyshevDistance(shape, a, b);
}

// Per-cell local training-word lists from the reference shard schedule (training_data.hpp:251-316
// and evaluation_ops.cuh:122-169): writes the CSR offsets [cells+1] and catalog indices.
int ref_shard_csr(uint64_t initial, uint64_t step, uint64_t period, uint64_t introduced, uint64_t generation,
                  uint64_t rows, uint64_t cols, uint64_t epi_rows, uint64_t radius0, uint64_t growth,
                  const uint64_t* release_generations, uint64_t release_count, uint32_t* offsets, uint32_t* ids,
                  uint64_t ids_capacity) {
    tf::WordCountSchedule schedule{initial, step, period};
    sp::CellularGridShape grid{}, epi{};
    if (!sp::TryMakeRectangularCellularGridShape(rows, cols, grid) ||
        !sp::TryMakeRectangularCellularGridShape(epi_rows, cols, epi))
        return -1;
    tf::TrainingDataShardReleaseHistory history{};
    const tf::TrainingDataShardReleaseHistory* hist_ptr = nullptr;
    if (release_generations != nullptr) {
        for (uint64_t i = 0; i < release_count; ++i) history.release_generations[i] = release_generations[i];
        history.release_count = release_count;
        hist_ptr = &history;
    }
    static tf::TrainingDataShardRuntimeSet set{};
    set = {};
    if (!tf::TryBuildTrainingDataShardRuntimeSet(schedule, introduced, generation, grid, epi, radius0, growth,
                                                 hist_ptr, set))
        return -1;
    uint64_t n = 0;
    for (uint64_t c = 0; c < grid.cell_count; ++c) {
        offsets[c] = static_cast<uint32_t>(n);
        for (std::size_t s = 0; s < set.shard_count; ++s) {
            if (!tf::DoesTrainingDataShardCoverCell(set.shards[s], grid, c)) continue;
            for (std::size_t k = 0; k < set.shards[s].word_count; ++k) {
                if (n >= ids_capacity) return -2;
                ids[n++] = static_cast<uint32_t>(set.shards[s].first_catalog_word_index + k);
            }
        }
    }
    offsets[grid.cell_count] = static_cast<uint32_t>(n);
    return static_cast<int>(set.shard_count);
}

// Byte images of the configuration structs the checkpoint writer stores whole
// (slab_runtime.cu:1962-1971), built from the reference's own types.
int ref_word_counts_image(const uint64_t* v, uint8_t* out) {
    neuroevolution::genetic_algorithm::device_common::RuntimeWordCounts w{};
    w.training_word_count = v[0];
    w.action_space_word_count = v[1];
    w.training_word_schedule.initial_word_count = v[2];
    w.training_word_schedule.word_count_step = v[3];
    w.training_word_schedule.word_count_step_period_generations = v[4];
    w.shard_initial_radius = v[5];
    w.shard_radius_growth_period_generations = v[6];
    std::memcpy(out, &w, sizeof(w));
    return static_cast<int>(sizeof(w));
}

int ref_assembly_image(uint64_t radius, const float* f, uint8_t* out) {
    neuroevolution::genetic_algorithm::GenerationAssemblyConfig c{};
    std::memset(&c, 0, sizeof(c));
    c.parent_selection.cellular_breeding_radius = radius;
    c.parent_selection.rank_exponent = f[0];
    c.breeding.crossover_temperature_level1 = f[1];
    c.breeding.crossover_temperature_level2 = f[2];
    c.breeding.crossover_temperature_level3 = f[3];
    c.mutation.mutation_probability = f[4];
    c.mutation.mutation_sigma = f[5];
    c.mutation.output_tail_row_scale_mutation_probability = f[6];
    std::memcpy(out, &c, sizeof(c));
    return static_cast<int>(sizeof(c));
}

// The defaults of a value-initialised GenerationAssemblyConfig: radius, then the 7 floats.
void ref_assembly_defaults(uint64_t* radius, float* f) {
    const neuroevolution::genetic_algorithm::GenerationAssemblyConfig c{};
    *radius = c.parent_selection.cellular_breeding_radius;
    f[0] = c.parent_selection.rank_exponent;
    f[1] = c.breeding.crossover_temperature_level1;
    f[2] = c.breeding.crossover_temperature_level2;
    f[3] = c.breeding.crossover_temperature_level3;
    f[4] = c.mutation.mutation_probability;
    f[5] = c.mutation.mutation_sigma;
    f[6] = c.mutation.output_tail_row_scale_mutation_probability;
}

int ref_grid_shape_image(uint64_t rows, uint64_t cols, uint8_t* out) {
    sp::CellularGridShape g{};
    if (!sp::TryMakeRectangularCellularGridShape(rows, cols, g)) return -1;
    std::memcpy(out, &g, sizeof(g));
    return static_cast<int>(sizeof(g));
}

// sizeof(wordle::Word), of a release-history entry (generation, p99), and the history capacity.
void ref_checkpoint_sizes(uint64_t* out) {
    const tf::TrainingDataShardReleaseHistory h{};
    out[0] = sizeof(wd::Word);
    out[1] = sizeof(h.release_generations[0]);
    out[2] = sizeof(h.release_fitness_p99[0]);
    out[3] = sizeof(h.release_count);
    out[4] = tf::kTrainingWordCatalogCapacity;
}

}  // extern "C"
