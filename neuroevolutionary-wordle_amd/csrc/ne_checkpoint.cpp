// ne_checkpoint.cpp -- checkpoint encoding, checksum, and atomic file write/read (host only; no
// device calls).  The file layout follows TryWriteRuntimeCheckpointAtomically and
// TryReadRuntimeCheckpoint (src/genetic_algorithm/device/slab_runtime.cu:1934-2081) field for field.
#include "ne_checkpoint.hpp"

#include <cstdio>
#include <filesystem>
#include <new>
#include <string>
#include <system_error>

namespace ne_ckpt {

void EncodeWordCounts(const ne_word_counts& w, uint8_t out[kWordCountsBytes]) {
    const uint64_t v[7] = {w.training_word_count,
                           w.action_space_word_count,
                           w.training_word_schedule.initial_word_count,
                           w.training_word_schedule.word_count_step,
                           w.training_word_schedule.word_count_step_period_generations,
                           w.shard_initial_radius,
                           w.shard_radius_growth_period_generations};
    std::memcpy(out, v, sizeof(v));
}

ne_word_counts DecodeWordCounts(const uint8_t in[kWordCountsBytes]) {
    ne_word_counts w{};
    w.training_word_count = LoadU64(in + 0);
    w.action_space_word_count = LoadU64(in + 8);
    w.training_word_schedule.initial_word_count = LoadU64(in + 16);
    w.training_word_schedule.word_count_step = LoadU64(in + 24);
    w.training_word_schedule.word_count_step_period_generations = LoadU64(in + 32);
    w.shard_initial_radius = LoadU64(in + 40);
    w.shard_radius_growth_period_generations = LoadU64(in + 48);
    return w;
}

// GenerationAssemblyConfig (generation_assembly.hpp:9-13): ParentSelectionConfig {size_t radius; float
// rank_exponent;} padded to 16, BreedingConfig {3 float} at 16, MutationConfig {3 float} at 28.
void EncodeAssembly(const ne_assembly_config& c, uint8_t out[kAssemblyBytes]) {
    std::memset(out, 0, kAssemblyBytes);
    StoreU64(out, c.cellular_breeding_radius);
    const float f[7] = {c.rank_exponent,
                        c.crossover_temperature_level1,
                        c.crossover_temperature_level2,
                        c.crossover_temperature_level3,
                        c.mutation_probability,
                        c.mutation_sigma,
                        c.output_tail_row_scale_mutation_probability};
    std::memcpy(out + 8, &f[0], 4);
    std::memcpy(out + 16, &f[1], 24);
}

ne_assembly_config DecodeAssembly(const uint8_t in[kAssemblyBytes]) {
    ne_assembly_config c{};
    c.cellular_breeding_radius = LoadU64(in);
    float f[6];
    std::memcpy(&c.rank_exponent, in + 8, 4);
    std::memcpy(f, in + 16, 24);
    c.crossover_temperature_level1 = f[0];
    c.crossover_temperature_level2 = f[1];
    c.crossover_temperature_level3 = f[2];
    c.mutation_probability = f[3];
    c.mutation_sigma = f[4];
    c.output_tail_row_scale_mutation_probability = f[5];
    return c;
}

// PendingOutputEmbeddingInjection (genotype_slab/device_runtime.hpp:26-30): bool at 0, two size_t at 8, 16.
void EncodeInjection(const ne_pending_injection& j, uint8_t out[kInjectionBytes]) {
    std::memset(out, 0, kInjectionBytes);
    out[0] = j.enabled ? 1 : 0;
    StoreU64(out + 8, j.first_catalog_word_index);
    StoreU64(out + 16, j.injection_count);
}

ne_pending_injection DecodeInjection(const uint8_t in[kInjectionBytes]) {
    ne_pending_injection j{};
    j.enabled = in[0] != 0 ? 1u : 0u;
    j.first_catalog_word_index = LoadU64(in + 8);
    j.injection_count = LoadU64(in + 16);
    return j;
}

static void MixHistory(Fnv1a64& f, const std::vector<uint64_t>& gens, const std::vector<float>& p99) {
    f.Value(uint64_t(gens.size()));
    for (size_t i = 0; i < gens.size(); ++i) {
        f.Value(gens[i]);
        f.Value(p99[i]);
    }
}

uint64_t IdentityHash(const uint8_t word_counts[kWordCountsBytes], bool with_history,
                      const std::vector<uint64_t>& release_generations, const std::vector<float>& release_p99,
                      const uint8_t* catalog_letters, uint64_t catalog_words) {
    Fnv1a64 f;
    f.Bytes(word_counts, kWordCountsBytes);
    if (with_history) MixHistory(f, release_generations, release_p99);
    if (catalog_letters) {
        f.Value(catalog_words);
        f.Bytes(catalog_letters, 5 * catalog_words);  // wordle::Word is 5 letter-index bytes (word.hpp:14-16)
    }
    return f.h;
}

uint64_t Checksum(const ne_checkpoint& c) {
    Fnv1a64 f;
    f.Value(c.schema_version);
    f.Value(c.genome_layout_version);
    f.Value(c.resume_phase);
    f.Value(c.identity_hash);
    f.Value(c.generation_seed);
    f.Bytes(c.word_counts, kWordCountsBytes);
    MixHistory(f, c.release_generations, c.release_p99);
    f.Bytes(c.assembly, kAssemblyBytes);
    f.Bytes(c.injection, kInjectionBytes);
    f.Bytes(c.runtime_config, kRuntimeConfigBytes);
    f.Bytes(c.slab_layout, kSlabLayoutBytes);
    for (const auto& s : c.shapes) f.Bytes(s, kGridShapeBytes);
    f.Value(c.generation_index);
    f.Value(uint64_t(c.slots.size()));
    for (size_t i = 0; i < c.slots.size(); ++i) {
        f.Value(c.slots[i]);
        f.Value(c.fitness[i]);
        f.Value(c.evals[i]);
        f.Value(c.has[i]);
    }
    f.Value(uint64_t(c.ChildCount()));
    f.Bytes(c.pairs.data(), c.pairs.size() * 4);
    f.Value(uint64_t(c.Records()));
    for (size_t r = 0; r < c.Records(); ++r) {
        f.Value(c.organisms[r]);
        f.Value(c.genome_bytes);
        f.Bytes(c.genomes.data() + r * c.genome_bytes, c.genome_bytes);
    }
    return f.h;
}

uint64_t PayloadBytes(const ne_checkpoint& c) {
    return c.ChildCount() * 8 + 8 + c.release_generations.size() * 12 + c.slots.size() * 13 +
           c.Records() * c.genome_bytes;
}

}  // namespace ne_ckpt

using namespace ne_ckpt;

namespace {

struct Writer {
    std::FILE* f;
    bool ok = true;
    void Bytes(const void* p, size_t n) {
        if (ok && n) ok = std::fwrite(p, 1, n, f) == n;
    }
    template <typename T>
    void Value(const T& v) {
        Bytes(&v, sizeof(T));
    }
};

struct Reader {
    std::FILE* f;
    uint64_t left;  // bytes not yet read: bounds every count taken from the file before allocating
    bool ok = true;
    void Bytes(void* p, size_t n) {
        if (!ok) return;
        if (n > left) {
            ok = false;
            return;
        }
        if (n) ok = std::fread(p, 1, n, f) == n;
        left -= n;
    }
    template <typename T>
    void Value(T& v) {
        Bytes(&v, sizeof(T));
    }
    bool Fits(uint64_t count, uint64_t unit) const { return ok && count <= left / unit; }
};

int Invalid(ne_checkpoint* c) {
    delete c;
    return NE_CHECKPOINT_INVALID;
}

}  // namespace

extern "C" {

int ne_write_checkpoint(const ne_checkpoint* c, const char* path) {
    if (!c || !path || !*path) return NE_INVALID_ARGUMENT;
    if (c->checksum != Checksum(*c)) return NE_CHECKPOINT_INVALID;
    namespace fs = std::filesystem;
    const fs::path target(path);
    const fs::path tmp = fs::path(target.string() + ".tmp");
    std::error_code ec;
    if (!target.parent_path().empty()) {
        fs::create_directories(target.parent_path(), ec);
        if (ec) return NE_IO_FAILURE;
    }
    std::FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return NE_IO_FAILURE;
    Writer w{f};
    w.Value(kMagic);
    w.Value(c->schema_version);
    w.Value(c->genome_layout_version);
    w.Value(c->checksum);
    w.Value(c->identity_hash);
    w.Value(c->resume_phase);
    w.Value(c->generation_seed);
    w.Bytes(c->word_counts, kWordCountsBytes);
    w.Value(uint64_t(c->release_generations.size()));
    for (size_t i = 0; i < c->release_generations.size(); ++i) {
        w.Value(c->release_generations[i]);
        w.Value(c->release_p99[i]);
    }
    w.Bytes(c->assembly, kAssemblyBytes);
    w.Bytes(c->injection, kInjectionBytes);
    w.Bytes(c->runtime_config, kRuntimeConfigBytes);
    w.Bytes(c->slab_layout, kSlabLayoutBytes);
    for (const auto& s : c->shapes) w.Bytes(s, kGridShapeBytes);
    w.Value(c->generation_index);
    w.Value(uint64_t(c->slots.size()));
    for (size_t i = 0; i < c->slots.size(); ++i) {
        w.Value(c->slots[i]);
        w.Value(c->fitness[i]);
        w.Value(c->evals[i]);
        w.Value(c->has[i]);
    }
    w.Value(uint64_t(c->ChildCount()));
    w.Bytes(c->pairs.data(), c->pairs.size() * 4);
    w.Value(uint64_t(c->Records()));
    for (size_t r = 0; r < c->Records(); ++r) {
        w.Value(c->organisms[r]);
        w.Value(c->genome_bytes);
        w.Bytes(c->genomes.data() + r * c->genome_bytes, c->genome_bytes);
    }
    const bool closed = std::fclose(f) == 0;
    if (!w.ok || !closed) {
        fs::remove(tmp, ec);
        return NE_IO_FAILURE;
    }
    fs::rename(tmp, target, ec);
    if (ec) {  // as the reference: remove the old file and rename again
        ec.clear();
        fs::remove(target, ec);
        ec.clear();
        fs::rename(tmp, target, ec);
    }
    if (ec) {
        fs::remove(tmp, ec);
        return NE_IO_FAILURE;
    }
    return NE_OK;
}

int ne_read_checkpoint(const char* path, ne_checkpoint** out) {
    if (!path || !out) return NE_INVALID_ARGUMENT;
    *out = nullptr;
    std::error_code ec;
    const uint64_t size = std::filesystem::file_size(path, ec);
    if (ec) return NE_IO_FAILURE;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return NE_IO_FAILURE;
    ne_checkpoint* c = new (std::nothrow) ne_checkpoint();
    if (!c) {
        std::fclose(f);
        return NE_DEVICE_FAILURE;
    }
    Reader r{f, size};
    uint64_t magic = 0;
    r.Value(magic);
    r.ok &= magic == kMagic;
    r.Value(c->schema_version);
    r.Value(c->genome_layout_version);
    r.Value(c->checksum);
    r.Value(c->identity_hash);
    r.Value(c->resume_phase);
    r.Value(c->generation_seed);
    r.Bytes(c->word_counts, kWordCountsBytes);
    uint64_t releases = 0;
    r.Value(releases);
    r.ok &= releases <= kReleaseCapacity && r.Fits(releases, 12);
    if (r.ok) {
        c->release_generations.resize(releases);
        c->release_p99.resize(releases);
    }
    for (uint64_t i = 0; r.ok && i < releases; ++i) {
        r.Value(c->release_generations[i]);
        r.Value(c->release_p99[i]);
    }
    r.Bytes(c->assembly, kAssemblyBytes);
    r.Bytes(c->injection, kInjectionBytes);
    r.Bytes(c->runtime_config, kRuntimeConfigBytes);
    r.Bytes(c->slab_layout, kSlabLayoutBytes);
    for (auto& s : c->shapes) r.Bytes(s, kGridShapeBytes);
    uint64_t count = 0;
    r.Value(c->generation_index);
    r.Value(count);
    r.ok &= count > 0 && r.Fits(count, 13);  // TryCreateSlabGeneration rejects an empty generation
    if (r.ok) {
        c->slots.resize(count);
        c->fitness.resize(count);
        c->evals.resize(count);
        c->has.resize(count);
    }
    for (uint64_t i = 0; r.ok && i < count; ++i) {
        r.Value(c->slots[i]);
        r.Value(c->fitness[i]);
        r.Value(c->evals[i]);
        r.Value(c->has[i]);
    }
    uint64_t children = 0;
    r.Value(children);
    r.ok &= children > 0 && r.Fits(children, 8);  // TryCreateSlabAssemblyPlan rejects an empty plan
    if (r.ok) c->pairs.resize(2 * children);
    if (r.ok) r.Bytes(c->pairs.data(), c->pairs.size() * 4);
    uint64_t live = 0;
    r.Value(live);
    r.ok &= live <= count;
    // the stride comes from the file: bounded by the bytes left before anything is sized from it, so a
    // corrupt stride can neither wrap `stride + 12` or `live * stride` nor divide by zero (whether it is
    // the stride of the file's action count is checked at restore, as TryRestorePrebreedingCheckpoint-
    // ToDevice does, slab_runtime.cu:1790-1813)
    c->genome_bytes = c->Stride();
    r.ok &= live == 0 || (c->genome_bytes > 0 && c->genome_bytes <= r.left && r.Fits(live, c->genome_bytes + 12));
    if (r.ok) {
        try {
            c->organisms.resize(live);
            c->genomes.resize(live * c->genome_bytes);  // live * stride <= bytes left in the file
        } catch (const std::exception&) {
            r.ok = false;
        }
    }
    for (uint64_t k = 0; r.ok && k < live; ++k) {
        uint64_t bytes = 0;
        r.Value(c->organisms[k]);
        r.Value(bytes);
        r.ok &= bytes == c->genome_bytes;
        r.Bytes(c->genomes.data() + k * c->genome_bytes, c->genome_bytes);
    }
    std::fclose(f);
    if (!r.ok || c->schema_version != kSchemaVersion || c->genome_layout_version != kGenomeLayoutVersion ||
        c->checksum != Checksum(*c))
        return Invalid(c);
    *out = c;
    return NE_OK;
}

int ne_clone_checkpoint(const ne_checkpoint* c, ne_checkpoint** out) {
    if (!c || !out) return NE_INVALID_ARGUMENT;
    *out = new (std::nothrow) ne_checkpoint(*c);
    return *out ? NE_OK : NE_DEVICE_FAILURE;
}

void ne_free_checkpoint(ne_checkpoint* c) { delete c; }

int ne_get_checkpoint_info(const ne_checkpoint* c, ne_checkpoint_info* out) {
    if (!c || !out) return NE_INVALID_ARGUMENT;
    ne_checkpoint_info i{};
    i.schema_version = c->schema_version;
    i.genome_layout_version = c->genome_layout_version;
    i.resume_phase = c->resume_phase;
    i.generation_seed = c->generation_seed;
    i.checksum = c->checksum;
    i.training_data_identity_hash = c->identity_hash;
    i.generation_index = c->generation_index;
    i.population = c->slots.size();
    i.child_count = c->ChildCount();
    i.live_genotypes = c->Records();
    i.payload_bytes = PayloadBytes(*c);
    i.action_count = c->ActionCount();
    i.slot_stride_bytes = c->Stride();
    i.slot_count = c->SlotCount();
    i.slab_bytes = c->SlabBytes();
    i.generation_byte_budget = c->RuntimeConfigField(1);
    i.grid_column_count = c->RuntimeConfigField(5);
    i.release_count = c->release_generations.size();
    i.word_counts = DecodeWordCounts(c->word_counts);
    i.assembly_config = DecodeAssembly(c->assembly);
    i.injection = DecodeInjection(c->injection);
    *out = i;
    return NE_OK;
}

int ne_get_checkpoint_generation(const ne_checkpoint* c, uint32_t* slot_indices, float* fitness,
                                 uint32_t* evaluation_counts, uint8_t* has_fitness, uint64_t capacity) {
    if (!c) return NE_INVALID_ARGUMENT;
    const uint64_t n = c->slots.size();
    if (capacity < n) return NE_INVALID_ARGUMENT;
    if (slot_indices) std::memcpy(slot_indices, c->slots.data(), n * 4);
    if (fitness) std::memcpy(fitness, c->fitness.data(), n * 4);
    if (evaluation_counts) std::memcpy(evaluation_counts, c->evals.data(), n * 4);
    if (has_fitness) std::memcpy(has_fitness, c->has.data(), n);
    return NE_OK;
}

int ne_get_checkpoint_plan(const ne_checkpoint* c, uint32_t* pairs, uint64_t capacity_pairs) {
    if (!c || !pairs || capacity_pairs < c->ChildCount()) return NE_INVALID_ARGUMENT;
    std::memcpy(pairs, c->pairs.data(), c->pairs.size() * 4);
    return NE_OK;
}

int ne_get_checkpoint_genome(const ne_checkpoint* c, uint64_t record, uint32_t* organism_index, void* dst,
                             uint64_t capacity) {
    if (!c || record >= c->Records()) return NE_INVALID_ARGUMENT;
    if (organism_index) *organism_index = c->organisms[record];
    if (dst) {
        if (capacity < c->genome_bytes) return NE_INVALID_ARGUMENT;
        std::memcpy(dst, c->genomes.data() + record * c->genome_bytes, c->genome_bytes);
    }
    return NE_OK;
}

uint64_t ne_training_data_identity_hash(const ne_word_counts* counts, const ne_shard_release_history* history,
                                        const uint8_t* catalog_letters, uint64_t catalog_word_count) {
    uint8_t wc[kWordCountsBytes] = {};
    if (counts) EncodeWordCounts(*counts, wc);
    std::vector<uint64_t> gens;
    std::vector<float> p99;
    if (history && history->release_count) {
        if (!history->release_generations) return 0;
        gens.assign(history->release_generations, history->release_generations + history->release_count);
        p99.assign(history->release_count, 0.0f);
        if (history->release_fitness_p99)
            p99.assign(history->release_fitness_p99, history->release_fitness_p99 + history->release_count);
    }
    return IdentityHash(wc, history != nullptr, gens, p99, catalog_letters, catalog_word_count);
}

}  // extern "C"
