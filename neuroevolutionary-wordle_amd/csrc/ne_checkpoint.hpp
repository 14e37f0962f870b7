// ne_checkpoint.hpp -- the in-memory checkpoint and its file format (host only).
//
// The container is the reference's RuntimeCheckpoint file (src/genetic_algorithm/device/
// slab_runtime.cu:1934-2081): the same header, field order, little-endian scalars, LP64 images of the
// reference's configuration structs written whole (padding included) and the same checksum.  The genome
// records differ: they hold this build's fp32 slots (the reference stores fp16 parameters), so the file
// carries its own genome layout version and each side's reader refuses the other's genomes.
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/ne_runtime.h"

namespace ne_ckpt {

constexpr uint64_t kMagic = 0x4B4350474147524Eull;  // slab_runtime.cu:1184 ("NRGAGPCK" as bytes)
constexpr uint32_t kSchemaVersion = 3;              // kRuntimeCheckpointSchemaVersion (slab_runtime.hpp:86)
// kRuntimeCheckpointGenomeLayoutVersion (slab_runtime.hpp:87) is 1 for the reference's fp16 genomes; fp32
// slots with the 145->64 encoder are "FP32" (these four bytes in file order).
constexpr uint32_t kGenomeLayoutVersion = 0x32335046u;
constexpr uint32_t kPhasePreRecombination = 1;      // RuntimeCheckpointResumePhase (slab_runtime.hpp:89-91)
constexpr uint64_t kReleaseCapacity = 4739;         // kTrainingWordCatalogCapacity (training_data.hpp:22)

// LP64 images of the reference structs the writer stores whole.
constexpr size_t kWordCountsBytes = 56;     // RuntimeWordCounts: 7 x size_t
constexpr size_t kAssemblyBytes = 40;       // GenerationAssemblyConfig: {size_t, float, pad 4}, 3 float, 3 float
constexpr size_t kInjectionBytes = 24;      // PendingOutputEmbeddingInjection: {bool, pad 7, size_t, size_t}
constexpr size_t kRuntimeConfigBytes = 48;  // DeviceSlabGARuntimeConfig: 6 x size_t
constexpr size_t kSlabLayoutBytes = 32;     // GenotypeSlabLayout: action_count, stride, slot_count, slab_bytes
constexpr size_t kGridShapeBytes = 24;      // CellularGridShape: rows, columns, cells

// FNV-1a, 64 bit (xor the byte, then multiply by the prime 1099511628211), as MixCheckpointBytes
// (slab_runtime.cu:1186-1196).  The reference starts from 1469598103934665603 (slab_runtime.cu:1207,
// 1283): the published offset basis 14695981039346656037 with its last digit dropped.  Checksums and
// identity hashes use that start so that files interoperate.
constexpr uint64_t kReferenceFnvStart = 1469598103934665603ull;
struct Fnv1a64 {
    uint64_t h = kReferenceFnvStart;
    void Bytes(const void* p, size_t n) {
        const auto* b = static_cast<const uint8_t*>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 1099511628211ull;
        }
    }
    template <typename T>
    void Value(const T& v) {
        Bytes(&v, sizeof(T));
    }
};

inline uint64_t LoadU64(const uint8_t* p) {
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}
inline void StoreU64(uint8_t* p, uint64_t v) { std::memcpy(p, &v, 8); }

void EncodeWordCounts(const ne_word_counts& w, uint8_t out[kWordCountsBytes]);
ne_word_counts DecodeWordCounts(const uint8_t in[kWordCountsBytes]);
void EncodeAssembly(const ne_assembly_config& c, uint8_t out[kAssemblyBytes]);
ne_assembly_config DecodeAssembly(const uint8_t in[kAssemblyBytes]);
void EncodeInjection(const ne_pending_injection& j, uint8_t out[kInjectionBytes]);
ne_pending_injection DecodeInjection(const uint8_t in[kInjectionBytes]);

// Identity hash of TryDownloadBoundaryCheckpointPayload (slab_runtime.cu:1282-1296): the history is mixed
// only when one is given (with_history), the catalog only when catalog_letters is not null.
uint64_t IdentityHash(const uint8_t word_counts[kWordCountsBytes], bool with_history,
                      const std::vector<uint64_t>& release_generations, const std::vector<float>& release_p99,
                      const uint8_t* catalog_letters, uint64_t catalog_words);

}  // namespace ne_ckpt

struct ne_checkpoint {
    uint32_t schema_version = ne_ckpt::kSchemaVersion;
    uint32_t genome_layout_version = ne_ckpt::kGenomeLayoutVersion;
    uint32_t resume_phase = ne_ckpt::kPhasePreRecombination;
    uint32_t generation_seed = 0;
    uint64_t checksum = 0;
    uint64_t identity_hash = 0;
    uint8_t word_counts[ne_ckpt::kWordCountsBytes] = {};
    std::vector<uint64_t> release_generations;
    std::vector<float> release_p99;
    uint8_t assembly[ne_ckpt::kAssemblyBytes] = {};
    uint8_t injection[ne_ckpt::kInjectionBytes] = {};
    uint8_t runtime_config[ne_ckpt::kRuntimeConfigBytes] = {};
    uint8_t slab_layout[ne_ckpt::kSlabLayoutBytes] = {};
    uint8_t shapes[3][ne_ckpt::kGridShapeBytes] = {};  // current, next, epicenter
    uint64_t generation_index = 0;
    std::vector<uint32_t> slots;
    std::vector<float> fitness;
    std::vector<uint32_t> evals;
    std::vector<uint8_t> has;
    std::vector<uint32_t> pairs;      // [child][2]
    std::vector<uint32_t> organisms;  // organism index of each genome record
    uint64_t genome_bytes = 0;        // bytes per record (the slab stride)
    std::vector<uint8_t> genomes;     // [record][genome_bytes]

    uint64_t ActionCount() const { return ne_ckpt::LoadU64(slab_layout + 0); }
    uint64_t Stride() const { return ne_ckpt::LoadU64(slab_layout + 8); }
    uint64_t SlotCount() const { return ne_ckpt::LoadU64(slab_layout + 16); }
    uint64_t SlabBytes() const { return ne_ckpt::LoadU64(slab_layout + 24); }
    uint64_t RuntimeConfigField(int i) const { return ne_ckpt::LoadU64(runtime_config + 8 * i); }
    uint64_t ShapeField(int shape, int i) const { return ne_ckpt::LoadU64(shapes[shape] + 8 * i); }
    uint64_t ChildCount() const { return pairs.size() / 2; }
    uint64_t Records() const { return organisms.size(); }
};

namespace ne_ckpt {
// ComputeRuntimeCheckpointChecksum (slab_runtime.cu:1206-1248).
uint64_t Checksum(const ne_checkpoint& c);
// CheckpointPayloadByteCount (run_genetic_algorithm.cu:1360-1372).
uint64_t PayloadBytes(const ne_checkpoint& c);
}  // namespace ne_ckpt
