// ne_kernels.hpp -- argument blocks and launch wrappers of the GA kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ne_device.hpp"

namespace ne {

struct SummaryRecord {
    float best_fitness;
    float average_fitness;
    uint32_t best_index;
    uint32_t best_slot;
};

// The persistent evaluation's work queues: one per XCD of the MI355X (workgroups are dispatched to the 8
// XCDs round-robin, so workgroup b runs on XCD b mod 8), each queue's counter on its own 128-byte line.
constexpr uint32_t kWorkQueues = 8;
constexpr uint32_t kWorkCounterStride = 32;

struct EvalArgs {
    const uint8_t* slab;
    uint64_t stride;
    uint32_t slot_count;
    uint32_t action_count;         // selectable actions: min(slab A, runtime A)
    const uint32_t* slot_indices;  // current generation
    const uint32_t* cell_offsets;  // CSR of local secrets per cell (= individual)
    const uint32_t* secret_ids;
    const uint32_t* codes;         // packed catalog words
    const uint32_t* canon;         // [catalog tiles * 16] lowest index holding the same word (padded)
    uint32_t canon_bytes;
    uint32_t canon_identity;       // 1: every catalog word is unique (canon[a] == a), so masks compare action indices
    const uint32_t* img;           // [catalog tiles][kImgTileWords] scoring image
    uint32_t img_bytes;
    const float* enc_t;            // [N][kEncTStride]: transposed encoder, then W0[:, 0], per genome
    const float* fixed_absmax;     // [26] max over the catalog of |fixed feature d| (BF16 bound)
    uint32_t force_fallback;       // BF16 mode: settle every decision with the FP32 scan (diagnostic)
    uint32_t trunk_bf16;           // BF16 mode: trunk layers 1 and 2 on BF16 MFMA (FP32 accumulate)
    const uint8_t* bf16_cat;       // BF16 mode: [catalog tiles][kBf16CatTileBytes] catalog screening operands
    uint32_t bf16_cat_bytes;
    const uint8_t* bf16_ops;       // BF16 mode: [N][a_tiles][kBf16GenTileBytes] per-genome screening operands
    uint64_t bf16_stride;          // bytes per genome in bf16_ops
    const float* tail_absmax;      // BF16 mode: [N][40] max |tail value| per tail dim
    float* raw_fitness;            // [N] summed episode scores
    uint32_t* local_w;             // [N]
    unsigned long long* counters;  // [0] decisions, [1] encoded turns, [2] BF16 candidates rescored, [3] FP32 fallbacks
    int* status;
    GameTrace* game_trace;         // optional (trace mode)
    float* policy_trace;           // optional: [games][6][64]
    const uint64_t* game_offsets;  // [N] first game index of each individual (trace mode)
    const uint32_t* work;          // [work_count][2] (genome, split) items in kWorkQueues queues
    const uint32_t* queue_begin;   // [kWorkQueues + 1]: queue q holds items queue_begin[q] .. queue_begin[q + 1] - 1
    uint32_t work_count;
    uint32_t* work_counter;        // [kWorkQueues][kWorkCounterStride]: next item of each queue (0 at launch)
    uint32_t splits;               // the largest split count of a genome
    float* acc_scratch;            // [workgroup][kEvalMaxWaves][2][16 KB] layer-0 accumulators between decisions
};

struct PlanArgs {
    const float* fitness;
    const uint8_t* has_fitness;
    uint32_t cur_size;
    uint32_t child_count;
    uint32_t cols;
    uint32_t gen_index;
    uint32_t radius;
    uint32_t generation_seed;
    const float* rank_weights;     // [rank] -> weight, rank 1..cur_size
    uint32_t* pairs;
    int* status;
    const uint32_t* step;          // optional: device {gen_index, generation_seed} in place of the fields
};

// Per child, the element positions where the donor parent can change, ascending: [0] = count, then
// the positions (layer, neuron-row, bias, tail-row and splice bounds, and the genome end).
constexpr uint32_t kBreakStride = 64;

struct XoverArgs {
    uint32_t child_count;
    uint32_t gen_index;
    uint32_t seed;
    uint32_t parent_action_count;
    float t1, t2, t3;
    XoverRecord* records;
    uint32_t* breaks;              // [child_count][kBreakStride]
    const uint32_t* step;          // optional: device {gen_index, seed} in place of the fields
};

struct BreedArgs {
    uint8_t* slab;
    uint64_t parent_stride;
    uint64_t child_stride;
    const uint32_t* pairs;
    const uint32_t* parent_slots;
    const XoverRecord* records;
    const uint32_t* breaks;        // [children][kBreakStride] from the crossover-decision kernel
    const uint32_t* free_stack;
    uint32_t free_count_before;
    uint32_t child_offset;
    uint32_t* child_slots;
    uint32_t* slot_state;
    uint32_t parent_action_count;
    uint32_t gen_index;
    uint32_t seed;
    float mutation_p;
    float mutation_sigma;
    float row_scale_p;
    uint32_t slot_count;
    int* status;
    uint32_t elem_begin;           // child elements written: [elem_begin, elem_end), multiples of 4
    uint32_t elem_end;
    const uint32_t* child_list;    // optional: child index of each batch entry (else child_offset + entry)
    const uint32_t* given_slots;   // optional: [child] slots already holding the child (completing it)
    const uint32_t* step;          // optional: device {gen_index, seed} in place of the fields
    const uint32_t* free_count_dev;  // optional: the free count read on the device in place of
                                     // free_count_before; fewer free slots than the batch -> slab full
    uint32_t interleave_cells;     // optional (brood of B x cells children, no child list): batch entry e
                                   // breeds child (e % B) x cells + e / B, so a cell's B brood members run
                                   // side by side and share their neighbourhood's parents in the caches;
                                   // each child keeps the slot and draws of its own index
};

// What a captured generation step leaves for one host read after its replay.
struct StepReadback {
    int status;
    uint32_t free_count;
    SummaryRecord summary;
    unsigned long long counters[4];
};

struct BootArgs {
    uint8_t* slab;
    uint64_t stride;
    uint32_t action_count;
    const uint32_t* free_stack;
    uint32_t free_count_before;
    uint32_t* slots;
    uint32_t* slot_state;
    uint32_t gen_index;
    uint32_t seed;
    float layer_std[kDenseLayers];
    float tail_std;
};

struct RepackArgs {
    uint32_t* cur_slots;
    uint32_t cur_size;
    const uint32_t* refcounts;
    uint32_t* slot_state;
    uint32_t* free_stack;
    uint32_t* free_count;
    uint32_t* owner;       // [old_slot_count], filled with kInvalidSlot by the caller
    uint32_t* move_src;    // [cur_size]
    uint32_t* survivors;   // out
    uint32_t old_slot_count;
    uint32_t new_slot_count;
    int* status;
};

constexpr uint32_t kMaxSortedNorms = 8192;  // largest parent action count the median sort handles

// One hint-grid state: turn count and, per turn, the packed guess and packed feedback.
struct HintState {
    uint32_t tc;
    uint32_t code[4];
    uint32_t fb[4];
};

struct InjectArgs {
    uint8_t* slab;
    uint64_t stride;
    const uint32_t* child_slots;
    uint32_t children;
    uint32_t parent_action_count;
    uint32_t count;              // new tail rows per child
    const HintState* states;     // [count][3]
    float* medians;              // [children]
    int* status;
};

// ---- surrogate screening --------------------------------------------------------------------
// The 64 probe game states of a child's fingerprint: 16 secrets x boards of 0..3 turns (probe
// p = 16 d + k holds the first d turns of board k).
constexpr int kProbeSecrets = 16;
constexpr int kProbeDepths = 4;
constexpr int kProbes = kProbeSecrets * kProbeDepths;
constexpr int kFingerprintDims = 3 * kProbes;  // 192
struct ProbeSet {
    uint32_t secret_code[kProbeSecrets];
    uint32_t turn_code[kProbeSecrets][kProbeDepths - 1];
    uint32_t turn_fb[kProbeSecrets][kProbeDepths - 1];
    uint32_t turn_canon[kProbeSecrets][kProbeDepths - 1];
};

struct FingerprintArgs {
    const uint8_t* slab;
    uint64_t stride;
    uint32_t slot_count;
    uint32_t action_count;
    const uint32_t* slots;         // [n] children
    const uint32_t* codes;
    const uint32_t* img;
    uint32_t img_bytes;
    const float* enc_t;            // [n][kEncTStride] (TransposeEncoderKernel over `slots`)
    const ProbeSet* probes;
    const uint32_t* pairs;         // [n][2] parent positions in the parent generation
    const float* parent_fitness;   // [parents]
    float parent_shift, parent_scale;  // parent fitness input = (f - shift) * scale
    float* out;                    // [n][kFingerprintDims]
    int* status;
    uint32_t trunk_bf16;           // 1: the trunk's layers 1 and 2 on split-BF16 MFMA (the BF16-trunk form)
};

void LaunchFingerprint(const FingerprintArgs& a, uint32_t n, hipStream_t s);

// Surrogate MLP 192 -> 128 -> 64 -> 1 (ReLU, ReLU, linear), fp32.  Parameters in one array: W1T
// [192][128] (input-major), b1 [128], W2T [128][64], b2 [64], w3 [64], b3.
constexpr uint32_t kSgIn = kFingerprintDims, kSgH1 = 128, kSgH2 = 64;
constexpr uint32_t kSgOffW1 = 0;
constexpr uint32_t kSgOffB1 = kSgOffW1 + kSgIn * kSgH1;
constexpr uint32_t kSgOffW2 = kSgOffB1 + kSgH1;
constexpr uint32_t kSgOffB2 = kSgOffW2 + kSgH1 * kSgH2;
constexpr uint32_t kSgOffW3 = kSgOffB2 + kSgH2;
constexpr uint32_t kSgOffB3 = kSgOffW3 + kSgH2;
constexpr uint32_t kSgParams = kSgOffB3 + 1;  // 33,025

void LaunchSurrogateInit(float* params, float* adam_m, float* adam_v, uint32_t seed, hipStream_t s);
// y[n] = f(x[n][192]); h1 [n][128] and h2 [n][64] (post-ReLU) are kept when non-null.
void LaunchSurrogateForward(const float* params, const float* x, uint32_t n, float* y, float* h1, float* h2,
                            hipStream_t s);
// Loss 1/(2n) sum (y - t)^2: dy [n], dh2 [n][64], dh1 [n][128].
void LaunchSurrogateBackprop(const float* params, const float* h1, const float* h2, const float* y, const float* t,
                             uint32_t n, float* dy, float* dh1, float* dh2, hipStream_t s);
// Gradient of every parameter over the batch, then one Adam step (beta 0.9 / 0.999, eps 1e-8); coef (device)
// = {1 / (1 - 0.9^step), 1 / (1 - 0.999^step)}, so a captured graph replays with new steps.
void LaunchSurrogateAdam(float* params, float* adam_m, float* adam_v, const float* x, const float* h1, const float* h2,
                         const float* dy, const float* dh1, const float* dh2, uint32_t n, float lr, const float* coef,
                         hipStream_t s);
// Replay buffer: rows (head + e) % cap <- fingerprint rows src_row[e] of fp, fitness[e].
void LaunchReplayAppend(const float* fp, const uint32_t* src_row, const float* fitness, uint32_t count, float* rx,
                        float* rf, uint32_t head, uint32_t cap, hipStream_t s);
// Minibatch: x[b] = rx[idx[b]], t[b] = (rf[idx[b]] - stats[0]) * stats[1] (stats on the device).
void LaunchReplayGather(const float* rx, const float* rf, const uint32_t* idx, uint32_t n, const float* stats,
                        float* x, float* t, hipStream_t s);
void LaunchInject(const InjectArgs& a, hipStream_t s);
void LaunchPackBf16(const uint8_t* slab, uint64_t stride, const uint32_t* slots, uint32_t n, uint32_t slot_count,
                    const uint32_t* img, uint32_t A, uint8_t* out, uint64_t out_stride, uint32_t* tail_absmax,
                    hipStream_t s);
void LaunchTransposeEncoder(const uint8_t* slab, uint64_t stride, const uint32_t* slots, uint32_t n,
                            uint32_t slot_count, float* enc_t, hipStream_t s);
// Persistent evaluation: `blocks` workgroups of `waves` waves drain the work list.
// The evaluation kernels of the untraced forms (identical results in each form); traced evaluations always
// use the first.
constexpr uint32_t kEvalKernelRegisters = 0;  // one tile per wave, layer-0 accumulators in registers
constexpr uint32_t kEvalKernelStreamed = 1;   // one tile per wave, accumulators in scratch, 4 waves per SIMD
constexpr uint32_t kEvalKernelPaired = 2;     // two tiles per wave, accumulators in scratch
void LaunchEvaluate(const EvalArgs& a, uint32_t blocks, uint32_t waves, uint32_t kernel, bool trace, bool bf16,
                    hipStream_t s);
// Workgroups of `waves` waves of the evaluation kernel that fit on one CU at once (0 if unknown).
uint32_t EvaluateResidentPerCu(uint32_t waves, uint32_t kernel, bool trace, bool bf16, bool t16 = false);
// bytes of layer-0 accumulator scratch one evaluation workgroup uses (either kernel)
uint64_t EvaluateAccScratchBytes();
// The most waves per evaluation workgroup this build supports (its LDS tables are sized for them).
uint32_t EvaluateMaxWaves();
// The most waves per workgroup of one evaluation kernel in one form (the BF16 form's streamed and two-tile
// kernels: 3).
uint32_t EvaluateMaxWavesFor(uint32_t kernel, bool trace, bool bf16);
// The kernel that runs a requested (kernel, waves) in a form: traced evaluations use the register kernel, and
// so does the BF16 form's streamed kernel at any size but 3 waves.
uint32_t EvaluateKernelFor(uint32_t kernel, uint32_t waves, bool trace, bool bf16);
void LaunchFinalize(const float* raw, const uint32_t* local_w, uint32_t n, float* fitness, uint32_t* evals,
                    uint8_t* has, int* status, hipStream_t s);
void LaunchSummary(const float* fitness, const uint8_t* has, const uint32_t* slots, uint32_t n, SummaryRecord* out,
                   int* status, hipStream_t s);
void LaunchPlan(const PlanArgs& a, hipStream_t s);
void LaunchXover(const XoverArgs& a, hipStream_t s);
void LaunchBreed(const BreedArgs& a, uint32_t batch, hipStream_t s);
void LaunchBootstrap(const BootArgs& a, uint32_t organisms, hipStream_t s);
void LaunchRefcount(const uint32_t* pairs, uint32_t child_count, uint32_t cur_size, uint32_t* refcounts, int* status,
                    hipStream_t s);
void LaunchCollectReleased(uint32_t* cur_slots, uint32_t cur_size, const uint32_t* refcounts, uint32_t* slot_state,
                           uint32_t* free_stack, uint32_t* free_count, uint32_t slot_count, int* status,
                           hipStream_t s);
void LaunchReleaseBatch(const uint32_t* pairs, uint32_t child_offset, uint32_t batch, uint32_t* refcounts,
                        int* status, hipStream_t s);
void LaunchPopSlots(uint32_t* free_count, uint32_t count, hipStream_t s);
void LaunchSpillRelease(uint32_t* child_slots, uint32_t k, uint32_t fc, uint32_t* slot_state, uint32_t* free_stack,
                        uint32_t* free_count, uint32_t slot_count, int* status, hipStream_t s);
void LaunchSpillRestore(uint32_t* child_slots, uint32_t k, uint32_t fc, uint32_t* slot_state,
                        const uint32_t* free_stack, uint32_t* free_count, int* status, hipStream_t s);
void LaunchInitSlabMetadata(uint32_t* slot_state, uint32_t* free_stack, uint32_t* free_count, uint32_t slot_count,
                            hipStream_t s);
void LaunchRepackPlan(const RepackArgs& a, hipStream_t s);
void LaunchRepackCopies(uint8_t* slab, uint64_t old_stride, uint64_t new_stride, const uint32_t* move_src,
                        uint32_t survivors, uint32_t base, uint8_t* staging, hipStream_t s);
void LaunchFillU32(uint32_t* dst, uint32_t value, uint64_t count, hipStream_t s);
void LaunchFillU8(uint8_t* dst, uint8_t value, uint64_t count, hipStream_t s);
void LaunchStepReadback(const int* status, const uint32_t* free_count, const SummaryRecord* summary,
                        const unsigned long long* counters, StepReadback* out, hipStream_t s);

// ---- generation telemetry (ne_telemetry.hip) ----
constexpr uint32_t kFitnessStatsMax = 16384;  // generation sizes the LDS sort holds (64 KB)
struct FitnessStatsRecord {
    double sum;             // ascending-order double sum of the fitness
    double squared_deltas;  // ascending-order double sum of (fitness - mean)^2
    float min, max, median_low, median_high, p90, p99;
    uint32_t distinct, missing, childless, one_child, multiple_children;
};
void LaunchFitnessStats(const float* fit, const uint8_t* has, uint32_t n, uint32_t pow2, uint32_t i90, uint32_t i99,
                        const uint32_t* refcounts, FitnessStatsRecord* out, hipStream_t s);
void LaunchConvergence(const uint8_t* slab, uint64_t stride, const uint32_t* slots, uint32_t n_org,
                       const uint32_t* weights, uint32_t n_w, uint64_t scalars, const uint32_t* pairs, uint32_t n_pairs,
                       float inverse, float* by_org, float* by_weight, float* centroid, float* squared, hipStream_t s);

}  // namespace ne
