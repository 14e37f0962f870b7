// ne_runtime.hip -- host orchestration of the GA generation step behind the C ABI in
// include/ne_runtime.h.  One handle owns its slab, generation buffers, catalog and stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <ctime>
#include <cstdarg>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <numeric>
#include <random>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include "../../include/ne_runtime.h"
#include "ne_checkpoint.hpp"
#include "ne_device.hpp"
#include "ne_kernels.hpp"

using namespace ne;

namespace {

constexpr uint64_t kCatalogCapacity = 1u << 20;

struct StageEvents {
    hipEvent_t begin = nullptr, end = nullptr;
};

// One evaluation's launch description: the kernel arguments, the workgroup shape and the games played.
struct EvalSetup {
    EvalArgs a;
    uint32_t blocks, waves, kernel;
    bool bf16, trace;
    uint64_t total_games;
};

// Everything a captured generation step bakes in: the kernel arguments (with the per-step values left
// to the device block) and the generation buffers the clears and the summary use.  Compared byte for
// byte: the structures are zero-filled before their fields are set.
struct StepGraphKey {
    EvalArgs a;
    PlanArgs pa;
    XoverArgs xa;
    BreedArgs b;
    uint64_t n, next_n, max_gen;
    uint32_t blocks, waves, bf16, pad;
    const void* buffers[8];
};

struct StepGraph {
    StepGraphKey key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t last_use = 0;
};

}  // namespace

struct ne_runtime {
    ne_runtime_config config{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;

    // slab
    uint8_t* slab = nullptr;
    uint64_t slab_bytes = 0;
    uint64_t action_count = 0;
    uint64_t stride = 0;
    uint64_t slot_count = 0;
    uint32_t* slot_state = nullptr;
    uint32_t* free_stack = nullptr;
    uint32_t* free_count = nullptr;  // device scalar

    // generations
    uint64_t max_gen = 0;
    uint64_t epi_rows = 0;  // rows of the epicenter grid the shard centres are drawn on
    uint64_t generation_byte_budget = 0;
    uint64_t cur_index = 0, cur_size = 0, next_index = 0, next_size = 0, planned_children = 0;
    uint32_t* cur_slots = nullptr;
    float* cur_fit = nullptr;
    uint32_t* cur_evals = nullptr;
    uint8_t* cur_has = nullptr;
    uint32_t* nxt_slots = nullptr;
    float* nxt_fit = nullptr;
    uint32_t* nxt_evals = nullptr;
    uint8_t* nxt_has = nullptr;

    // plan and assembly
    uint32_t* pairs = nullptr;
    uint32_t* refcounts = nullptr;
    XoverRecord* xover = nullptr;
    uint32_t* breaks = nullptr;  // [max_gen][kBreakStride] donor-change positions per child
    float* rank_weights = nullptr;
    uint64_t rank_weights_len = 0;
    float rank_weights_exponent = -1.0f;

    // evaluation
    float* enc_t = nullptr;  // [max_gen][145][64] transposed encoders of the generation being evaluated
    uint8_t* bf16_ops = nullptr;   // BF16 form: packed screening operands of the generation
    uint64_t bf16_ops_bytes = 0;
    uint32_t* tail_absmax = nullptr;  // BF16 form: [max_gen][40] max |tail value| per dim (float bits)
    float* raw = nullptr;
    uint32_t* local_w = nullptr;
    unsigned long long* counters = nullptr;
    int* status = nullptr;
    SummaryRecord* summary = nullptr;
    uint32_t* d_offsets = nullptr;
    uint32_t* d_secrets = nullptr;
    uint64_t d_secrets_cap = 0;
    std::vector<uint32_t> h_offsets, h_secrets;
    bool csr_from_caller = false;

    // catalog
    uint32_t* codes = nullptr;
    uint32_t* canon = nullptr;  // lowest catalog index holding the same word
    bool canon_identity = false;  // no word occurs twice in the catalog (canon[a] == a)
    uint32_t* img = nullptr;    // scoring image, kImgTileWords per tile of 16 words
    float* fixed_absmax = nullptr;  // [26] max over the catalog of |fixed feature|
    uint8_t* bf16_cat = nullptr;    // [tiles][kBf16CatTileBytes] BF16 screening operands of fixed dims 0..23
    uint64_t bf16_cat_bytes = 0;
    uint64_t img_bytes = 0;
    uint64_t catalog_count = 0;
    std::vector<uint8_t> h_catalog;

    // repack
    uint32_t* owner = nullptr;
    uint64_t owner_cap = 0;
    uint32_t* move_src = nullptr;
    uint32_t* d_survivors = nullptr;
    uint8_t* staging = nullptr;
    uint64_t staging_cap = 0;

    // trace
    uint32_t trace_flags = 0;
    uint32_t eval_mode = NE_EVAL_FP32;
    GameTrace* d_trace = nullptr;
    float* d_policy_trace = nullptr;
    uint64_t* d_game_offsets = nullptr;
    uint64_t trace_games = 0;
    uint64_t trace_capacity = 0;

    // timing
    StageEvents ev[14];  // 12: the repack's gather and scatter copies; 13: the brood's second phase
    ne_stage_times times{};
    ne_fitness_summary last_summary{};
    bool have_summary = false;
    int last_status = NE_OK;
    bool cur_evaluated = false;  // the current generation's fitness is that of its own evaluation
    bool verbose = false;        // progress lines of the step being run (the caller's `verbose`)
    // what that fitness was measured on: a screened step reuses it only while these still hold
    uint64_t upload_epoch = 1;          // bumped by every catalog or secrets upload (not by evaluations)
    uint64_t fit_upload_epoch = 0;      // upload_epoch when the fitness was measured (0: unknown)
    uint64_t fit_action_space = 0;
    std::vector<uint32_t> fit_offsets, fit_secrets;  // the per-cell secrets it was measured on
    uint32_t* d_work = nullptr;         // persistent evaluation's work list
    uint64_t work_cap = 0;
    uint32_t* d_work_counter = nullptr;
    uint64_t cu_count = 256;
    uint32_t tuned_waves = 0;           // 0: the idle-slot heuristic picks the workgroup size
    uint32_t tuned_rounds = 1;          // rounds of `waves` tiles one work item plays
    uint32_t tuned_wg_per_cu = 0;       // persistent workgroups per CU (0: as many as fit)
    uint32_t tuned_kernel = 0;          // evaluation kernel of the untraced forms (kEvalKernel*)
    float* eval_acc = nullptr;          // the evaluation's layer-0 scratch, EvaluateAccScratchBytes() per workgroup
    uint64_t eval_acc_bytes = 0;

    // captured generation steps (ne_set_step_graph): one graph per buffer parity, the evaluation's host
    // preparation cached while its inputs stay the same
    bool step_graph_on = false;
    StepGraph step_graphs[2];
    uint64_t step_graph_uses = 0;
    uint32_t* d_step = nullptr;             // {gen_index, generation_seed} of the step being replayed
    uint32_t* h_step = nullptr;             // pinned
    StepReadback* d_readback = nullptr;
    StepReadback* h_readback = nullptr;     // pinned
    uint64_t input_epoch = 1;               // bumped by every catalog or secrets upload
    EvalSetup cached_eval{};
    uint64_t cached_eval_key[10] = {};
    // the same for eager evaluations (EvaluateCore): the key, and the per-cell lists the buffers hold
    EvalSetup eager_eval{};
    uint64_t eager_eval_key[10] = {};
    std::vector<uint32_t> eager_offsets, eager_secrets;

    // surrogate screening
    bool screen_on = false;
    bool screen_active = false;  // the next screened step assembles a brood
    ne_screen_config screen{};
    ProbeSet h_probes{};
    uint32_t h_probe_words[kProbeSecrets][kProbeDepths] = {};  // secret index, then the 3 guesses' indices
    ProbeSet* d_probes = nullptr;
    float* sg_params = nullptr;
    float* sg_m = nullptr;
    float* sg_v = nullptr;
    uint64_t sg_step = 0;
    float* fp = nullptr;    // [max_gen][192] fingerprints of the brood
    float* pred = nullptr;  // [max_gen]
    float* e_fit = nullptr;
    uint32_t* e_evals = nullptr;
    uint8_t* e_has = nullptr;
    uint32_t* e_slots = nullptr;
    uint32_t* e_rows = nullptr;
    float* replay_x = nullptr;
    float* replay_f = nullptr;
    std::vector<float> h_replay_f;
    uint64_t replay_count = 0, replay_head = 0;
    float *bx = nullptr, *bt = nullptr, *by = nullptr, *bh1 = nullptr, *bh2 = nullptr;
    float *bdy = nullptr, *bdh1 = nullptr, *bdh2 = nullptr;
    uint32_t* bidx = nullptr;
    float* sg_coef = nullptr;   // [0..1] target standardisation, then per training step {c1, c2}
    float* sg_coef1 = nullptr;  // {c1, c2} of a direct ne_surrogate_train_step
    hipGraph_t train_graph = nullptr;       // one generation's training steps, captured once
    hipGraphExec_t train_exec = nullptr;
    // The training runs on its own stream (default priority), overlapping the next step's planning, brood
    // assembly and fingerprints; the next prediction (and every surrogate call) waits for it on the device.
    hipStream_t train_stream = nullptr;
    hipEvent_t train_ready = nullptr;       // on the main stream: the replay rows of this step are in
    hipEvent_t train_t0 = nullptr, train_done = nullptr;  // on the training stream
    bool train_pending = false;             // a training not yet ordered before the main stream's later work
    bool train_unread = false;              // its loss and time not yet read on the host
    uint32_t* h_bidx = nullptr;             // pinned: minibatch indices, standardisation and bias corrections,
    float* h_coef = nullptr;                //   and the last minibatch's outputs and targets (the loss)
    float* h_loss = nullptr;
    float last_loss = std::nanf("");        // of the most recently completed training
    float last_train_ms = 0.0f;

    // telemetry
    FitnessStatsRecord* d_fstats = nullptr;
    float* cv_by_org = nullptr;     // [256][8192] sampled scalars, organism-major
    float* cv_by_weight = nullptr;  // the same, weight-major
    float* cv_centroid = nullptr;   // [8192]
    float* cv_squared = nullptr;    // [256 + 512] squared distances
    uint32_t* cv_index = nullptr;   // [256] slots, [8192] scalar indices, [512][2] pairs
};

namespace {

bool Ok(hipError_t e) { return e == hipSuccess; }

template <typename T>
bool Alloc(T** p, uint64_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return Ok(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
}

template <typename T>
void Free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// float -> bfloat16 bits, round to nearest even (NaN stays a quiet NaN)
uint16_t Bf16Bits(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return uint16_t((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

static void DestroyStepGraph(StepGraph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
}

int Fail(ne_runtime* rt, int code) {
    if (rt) rt->last_status = code;
    return code;
}

int Sync(ne_runtime* rt) {
    if (!Ok(hipStreamSynchronize(rt->stream))) return Fail(rt, NE_DEVICE_FAILURE);
    return NE_OK;
}

// Synchronise, then read and clear the sticky device status (first error wins).
int Finish(ne_runtime* rt) {
    if (Sync(rt) != NE_OK) return NE_DEVICE_FAILURE;
    int st = 0;
    if (!Ok(hipMemcpy(&st, rt->status, sizeof(int), hipMemcpyDeviceToHost))) return Fail(rt, NE_DEVICE_FAILURE);
    if (st != 0) {
        (void)hipMemset(rt->status, 0, sizeof(int));
        return Fail(rt, st);
    }
    return NE_OK;
}

uint32_t HostFreeCount(ne_runtime* rt) {
    uint32_t v = 0;
    (void)hipMemcpy(&v, rt->free_count, sizeof(uint32_t), hipMemcpyDeviceToHost);
    return v;
}

uint64_t SlotCountForBudget(uint64_t budget, uint64_t action_count, uint64_t ceiling) {
    // PopulationSizeForGenotypeBudgetBytes (genome/dynamic_layout.hpp:37-52) on this build's stride.
    const uint64_t stride = SlotStrideBytes(action_count);
    if (stride == 0 || budget < stride) return 0;
    const uint64_t n = budget / stride;
    return (ceiling == 0 || n < ceiling) ? n : ceiling;
}

// Verbose progress lines in the reference's form (src/common/progress_log.hpp:16-48): "[local time]
// message" on stdout, a stage's duration appended as " in X.X ms".
using ProgressClock = std::chrono::steady_clock;

void Progress(const ne_runtime* rt, const char* fmt, ...) {
    if (!rt->verbose) return;
    char msg[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    const std::time_t now = std::time(nullptr);
    std::tm local{};
    localtime_r(&now, &local);
    char stamp[32];
    std::strftime(stamp, sizeof(stamp), "%Y-%m-%d %H:%M:%S", &local);
    std::printf("[%s] %s\n", stamp, msg);
    std::fflush(stdout);
}

double MsSince(ProgressClock::time_point t0) {
    return std::chrono::duration<double, std::milli>(ProgressClock::now() - t0).count();
}

void EventBegin(ne_runtime* rt, int i) { (void)hipEventRecord(rt->ev[i].begin, rt->stream); }
void EventEnd(ne_runtime* rt, int i) { (void)hipEventRecord(rt->ev[i].end, rt->stream); }
float EventMs(ne_runtime* rt, int i) {
    float ms = 0.0f;
    if (!Ok(hipEventElapsedTime(&ms, rt->ev[i].begin, rt->ev[i].end))) return 0.0f;
    return ms;
}

// ---- reference shard schedule -> per-cell CSR -------------------------------------------------
// TryBuildTrainingDataShardRuntimeSet and DoesTrainingDataShardCoverCell
// (src/training_folder/training_data.hpp:150-316): a global foundation shard of the first
// initial_word_count catalog words, then shards of word_count_step words centred on golden-ratio
// cells whose radius grows every growth period; a cell's local list is the covering shards' words
// in shard order.
struct Shard {
    uint64_t first, count, row, col, radius;
    bool global;
};

uint64_t RadiusAt(uint64_t intro, uint64_t gen, uint64_t initial, uint64_t period) {
    if (period == 0 || gen < intro) return 0;
    const uint64_t steps = (gen - intro) / period;
    if (steps > (~uint64_t(0)) - initial) return ~uint64_t(0);
    return initial + steps;
}

uint64_t AxisDistance(uint64_t a, uint64_t b, uint64_t side) {
    const uint64_t f = a >= b ? a - b : b - a;
    const uint64_t w = side - f;
    return f < w ? f : w;
}

int BuildShardCsrCore(const ne_word_counts* wc, const ne_shard_release_history* hist, uint64_t generation,
                      uint64_t rows, uint64_t cols, uint64_t epi_rows, uint64_t catalog_count,
                      std::vector<uint32_t>& offsets, std::vector<uint32_t>& secrets) {
    const ne_word_count_schedule& sc = wc->training_word_schedule;
    const uint64_t introduced = wc->training_word_count;
    if (sc.initial_word_count == 0 || sc.word_count_step_period_generations == 0 || introduced == 0 ||
        introduced > catalog_count || wc->shard_radius_growth_period_generations == 0 || rows == 0 || cols == 0 ||
        epi_rows == 0)
        return NE_INVALID_TRAINING_SHARD;
    const uint64_t epi_cells = epi_rows * cols;
    std::vector<Shard> shards;
    const uint64_t foundation = std::min(sc.initial_word_count, introduced);
    shards.push_back({0, foundation, 0, 0, 0, true});
    if (sc.word_count_step != 0 && foundation != introduced) {
        uint64_t next = foundation, local = 0;
        while (next < introduced) {
            const uint64_t cnt = std::min(introduced - next, sc.word_count_step);
            const uint64_t hist_index = shards.size();
            uint64_t intro = (local + 1) * sc.word_count_step_period_generations;
            if (hist) {
                if (hist_index >= hist->release_count) return NE_INVALID_TRAINING_SHARD;
                intro = hist->release_generations[hist_index];
            }
            // DeterministicTrainingShardCenterCoordinate (training_data.hpp:164-178): wrapping
            // 64-bit product with the golden-ratio stride, then modulo the epicenter cell count
            const uint64_t cell = ((local + 1) * 11400714819323198485ull) % epi_cells;
            uint64_t crow = cell / cols;
            if (crow >= rows) crow = rows - 1;
            shards.push_back({next, cnt, crow, cell % cols,
                              RadiusAt(intro, generation, wc->shard_initial_radius,
                                       wc->shard_radius_growth_period_generations),
                              false});
            next += cnt;
            ++local;
        }
    }
    const uint64_t cells = rows * cols;
    offsets.assign(cells + 1, 0);
    secrets.clear();
    for (uint64_t c = 0; c < cells; ++c) {
        offsets[c] = uint32_t(secrets.size());
        for (const Shard& s : shards) {
            bool covers = s.global;
            if (!covers) {
                const uint64_t d = std::max(AxisDistance(s.row, c / cols, rows), AxisDistance(s.col, c % cols, cols));
                covers = d <= s.radius;
            }
            if (!covers) continue;
            for (uint64_t k = 0; k < s.count; ++k) secrets.push_back(uint32_t(s.first + k));
        }
    }
    offsets[cells] = uint32_t(secrets.size());
    return NE_OK;
}

int BuildShardCsr(ne_runtime* rt, const ne_word_counts* wc, const ne_shard_release_history* hist, uint64_t rows,
                  uint64_t cols, std::vector<uint32_t>& offsets, std::vector<uint32_t>& secrets) {
    return BuildShardCsrCore(wc, hist, rt->cur_index, rows, cols, rt->epi_rows, rt->catalog_count, offsets, secrets);
}

int EnsureRankWeights(ne_runtime* rt, float exponent, uint64_t n) {
    // rank^exponent (selection_ops.cuh:59-61) tabulated once on the host for ranks 1..n.
    if (rt->rank_weights && rt->rank_weights_len >= n + 1 && rt->rank_weights_exponent == exponent) return NE_OK;
    Free(rt->rank_weights);
    std::vector<float> w(n + 1, 0.0f);
    for (uint64_t r = 1; r <= n; ++r) w[r] = float(std::pow(double(r), double(exponent)));
    if (!Alloc(&rt->rank_weights, n + 1) ||
        !Ok(hipMemcpy(rt->rank_weights, w.data(), (n + 1) * sizeof(float), hipMemcpyHostToDevice)))
        return Fail(rt, NE_DEVICE_FAILURE);
    rt->rank_weights_len = n + 1;
    rt->rank_weights_exponent = exponent;
    return NE_OK;
}

ne_assembly_config DefaultAssembly() {
    ne_assembly_config c{};
    c.cellular_breeding_radius = 2;
    c.rank_exponent = 0.5f;
    c.crossover_temperature_level1 = 0.02f;
    c.crossover_temperature_level2 = 0.01f;
    c.crossover_temperature_level3 = 0.005f;
    c.mutation_probability = 0.0001f;
    c.mutation_sigma = 0.02f;
    c.output_tail_row_scale_mutation_probability = 0.0f;
    return c;
}

bool ValidAssembly(const ne_assembly_config& c) {
    // IsValidGenerationAssemblyConfig (generation_assembly.hpp:15-19).
    auto unit = [](float v) { return v >= 0.0f && v <= 1.0f; };
    return c.cellular_breeding_radius > 0 && c.rank_exponent >= 0.0f && c.rank_exponent <= 4.0f &&
           unit(c.crossover_temperature_level1) && unit(c.crossover_temperature_level2) &&
           unit(c.crossover_temperature_level3) && unit(c.mutation_probability) && c.mutation_sigma >= 0.0f &&
           unit(c.output_tail_row_scale_mutation_probability);
}

// Fill kernels rather than memsets, so that a captured step holds no memset node.
int ClearGenerationBuffers(ne_runtime* rt, bool next) {
    uint32_t* slots = next ? rt->nxt_slots : rt->cur_slots;
    LaunchFillU32(slots, kInvalidSlot, rt->max_gen, rt->stream);
    LaunchFillU32(reinterpret_cast<uint32_t*>(next ? rt->nxt_fit : rt->cur_fit), 0u, rt->max_gen, rt->stream);
    LaunchFillU32(next ? rt->nxt_evals : rt->cur_evals, 0u, rt->max_gen, rt->stream);
    LaunchFillU8(next ? rt->nxt_has : rt->cur_has, 0, rt->max_gen, rt->stream);
    return Ok(hipGetLastError()) ? NE_OK : Fail(rt, NE_DEVICE_FAILURE);
}

int EnsureTrace(ne_runtime* rt, uint64_t games) {
    if (rt->trace_capacity >= games && rt->d_trace) return NE_OK;
    Free(rt->d_trace);
    Free(rt->d_policy_trace);
    if (!Alloc(&rt->d_trace, games)) return Fail(rt, NE_DEVICE_FAILURE);
    if ((rt->trace_flags & NE_TRACE_POLICY) && !Alloc(&rt->d_policy_trace, games * uint64_t(kMaxTurns) * kPol))
        return Fail(rt, NE_DEVICE_FAILURE);
    rt->trace_capacity = games;
    return NE_OK;
}

}  // namespace

// TryEvaluateCurrentGenerationFitnessOnDevice (slab_runtime.cu:1086-1175): per-cell secrets, one
// game-play kernel, then fitness normalisation and the population summary.
// Evaluates n genomes: slots[e] (device) plays the 3 episode variants of every secret in
// secs[offs[e] .. offs[e+1]).  Fitness, evaluation counts and has-fitness go to out_* (device, n each);
// rt->times and rt->last_summary describe this evaluation.
// Host half of an evaluation: validation, the CSR and work-list uploads, buffers and the kernel arguments.
// Nothing here is captured into a graph.
static int PrepareEvaluation(ne_runtime* rt, uint64_t action_space, const uint32_t* slots, uint64_t n,
                             const std::vector<uint32_t>& offs, const std::vector<uint32_t>& secs, EvalSetup* out) {
    ++rt->input_epoch;  // the CSR and work-list buffers are about to hold this evaluation's lists
    if (n == 0 || n > rt->max_gen || offs.size() < n + 1) return Fail(rt, NE_INVALID_GENERATION);
    for (uint64_t i = 0; i < offs[n]; ++i)
        if (secs[i] >= rt->catalog_count) return Fail(rt, NE_INVALID_TRAINING_SHARD);
    const uint64_t n_secrets = offs[n];
    if (n_secrets > rt->d_secrets_cap) {
        Free(rt->d_secrets);
        if (!Alloc(&rt->d_secrets, n_secrets)) return Fail(rt, NE_DEVICE_FAILURE);
        rt->d_secrets_cap = n_secrets;
    }
    bool ok = Ok(hipMemcpyAsync(rt->d_offsets, offs.data(), (n + 1) * sizeof(uint32_t),
                                hipMemcpyHostToDevice, rt->stream));
    if (n_secrets)
        ok &= Ok(hipMemcpyAsync(rt->d_secrets, secs.data(), n_secrets * sizeof(uint32_t), hipMemcpyHostToDevice,
                                rt->stream));
    uint64_t max_tiles = 0, total_games = 0;
    std::vector<uint64_t> game_offsets(n, 0);
    for (uint64_t c = 0; c < n; ++c) {
        const uint64_t W = offs[c + 1] - offs[c];
        if (W == 0) return Fail(rt, NE_INVALID_TRAINING_SHARD);
        max_tiles = std::max(max_tiles, (W + 15) / 16 + (2 * W + 15) / 16);
        game_offsets[c] = total_games;
        total_games += 3 * W;
    }
    const bool trace = (rt->trace_flags & (NE_TRACE_GAMES | NE_TRACE_POLICY)) != 0;
    if (trace) {
        if (EnsureTrace(rt, total_games) != NE_OK) return rt->last_status;
        Free(rt->d_game_offsets);
        if (!Alloc(&rt->d_game_offsets, n)) return Fail(rt, NE_DEVICE_FAILURE);
        ok &= Ok(hipMemcpyAsync(rt->d_game_offsets, game_offsets.data(), n * sizeof(uint64_t),
                                hipMemcpyHostToDevice, rt->stream));
        ok &= Ok(hipMemsetAsync(rt->d_trace, 0xFF, total_games * sizeof(GameTrace), rt->stream));
        rt->trace_games = total_games;
    }
    if (!ok) return Fail(rt, NE_DEVICE_FAILURE);

    EvalArgs a;
    std::memset(&a, 0, sizeof(a));  // the arguments are compared byte for byte by the step graph
    a.slab = rt->slab;
    a.stride = rt->stride;
    a.slot_count = uint32_t(rt->slot_count);
    a.action_count = uint32_t(std::min(rt->action_count, action_space));
    if (a.action_count == 0) return Fail(rt, NE_INVALID_TRAINING_SHARD);
    a.slot_indices = slots;
    a.cell_offsets = rt->d_offsets;
    a.secret_ids = rt->d_secrets;
    a.codes = rt->codes;
    a.canon = rt->canon;
    a.canon_bytes = uint32_t(rt->img_bytes / kImgTileBytes * 64);
    a.canon_identity = rt->canon_identity ? 1u : 0u;
    a.img = rt->img;
    a.img_bytes = uint32_t(rt->img_bytes);
    a.fixed_absmax = rt->fixed_absmax;
    a.bf16_cat = rt->bf16_cat;
    a.bf16_cat_bytes = uint32_t(rt->bf16_cat_bytes);
    a.force_fallback = (rt->eval_mode & NE_EVAL_FORCE_FALLBACK) ? 1u : 0u;
    const bool bf16 = (rt->eval_mode & 0xFFu) == NE_EVAL_BF16_RESCORED || (rt->eval_mode & 0xFFu) == NE_EVAL_BF16_TRUNK;
    a.trunk_bf16 = (rt->eval_mode & 0xFFu) == NE_EVAL_BF16_TRUNK ? 1u : 0u;
    a.raw_fitness = rt->raw;
    a.local_w = rt->local_w;
    a.counters = rt->counters;
    a.status = rt->status;
    a.game_trace = trace ? rt->d_trace : nullptr;
    a.policy_trace = (rt->trace_flags & NE_TRACE_POLICY) ? rt->d_policy_trace : nullptr;
    a.game_offsets = trace ? rt->d_game_offsets : nullptr;
    // Kernel (untraced forms): the paired kernel plays two game tiles per wave, a workgroup's 2 x waves tiles
    // with one operand stream per decision; the others one tile per wave (the BF16 form's streamed kernel only
    // at 3 waves: at other sizes the register kernel runs, EvaluateKernelFor below).
    uint32_t kernel = !trace ? rt->tuned_kernel : kEvalKernelRegisters;
    const uint32_t max_waves = EvaluateMaxWavesFor(kernel, trace, bf16);
    const uint32_t tpw = kernel == kEvalKernelPaired ? 2u : 1u;
    // Waves per workgroup: 6-wave workgroups of this kernel's register footprint fit once per CU, 2-4
    // wave ones fill a CU (12 waves); pick the size that leaves the fewest idle tile slots over all
    // genomes (ties: the larger).  A genome's tiles then spread over ceil(max_tiles / (tpw x waves))
    // workgroups.
    std::vector<uint32_t> tiles(n);
    for (uint64_t c = 0; c < n; ++c) {
        const uint64_t W = offs[c + 1] - offs[c];
        tiles[c] = uint32_t((W + 15) / 16 + (2 * W + 15) / 16);
    }
    uint32_t waves = 1;
    uint64_t best_cost = ~0ull;
    for (uint32_t w = std::min<uint32_t>(4, max_waves); w >= 1; --w) {
        const uint64_t per = uint64_t(w) * tpw;
        uint64_t cost = 0;
        for (uint64_t c = 0; c < n; ++c) cost += (tiles[c] + per - 1) / per * per;
        if (cost < best_cost) {
            best_cost = cost;
            waves = w;
        }
    }
    // the autotuner's choice (ne_autotune_evaluation), at most what this kernel's LDS tables hold (results do
    // not depend on the workgroup size)
    if (rt->tuned_waves != 0) waves = std::min(rt->tuned_waves, max_waves);
    kernel = EvaluateKernelFor(kernel, waves, trace, bf16);
    // Work list for the persistent kernel, in one queue per XCD (a workgroup drains its own XCD's queue,
    // then helps with the others): genomes by descending tile count, dealt round-robin over the queues, so
    // the longest genomes start first and the queues carry equal work.  Within a queue the FP32 form goes
    // split-major (every genome's first split, then the second ...: the longest items first), the BF16
    // form genome-major (a genome's splits one after another, so they run side by side on one XCD and
    // read its screening operands through one L2: 6% less HBM traffic, 64x64 / 4,739 evaluation 9.4 ->
    // 8.9-9.1 ms; the same order makes the FP32 streamed kernel ~2% slower).  Neither skips an empty
    // split.  An item plays `rounds` rounds of tpw x waves tiles, so a genome's tiles spread over
    // ceil(tiles / (rounds x tpw x waves)) items.  Layout: the kWorkQueues + 1 queue bounds, then the
    // items from word kWorkHeader on.
    const uint32_t rounds = std::max<uint32_t>(1, rt->tuned_rounds);
    const uint64_t per_item = uint64_t(waves) * tpw;
    const uint32_t splits = uint32_t((max_tiles + per_item * rounds - 1) / (per_item * rounds));
    std::vector<uint32_t> order(n);
    for (uint64_t c = 0; c < n; ++c) order[c] = uint32_t(c);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return tiles[x] > tiles[y]; });
    constexpr uint64_t kWorkHeader = 16;
    static_assert(kWorkQueues + 1 <= kWorkHeader, "queue bounds");
    std::vector<uint32_t> work(kWorkHeader, 0u);
    auto push = [&](uint32_t g, uint32_t sp) {
        if (uint64_t(sp) * per_item >= tiles[g]) return;
        work.push_back(g);
        work.push_back(sp);
    };
    for (uint32_t q = 0; q < kWorkQueues; ++q) {
        work[q] = uint32_t((work.size() - kWorkHeader) / 2);
        if (bf16) {
            for (uint64_t r = q; r < n; r += kWorkQueues)
                for (uint32_t sp = 0; sp < splits; ++sp) push(order[r], sp);
        } else {
            for (uint32_t sp = 0; sp < splits; ++sp)
                for (uint64_t r = q; r < n; r += kWorkQueues) push(order[r], sp);
        }
    }
    const uint64_t items = (work.size() - kWorkHeader) / 2;
    work[kWorkQueues] = uint32_t(items);
    if (items > rt->work_cap || rt->d_work == nullptr) {
        Free(rt->d_work);
        rt->work_cap = 0;
        if (!Alloc(&rt->d_work, kWorkHeader + 2 * items)) return Fail(rt, NE_DEVICE_FAILURE);
        rt->work_cap = items;
    }
    if (!Ok(hipMemcpyAsync(rt->d_work, work.data(), work.size() * 4, hipMemcpyHostToDevice, rt->stream)))
        return Fail(rt, NE_DEVICE_FAILURE);
    a.queue_begin = rt->d_work;
    a.work = rt->d_work + kWorkHeader;
    a.work_count = uint32_t(items);
    a.work_counter = rt->d_work_counter;
    a.splits = splits;
    // resident workgroups per CU, as the kernel's registers and LDS allow (3 waves per SIMD at 168 VGPRs
    // and 4 workgroups per CU by LDS in the FP32 form)
    const uint32_t resident = EvaluateResidentPerCu(waves, kernel, trace, bf16, a.trunk_bf16 != 0);
    uint64_t per_cu = resident != 0 ? resident : std::min<uint64_t>(12 / waves, 4);
    if (rt->tuned_wg_per_cu != 0) per_cu = std::min<uint64_t>(per_cu, rt->tuned_wg_per_cu);
    const uint32_t blocks = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(items, rt->cu_count * per_cu)));
    {
        // where the layer-0 accumulators of every resident workgroup's tiles wait between decisions (kept:
        // the captured step bakes the pointer in)
        const uint64_t need = uint64_t(blocks) * EvaluateAccScratchBytes();
        if (need > rt->eval_acc_bytes) {
            Free(rt->eval_acc);
            rt->eval_acc_bytes = 0;
            if (!Alloc(&rt->eval_acc, need / sizeof(float))) return Fail(rt, NE_DEVICE_FAILURE);
            rt->eval_acc_bytes = need;
        }
        a.acc_scratch = rt->eval_acc;
    }
    if (rt->enc_t == nullptr && !Alloc(&rt->enc_t, rt->max_gen * kEncTStride))
        return Fail(rt, NE_DEVICE_FAILURE);
    a.enc_t = rt->enc_t;
    if (bf16) {
        const uint64_t per_genome = uint64_t((a.action_count + 15) / 16) * kBf16GenTileBytes;
        if (per_genome * rt->max_gen > rt->bf16_ops_bytes) {
            Free(rt->bf16_ops);
            rt->bf16_ops_bytes = 0;
            if (!Alloc(&rt->bf16_ops, per_genome * rt->max_gen)) return Fail(rt, NE_DEVICE_FAILURE);
            rt->bf16_ops_bytes = per_genome * rt->max_gen;
        }
        if (rt->tail_absmax == nullptr && !Alloc(&rt->tail_absmax, rt->max_gen * 40)) return Fail(rt, NE_DEVICE_FAILURE);
        a.bf16_ops = rt->bf16_ops;
        a.bf16_stride = per_genome;
        a.tail_absmax = reinterpret_cast<const float*>(rt->tail_absmax);
    }
    out->a = a;
    out->blocks = blocks;
    out->waves = waves;
    out->kernel = kernel;
    out->bf16 = bf16;
    out->trace = trace;
    out->total_games = total_games;
    return NE_OK;
}

// Device half, part 1: clears (fill kernels, no memset nodes, so the sequence can be captured), the
// encoder transpose (+ BF16 pack) and the game-play kernel.  `timed`: stage events around the launches.
static void LaunchEvaluation(ne_runtime* rt, const EvalSetup& s, uint64_t n, float* out_fit, uint32_t* out_evals,
                             uint8_t* out_has, bool timed) {
    LaunchFillU32(reinterpret_cast<uint32_t*>(out_fit), 0u, n, rt->stream);
    LaunchFillU32(out_evals, 0u, n, rt->stream);
    LaunchFillU8(out_has, 0, n, rt->stream);
    LaunchFillU32(reinterpret_cast<uint32_t*>(rt->raw), 0u, rt->max_gen, rt->stream);
    LaunchFillU32(rt->local_w, 0u, rt->max_gen, rt->stream);
    LaunchFillU32(reinterpret_cast<uint32_t*>(rt->counters), 0u, 8, rt->stream);
    LaunchFillU32(reinterpret_cast<uint32_t*>(rt->status), 0u, 1, rt->stream);
    LaunchFillU32(rt->d_work_counter, 0u, kWorkQueues * kWorkCounterStride, rt->stream);
    if (s.bf16) LaunchFillU32(rt->tail_absmax, 0u, n * 40, rt->stream);
    if (timed) EventBegin(rt, 8);
    LaunchTransposeEncoder(rt->slab, rt->stride, s.a.slot_indices, uint32_t(n), uint32_t(rt->slot_count), rt->enc_t,
                           rt->stream);
    if (s.bf16)
        LaunchPackBf16(rt->slab, rt->stride, s.a.slot_indices, uint32_t(n), uint32_t(rt->slot_count), rt->img,
                       s.a.action_count, rt->bf16_ops, s.a.bf16_stride, rt->tail_absmax, rt->stream);
    if (timed) EventEnd(rt, 8);
    if (timed) EventBegin(rt, 0);
    LaunchEvaluate(s.a, s.blocks, s.waves, s.kernel, s.trace, s.bf16, rt->stream);
    if (timed) EventEnd(rt, 0);
}

// Device half, part 2: fitness normalisation and the generation summary.
static void LaunchFinalizeSummary(ne_runtime* rt, const uint32_t* slots, uint64_t n, float* out_fit,
                                  uint32_t* out_evals, uint8_t* out_has, bool timed) {
    if (timed) EventBegin(rt, 1);
    LaunchFinalize(rt->raw, rt->local_w, uint32_t(n), out_fit, out_evals, out_has, rt->status, rt->stream);
    LaunchSummary(out_fit, out_has, slots, uint32_t(n), rt->summary, rt->status, rt->stream);
    if (timed) EventEnd(rt, 1);
}

// The host record of an evaluation: summary, counters, stage times.
static void RecordEvaluation(ne_runtime* rt, const SummaryRecord& sr, const unsigned long long (&counters)[4],
                             uint64_t total_games, uint64_t n, uint64_t generation_index) {
    rt->times.decisions = counters[0];
    rt->times.rescored = counters[2];
    rt->times.fallbacks = counters[3];
    rt->times.games = total_games;
    // encoded turns: a game with d >= 1 decisions encodes every turn present at its last decision
    rt->times.encoded_turns = counters[1];
    rt->last_summary.best_fitness = sr.best_fitness;
    rt->last_summary.average_fitness = sr.average_fitness;
    rt->last_summary.best_index = sr.best_index;
    rt->last_summary.best_slot_index = sr.best_slot;
    rt->last_summary.generation_index = generation_index;
    rt->last_summary.action_count = rt->action_count;
    rt->last_summary.population_size = n;
    rt->have_summary = true;
}

// The device half of a prepared evaluation and its host record: every launch, the status, summary and counters
// gathered into one block, then one wait and one read (a failing kernel's status is the first one set, so
// finalisation behind it cannot hide it).
static int RunEvaluation(ne_runtime* rt, const EvalSetup& s, const uint32_t* slots, uint64_t n, float* out_fit,
                         uint32_t* out_evals, uint8_t* out_has, uint64_t generation_index) {
    if (!rt->d_readback && (!Alloc(&rt->d_readback, 1) ||
                            !Ok(hipHostMalloc((void**)&rt->h_readback, sizeof(StepReadback)))))
        return Fail(rt, NE_DEVICE_FAILURE);
    LaunchEvaluation(rt, s, n, out_fit, out_evals, out_has, true);
    LaunchFinalizeSummary(rt, slots, n, out_fit, out_evals, out_has, true);
    LaunchStepReadback(rt->status, rt->free_count, rt->summary, rt->counters, rt->d_readback, rt->stream);
    if (!Ok(hipGetLastError()) ||
        !Ok(hipMemcpyAsync(rt->h_readback, rt->d_readback, sizeof(StepReadback), hipMemcpyDeviceToHost, rt->stream)))
        return Fail(rt, NE_DEVICE_FAILURE);
    if (Sync(rt) != NE_OK) return NE_DEVICE_FAILURE;
    const StepReadback rb = *rt->h_readback;
    if (rb.status != 0) {
        (void)hipMemset(rt->status, 0, sizeof(int));
        return Fail(rt, rb.status);
    }
    unsigned long long counters[4];
    std::memcpy(counters, rb.counters, sizeof(counters));
    const SummaryRecord sr = rb.summary;
    rt->times.evaluate_ms = EventMs(rt, 0);
    rt->times.prepare_ms = EventMs(rt, 8);
    rt->times.finalize_ms = EventMs(rt, 1);
    RecordEvaluation(rt, sr, counters, s.total_games, n, generation_index);
    rt->last_status = NE_OK;
    return NE_OK;
}

// What an evaluation's host preparation depends on besides the per-cell lists: the inputs' epoch (any other
// preparation or upload since changes it), the size, action space, form and launch shape, and the slab.
static void EagerEvalKey(const ne_runtime* rt, uint64_t action_space, uint64_t n, uint64_t (&k)[10]) {
    const uint64_t v[10] = {rt->input_epoch, n, std::min(rt->action_count, action_space), rt->eval_mode,
                            rt->tuned_waves | (uint64_t(rt->tuned_rounds) << 16) |
                                (uint64_t(rt->tuned_wg_per_cu) << 32) | (uint64_t(rt->tuned_kernel) << 48),
                            uint64_t(uintptr_t(rt->slab)), rt->stride, rt->slot_count, rt->action_count, rt->max_gen};
    std::memcpy(k, v, sizeof(v));
}

// An evaluation of n genomes: the host preparation (validation, the CSR and work-list uploads) is kept while
// nothing else has prepared an evaluation or uploaded inputs since, the key is the same and the per-cell
// lists are equal (a generation step evaluates the same cells' secrets generation after generation); traced
// evaluations always prepare (their trace buffers are cleared there).
static int EvaluateCore(ne_runtime* rt, uint64_t action_space, const uint32_t* slots, uint64_t n,
                        const std::vector<uint32_t>& offs, const std::vector<uint32_t>& secs, float* out_fit,
                        uint32_t* out_evals, uint8_t* out_has, uint64_t generation_index) {
    const bool trace = (rt->trace_flags & (NE_TRACE_GAMES | NE_TRACE_POLICY)) != 0;
    uint64_t key[10];
    EagerEvalKey(rt, action_space, n, key);
    const bool same_lists = offs.size() >= n + 1 && rt->eager_offsets.size() == n + 1 &&
                            std::equal(offs.begin(), offs.begin() + n + 1, rt->eager_offsets.begin()) &&
                            secs.size() >= offs[n] && rt->eager_secrets.size() == offs[n] &&
                            std::equal(secs.begin(), secs.begin() + offs[n], rt->eager_secrets.begin());
    EvalSetup s;
    if (!trace && same_lists && std::memcmp(key, rt->eager_eval_key, sizeof(key)) == 0) {
        s = rt->eager_eval;
        s.a.slot_indices = slots;
    } else {
        std::memset(rt->eager_eval_key, 0, sizeof(rt->eager_eval_key));
        const int st = PrepareEvaluation(rt, action_space, slots, n, offs, secs, &s);
        if (st != NE_OK) return st;
        if (!trace) {
            rt->eager_eval = s;
            EagerEvalKey(rt, action_space, n, rt->eager_eval_key);  // after the preparation's own epoch bump
            rt->eager_offsets.assign(offs.begin(), offs.begin() + n + 1);
            rt->eager_secrets.assign(secs.begin(), secs.begin() + offs[n]);
        }
    }
    return RunEvaluation(rt, s, slots, n, out_fit, out_evals, out_has, generation_index);
}


// Records the inputs the current generation's fitness was measured on (n cells of a per-cell CSR).
static void RecordFitnessInputs(ne_runtime* rt, uint64_t action_space, const std::vector<uint32_t>& offs,
                                const std::vector<uint32_t>& secs, uint64_t n) {
    rt->fit_upload_epoch = rt->upload_epoch;
    rt->fit_action_space = action_space;
    rt->fit_offsets.assign(offs.begin(), offs.begin() + n + 1);
    rt->fit_secrets.assign(secs.begin() + offs[0], secs.begin() + offs[n]);
}

// The current generation's fitness still describes it for this step's inputs: no catalog or secrets
// upload since it was measured, the same action space and the same per-cell secrets (the caller's CSR,
// or the shard schedule's lists for this generation under `wc` and `hist`).  The reference evaluates at
// every generation boundary (slab_runtime.cu:1374-1424); a reuse is valid only when that evaluation
// would repeat the recorded one.
static bool FitnessInputsHold(ne_runtime* rt, const ne_word_counts* wc, const ne_shard_release_history* hist) {
    if (!rt->cur_evaluated || rt->fit_upload_epoch != rt->upload_epoch || !wc ||
        rt->fit_action_space != wc->action_space_word_count)
        return false;
    const uint64_t cols = rt->config.grid_column_count, n = rt->cur_size;
    if (cols == 0 || n == 0 || n % cols != 0 || rt->fit_offsets.size() != n + 1) return false;
    std::vector<uint32_t> built_offsets, built_secrets;
    const std::vector<uint32_t>* offs = &rt->h_offsets;
    const std::vector<uint32_t>* secs = &rt->h_secrets;
    if (!rt->csr_from_caller) {
        if (BuildShardCsr(rt, wc, hist, n / cols, cols, built_offsets, built_secrets) != NE_OK) return false;
        offs = &built_offsets;
        secs = &built_secrets;
    }
    if (offs->size() < n + 1) return false;
    if (!std::equal(offs->begin(), offs->begin() + n + 1, rt->fit_offsets.begin())) return false;
    const uint64_t k = (*offs)[n] - (*offs)[0];
    return k == rt->fit_secrets.size() &&
           std::equal(secs->begin() + (*offs)[0], secs->begin() + (*offs)[n], rt->fit_secrets.begin());
}

static int EvaluateFull(ne_runtime* rt, const ne_word_counts* wc, const ne_shard_release_history* hist) {
    const uint64_t cols = rt->config.grid_column_count;
    if (!wc) return Fail(rt, NE_INVALID_ARGUMENT);
    if (rt->cur_size == 0 || cols == 0 || rt->cur_size % cols != 0) return Fail(rt, NE_INVALID_GENERATION);
    if (rt->catalog_count == 0) return Fail(rt, NE_INVALID_TRAINING_SHARD);
    if (wc->training_word_count > rt->catalog_count || wc->action_space_word_count > rt->catalog_count ||
        wc->training_word_count > wc->action_space_word_count)
        return Fail(rt, NE_INVALID_TRAINING_SHARD);
    const uint64_t rows = rt->cur_size / cols;
    std::vector<uint32_t> built_offsets, built_secrets;
    const std::vector<uint32_t>* offs = &rt->h_offsets;
    const std::vector<uint32_t>* secs = &rt->h_secrets;
    if (!rt->csr_from_caller) {
        const int st = BuildShardCsr(rt, wc, hist, rows, cols, built_offsets, built_secrets);
        if (st != NE_OK) return Fail(rt, st);
        offs = &built_offsets;
        secs = &built_secrets;
    } else if (rt->h_offsets.size() < rt->cur_size + 1) {
        return Fail(rt, NE_INVALID_TRAINING_SHARD);
    }
    if (offs->size() < rt->cur_size + 1) return Fail(rt, NE_INVALID_TRAINING_SHARD);
    const int st = EvaluateCore(rt, wc->action_space_word_count, rt->cur_slots, rt->cur_size, *offs, *secs, rt->cur_fit,
                                rt->cur_evals, rt->cur_has, rt->cur_index);
    if (st == NE_OK) {
        rt->cur_evaluated = true;
        RecordFitnessInputs(rt, wc->action_space_word_count, *offs, *secs, rt->cur_size);
    }
    return st;
}

// TryPlanNextGenerationShape (slab_runtime.cu:919-939).
static bool PlanShape(ne_runtime* rt, const ne_pending_injection& inj, uint64_t* next_a, uint64_t* next_n) {
    *next_a = rt->action_count;
    if (inj.enabled && (inj.injection_count == 0 || inj.first_catalog_word_index != rt->action_count)) return false;
    if (inj.enabled) *next_a += inj.injection_count;
    uint64_t n = SlotCountForBudget(rt->generation_byte_budget, *next_a, rt->max_gen);
    n = (n / rt->config.grid_column_count) * rt->config.grid_column_count;
    if (n > rt->cur_size) n = rt->cur_size;
    *next_n = n;
    return n > 0;
}

static int EnsureRepackBuffers(ne_runtime* rt, uint64_t staging_bytes) {
    if (rt->owner_cap < rt->slot_count) {
        Free(rt->owner);
        if (!Alloc(&rt->owner, rt->slot_count)) return Fail(rt, NE_DEVICE_FAILURE);
        rt->owner_cap = rt->slot_count;
    }
    if (!rt->move_src && !Alloc(&rt->move_src, rt->max_gen)) return Fail(rt, NE_DEVICE_FAILURE);
    if (!rt->d_survivors && !Alloc(&rt->d_survivors, 1)) return Fail(rt, NE_DEVICE_FAILURE);
    if (rt->staging_cap < staging_bytes) {
        Free(rt->staging);
        if (!Alloc(&rt->staging, staging_bytes)) return Fail(rt, NE_DEVICE_FAILURE);
        rt->staging_cap = staging_bytes;
    }
    return NE_OK;
}

// TryPrepareSlabForExpandedActionCountOnDevice (genotype_slab/device_runtime.cu:1338-1432).
static int Widen(ne_runtime* rt, uint64_t next_a) {
    const uint64_t new_stride = SlotStrideBytes(next_a);
    const uint64_t new_slots = rt->slab_bytes / new_stride;
    if (next_a <= rt->action_count || new_slots == 0 || rt->planned_children == 0)
        return Fail(rt, NE_SLAB_REPACK_FAILED);
    if (EnsureRepackBuffers(rt, rt->cur_size * rt->stride) != NE_OK) return rt->last_status;
    bool ok = Ok(hipMemsetAsync(rt->refcounts, 0, rt->max_gen * sizeof(uint32_t), rt->stream));
    if (!ok) return Fail(rt, NE_DEVICE_FAILURE);
    EventBegin(rt, 4);
    LaunchRefcount(rt->pairs, uint32_t(rt->planned_children), uint32_t(rt->cur_size), rt->refcounts, rt->status,
                   rt->stream);
    LaunchCollectReleased(rt->cur_slots, uint32_t(rt->cur_size), rt->refcounts, rt->slot_state, rt->free_stack,
                          rt->free_count, uint32_t(rt->slot_count), rt->status, rt->stream);
    LaunchFillU32(rt->owner, kInvalidSlot, rt->slot_count, rt->stream);
    RepackArgs ra{};
    ra.cur_slots = rt->cur_slots;
    ra.cur_size = uint32_t(rt->cur_size);
    ra.refcounts = rt->refcounts;
    ra.slot_state = rt->slot_state;
    ra.free_stack = rt->free_stack;
    ra.free_count = rt->free_count;
    ra.owner = rt->owner;
    ra.move_src = rt->move_src;
    ra.survivors = rt->d_survivors;
    ra.old_slot_count = uint32_t(rt->slot_count);
    ra.new_slot_count = uint32_t(new_slots);
    ra.status = rt->status;
    // slot_state/free_stack must hold new_slots entries: they were sized for the largest slot count
    if (new_slots > rt->slot_count) return Fail(rt, NE_SLAB_REPACK_FAILED);
    LaunchRepackPlan(ra, rt->stream);
    int st = Finish(rt);
    if (st != NE_OK) return Fail(rt, st == NE_INVALID_SLAB ? NE_INVALID_SLAB : NE_SLAB_REPACK_FAILED);
    uint32_t survivors = 0;
    if (!Ok(hipMemcpy(&survivors, rt->d_survivors, sizeof(uint32_t), hipMemcpyDeviceToHost)))
        return Fail(rt, NE_DEVICE_FAILURE);
    EventBegin(rt, 12);
    LaunchRepackCopies(rt->slab, rt->stride, new_stride, rt->move_src, survivors, uint32_t(new_slots - survivors),
                       rt->staging, rt->stream);
    EventEnd(rt, 12);
    EventEnd(rt, 4);
    st = Finish(rt);
    if (st != NE_OK) return st;
    rt->times.repack_ms = EventMs(rt, 4);
    rt->times.repack_copy_ms = EventMs(rt, 12);
    rt->times.repack_survivors = survivors;
    rt->times.repack_old_stride = rt->stride;
    rt->action_count = next_a;
    rt->stride = new_stride;
    rt->slot_count = new_slots;
    return NE_OK;
}

// Host spill (TrySpillAssembledChildrenToHost / TryRestoreSpilledChildrenToDevice, slab_runtime.cu:
// 809-917): the assembled children 0..k-1 go to pinned host memory and their slots back onto the free
// stack; after the last batch they take slots again and their bytes return.
static int SpillPrefix(ne_runtime* rt, uint64_t k, std::vector<uint8_t*>& chunks) {
    std::vector<uint32_t> slots(k);
    if (!Ok(hipMemcpyAsync(slots.data(), rt->nxt_slots, k * 4, hipMemcpyDeviceToHost, rt->stream)) ||
        !Ok(hipStreamSynchronize(rt->stream)))
        return Fail(rt, NE_DEVICE_FAILURE);
    uint8_t* host = nullptr;
    if (!Ok(hipHostMalloc(reinterpret_cast<void**>(&host), k * rt->stride, hipHostMallocDefault)))
        return Fail(rt, NE_DEVICE_FAILURE);
    chunks.push_back(host);
    bool ok = true;
    for (uint64_t i = 0; ok && i < k; ++i) {
        if (slots[i] >= rt->slot_count) return Fail(rt, NE_INVALID_GENERATION);
        ok = Ok(hipMemcpyAsync(host + i * rt->stride, rt->slab + uint64_t(slots[i]) * rt->stride, rt->stride,
                               hipMemcpyDeviceToHost, rt->stream));
    }
    if (!ok) return Fail(rt, NE_DEVICE_FAILURE);
    LaunchSpillRelease(rt->nxt_slots, uint32_t(k), HostFreeCount(rt), rt->slot_state, rt->free_stack, rt->free_count,
                       uint32_t(rt->slot_count), rt->status, rt->stream);
    return Finish(rt);
}

static int RestorePrefix(ne_runtime* rt, uint64_t k, const uint8_t* host) {
    const uint32_t fc = HostFreeCount(rt);
    if (fc < k) return Fail(rt, NE_DEVICE_FAILURE);  // as the reference reports a failed restore
    LaunchSpillRestore(rt->nxt_slots, uint32_t(k), fc, rt->slot_state, rt->free_stack, rt->free_count, rt->status,
                       rt->stream);
    int st = Finish(rt);
    if (st != NE_OK) return st;
    std::vector<uint32_t> slots(k);
    bool ok = Ok(hipMemcpyAsync(slots.data(), rt->nxt_slots, k * 4, hipMemcpyDeviceToHost, rt->stream)) &&
              Ok(hipStreamSynchronize(rt->stream));
    for (uint64_t i = 0; ok && i < k; ++i)
        ok = Ok(hipMemcpyAsync(rt->slab + uint64_t(slots[i]) * rt->stride, host + i * rt->stride, rt->stride,
                               hipMemcpyHostToDevice, rt->stream));
    if (!ok) return Fail(rt, NE_DEVICE_FAILURE);
    return Finish(rt);
}

// Assembly: validate, reference counts, GC of unreferenced parents, batches of children bounded by
// free slots, parent release after each batch (device_runtime.cu:1434-1557); one host spill of the
// assembled prefix when the slab fills (slab_runtime.cu:1554-1661).
static int Assemble(ne_runtime* rt, uint32_t seed, const ne_assembly_config& cfg, uint64_t parent_a) {
    rt->next_index = rt->cur_index + 1;
    rt->next_size = rt->planned_children;
    int st = ClearGenerationBuffers(rt, true);
    if (st != NE_OK) return st;
    if (!Ok(hipMemsetAsync(rt->refcounts, 0, rt->max_gen * sizeof(uint32_t), rt->stream))) return Fail(rt, NE_DEVICE_FAILURE);
    EventBegin(rt, 6);
    LaunchRefcount(rt->pairs, uint32_t(rt->next_size), uint32_t(rt->cur_size), rt->refcounts, rt->status, rt->stream);
    LaunchCollectReleased(rt->cur_slots, uint32_t(rt->cur_size), rt->refcounts, rt->slot_state, rt->free_stack,
                          rt->free_count, uint32_t(rt->slot_count), rt->status, rt->stream);
    XoverArgs xa{};
    xa.child_count = uint32_t(rt->next_size);
    xa.gen_index = uint32_t(rt->cur_index);
    xa.seed = seed;
    xa.parent_action_count = uint32_t(parent_a);
    xa.t1 = cfg.crossover_temperature_level1;
    xa.t2 = cfg.crossover_temperature_level2;
    xa.t3 = cfg.crossover_temperature_level3;
    xa.records = rt->xover;
    xa.breaks = rt->breaks;
    LaunchXover(xa, rt->stream);
    EventEnd(rt, 6);
    st = Finish(rt);
    if (st != NE_OK) return st;
    float book_ms = EventMs(rt, 6);
    float breed_ms = 0.0f;
    uint64_t offset = 0, spilled = 0;
    std::vector<uint8_t*> spill;  // the pinned host copy of a spilled prefix
    struct FreeSpill {
        std::vector<uint8_t*>& v;
        hipStream_t stream;
        ~FreeSpill() {
            if (v.empty()) return;
            (void)hipStreamSynchronize(stream);  // no copy may still use the pinned memory
            for (uint8_t* p : v) (void)hipHostFree(p);
        }
    } free_spill{spill, rt->stream};
    while (offset < rt->next_size) {
        const uint32_t fc = HostFreeCount(rt);
        if (fc == 0) {
            if (spilled != 0 || offset == 0 || rt->config.host_spillover_byte_budget == 0)
                return Fail(rt, NE_SLAB_FULL);
            if (offset > rt->config.host_spillover_byte_budget / rt->stride)
                return Fail(rt, NE_DEVICE_FAILURE);  // the reference's outcome when the spill slab is too small
            Progress(rt, "Generation %llu: slab full, spilling %llu assembled children to host memory",
                     (unsigned long long)rt->next_index, (unsigned long long)offset);
            const auto ts = ProgressClock::now();
            st = SpillPrefix(rt, offset, spill);
            if (st != NE_OK) return st;
            Progress(rt, "Generation %llu: spilled %llu children in %.1f ms", (unsigned long long)rt->next_index,
                     (unsigned long long)offset, MsSince(ts));
            spilled = offset;
            continue;
        }
        const uint64_t batch = std::min<uint64_t>(rt->next_size - offset, fc);
        const auto tb = ProgressClock::now();
        Progress(rt, "Generation %llu: assembly batch from child %llu, %llu children (%u free slots)",
                 (unsigned long long)rt->next_index, (unsigned long long)offset, (unsigned long long)batch, fc);
        BreedArgs b{};
        b.slab = rt->slab;
        b.parent_stride = rt->stride;
        b.child_stride = rt->stride;
        b.pairs = rt->pairs;
        b.parent_slots = rt->cur_slots;
        b.records = rt->xover;
        b.breaks = rt->breaks;
        b.free_stack = rt->free_stack;
        b.free_count_before = fc;
        b.child_offset = uint32_t(offset);
        b.child_slots = rt->nxt_slots;
        b.slot_state = rt->slot_state;
        b.parent_action_count = uint32_t(parent_a);
        b.gen_index = uint32_t(rt->cur_index);
        b.seed = seed;
        b.mutation_p = cfg.mutation_probability;
        b.mutation_sigma = cfg.mutation_sigma;
        b.row_scale_p = cfg.output_tail_row_scale_mutation_probability;
        b.slot_count = uint32_t(rt->slot_count);
        b.status = rt->status;
        b.elem_begin = 0;
        b.elem_end = uint32_t(rt->stride / 4);
        EventBegin(rt, 3);
        LaunchBreed(b, uint32_t(batch), rt->stream);
        EventEnd(rt, 3);
        EventBegin(rt, 7);
        LaunchPopSlots(rt->free_count, uint32_t(batch), rt->stream);
        LaunchReleaseBatch(rt->pairs, uint32_t(offset), uint32_t(batch), rt->refcounts, rt->status, rt->stream);
        LaunchCollectReleased(rt->cur_slots, uint32_t(rt->cur_size), rt->refcounts, rt->slot_state, rt->free_stack,
                              rt->free_count, uint32_t(rt->slot_count), rt->status, rt->stream);
        EventEnd(rt, 7);
        st = Finish(rt);
        if (st != NE_OK) return st;
        Progress(rt, "Generation %llu: assembly batch from child %llu done in %.1f ms", (unsigned long long)rt->next_index,
                 (unsigned long long)offset, MsSince(tb));
        breed_ms += EventMs(rt, 3);
        book_ms += EventMs(rt, 7);
        offset += batch;
    }
    if (spilled != 0) {
        st = RestorePrefix(rt, spilled, spill.front());
        if (st != NE_OK) return st;
    }
    rt->times.spilled = spilled;
    rt->times.assemble_ms = breed_ms;
    rt->times.bookkeeping_ms = book_ms;
    rt->times.children = rt->next_size;
    return NE_OK;
}

// TryBuildHintGridGroup (src/wordle/hint_grid.hpp:76-164) on packed words: grid 0 holds every
// distinct left rotation of the word whose feedback is all yellow; grids 1 and 2 each hold one swap
// of two differing positions whose feedback is 3 green + 2 yellow, the first two distinct ones in
// (i < j) order.  Fails when fewer than two such swaps exist.
static bool BuildHintStates(const uint8_t* w, HintState out[3]) {
    auto pack = [](const uint8_t* l) {
        uint32_t c = 0;
        for (int q = 0; q < 5; ++q) c |= uint32_t(l[q]) << (5 * q);
        return c;
    };
    const uint32_t sol = pack(w);
    for (int g = 0; g < 3; ++g) out[g] = HintState{0, {0, 0, 0, 0}, {0, 0, 0, 0}};
    constexpr uint32_t kAllYellow = 0x155;  // tile 1 in each 2-bit field
    for (int rot = 1; rot < 5; ++rot) {
        uint8_t g[5];
        for (int q = 0; q < 5; ++q) g[q] = w[(q + rot) % 5];
        const uint32_t code = pack(g);
        if (code == sol) continue;
        bool dup = false;
        for (uint32_t k = 0; k < out[0].tc; ++k) dup |= out[0].code[k] == code;
        if (dup) continue;
        const uint32_t fb = FeedbackPacked(code, sol);
        if (fb != kAllYellow) continue;
        out[0].code[out[0].tc] = code;
        out[0].fb[out[0].tc] = fb;
        ++out[0].tc;
    }
    int chosen = 0;
    for (int a = 0; a < 5; ++a)
        for (int b = a + 1; b < 5; ++b) {
            if (w[a] == w[b]) continue;
            uint8_t g[5];
            for (int q = 0; q < 5; ++q) g[q] = w[q];
            std::swap(g[a], g[b]);
            const uint32_t code = pack(g);
            const uint32_t fb = FeedbackPacked(code, sol);
            int greens = 0, yellows = 0;
            for (int q = 0; q < 5; ++q) {
                const uint32_t t = (fb >> (2 * q)) & 3u;
                greens += t == 0;
                yellows += t == 1;
            }
            if (greens != 3 || yellows != 2) continue;
            bool dup = false;
            for (int k = 0; k < chosen; ++k) dup |= out[1 + k].code[0] == code;
            if (dup) continue;
            out[1 + chosen] = HintState{1, {code, 0, 0, 0}, {fb, 0, 0, 0}};
            if (++chosen == 2) return true;
        }
    return false;
}

// Injection of the new output-embedding rows into every assembled child
// (AssembleInjectedChildBatchKernel's injection half, genotype_slab/device_runtime.cu:518-617).
static int Inject(ne_runtime* rt, uint64_t parent_a, const ne_pending_injection& inj) {
    if (parent_a == 0 || parent_a > kMaxSortedNorms) return Fail(rt, NE_UNSUPPORTED);
    std::vector<HintState> states(inj.injection_count * 3);
    for (uint64_t k = 0; k < inj.injection_count; ++k)
        if (!BuildHintStates(rt->h_catalog.data() + 5 * (inj.first_catalog_word_index + k), &states[3 * k]))
            return Fail(rt, NE_OUTPUT_EMBEDDING_INJECTION_FAILED);
    HintState* d_states = nullptr;
    float* d_medians = nullptr;
    if (!Alloc(&d_states, states.size()) || !Alloc(&d_medians, rt->next_size) ||
        !Ok(hipMemcpy(d_states, states.data(), states.size() * sizeof(HintState), hipMemcpyHostToDevice))) {
        Free(d_states);
        Free(d_medians);
        return Fail(rt, NE_DEVICE_FAILURE);
    }
    InjectArgs a{};
    a.slab = rt->slab;
    a.stride = rt->stride;
    a.child_slots = rt->nxt_slots;
    a.children = uint32_t(rt->next_size);
    a.parent_action_count = uint32_t(parent_a);
    a.count = uint32_t(inj.injection_count);
    a.states = d_states;
    a.medians = d_medians;
    a.status = rt->status;
    EventBegin(rt, 5);
    LaunchInject(a, rt->stream);
    EventEnd(rt, 5);
    const int st = Finish(rt);
    Free(d_states);
    Free(d_medians);
    if (st != NE_OK) return st;
    rt->times.inject_ms = EventMs(rt, 5);
    return NE_OK;
}

// SwapDeviceSlabGenerationBuffers (slab_runtime.cu:2128-2135).
static void SwapGenerations(ne_runtime* rt) {
    std::swap(rt->cur_slots, rt->nxt_slots);
    std::swap(rt->cur_fit, rt->nxt_fit);
    std::swap(rt->cur_evals, rt->nxt_evals);
    std::swap(rt->cur_has, rt->nxt_has);
    rt->cur_index = rt->next_index;
    rt->cur_size = rt->next_size;
    rt->next_index = 0;
    rt->next_size = 0;
    rt->planned_children = 0;
    rt->cur_evaluated = false;
}

// ---- surrogate screening (host side) ----------------------------------------------------------
constexpr uint32_t kTagProbe = 0x50524F42u;
constexpr uint32_t kTagAudit = 0x41554454u;
constexpr uint32_t kTagBatch = 0x42415443u;
constexpr float kParentFitnessScale = 10.0f;
// Probe boards are answered from the first kProbeVocabulary catalog words (the whole action space when
// smaller): the guess characterises the policy at a ninth of the scoring cost of 4,739 actions.
constexpr uint64_t kProbeVocabulary = 512;

// Average ranks (ties share the mean of their positions), 1-based.
static void AverageRanks(const float* v, uint64_t n, std::vector<double>& r) {
    std::vector<uint64_t> idx(n);
    for (uint64_t i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return v[a] < v[b]; });
    r.assign(n, 0.0);
    for (uint64_t i = 0; i < n;) {
        uint64_t k = i;
        while (k + 1 < n && v[idx[k + 1]] == v[idx[i]]) ++k;
        const double rank = 0.5 * double(i + k) + 1.0;
        for (uint64_t q = i; q <= k; ++q) r[idx[q]] = rank;
        i = k + 1;
    }
}

static double RankCorrelation(const float* a, const float* b, uint64_t n) {
    if (n < 3) return std::nan("");
    for (uint64_t i = 0; i < n; ++i)
        if (!std::isfinite(a[i]) || !std::isfinite(b[i])) return std::nan("");
    std::vector<double> ra, rb;
    AverageRanks(a, n, ra);
    AverageRanks(b, n, rb);
    double ma = 0.0, mb = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        ma += ra[i];
        mb += rb[i];
    }
    ma /= double(n);
    mb /= double(n);
    double sab = 0.0, saa = 0.0, sbb = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        sab += (ra[i] - ma) * (rb[i] - mb);
        saa += (ra[i] - ma) * (ra[i] - ma);
        sbb += (rb[i] - mb) * (rb[i] - mb);
    }
    if (saa == 0.0 || sbb == 0.0) return std::nan("");
    return sab / std::sqrt(saa * sbb);
}

// 16 probe secrets and, per secret, 3 opening guesses (distinct words, none the secret), drawn from
// the first `words` catalog words with Philox(k, purpose, attempt; seed, tag).
static bool DrawProbes(ne_runtime* rt, uint64_t words, uint32_t seed) {
    if (words < kProbeSecrets + kProbeDepths) return false;
    std::vector<uint32_t> code(rt->catalog_count), canon(rt->catalog_count);
    std::unordered_map<uint32_t, uint32_t> first;
    for (uint64_t a = 0; a < rt->catalog_count; ++a) {
        uint32_t c = 0;
        for (int q = 0; q < 5; ++q) c |= uint32_t(rt->h_catalog[5 * a + q]) << (5 * q);
        code[a] = c;
        canon[a] = first.emplace(c, uint32_t(a)).first->second;
    }
    ProbeSet ps{};
    std::vector<uint32_t> secret_codes;
    for (int k = 0; k < kProbeSecrets; ++k) {
        uint32_t idx = 0;
        for (uint32_t attempt = 0;; ++attempt) {
            if (attempt > 100000) return false;
            idx = Philox4x32(uint32_t(k), 0, attempt, 0, seed, kTagProbe).x % uint32_t(words);
            if (std::find(secret_codes.begin(), secret_codes.end(), code[idx]) == secret_codes.end()) break;
        }
        secret_codes.push_back(code[idx]);
        ps.secret_code[k] = code[idx];
        rt->h_probe_words[k][0] = idx;
        for (int t = 0; t < kProbeDepths - 1; ++t) {
            uint32_t g = 0;
            for (uint32_t attempt = 0;; ++attempt) {
                if (attempt > 100000) return false;
                g = Philox4x32(uint32_t(k), uint32_t(1 + t), attempt, 0, seed, kTagProbe).x % uint32_t(words);
                bool clash = code[g] == code[idx];
                for (int q = 0; q < t; ++q) clash = clash || code[g] == ps.turn_code[k][q];
                if (!clash) break;
            }
            ps.turn_code[k][t] = code[g];
            ps.turn_fb[k][t] = FeedbackPacked(code[g], code[idx]);
            ps.turn_canon[k][t] = canon[g];
            rt->h_probe_words[k][1 + t] = g;
        }
    }
    rt->h_probes = ps;
    return true;
}

// Adam's bias corrections of step t (counted from 1).
static void AdamCoef(uint64_t t, float* c) {
    c[0] = float(1.0 / (1.0 - std::pow(0.9, double(t))));
    c[1] = float(1.0 / (1.0 - std::pow(0.999, double(t))));
}

// One Adam step of the surrogate on a device minibatch (x [n][192], t [n]); coef (device) holds its bias
// corrections.  Launches only, so it can be captured.
static void SurrogateStep(ne_runtime* rt, const float* x, const float* t, uint32_t n, const float* coef,
                          hipStream_t stream) {
    LaunchSurrogateForward(rt->sg_params, x, n, rt->by, rt->bh1, rt->bh2, stream);
    LaunchSurrogateBackprop(rt->sg_params, rt->bh1, rt->bh2, rt->by, t, n, rt->bdy, rt->bdh1, rt->bdh2, stream);
    LaunchSurrogateAdam(rt->sg_params, rt->sg_m, rt->sg_v, x, rt->bh1, rt->bh2, rt->bdy, rt->bdh1, rt->bdh2, n,
                        rt->screen.learning_rate, coef, stream);
}

// The launches of one generation's training (steps x {gather, forward, backprop, Adam}), reading their
// indices, standardisation and bias corrections from device buffers filled before each run.
static void TrainingLaunches(ne_runtime* rt, hipStream_t stream) {
    const uint32_t bs = rt->screen.batch_size;
    for (uint32_t st = 0; st < rt->screen.train_steps; ++st) {
        LaunchReplayGather(rt->replay_x, rt->replay_f, rt->bidx + uint64_t(st) * bs, bs, rt->sg_coef, rt->bx, rt->bt,
                           stream);
        SurrogateStep(rt, rt->bx, rt->bt, bs, rt->sg_coef + 2 + 2 * uint64_t(st), stream);
    }
}

// Orders the main stream's later work after the training in flight (device side).
static int JoinTraining(ne_runtime* rt) {
    if (!rt->train_pending) return NE_OK;
    if (!Ok(hipStreamWaitEvent(rt->stream, rt->train_done, 0))) return Fail(rt, NE_DEVICE_FAILURE);
    rt->train_pending = false;
    return NE_OK;
}

// Waits on the host for the training in flight and reads its loss (1/2 MSE of its last minibatch) and time.
static int ReadTraining(ne_runtime* rt) {
    if (JoinTraining(rt) != NE_OK) return NE_DEVICE_FAILURE;
    if (!rt->train_unread) return NE_OK;
    if (!Ok(hipEventSynchronize(rt->train_done))) return Fail(rt, NE_DEVICE_FAILURE);
    const uint32_t bs = rt->screen.batch_size;
    double l = 0.0;
    for (uint32_t b = 0; b < bs; ++b) {
        const double d = double(rt->h_loss[b]) - double(rt->h_loss[bs + b]);
        l += 0.5 * d * d;
    }
    rt->last_loss = float(l / double(bs));
    float ms = 0.0f;
    rt->last_train_ms = Ok(hipEventElapsedTime(&ms, rt->train_t0, rt->train_done)) ? ms : 0.0f;
    rt->train_unread = false;
    return NE_OK;
}

// Mean and standard deviation of the replay buffer's fitness (the training targets' standardisation).
static void ReplayStats(const ne_runtime* rt, float* mu, float* sigma) {
    double m = 0.0, q = 0.0;
    const uint64_t n = rt->replay_count;
    for (uint64_t i = 0; i < n; ++i) m += rt->h_replay_f[i];
    m = n ? m / double(n) : 0.0;
    for (uint64_t i = 0; i < n; ++i) q += (rt->h_replay_f[i] - m) * (rt->h_replay_f[i] - m);
    const double sd = n > 1 ? std::sqrt(q / double(n - 1)) : 0.0;
    *mu = float(m);
    *sigma = sd > 1e-12 ? float(sd) : 1.0f;
}

// train_steps minibatches drawn uniformly from the replay buffer, on the training stream after the main
// stream's work so far (the replay rows appended): started here, not waited for (JoinTraining, ReadTraining).
static int StartTraining(ne_runtime* rt) {
    const uint32_t bs = rt->screen.batch_size;
    if (rt->replay_count < bs) return NE_OK;
    if (ReadTraining(rt) != NE_OK) return NE_DEVICE_FAILURE;  // the pinned buffers are free again
    float mu = 0.0f, sigma = 1.0f;
    ReplayStats(rt, &mu, &sigma);
    // every step's minibatch indices in one upload (draw of row b of step s: Philox(s, b; seed, tag)),
    // then the steps back to back on the training stream
    const uint32_t steps = rt->screen.train_steps;
    for (uint32_t st = 0; st < steps; ++st) {
        const uint64_t s_abs = rt->sg_step + 1 + st;
        for (uint32_t b = 0; b < bs; ++b)
            rt->h_bidx[uint64_t(st) * bs + b] = Philox4x32(uint32_t(s_abs), b, uint32_t(s_abs >> 32), 0,
                                                           rt->screen.seed, kTagBatch).x % uint32_t(rt->replay_count);
    }
    rt->h_coef[0] = mu;
    rt->h_coef[1] = 1.0f / sigma;
    for (uint32_t st = 0; st < steps; ++st) AdamCoef(rt->sg_step + 1 + st, &rt->h_coef[2 + 2 * st]);
    hipStream_t ts = rt->train_stream;
    if (!Ok(hipEventRecord(rt->train_ready, rt->stream)) || !Ok(hipStreamWaitEvent(ts, rt->train_ready, 0)) ||
        !Ok(hipEventRecord(rt->train_t0, ts)) ||
        !Ok(hipMemcpyAsync(rt->bidx, rt->h_bidx, uint64_t(steps) * bs * 4, hipMemcpyHostToDevice, ts)) ||
        !Ok(hipMemcpyAsync(rt->sg_coef, rt->h_coef, (2 + 2 * uint64_t(steps)) * 4, hipMemcpyHostToDevice, ts)))
        return Fail(rt, NE_DEVICE_FAILURE);
    if (rt->screen.flags & NE_SCREEN_EAGER_TRAINING) {
        TrainingLaunches(rt, ts);
    } else {
        // the steps as one hipGraph: captured on first use, replayed every generation (kernels only)
        if (!rt->train_exec) {
            if (!Ok(hipStreamBeginCapture(ts, hipStreamCaptureModeThreadLocal))) return Fail(rt, NE_DEVICE_FAILURE);
            TrainingLaunches(rt, ts);
            if (!Ok(hipStreamEndCapture(ts, &rt->train_graph)) ||
                !Ok(hipGraphInstantiate(&rt->train_exec, rt->train_graph, nullptr, nullptr, 0)))
                return Fail(rt, NE_DEVICE_FAILURE);
        }
        if (!Ok(hipGraphLaunch(rt->train_exec, ts))) return Fail(rt, NE_DEVICE_FAILURE);
    }
    rt->sg_step += steps;
    if (!Ok(hipMemcpyAsync(rt->h_loss, rt->by, bs * 4, hipMemcpyDeviceToHost, ts)) ||
        !Ok(hipMemcpyAsync(rt->h_loss + bs, rt->bt, bs * 4, hipMemcpyDeviceToHost, ts)) ||
        !Ok(hipEventRecord(rt->train_done, ts)))
        return Fail(rt, NE_DEVICE_FAILURE);
    rt->train_pending = true;
    rt->train_unread = true;
    return NE_OK;
}

// Brood assembly of screened evolution, in two phases over the planned children (rt->pairs): phase 1
// breeds every child's policy and probe-vocabulary tail rows (all the fingerprint reads) into newly
// popped slots; phase 2 (BreedRest) completes the children that will be evaluated.  Draws are keyed by
// (element, child), so a completed child equals a one-pass breed.  The parents stay alive until phase 2.
static BreedArgs BroodBreedArgs(ne_runtime* rt, uint32_t seed, const ne_assembly_config& cfg) {
    BreedArgs b{};
    b.slab = rt->slab;
    b.parent_stride = rt->stride;
    b.child_stride = rt->stride;
    b.pairs = rt->pairs;
    b.parent_slots = rt->cur_slots;
    b.records = rt->xover;
    b.breaks = rt->breaks;
    b.free_stack = rt->free_stack;
    b.child_slots = rt->nxt_slots;
    b.slot_state = rt->slot_state;
    b.parent_action_count = uint32_t(rt->action_count);
    b.gen_index = uint32_t(rt->cur_index);
    b.seed = seed;
    b.mutation_p = cfg.mutation_probability;
    b.mutation_sigma = cfg.mutation_sigma;
    b.row_scale_p = cfg.output_tail_row_scale_mutation_probability;
    b.slot_count = uint32_t(rt->slot_count);
    b.status = rt->status;
    return b;
}

static uint32_t BroodPhase1End(const ne_runtime* rt) {
    const uint64_t e = kPolicyParams + std::min<uint64_t>(rt->action_count, kProbeVocabulary) * kTail;
    return uint32_t(std::min<uint64_t>((e + 3) & ~uint64_t(3), rt->stride / 4));
}

static int AssembleBrood(ne_runtime* rt, uint32_t seed, const ne_assembly_config& cfg, uint64_t BN) {
    rt->next_index = rt->cur_index + 1;
    rt->next_size = BN;
    int st = ClearGenerationBuffers(rt, true);
    if (st != NE_OK) return st;
    XoverArgs xa{};
    xa.child_count = uint32_t(BN);
    xa.gen_index = uint32_t(rt->cur_index);
    xa.seed = seed;
    xa.parent_action_count = uint32_t(rt->action_count);
    xa.t1 = cfg.crossover_temperature_level1;
    xa.t2 = cfg.crossover_temperature_level2;
    xa.t3 = cfg.crossover_temperature_level3;
    xa.records = rt->xover;
    xa.breaks = rt->breaks;
    LaunchXover(xa, rt->stream);
    st = Finish(rt);
    if (st != NE_OK) return st;
    const uint32_t fc = HostFreeCount(rt);
    if (fc < BN) return Fail(rt, NE_SLAB_FULL);  // parents and the whole brood are alive together
    BreedArgs b = BroodBreedArgs(rt, seed, cfg);
    b.free_count_before = fc;
    b.child_offset = 0;
    b.elem_begin = 0;
    b.elem_end = BroodPhase1End(rt);
    // a cell's brood members side by side (child c's cell is c mod n): the parents of the children in flight
    // then fit the caches better
    const uint64_t n_cells = rt->cur_size;
    if (n_cells != 0 && BN % n_cells == 0 && BN / n_cells > 1) b.interleave_cells = uint32_t(n_cells);
    EventBegin(rt, 3);
    LaunchBreed(b, uint32_t(BN), rt->stream);
    EventEnd(rt, 3);
    LaunchPopSlots(rt->free_count, uint32_t(BN), rt->stream);
    // not waited for: the caller's next wait checks the status and reads event 3
    rt->times.children = BN;
    return Ok(hipGetLastError()) ? NE_OK : Fail(rt, NE_DEVICE_FAILURE);
}

// Phase 2: the rest of each listed child (device list of `count` brood rows), in its own slot.  Launched
// inside event 13, not waited for.
static int BreedRest(ne_runtime* rt, uint32_t seed, const ne_assembly_config& cfg, const uint32_t* rows,
                     uint64_t count) {
    const uint32_t begin = BroodPhase1End(rt), end = uint32_t(rt->stride / 4);
    EventBegin(rt, 13);
    if (begin < end && count != 0) {
        BreedArgs b = BroodBreedArgs(rt, seed, cfg);
        b.elem_begin = begin;
        b.elem_end = end;
        b.child_list = rows;
        b.given_slots = rt->nxt_slots;
        LaunchBreed(b, uint32_t(count), rt->stream);
    }
    EventEnd(rt, 13);
    return Ok(hipGetLastError()) ? NE_OK : Fail(rt, NE_DEVICE_FAILURE);
}

// Fingerprints of n genomes (device slot list) into out (device [n][192]).
static void LaunchFingerprints(ne_runtime* rt, const uint32_t* slots, uint32_t n, const uint32_t* pairs,
                               const float* parent_fitness, float* out) {
    LaunchTransposeEncoder(rt->slab, rt->stride, slots, n, uint32_t(rt->slot_count), rt->enc_t, rt->stream);
    FingerprintArgs f{};
    f.slab = rt->slab;
    f.stride = rt->stride;
    f.slot_count = uint32_t(rt->slot_count);
    f.action_count = uint32_t(std::min<uint64_t>(rt->action_count, kProbeVocabulary));
    f.slots = slots;
    f.codes = rt->codes;
    f.img = rt->img;
    f.img_bytes = uint32_t(rt->img_bytes);
    f.enc_t = rt->enc_t;
    f.probes = rt->d_probes;
    f.pairs = pairs;
    f.parent_fitness = parent_fitness;
    f.parent_shift = 0.0f;
    f.parent_scale = kParentFitnessScale;
    f.out = out;
    f.status = rt->status;
    f.trunk_bf16 = (rt->eval_mode & 0xFFu) == NE_EVAL_BF16_TRUNK ? 1u : 0u;
    LaunchFingerprint(f, n, rt->stream);
}

// ================================================================================================
extern "C" {

}  // extern "C"

// The runtime over a slab of `slots` slots of config.action_count actions and generation buffers of
// max_gen entries (TryCreateDeviceSlabGARuntimeBuffers, slab_runtime.cu:943-1000: the slab layout's
// byte size is slots x stride).
static int CreateCore(const ne_runtime_config& c, uint64_t slots, uint64_t max_gen, uint64_t epi_rows,
                      ne_runtime** out_runtime) {
    ne_runtime* rt = new (std::nothrow) ne_runtime();
    if (!rt) return NE_DEVICE_FAILURE;
    rt->config = c;
    (void)hipGetDevice(&rt->device);
    {
        int cus = 0;
        if (Ok(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, rt->device)) && cus > 0)
            rt->cu_count = uint64_t(cus);
    }
    if (c.stream) {
        rt->stream = reinterpret_cast<hipStream_t>(c.stream);
    } else {
        if (!Ok(hipStreamCreateWithFlags(&rt->stream, hipStreamNonBlocking))) {
            delete rt;
            return NE_DEVICE_FAILURE;
        }
        rt->own_stream = true;
    }
    rt->action_count = c.action_count;
    rt->stride = SlotStrideBytes(c.action_count);
    rt->slot_count = slots;
    rt->slab_bytes = slots * rt->stride;
    rt->max_gen = max_gen;
    rt->epi_rows = epi_rows;
    rt->generation_byte_budget = c.generation_byte_budget;
    bool ok = Alloc(&rt->slab, slots * rt->stride);
    ok &= Alloc(&rt->slot_state, slots) && Alloc(&rt->free_stack, slots) && Alloc(&rt->free_count, 1);
    ok &= Alloc(&rt->cur_slots, max_gen) && Alloc(&rt->cur_fit, max_gen) && Alloc(&rt->cur_evals, max_gen) &&
          Alloc(&rt->cur_has, max_gen);
    ok &= Alloc(&rt->nxt_slots, max_gen) && Alloc(&rt->nxt_fit, max_gen) && Alloc(&rt->nxt_evals, max_gen) &&
          Alloc(&rt->nxt_has, max_gen);
    ok &= Alloc(&rt->pairs, 2 * max_gen) && Alloc(&rt->refcounts, max_gen) && Alloc(&rt->xover, max_gen) &&
          Alloc(&rt->breaks, max_gen * kBreakStride);
    ok &= Alloc(&rt->raw, max_gen) && Alloc(&rt->local_w, max_gen) && Alloc(&rt->counters, 4) &&
          Alloc(&rt->status, 1) && Alloc(&rt->summary, 1) && Alloc(&rt->d_offsets, max_gen + 1) &&
          Alloc(&rt->d_work_counter, kWorkQueues * kWorkCounterStride);
    for (auto& e : rt->ev) ok &= Ok(hipEventCreate(&e.begin)) && Ok(hipEventCreate(&e.end));
    if (ok) {
        ok &= Ok(hipMemsetAsync(rt->slab, 0, slots * rt->stride, rt->stream));
        ok &= Ok(hipMemsetAsync(rt->status, 0, sizeof(int), rt->stream));
        ok &= Ok(hipMemsetAsync(rt->pairs, 0, 2 * max_gen * sizeof(uint32_t), rt->stream));
        LaunchInitSlabMetadata(rt->slot_state, rt->free_stack, rt->free_count, uint32_t(slots), rt->stream);
        rt->times = ne_stage_times{};
    }
    if (ok) ok = ClearGenerationBuffers(rt, false) == NE_OK && ClearGenerationBuffers(rt, true) == NE_OK;
    if (ok) ok = Sync(rt) == NE_OK;
    if (!ok) {
        ne_destroy(rt);
        return NE_DEVICE_FAILURE;
    }
    *out_runtime = rt;
    return NE_OK;
}

extern "C" {

int ne_create(const ne_runtime_config* config, ne_runtime** out_runtime) {
    if (!config || !out_runtime) return NE_INVALID_ARGUMENT;
    *out_runtime = nullptr;
    ne_runtime_config c = *config;
    if (c.encoder_hidden != 0) return NE_UNSUPPORTED;
    // IsValidDeviceSlabGARuntimeConfig (slab_runtime.hpp:50-60) on this build's stride.
    if (c.genotype_slab_byte_budget < c.generation_byte_budget || c.action_count == 0 || c.grid_column_count == 0 ||
        (c.population_size_ceiling != 0 && c.population_size_ceiling % c.grid_column_count != 0))
        return NE_INVALID_RUNTIME_CONFIG;
    const uint64_t slots = SlotCountForBudget(c.genotype_slab_byte_budget, c.action_count, 0);
    uint64_t max_gen = SlotCountForBudget(c.generation_byte_budget, c.action_count, c.population_size_ceiling);
    max_gen = (max_gen / c.grid_column_count) * c.grid_column_count;
    if (slots == 0 || max_gen == 0 || slots < max_gen || slots >= 0xFFFFFFFFull) return NE_INVALID_RUNTIME_CONFIG;
    // TryComputeGenerationByteBudgetBytes (slab_runtime.cu:683-696): capacity x stride
    c.generation_byte_budget = max_gen * SlotStrideBytes(c.action_count);
    return CreateCore(c, slots, max_gen, max_gen / c.grid_column_count, out_runtime);
}

void ne_destroy(ne_runtime* rt) {
    if (!rt) return;
    if (rt->stream) (void)hipStreamSynchronize(rt->stream);
    if (rt->train_stream) (void)hipStreamSynchronize(rt->train_stream);
    for (StepGraph& g : rt->step_graphs) DestroyStepGraph(g);
    Free(rt->d_step);
    Free(rt->d_readback);
    if (rt->h_step) (void)hipHostFree(rt->h_step);
    if (rt->h_readback) (void)hipHostFree(rt->h_readback);
    Free(rt->slab);
    Free(rt->slot_state);
    Free(rt->free_stack);
    Free(rt->free_count);
    Free(rt->cur_slots);
    Free(rt->cur_fit);
    Free(rt->cur_evals);
    Free(rt->cur_has);
    Free(rt->nxt_slots);
    Free(rt->nxt_fit);
    Free(rt->nxt_evals);
    Free(rt->nxt_has);
    Free(rt->pairs);
    Free(rt->refcounts);
    Free(rt->xover);
    Free(rt->breaks);
    Free(rt->rank_weights);
    Free(rt->enc_t);
    Free(rt->eval_acc);
    Free(rt->bf16_ops);
    Free(rt->tail_absmax);
    Free(rt->raw);
    Free(rt->local_w);
    Free(rt->counters);
    Free(rt->status);
    Free(rt->d_work_counter);
    Free(rt->d_work);
    Free(rt->summary);
    Free(rt->d_offsets);
    Free(rt->d_secrets);
    Free(rt->codes);
    Free(rt->canon);
    Free(rt->img);
    Free(rt->fixed_absmax);
    Free(rt->bf16_cat);
    Free(rt->d_probes);
    Free(rt->sg_params);
    Free(rt->sg_m);
    Free(rt->sg_v);
    Free(rt->fp);
    Free(rt->pred);
    Free(rt->e_fit);
    Free(rt->e_evals);
    Free(rt->e_has);
    Free(rt->e_slots);
    Free(rt->e_rows);
    Free(rt->replay_x);
    Free(rt->replay_f);
    Free(rt->bx);
    Free(rt->bt);
    Free(rt->by);
    Free(rt->bh1);
    Free(rt->bh2);
    Free(rt->bdy);
    Free(rt->bdh1);
    Free(rt->bdh2);
    Free(rt->bidx);
    Free(rt->sg_coef);
    Free(rt->sg_coef1);
    Free(rt->d_fstats);
    Free(rt->cv_by_org);
    Free(rt->cv_by_weight);
    Free(rt->cv_centroid);
    Free(rt->cv_squared);
    Free(rt->cv_index);
    if (rt->train_exec) (void)hipGraphExecDestroy(rt->train_exec);
    if (rt->train_graph) (void)hipGraphDestroy(rt->train_graph);
    Free(rt->owner);
    Free(rt->move_src);
    Free(rt->d_survivors);
    Free(rt->staging);
    Free(rt->d_trace);
    Free(rt->d_policy_trace);
    Free(rt->d_game_offsets);
    for (auto& e : rt->ev) {
        if (e.begin) (void)hipEventDestroy(e.begin);
        if (e.end) (void)hipEventDestroy(e.end);
    }
    if (rt->train_stream) (void)hipStreamDestroy(rt->train_stream);
    for (hipEvent_t e : {rt->train_ready, rt->train_t0, rt->train_done})
        if (e) (void)hipEventDestroy(e);
    for (void* h : {(void*)rt->h_bidx, (void*)rt->h_coef, (void*)rt->h_loss})
        if (h) (void)hipHostFree(h);
    if (rt->own_stream && rt->stream) (void)hipStreamDestroy(rt->stream);
    delete rt;
}

int ne_upload_catalog(ne_runtime* rt, const uint8_t* letters, uint64_t word_count) {
    if (!rt || !letters || word_count == 0 || word_count > kCatalogCapacity) return Fail(rt, NE_INVALID_ARGUMENT);
    ++rt->input_epoch;
    ++rt->upload_epoch;
    std::vector<uint32_t> codes(word_count), canon(word_count);
    std::vector<float> fixed(word_count * kFixed, 0.0f);
    std::unordered_map<uint32_t, uint32_t> first_index;
    for (uint64_t a = 0; a < word_count; ++a) {
        uint32_t code = 0;
        int counts[kAlpha] = {0};
        for (int p = 0; p < 5; ++p) {
            const uint8_t l = letters[5 * a + p];
            if (l >= kAlpha) return Fail(rt, NE_INVALID_ARGUMENT);
            code |= uint32_t(l) << (5 * p);
            ++counts[l];
        }
        codes[a] = code;
        canon[a] = first_index.emplace(code, uint32_t(a)).first->second;
        // MaterializeFixedWordFeatures (output_embedding.hpp:59-73): count, or -1 when absent
        for (int l = 0; l < kAlpha; ++l) fixed[a * kFixed + l] = counts[l] == 0 ? -1.0f : float(counts[l]);
    }
    float absmax[kFixed] = {0.0f};
    for (uint64_t a = 0; a < word_count; ++a)
        for (int d = 0; d < kFixed; ++d) absmax[d] = std::max(absmax[d], std::fabs(fixed[a * kFixed + d]));
    // scoring image: lane i + 16g of a tile holds the fixed steps of action row i for lane group g
    const uint64_t tiles = (word_count + 15) / 16;
    std::vector<uint32_t> img(tiles * kImgTileWords, 0u);
    for (uint64_t t = 0; t < tiles; ++t) {
        uint32_t* tile = img.data() + t * kImgTileWords;
        for (int lane = 0; lane < 64; ++lane) {
            const uint64_t a = t * 16 + (lane & 15);
            if (a >= word_count) continue;
            for (int s = 0; s < ScoreFixedCount(lane >> 4); ++s) {
                const float v = fixed[a * kFixed + ScoreDim(s, lane >> 4)];
                std::memcpy(&tile[(s / 4) * 256 + 4 * lane + (s % 4)], &v, 4);
            }
        }
        for (int i = 0; i < 16; ++i) tile[512 + i] = t * 16 + i < word_count ? canon[t * 16 + i] : kInvalidSlot;
    }
    bool identity = true;
    for (uint64_t a = 0; a < word_count; ++a) identity = identity && canon[a] == uint32_t(a);
    canon.resize(tiles * 16, kInvalidSlot);  // whole tiles, for the screening stream's transfers
    // BF16 screening operands of instruction 0, lanes 0..47 (fixed dims 0..23), round-to-nearest-even
    std::vector<uint16_t> bf16_cat(tiles * kBf16CatTileBytes / 2, 0);
    for (uint64_t t = 0; t < tiles; ++t)
        for (int lane = 0; lane < 48; ++lane) {
            const uint64_t a = t * 16 + (lane & 15);
            if (a >= word_count) continue;
            for (int e = 0; e < 8; ++e)
                bf16_cat[(t * 48 + lane) * 8 + e] = Bf16Bits(fixed[a * kFixed + Bf16SlotDim(0, 8 * (lane >> 4) + e)]);
        }
    Free(rt->codes);
    Free(rt->canon);
    Free(rt->img);
    Free(rt->fixed_absmax);
    Free(rt->bf16_cat);
    rt->img_bytes = tiles * kImgTileBytes;
    rt->bf16_cat_bytes = bf16_cat.size() * 2;
    if (!Alloc(&rt->fixed_absmax, kFixed) ||
        !Ok(hipMemcpy(rt->fixed_absmax, absmax, sizeof(absmax), hipMemcpyHostToDevice)))
        return Fail(rt, NE_DEVICE_FAILURE);
    if (!Alloc(&rt->codes, word_count) || !Alloc(&rt->canon, canon.size()) || !Alloc(&rt->img, img.size()) ||
        !Ok(hipMemcpy(rt->codes, codes.data(), word_count * 4, hipMemcpyHostToDevice)) ||
        !Ok(hipMemcpy(rt->canon, canon.data(), canon.size() * 4, hipMemcpyHostToDevice)) ||
        !Ok(hipMemcpy(rt->img, img.data(), img.size() * 4, hipMemcpyHostToDevice)) ||
        !Alloc(&rt->bf16_cat, rt->bf16_cat_bytes) ||
        !Ok(hipMemcpy(rt->bf16_cat, bf16_cat.data(), rt->bf16_cat_bytes, hipMemcpyHostToDevice)))
        return Fail(rt, NE_DEVICE_FAILURE);
    rt->catalog_count = word_count;
#ifdef NE_NO_CANON_IDENTITY
    identity = false;
#endif
    rt->canon_identity = identity;
    rt->h_catalog.assign(letters, letters + 5 * word_count);
    return NE_OK;
}

int ne_upload_secrets_csr(ne_runtime* rt, const uint32_t* cell_offsets, uint64_t cell_count, const uint32_t* secret_ids,
                          uint64_t secret_count) {
    if (!rt || !cell_offsets || cell_count == 0 || cell_count > rt->max_gen) return Fail(rt, NE_INVALID_ARGUMENT);
    if (cell_offsets[0] != 0 || cell_offsets[cell_count] != secret_count || (secret_count && !secret_ids))
        return Fail(rt, NE_INVALID_ARGUMENT);
    for (uint64_t c = 0; c < cell_count; ++c)
        if (cell_offsets[c + 1] < cell_offsets[c]) return Fail(rt, NE_INVALID_ARGUMENT);
    rt->h_offsets.assign(cell_offsets, cell_offsets + cell_count + 1);
    rt->h_secrets.assign(secret_ids, secret_ids + secret_count);
    rt->csr_from_caller = true;
    ++rt->input_epoch;
    ++rt->upload_epoch;
    return NE_OK;
}

int ne_clear_secrets(ne_runtime* rt) {
    if (!rt) return NE_INVALID_ARGUMENT;
    rt->h_offsets.clear();
    rt->h_secrets.clear();
    rt->csr_from_caller = false;
    ++rt->input_epoch;
    ++rt->upload_epoch;
    return NE_OK;
}

int ne_bootstrap(ne_runtime* rt, uint64_t generation_size, uint32_t generation_seed, uint64_t generation_index,
                 const ne_init_config* init) {
    if (!rt) return NE_INVALID_ARGUMENT;
    if (generation_size == 0 || generation_size > rt->max_gen || generation_size % rt->config.grid_column_count != 0)
        return Fail(rt, NE_INVALID_GENERATION);
    ne_init_config ic{2.0f, 1.0f, 0.05f, 0.0f};
    if (init) ic = *init;
    bool ok = ClearGenerationBuffers(rt, false) == NE_OK && ClearGenerationBuffers(rt, true) == NE_OK;
    ok &= Ok(hipMemsetAsync(rt->pairs, 0, 2 * rt->max_gen * sizeof(uint32_t), rt->stream));
    ok &= Ok(hipMemsetAsync(rt->refcounts, 0, rt->max_gen * sizeof(uint32_t), rt->stream));
    ok &= Ok(hipMemsetAsync(rt->status, 0, sizeof(int), rt->stream));
    if (!ok) return Fail(rt, NE_DEVICE_FAILURE);
    LaunchInitSlabMetadata(rt->slot_state, rt->free_stack, rt->free_count, uint32_t(rt->slot_count), rt->stream);
    BootArgs b{};
    b.slab = rt->slab;
    b.stride = rt->stride;
    b.action_count = uint32_t(rt->action_count);
    b.free_stack = rt->free_stack;
    b.free_count_before = uint32_t(rt->slot_count);
    b.slots = rt->cur_slots;
    b.slot_state = rt->slot_state;
    b.gen_index = uint32_t(generation_index);
    b.seed = generation_seed;
    for (int l = 0; l < kDenseLayers; ++l) {
        // HeNormalStddev (genome_ops.cuh:61-63): gain * sqrt(scale / fan_in)
        b.layer_std[l] = ic.weight_gain * std::sqrt(ic.weight_variance_scale / float(DenseLayerAt(l).in));
    }
    b.tail_std = ic.tail_stddev;
    LaunchBootstrap(b, uint32_t(generation_size), rt->stream);
    LaunchPopSlots(rt->free_count, uint32_t(generation_size), rt->stream);
    const int st = Finish(rt);
    if (st != NE_OK) return st;
    rt->cur_index = generation_index;
    rt->cur_size = generation_size;
    rt->cur_evaluated = false;
    rt->next_index = rt->next_size = rt->planned_children = 0;
    rt->have_summary = false;
    return NE_OK;
}

int ne_evaluate(ne_runtime* rt, const ne_word_counts* counts, const ne_shard_release_history* history) {
    if (!rt) return NE_INVALID_ARGUMENT;
    return EvaluateFull(rt, counts, history);
}

int ne_read_summary(ne_runtime* rt, ne_fitness_summary* out) {
    if (!rt || !out) return NE_INVALID_ARGUMENT;
    if (!rt->have_summary) return Fail(rt, NE_POPULATION_NOT_EVALUATED);
    *out = rt->last_summary;
    return NE_OK;
}

}  // extern "C"

// The first half of a generation step, up to the pre-recombination boundary
// (TryPreparePrebreedingBoundaryOnDevice, slab_runtime.cu:1363-1510): evaluation, the shard-release
// callback (which may set the injection), the next generation's shape, the parent plan, the
// post-fitness callback, and the widening repack when an injection is pending.  *parent_a receives
// the action count the parents were bred with.
static int PrepareBoundary(ne_runtime* rt, uint32_t generation_seed, const ne_word_counts* counts,
                           const ne_assembly_config& cfg, ne_pending_injection& inj, const ne_callbacks* callbacks,
                           const ne_shard_release_history* history, uint64_t* parent_a) {
    rt->times = ne_stage_times{};
    const uint64_t g = rt->cur_index + 1;
    Progress(rt, "Generation %llu: starting advancement from generation %llu (%llu genomes, %llu actions)",
             (unsigned long long)g, (unsigned long long)rt->cur_index, (unsigned long long)rt->cur_size,
             (unsigned long long)rt->action_count);
    Progress(rt, "Generation %llu: evaluating current generation fitness", (unsigned long long)g);
    auto t0 = ProgressClock::now();
    int st = EvaluateFull(rt, counts, history);
    if (st != NE_OK) return st;
    Progress(rt, "Generation %llu: current generation fitness evaluated (%llu games, %llu decisions, best %.6f) in %.1f ms",
             (unsigned long long)g, (unsigned long long)rt->times.games, (unsigned long long)rt->times.decisions,
             double(rt->last_summary.best_fitness), MsSince(t0));
    if (callbacks && callbacks->training_shard_release &&
        callbacks->training_shard_release(rt, counts, &inj, callbacks->user) != 0)
        return Fail(rt, NE_CALLBACK_FAILED);
    Progress(rt, "Generation %llu: planning next generation shape", (unsigned long long)g);
    uint64_t next_a = 0, next_n = 0;
    if (!PlanShape(rt, inj, &next_a, &next_n)) return Fail(rt, NE_INVALID_ASSEMBLY_CONFIG);
    Progress(rt, "Generation %llu: next generation shape planned (%llu genomes, %llu actions%s)", (unsigned long long)g,
             (unsigned long long)next_n, (unsigned long long)next_a, inj.enabled ? ", widening" : "");
    if (inj.enabled && inj.first_catalog_word_index + inj.injection_count > rt->catalog_count)
        return Fail(rt, NE_OUTPUT_EMBEDDING_INJECTION_FAILED);
    const uint64_t radius = cfg.cellular_breeding_radius;
    st = EnsureRankWeights(rt, cfg.rank_exponent, std::max<uint64_t>(rt->cur_size, 25));
    if (st != NE_OK) return st;
    PlanArgs pa{};
    pa.fitness = rt->cur_fit;
    pa.has_fitness = rt->cur_has;
    pa.cur_size = uint32_t(rt->cur_size);
    pa.child_count = uint32_t(next_n);
    pa.cols = uint32_t(rt->config.grid_column_count);
    pa.gen_index = uint32_t(rt->cur_index);
    pa.radius = uint32_t(std::min<uint64_t>(radius, 0xFFFFFFFFu));
    pa.generation_seed = generation_seed;
    pa.rank_weights = rt->rank_weights;
    pa.pairs = rt->pairs;
    pa.status = rt->status;
    Progress(rt, "Generation %llu: building next-generation parent plan", (unsigned long long)g);
    t0 = ProgressClock::now();
    EventBegin(rt, 2);
    LaunchPlan(pa, rt->stream);
    EventEnd(rt, 2);
    st = Finish(rt);
    if (st != NE_OK) return st;
    rt->times.plan_ms = EventMs(rt, 2);
    rt->planned_children = next_n;
    Progress(rt, "Generation %llu: parent plan built (%llu children) in %.1f ms", (unsigned long long)g,
             (unsigned long long)next_n, MsSince(t0));
    if (callbacks && callbacks->post_fitness_evaluation &&
        callbacks->post_fitness_evaluation(rt, counts, callbacks->user) != 0)
        return Fail(rt, NE_CALLBACK_FAILED);
    *parent_a = rt->action_count;
    if (inj.enabled) {
        Progress(rt, "Generation %llu: widening the slab from %llu to %llu actions", (unsigned long long)g,
                 (unsigned long long)rt->action_count, (unsigned long long)next_a);
        t0 = ProgressClock::now();
        st = Widen(rt, next_a);
        if (st != NE_OK) return st;
        Progress(rt, "Generation %llu: slab widened (%llu slots of %llu bytes) in %.1f ms", (unsigned long long)g,
                 (unsigned long long)rt->slot_count, (unsigned long long)rt->stride, MsSince(t0));
    }
    return NE_OK;
}

// The second half (TryAssemblePreparedGenerationOnDevice, slab_runtime.cu:1512-1686): assembly in
// batches, the output-embedding injection, and the generation swap.
static int AssemblePrepared(ne_runtime* rt, uint32_t generation_seed, const ne_assembly_config& cfg,
                            uint64_t parent_a, const ne_pending_injection& inj) {
    const uint64_t g = rt->cur_index + 1;
    const auto t0 = ProgressClock::now();
    Progress(rt, "Generation %llu: assembling %llu children", (unsigned long long)g,
             (unsigned long long)rt->planned_children);
    int st = Assemble(rt, generation_seed, cfg, parent_a);
    if (st != NE_OK) return st;
    if (inj.enabled) {
        Progress(rt, "Generation %llu: injecting %llu output-embedding rows", (unsigned long long)g,
                 (unsigned long long)inj.injection_count);
        const auto ti = ProgressClock::now();
        st = Inject(rt, parent_a, inj);
        if (st != NE_OK) return st;
        Progress(rt, "Generation %llu: output-embedding rows injected in %.1f ms", (unsigned long long)g, MsSince(ti));
    }
    SwapGenerations(rt);
    rt->last_status = NE_OK;
    Progress(rt, "Generation %llu: advancement finished in %.1f ms (assembly and swap)", (unsigned long long)g,
             MsSince(t0));
    return NE_OK;
}

// ---- captured generation step ---------------------------------------------------------------
// A plain step that needs no host decision between its stages -- no injection or widening, the same
// population size, every child fitting the free slots in one assembly batch, the secrets given as a CSR,
// no trace and no callbacks -- is one fixed kernel sequence: clears, encoder transpose (+ BF16 pack),
// evaluation, finalize + summary, parent selection, refcounts and the release of unreferenced parents,
// crossover decisions, one assembly batch, the slot pop, the batch's releases, and a gather of what the
// host reads back.  It is captured once per buffer parity (the current and next generation buffers
// swap every step) and replayed with the step's generation index and seed in a device block, so a step
// costs one graph launch and one host read instead of ~25 launches and 6 synchronisations.  The kernels
// and their arguments are those of the eager step; only the per-step values come from the device block
// and the assembly batch takes its free count on the device.

static void LaunchStepSequence(ne_runtime* rt, const EvalSetup& es, const PlanArgs& pa, const XoverArgs& xa,
                               const BreedArgs& b, uint64_t n, uint64_t next_n) {
    LaunchEvaluation(rt, es, n, rt->cur_fit, rt->cur_evals, rt->cur_has, false);
    LaunchFinalizeSummary(rt, rt->cur_slots, n, rt->cur_fit, rt->cur_evals, rt->cur_has, false);
    LaunchPlan(pa, rt->stream);
    (void)ClearGenerationBuffers(rt, true);
    LaunchFillU32(rt->refcounts, 0u, rt->max_gen, rt->stream);
    LaunchRefcount(rt->pairs, uint32_t(next_n), uint32_t(n), rt->refcounts, rt->status, rt->stream);
    LaunchCollectReleased(rt->cur_slots, uint32_t(n), rt->refcounts, rt->slot_state, rt->free_stack, rt->free_count,
                          uint32_t(rt->slot_count), rt->status, rt->stream);
    LaunchXover(xa, rt->stream);
    LaunchBreed(b, uint32_t(next_n), rt->stream);
    LaunchPopSlots(rt->free_count, uint32_t(next_n), rt->stream);
    LaunchReleaseBatch(rt->pairs, 0u, uint32_t(next_n), rt->refcounts, rt->status, rt->stream);
    LaunchCollectReleased(rt->cur_slots, uint32_t(n), rt->refcounts, rt->slot_state, rt->free_stack, rt->free_count,
                          uint32_t(rt->slot_count), rt->status, rt->stream);
    LaunchStepReadback(rt->status, rt->free_count, rt->summary, rt->counters, rt->d_readback, rt->stream);
}

// Runs the step as a graph when it qualifies (*done = true, the step's status returned); otherwise
// leaves everything untouched for the eager step (*done = false).
static int TryGraphStep(ne_runtime* rt, uint32_t seed, const ne_word_counts* counts, const ne_assembly_config& cfg,
                        bool* done) {
    *done = false;
    const uint64_t cols = rt->config.grid_column_count, n = rt->cur_size;
    if (!rt->step_graph_on || rt->trace_flags != 0 || !rt->csr_from_caller || n == 0 || cols == 0 || n % cols != 0 ||
        rt->catalog_count == 0 || rt->h_offsets.size() < n + 1)
        return NE_OK;
    if (counts->training_word_count > rt->catalog_count || counts->action_space_word_count > rt->catalog_count ||
        counts->training_word_count > counts->action_space_word_count)
        return NE_OK;  // the eager step reports it
    const ne_pending_injection none{};
    uint64_t next_a = 0, next_n = 0;
    if (!PlanShape(rt, none, &next_a, &next_n) || next_n != n || next_a != rt->action_count) return NE_OK;
    // every child must find a free slot in one batch: the free count at the start of the step (it only
    // grows before the batch, as unreferenced parents are released) covers the whole generation
    if (HostFreeCount(rt) < next_n) return NE_OK;
    if (!rt->d_step && (!Alloc(&rt->d_step, 4) || !Ok(hipHostMalloc((void**)&rt->h_step, 16))))
        return Fail(rt, NE_DEVICE_FAILURE);
    if (!rt->d_readback && (!Alloc(&rt->d_readback, 1) ||
                            !Ok(hipHostMalloc((void**)&rt->h_readback, sizeof(StepReadback)))))
        return Fail(rt, NE_DEVICE_FAILURE);
    *done = true;
    rt->times = ne_stage_times{};

    // the evaluation's host preparation (the CSR and work-list uploads), kept while nothing else has
    // prepared an evaluation or uploaded inputs since (input_epoch) and its shape is the same; only the
    // slot list changes with the buffer parity
    auto eval_key = [&](uint64_t (&k)[10]) {
        const uint64_t v[10] = {rt->input_epoch, n, std::min(rt->action_count, counts->action_space_word_count),
                                rt->eval_mode,
                                rt->tuned_waves | (uint64_t(rt->tuned_rounds) << 16) | (uint64_t(rt->tuned_wg_per_cu) << 32) |
                                    (uint64_t(rt->tuned_kernel) << 48),
                                uint64_t(uintptr_t(rt->slab)), rt->stride,
                                rt->slot_count, rt->action_count, rt->max_gen};
        std::memcpy(k, v, sizeof(v));
    };
    uint64_t ek[10];
    eval_key(ek);
    if (std::memcmp(ek, rt->cached_eval_key, sizeof(ek)) != 0) {
        std::memset(rt->cached_eval_key, 0, sizeof(rt->cached_eval_key));
        const int st = PrepareEvaluation(rt, counts->action_space_word_count, rt->cur_slots, n, rt->h_offsets,
                                         rt->h_secrets, &rt->cached_eval);
        if (st != NE_OK) return st;
        eval_key(rt->cached_eval_key);  // after the preparation's own epoch bump
    }
    EvalSetup es = rt->cached_eval;
    es.a.slot_indices = rt->cur_slots;
    int st = EnsureRankWeights(rt, cfg.rank_exponent, std::max<uint64_t>(n, 25));
    if (st != NE_OK) return st;

    StepGraphKey key;
    std::memset(&key, 0, sizeof(key));
    key.a = es.a;
    key.blocks = es.blocks;
    key.waves = es.waves;
    key.bf16 = es.bf16 ? 1u : 0u;
    PlanArgs& pa = key.pa;
    pa.fitness = rt->cur_fit;
    pa.has_fitness = rt->cur_has;
    pa.cur_size = uint32_t(n);
    pa.child_count = uint32_t(next_n);
    pa.cols = uint32_t(cols);
    pa.radius = uint32_t(std::min<uint64_t>(cfg.cellular_breeding_radius, 0xFFFFFFFFu));
    pa.rank_weights = rt->rank_weights;
    pa.pairs = rt->pairs;
    pa.status = rt->status;
    pa.step = rt->d_step;
    XoverArgs& xa = key.xa;
    xa.child_count = uint32_t(next_n);
    xa.parent_action_count = uint32_t(rt->action_count);
    xa.t1 = cfg.crossover_temperature_level1;
    xa.t2 = cfg.crossover_temperature_level2;
    xa.t3 = cfg.crossover_temperature_level3;
    xa.records = rt->xover;
    xa.breaks = rt->breaks;
    xa.step = rt->d_step;
    BreedArgs& b = key.b;
    b.slab = rt->slab;
    b.parent_stride = rt->stride;
    b.child_stride = rt->stride;
    b.pairs = rt->pairs;
    b.parent_slots = rt->cur_slots;
    b.records = rt->xover;
    b.breaks = rt->breaks;
    b.free_stack = rt->free_stack;
    b.child_slots = rt->nxt_slots;
    b.slot_state = rt->slot_state;
    b.parent_action_count = uint32_t(rt->action_count);
    b.mutation_p = cfg.mutation_probability;
    b.mutation_sigma = cfg.mutation_sigma;
    b.row_scale_p = cfg.output_tail_row_scale_mutation_probability;
    b.slot_count = uint32_t(rt->slot_count);
    b.status = rt->status;
    b.elem_begin = 0;
    b.elem_end = uint32_t(rt->stride / 4);
    b.step = rt->d_step;
    b.free_count_dev = rt->free_count;
    key.n = n;
    key.next_n = next_n;
    key.max_gen = rt->max_gen;
    const void* bufs[8] = {rt->cur_evals, rt->nxt_fit, rt->nxt_evals, rt->nxt_has, rt->refcounts, rt->summary,
                           rt->d_readback, rt->enc_t};
    std::memcpy(key.buffers, bufs, sizeof(bufs));

    StepGraph* g = nullptr;
    for (StepGraph& c : rt->step_graphs)
        if (c.exec && std::memcmp(&c.key, &key, sizeof(key)) == 0) g = &c;
    const bool capture = g == nullptr;
    if (capture) {
        g = rt->step_graphs[0].last_use <= rt->step_graphs[1].last_use ? &rt->step_graphs[0] : &rt->step_graphs[1];
        DestroyStepGraph(*g);
        // nothing runs during a capture: when it cannot be made, the step runs eagerly and later steps
        // stay eager (ne_stage_times.step_graph reads 0)
        bool made = Ok(hipStreamBeginCapture(rt->stream, hipStreamCaptureModeThreadLocal));
        if (made) {
            LaunchStepSequence(rt, es, key.pa, key.xa, key.b, n, next_n);
            made = Ok(hipStreamEndCapture(rt->stream, &g->graph));
            made = made && Ok(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        }
        if (!made) {
            DestroyStepGraph(*g);
            (void)hipGetLastError();
            rt->step_graph_on = false;
            rt->times = ne_stage_times{};
            *done = false;
            return NE_OK;
        }
        g->key = key;
    }
    g->last_use = ++rt->step_graph_uses;
    rt->h_step[0] = uint32_t(rt->cur_index);
    rt->h_step[1] = seed;
    bool ok = Ok(hipMemcpyAsync(rt->d_step, rt->h_step, 8, hipMemcpyHostToDevice, rt->stream));
    EventBegin(rt, 11);
    ok &= Ok(hipGraphLaunch(g->exec, rt->stream));
    EventEnd(rt, 11);
    ok &= Ok(hipMemcpyAsync(rt->h_readback, rt->d_readback, sizeof(StepReadback), hipMemcpyDeviceToHost, rt->stream));
    if (!ok) return Fail(rt, NE_DEVICE_FAILURE);
    st = Sync(rt);
    if (st != NE_OK) return st;
    const StepReadback rb = *rt->h_readback;
    if (rb.status != 0) {
        (void)hipMemset(rt->status, 0, sizeof(int));
        return Fail(rt, rb.status);
    }
    RecordEvaluation(rt, rb.summary, rb.counters, es.total_games, n, rt->cur_index);
    rt->times.step_graph_ms = EventMs(rt, 11);
    rt->times.step_graph = capture ? 2u : 1u;
    rt->times.children = next_n;
    rt->cur_evaluated = true;
    rt->planned_children = next_n;
    rt->next_index = rt->cur_index + 1;
    rt->next_size = next_n;
    SwapGenerations(rt);
    rt->last_status = NE_OK;
    return NE_OK;
}

extern "C" {

int ne_advance_generation(ne_runtime* rt, uint32_t generation_seed, const ne_word_counts* counts,
                          const ne_assembly_config* config, const ne_pending_injection* injection, uint32_t verbose,
                          const ne_callbacks* callbacks, const ne_shard_release_history* history) {
    if (!rt || !counts) return Fail(rt, NE_INVALID_ARGUMENT);
    const ne_assembly_config cfg = config ? *config : DefaultAssembly();
    if (!ValidAssembly(cfg)) return Fail(rt, NE_INVALID_ASSEMBLY_CONFIG);
    ne_pending_injection inj{};
    if (injection) inj = *injection;
    struct VerboseScope {
        ne_runtime* rt;
        ~VerboseScope() { rt->verbose = false; }
    } verbose_scope{rt};
    rt->verbose = verbose != 0;
    const bool plain = !inj.enabled && (!callbacks || (!callbacks->training_shard_release &&
                                                       !callbacks->post_fitness_evaluation));
    if (plain && rt->step_graph_on) {
        bool done = false;
        const uint64_t g = rt->cur_index + 1;
        const auto t0 = ProgressClock::now();
        const int gst = TryGraphStep(rt, generation_seed, counts, cfg, &done);
        if (done && gst == NE_OK)
            Progress(rt, "Generation %llu: advanced as one captured step (%llu games, best %.6f) in %.1f ms",
                     (unsigned long long)g, (unsigned long long)rt->times.games,
                     double(rt->last_summary.best_fitness), MsSince(t0));
        if (done || gst != NE_OK) return gst;
    }
    uint64_t parent_a = 0;
    const int st = PrepareBoundary(rt, generation_seed, counts, cfg, inj, callbacks, history, &parent_a);
    if (st != NE_OK) return st;
    return AssemblePrepared(rt, generation_seed, cfg, parent_a, inj);
}

int ne_get_slab_info(ne_runtime* rt, ne_slab_info* out) {
    if (!rt || !out) return NE_INVALID_ARGUMENT;
    out->action_count = rt->action_count;
    out->slot_stride_bytes = rt->stride;
    out->slot_count = rt->slot_count;
    out->slab_bytes = rt->slab_bytes;
    out->policy_parameter_count = kPolicyParams;
    out->max_generation_size = rt->max_gen;
    out->current_generation_index = rt->cur_index;
    out->current_generation_size = rt->cur_size;
    out->next_generation_size = rt->next_size;
    out->free_slot_count = HostFreeCount(rt);
    return NE_OK;
}

int ne_download_slot(ne_runtime* rt, uint32_t slot_index, void* dst, uint64_t capacity, uint64_t* written) {
    if (!rt || !dst || slot_index >= rt->slot_count || capacity < rt->stride) return Fail(rt, NE_INVALID_ARGUMENT);
    uint32_t state = 0;
    if (!Ok(hipMemcpy(&state, rt->slot_state + slot_index, 4, hipMemcpyDeviceToHost))) return Fail(rt, NE_DEVICE_FAILURE);
    if (state == 0) return Fail(rt, NE_INVALID_SLAB);
    if (!Ok(hipMemcpy(dst, rt->slab + uint64_t(slot_index) * rt->stride, rt->stride, hipMemcpyDeviceToHost)))
        return Fail(rt, NE_DEVICE_FAILURE);
    if (written) *written = rt->stride;
    return NE_OK;
}

int ne_download_slab(ne_runtime* rt, void* storage, uint64_t storage_capacity, uint32_t* slot_states,
                     uint32_t* free_stack, uint64_t slot_capacity, uint32_t* free_count) {
    if (!rt) return NE_INVALID_ARGUMENT;
    if (storage && storage_capacity < rt->slot_count * rt->stride) return Fail(rt, NE_INVALID_ARGUMENT);
    if ((slot_states || free_stack) && slot_capacity < rt->slot_count) return Fail(rt, NE_INVALID_ARGUMENT);
    bool ok = Ok(hipStreamSynchronize(rt->stream));
    if (ok && storage) ok = Ok(hipMemcpy(storage, rt->slab, rt->slot_count * rt->stride, hipMemcpyDeviceToHost));
    if (ok && slot_states)
        ok = Ok(hipMemcpy(slot_states, rt->slot_state, rt->slot_count * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (ok && free_stack)
        ok = Ok(hipMemcpy(free_stack, rt->free_stack, rt->slot_count * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (ok && free_count) ok = Ok(hipMemcpy(free_count, rt->free_count, sizeof(uint32_t), hipMemcpyDeviceToHost));
    return ok ? NE_OK : Fail(rt, NE_DEVICE_FAILURE);
}

// SwapDeviceSlabGenerationBuffers (slab_runtime.cu:2128-2135): the current and next generation exchange
// their arrays, indices and sizes.  The new current generation's fitness is not taken as measured for the
// screened step (it re-evaluates).
int ne_swap_generations(ne_runtime* rt) {
    if (!rt) return NE_INVALID_ARGUMENT;
    std::swap(rt->cur_slots, rt->nxt_slots);
    std::swap(rt->cur_fit, rt->nxt_fit);
    std::swap(rt->cur_evals, rt->nxt_evals);
    std::swap(rt->cur_has, rt->nxt_has);
    std::swap(rt->cur_index, rt->next_index);
    std::swap(rt->cur_size, rt->next_size);
    rt->cur_evaluated = false;
    rt->have_summary = false;
    return NE_OK;
}

int ne_upload_slot(ne_runtime* rt, uint32_t slot_index, const void* src, uint64_t byte_count) {
    if (!rt || !src || slot_index >= rt->slot_count || byte_count > rt->stride) return Fail(rt, NE_INVALID_ARGUMENT);
    uint32_t state = 0;
    if (!Ok(hipMemcpy(&state, rt->slot_state + slot_index, 4, hipMemcpyDeviceToHost))) return Fail(rt, NE_DEVICE_FAILURE);
    if (state == 0) return Fail(rt, NE_INVALID_SLAB);
    if (!Ok(hipMemcpy(rt->slab + uint64_t(slot_index) * rt->stride, src, byte_count, hipMemcpyHostToDevice)))
        return Fail(rt, NE_DEVICE_FAILURE);
    return NE_OK;
}

int ne_download_generation(ne_runtime* rt, int which, uint32_t* slot_indices, float* fitness,
                           uint32_t* evaluation_counts, uint8_t* has_fitness, uint64_t capacity, uint64_t* count) {
    if (!rt || (which != 0 && which != 1)) return Fail(rt, NE_INVALID_ARGUMENT);
    const uint64_t n = which == 0 ? rt->cur_size : rt->next_size;
    if (count) *count = n;
    if (n == 0) return Fail(rt, NE_INVALID_GENERATION);
    if (capacity < n) return Fail(rt, NE_INVALID_ARGUMENT);
    bool ok = true;
    if (slot_indices)
        ok &= Ok(hipMemcpy(slot_indices, which == 0 ? rt->cur_slots : rt->nxt_slots, n * 4, hipMemcpyDeviceToHost));
    if (fitness) ok &= Ok(hipMemcpy(fitness, which == 0 ? rt->cur_fit : rt->nxt_fit, n * 4, hipMemcpyDeviceToHost));
    if (evaluation_counts)
        ok &= Ok(hipMemcpy(evaluation_counts, which == 0 ? rt->cur_evals : rt->nxt_evals, n * 4, hipMemcpyDeviceToHost));
    if (has_fitness) ok &= Ok(hipMemcpy(has_fitness, which == 0 ? rt->cur_has : rt->nxt_has, n, hipMemcpyDeviceToHost));
    return ok ? NE_OK : Fail(rt, NE_DEVICE_FAILURE);
}

int ne_download_plan(ne_runtime* rt, uint32_t* pairs, uint64_t capacity_pairs, uint64_t* count) {
    if (!rt || !pairs) return Fail(rt, NE_INVALID_ARGUMENT);
    const uint64_t n = rt->planned_children ? rt->planned_children : rt->cur_size;
    if (count) *count = n;
    if (capacity_pairs < n) return Fail(rt, NE_INVALID_ARGUMENT);
    return Ok(hipMemcpy(pairs, rt->pairs, n * 8, hipMemcpyDeviceToHost)) ? NE_OK : Fail(rt, NE_DEVICE_FAILURE);
}

int ne_download_local_word_counts(ne_runtime* rt, uint32_t* dst, uint64_t capacity, uint64_t* count) {
    if (!rt || !dst || capacity < rt->cur_size) return Fail(rt, NE_INVALID_ARGUMENT);
    if (count) *count = rt->cur_size;
    return Ok(hipMemcpy(dst, rt->local_w, rt->cur_size * 4, hipMemcpyDeviceToHost)) ? NE_OK
                                                                                    : Fail(rt, NE_DEVICE_FAILURE);
}

int ne_build_shard_csr(const ne_word_counts* counts, const ne_shard_release_history* history,
                       uint64_t generation_index, uint64_t rows, uint64_t cols, uint64_t epicenter_rows,
                       uint32_t* offsets, uint32_t* ids, uint64_t ids_capacity, uint64_t* id_count) {
    if (!counts || !offsets || !id_count) return NE_INVALID_ARGUMENT;
    std::vector<uint32_t> offs, secs;
    const int st = BuildShardCsrCore(counts, history, generation_index, rows, cols, epicenter_rows, 4739, offs, secs);
    if (st != NE_OK) return st;
    *id_count = secs.size();
    std::memcpy(offsets, offs.data(), offs.size() * sizeof(uint32_t));
    if (secs.size() > ids_capacity || (!ids && !secs.empty())) return NE_INVALID_ARGUMENT;
    if (!secs.empty()) std::memcpy(ids, secs.data(), secs.size() * sizeof(uint32_t));
    return NE_OK;
}

int ne_read_status(ne_runtime* rt) { return rt ? rt->last_status : NE_INVALID_ARGUMENT; }

const char* ne_status_string(int status) {
    switch (status) {
    case NE_OK: return "ok";
    case NE_DEVICE_FAILURE: return "device failure";
    case NE_INVALID_RUNTIME_CONFIG: return "invalid runtime config";
    case NE_INVALID_SLAB: return "invalid slab";
    case NE_INVALID_GENERATION: return "invalid generation";
    case NE_INVALID_TRAINING_SHARD: return "invalid training shard";
    case NE_GUESS_APPEND_FAILED: return "guess append failed";
    case NE_POLICY_FORWARD_FAILED: return "policy forward failed";
    case NE_ACTION_SELECTION_FAILED: return "action selection failed";
    case NE_POPULATION_NOT_EVALUATED: return "population not evaluated";
    case NE_INVALID_ASSEMBLY_CONFIG: return "invalid assembly config";
    case NE_PARENT_SELECTION_FAILED: return "parent selection failed";
    case NE_INVALID_ASSEMBLY_PLAN: return "invalid assembly plan";
    case NE_INVALID_PARENT_INDEX: return "invalid parent index";
    case NE_SLAB_FULL: return "slab full";
    case NE_OUTPUT_EMBEDDING_INJECTION_FAILED: return "output embedding injection failed";
    case NE_SLAB_REPACK_FAILED: return "slab repack failed";
    case NE_INVALID_ARGUMENT: return "invalid argument";
    case NE_UNSUPPORTED: return "unsupported configuration";
    case NE_CALLBACK_FAILED: return "callback failed";
    default: return "unknown status";
    }
}

int ne_get_stage_times(ne_runtime* rt, ne_stage_times* out) {
    if (!rt || !out) return NE_INVALID_ARGUMENT;
    *out = rt->times;
    return NE_OK;
}

int ne_set_eval_precision(ne_runtime* rt, uint32_t mode) {
    if (!rt) return NE_INVALID_ARGUMENT;
    const uint32_t base = mode & 0xFFu;
    if ((base != NE_EVAL_FP32 && base != NE_EVAL_BF16_RESCORED && base != NE_EVAL_BF16_TRUNK) || (mode & ~0x1FFu) != 0)
        return Fail(rt, NE_INVALID_ARGUMENT);
    rt->eval_mode = mode;
    return NE_OK;
}

int ne_set_trace(ne_runtime* rt, uint32_t flags) {
    if (!rt) return NE_INVALID_ARGUMENT;
    if ((flags & NE_TRACE_POLICY) && !(rt->trace_flags & NE_TRACE_POLICY)) {
        Free(rt->d_trace);
        Free(rt->d_policy_trace);
        rt->trace_capacity = 0;
    }
    rt->trace_flags = flags;
    return NE_OK;
}

int ne_download_trace(ne_runtime* rt, uint32_t kind, void* dst, uint64_t capacity_bytes, uint64_t* written) {
    if (!rt || !dst) return Fail(rt, NE_INVALID_ARGUMENT);
    const void* src = nullptr;
    uint64_t bytes = 0;
    if (kind == NE_TRACE_GAMES) {
        src = rt->d_trace;
        bytes = rt->trace_games * sizeof(GameTrace);
    } else if (kind == NE_TRACE_POLICY) {
        src = rt->d_policy_trace;
        bytes = rt->trace_games * uint64_t(kMaxTurns) * kPol * sizeof(float);
    } else if (kind == NE_TRACE_XOVER) {
        src = rt->xover;
        bytes = (rt->cur_size) * sizeof(XoverRecord);
    } else {
        return Fail(rt, NE_INVALID_ARGUMENT);
    }
    if (!src || bytes == 0) return Fail(rt, NE_INVALID_ARGUMENT);
    if (capacity_bytes < bytes) return Fail(rt, NE_INVALID_ARGUMENT);
    if (!Ok(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost))) return Fail(rt, NE_DEVICE_FAILURE);
    if (written) *written = bytes;
    return NE_OK;
}


// ---- evaluation shape ---------------------------------------------------------------------------
int ne_set_eval_shape(ne_runtime* rt, uint32_t waves) {
    if (!rt || waves > EvaluateMaxWaves()) return Fail(rt, NE_INVALID_ARGUMENT);
    rt->tuned_waves = waves;
    return NE_OK;
}

int ne_set_step_graph(ne_runtime* rt, uint32_t enabled) {
    if (!rt || enabled > 1) return Fail(rt, NE_INVALID_ARGUMENT);
    rt->step_graph_on = enabled != 0;
    if (!rt->step_graph_on)
        for (StepGraph& g : rt->step_graphs) DestroyStepGraph(g);
    return NE_OK;
}

int ne_autotune_evaluation(ne_runtime* rt, const ne_word_counts* counts, const ne_shard_release_history* history,
                           uint32_t* chosen_waves, float* timings_ms) {
    if (!rt || !counts) return Fail(rt, NE_INVALID_ARGUMENT);
    // coordinate search, one knob at a time, each candidate timed on one evaluation of the current generation
    // (results do not depend on the shape): waves per workgroup, then (untraced forms) the streamed kernel
    // and two game tiles per wave at every workgroup size they allow, then rounds per work item, then
    // workgroups per CU
    const uint32_t saved[4] = {rt->tuned_waves, rt->tuned_rounds, rt->tuned_wg_per_cu, rt->tuned_kernel};
    auto time_shape = [&](uint32_t w, uint32_t kern, uint32_t r, uint32_t g, float* ms) -> int {
        rt->tuned_waves = w;
        rt->tuned_kernel = kern;
        rt->tuned_rounds = r;
        rt->tuned_wg_per_cu = g;
        rt->times = ne_stage_times{};
        const int st = EvaluateFull(rt, counts, history);
        if (st != NE_OK) {
            rt->tuned_waves = saved[0];
            rt->tuned_rounds = saved[1];
            rt->tuned_wg_per_cu = saved[2];
            rt->tuned_kernel = saved[3];
            return st;
        }
        *ms = rt->times.evaluate_ms;
        return NE_OK;
    };
    uint32_t bw = 0, bt = kEvalKernelRegisters, br = 1, bg = 0;
    float best_ms = 0.0f;
    for (uint32_t w = 1; w <= 4; ++w) {
        float ms = 0.0f;
        const int st = time_shape(w, kEvalKernelRegisters, 1, 0, &ms);
        if (st != NE_OK) return st;
        if (timings_ms) timings_ms[w - 1] = ms;
        if (bw == 0 || ms < best_ms) {
            bw = w;
            best_ms = ms;
        }
    }
    const bool trace = rt->trace_flags != 0,
               bf16 = (rt->eval_mode & 0xFFu) == NE_EVAL_BF16_RESCORED || (rt->eval_mode & 0xFFu) == NE_EVAL_BF16_TRUNK;
    const bool t16 = (rt->eval_mode & 0xFFu) == NE_EVAL_BF16_TRUNK;
    if (!trace) {
        for (const uint32_t kern : {kEvalKernelStreamed, kEvalKernelPaired})
            for (uint32_t w = 1; w <= EvaluateMaxWavesFor(kern, trace, bf16); ++w) {
                if (EvaluateKernelFor(kern, w, trace, bf16) != kern) continue;  // runs another kernel, timed above
                float ms = 0.0f;
                const int st = time_shape(w, kern, 1, 0, &ms);
                if (st != NE_OK) return st;
                if (ms < best_ms) {
                    bw = w;
                    bt = kern;
                    best_ms = ms;
                }
            }
    }
    for (uint32_t r = 2; r <= 4; r *= 2) {
        float ms = 0.0f;
        const int st = time_shape(bw, bt, r, 0, &ms);
        if (st != NE_OK) return st;
        if (ms < best_ms) {
            br = r;
            best_ms = ms;
        }
    }
    const uint32_t resident = std::max<uint32_t>(1, EvaluateResidentPerCu(bw, bt, trace, bf16, t16));
    for (uint32_t g = resident - 1; g >= 1 && g + 2 >= resident; --g) {
        float ms = 0.0f;
        const int st = time_shape(bw, bt, br, g, &ms);
        if (st != NE_OK) return st;
        if (ms < best_ms) {
            bg = g;
            best_ms = ms;
        }
    }
    rt->tuned_waves = bw;
    rt->tuned_kernel = bt;
    rt->tuned_rounds = br;
    rt->tuned_wg_per_cu = bg;
    if (chosen_waves) *chosen_waves = bw;
    return EvaluateFull(rt, counts, history);  // leave the generation evaluated with the chosen shape
}

int ne_get_eval_shape(ne_runtime* rt, ne_eval_shape* out) {
    if (!rt || !out) return NE_INVALID_ARGUMENT;
    out->waves = rt->tuned_waves;
    out->rounds_per_item = rt->tuned_rounds;
    out->workgroups_per_cu = rt->tuned_wg_per_cu;
    out->kernel = rt->tuned_kernel;
    return NE_OK;
}

int ne_set_eval_shape_ex(ne_runtime* rt, const ne_eval_shape* shape) {
    if (!rt || !shape || shape->waves > EvaluateMaxWaves() || shape->rounds_per_item > 64 ||
        shape->workgroups_per_cu > 64 || shape->kernel > kEvalKernelPaired)
        return Fail(rt, NE_INVALID_ARGUMENT);
    rt->tuned_kernel = shape->kernel;
    rt->tuned_waves = shape->waves;
    rt->tuned_rounds = std::max<uint32_t>(1, shape->rounds_per_item);
    rt->tuned_wg_per_cu = shape->workgroups_per_cu;
    return NE_OK;
}

// ---- surrogate screening ----------------------------------------------------------------------
int ne_enable_screening(ne_runtime* rt, const ne_screen_config* config) {
    if (!rt || !config) return Fail(rt, NE_INVALID_ARGUMENT);
    const ne_screen_config& c = *config;
    if (c.brood_factor == 0 || c.brood_factor > 16 || c.batch_size == 0 || c.batch_size > 1024 ||
        c.replay_capacity < c.batch_size || !(c.audit_fraction >= 0.0f && c.audit_fraction <= 1.0f) ||
        !(c.learning_rate > 0.0f) || !(c.spearman_threshold >= -1.0f && c.spearman_threshold <= 1.0f))
        return Fail(rt, NE_INVALID_ARGUMENT);
    if (rt->catalog_count == 0) return Fail(rt, NE_INVALID_TRAINING_SHARD);
    if (!DrawProbes(rt, std::min<uint64_t>({rt->action_count, rt->catalog_count, kProbeVocabulary}), c.seed))
        return Fail(rt, NE_INVALID_TRAINING_SHARD);
    if (rt->train_stream && !Ok(hipStreamSynchronize(rt->train_stream))) return Fail(rt, NE_DEVICE_FAILURE);
    rt->train_pending = rt->train_unread = false;
    if (!rt->train_stream &&
        (!Ok(hipStreamCreateWithFlags(&rt->train_stream, hipStreamNonBlocking)) ||
         !Ok(hipEventCreateWithFlags(&rt->train_ready, hipEventDisableTiming)) || !Ok(hipEventCreate(&rt->train_t0)) ||
         !Ok(hipEventCreate(&rt->train_done))))
        return Fail(rt, NE_DEVICE_FAILURE);
    for (void* h : {(void*)rt->h_bidx, (void*)rt->h_coef, (void*)rt->h_loss})
        if (h) (void)hipHostFree(h);
    rt->h_bidx = nullptr;
    rt->h_coef = nullptr;
    rt->h_loss = nullptr;
    if (!Ok(hipHostMalloc(reinterpret_cast<void**>(&rt->h_bidx),
                          std::max<uint64_t>(uint64_t(c.train_steps) * c.batch_size, 1) * 4)) ||
        !Ok(hipHostMalloc(reinterpret_cast<void**>(&rt->h_coef), (2 + 2 * uint64_t(c.train_steps)) * 4)) ||
        !Ok(hipHostMalloc(reinterpret_cast<void**>(&rt->h_loss), 2 * uint64_t(c.batch_size) * 4)))
        return Fail(rt, NE_DEVICE_FAILURE);
    rt->last_loss = std::nanf("");
    rt->last_train_ms = 0.0f;
    const bool realloc = !rt->screen_on || rt->screen.replay_capacity != c.replay_capacity;
    const uint64_t bmax = std::max<uint64_t>(c.batch_size, 1024);
    if (realloc) {
        Free(rt->replay_x);
        Free(rt->replay_f);
        if (!Alloc(&rt->replay_x, uint64_t(c.replay_capacity) * kSgIn) || !Alloc(&rt->replay_f, c.replay_capacity))
            return Fail(rt, NE_DEVICE_FAILURE);
    }
    if (!rt->screen_on) {
        bool ok = Alloc(&rt->d_probes, 1) && Alloc(&rt->sg_params, kSgParams) && Alloc(&rt->sg_m, kSgParams) &&
                  Alloc(&rt->sg_v, kSgParams) && Alloc(&rt->fp, rt->max_gen * kSgIn) && Alloc(&rt->pred, rt->max_gen) &&
                  Alloc(&rt->e_fit, rt->max_gen) && Alloc(&rt->e_evals, rt->max_gen) && Alloc(&rt->e_has, rt->max_gen) &&
                  Alloc(&rt->e_slots, rt->max_gen) && Alloc(&rt->e_rows, rt->max_gen) && Alloc(&rt->bx, bmax * kSgIn) &&
                  Alloc(&rt->bt, bmax) && Alloc(&rt->by, bmax) && Alloc(&rt->bh1, bmax * kSgH1) &&
                  Alloc(&rt->bh2, bmax * kSgH2) && Alloc(&rt->bdy, bmax) && Alloc(&rt->bdh1, bmax * kSgH1) &&
                  Alloc(&rt->bdh2, bmax * kSgH2);
        if (!ok) return Fail(rt, NE_DEVICE_FAILURE);
    }
    if (rt->enc_t == nullptr && !Alloc(&rt->enc_t, rt->max_gen * kEncTStride)) return Fail(rt, NE_DEVICE_FAILURE);
    Free(rt->bidx);
    Free(rt->sg_coef);
    if (rt->train_exec) (void)hipGraphExecDestroy(rt->train_exec);
    if (rt->train_graph) (void)hipGraphDestroy(rt->train_graph);
    rt->train_exec = nullptr;
    rt->train_graph = nullptr;
    if (!Alloc(&rt->bidx, std::max<uint64_t>(uint64_t(c.train_steps) * c.batch_size, 1)) ||
        !Alloc(&rt->sg_coef, 2 + 2 * uint64_t(c.train_steps)) || (!rt->sg_coef1 && !Alloc(&rt->sg_coef1, 2)))
        return Fail(rt, NE_DEVICE_FAILURE);
    rt->screen = c;
    rt->screen_on = true;
    rt->screen_active = false;
    rt->sg_step = 0;
    rt->replay_count = 0;
    rt->replay_head = 0;
    rt->h_replay_f.assign(c.replay_capacity, 0.0f);
    LaunchSurrogateInit(rt->sg_params, rt->sg_m, rt->sg_v, c.seed, rt->stream);
    if (!Ok(hipMemcpyAsync(rt->d_probes, &rt->h_probes, sizeof(ProbeSet), hipMemcpyHostToDevice, rt->stream)))
        return Fail(rt, NE_DEVICE_FAILURE);
    return Finish(rt);
}

int ne_advance_generation_screened(ne_runtime* rt, uint32_t generation_seed, const ne_word_counts* counts,
                                   const ne_assembly_config* config, const ne_shard_release_history* history,
                                   ne_screen_report* report) {
    if (!rt || !counts) return Fail(rt, NE_INVALID_ARGUMENT);
    if (!rt->screen_on) return Fail(rt, NE_UNSUPPORTED);
    const ne_assembly_config cfg = config ? *config : DefaultAssembly();
    if (!ValidAssembly(cfg)) return Fail(rt, NE_INVALID_ASSEMBLY_CONFIG);
    ne_screen_report rep{};
    rt->times = ne_stage_times{};
    int st = NE_OK;
    if (!FitnessInputsHold(rt, counts, history)) {
        st = EvaluateFull(rt, counts, history);
        if (st != NE_OK) return st;
    }
    const uint64_t N = rt->cur_size, cols = rt->config.grid_column_count, rows = N / cols;
    const bool screened = rt->screen_active;
    const uint64_t B = screened ? rt->screen.brood_factor : 1;
    const uint64_t BN = B * N;
    if (BN > rt->max_gen) return Fail(rt, NE_INVALID_ASSEMBLY_CONFIG);
    // the cells' secrets, for the children (generation cur_index + 1)
    std::vector<uint32_t> built_o, built_s;
    const std::vector<uint32_t>* offs = &rt->h_offsets;
    const std::vector<uint32_t>* secs = &rt->h_secrets;
    if (!rt->csr_from_caller) {
        st = BuildShardCsrCore(counts, history, rt->cur_index + 1, rows, cols, rt->epi_rows, rt->catalog_count,
                               built_o, built_s);
        if (st != NE_OK) return Fail(rt, st);
        offs = &built_o;
        secs = &built_s;
    }
    if (offs->size() < N + 1) return Fail(rt, NE_INVALID_TRAINING_SHARD);

    // plan: child b*N + c belongs to cell c (its draws are keyed by b*N + c, so b = 0 is the plain plan)
    st = EnsureRankWeights(rt, cfg.rank_exponent, std::max<uint64_t>(N, 25));
    if (st != NE_OK) return st;
    PlanArgs pa{};
    pa.fitness = rt->cur_fit;
    pa.has_fitness = rt->cur_has;
    pa.cur_size = uint32_t(N);
    pa.child_count = uint32_t(BN);
    pa.cols = uint32_t(cols);
    pa.gen_index = uint32_t(rt->cur_index);
    pa.radius = uint32_t(std::min<uint64_t>(cfg.cellular_breeding_radius, 0xFFFFFFFFu));
    pa.generation_seed = generation_seed;
    pa.rank_weights = rt->rank_weights;
    pa.pairs = rt->pairs;
    pa.status = rt->status;
    EventBegin(rt, 2);
    LaunchPlan(pa, rt->stream);
    EventEnd(rt, 2);
    st = Finish(rt);
    if (st != NE_OK) return st;
    const float plan_ms = EventMs(rt, 2);
    rt->planned_children = BN;
    // screening: the brood's first phase only (every child's policy and probe-vocabulary rows); otherwise
    // the plain assembly (which also releases the parents as their children are done)
    st = screened ? AssembleBrood(rt, generation_seed, cfg, BN) : Assemble(rt, generation_seed, cfg, rt->action_count);
    if (st != NE_OK) return st;

    // fingerprints of the whole brood, then the surrogate's predictions (parent fitness per position
    // stays in cur_fit after the parents' slots are released)
    EventBegin(rt, 9);
    LaunchFingerprints(rt, rt->nxt_slots, uint32_t(BN), rt->pairs, rt->cur_fit, rt->fp);
    EventEnd(rt, 9);
    // the previous step's training (on the training stream, overlapping this step so far) before the
    // predictions
    if (JoinTraining(rt) != NE_OK) return NE_DEVICE_FAILURE;
    EventBegin(rt, 10);
    LaunchSurrogateForward(rt->sg_params, rt->fp, uint32_t(BN), rt->pred, nullptr, nullptr, rt->stream);
    EventEnd(rt, 10);
    // the audit: a Philox draw per brood member (keyed by member and parent generation), drawn while the
    // device assembles and fingerprints the brood
    std::vector<uint8_t> audited(BN, 0);
    for (uint64_t i = 0; i < BN; ++i) {
        const U4 w = Philox4x32(uint32_t(i), uint32_t(rt->cur_index), uint32_t(rt->cur_index >> 32), 0,
                                rt->screen.seed, kTagAudit);
        audited[i] = UniformOpenClosed(w.x) <= rt->screen.audit_fraction ? 1 : 0;
    }
    st = Finish(rt);
    if (st != NE_OK) return st;
    if (screened) rt->times.assemble_ms = EventMs(rt, 3);  // the brood's first phase
    if (ReadTraining(rt) != NE_OK) return NE_DEVICE_FAILURE;  // done by now: its loss and time
    rep.fingerprint_ms = EventMs(rt, 9);
    float sg_ms = EventMs(rt, 10) + rt->last_train_ms;
    std::vector<float> pred(BN);
    std::vector<uint32_t> brood_slots(BN);
    if (!Ok(hipMemcpy(pred.data(), rt->pred, BN * 4, hipMemcpyDeviceToHost)) ||
        !Ok(hipMemcpy(brood_slots.data(), rt->nxt_slots, BN * 4, hipMemcpyDeviceToHost)))
        return Fail(rt, NE_DEVICE_FAILURE);

    // selection: each cell's best-predicted child (lowest brood index on ties), then the audited members
    std::vector<uint32_t> rows_e;
    std::vector<int64_t> entry_of(BN, -1);
    rows_e.reserve(BN);
    for (uint64_t c = 0; c < N; ++c) {
        uint64_t best = c;
        for (uint64_t b = 1; b < B; ++b)
            if (pred[b * N + c] > pred[best]) best = b * N + c;
        entry_of[best] = int64_t(rows_e.size());
        rows_e.push_back(uint32_t(best));
    }
    std::vector<uint32_t> audit;
    for (uint64_t i = 0; i < BN; ++i) {
        if (!audited[i]) continue;
        audit.push_back(uint32_t(i));
        if (entry_of[i] < 0) {
            entry_of[i] = int64_t(rows_e.size());
            rows_e.push_back(uint32_t(i));
        }
    }
    const uint64_t E = rows_e.size();
    std::vector<uint32_t> e_slots(E), eo(E + 1, 0), es;
    for (uint64_t e = 0; e < E; ++e) {
        e_slots[e] = brood_slots[rows_e[e]];
        const uint64_t cell = rows_e[e] % N;
        es.insert(es.end(), secs->begin() + (*offs)[cell], secs->begin() + (*offs)[cell + 1]);
        eo[e + 1] = uint32_t(es.size());
    }
    if (!Ok(hipMemcpyAsync(rt->e_slots, e_slots.data(), E * 4, hipMemcpyHostToDevice, rt->stream)) ||
        !Ok(hipMemcpyAsync(rt->e_rows, rows_e.data(), E * 4, hipMemcpyHostToDevice, rt->stream)))
        return Fail(rt, NE_DEVICE_FAILURE);
    if (screened) {
        // complete the children that are evaluated, then every parent is free
        st = BreedRest(rt, generation_seed, cfg, rt->e_rows, E);
        if (st != NE_OK) return st;
        if (!Ok(hipMemsetAsync(rt->refcounts, 0, N * sizeof(uint32_t), rt->stream))) return Fail(rt, NE_DEVICE_FAILURE);
        LaunchCollectReleased(rt->cur_slots, uint32_t(N), rt->refcounts, rt->slot_state, rt->free_stack,
                              rt->free_count, uint32_t(rt->slot_count), rt->status, rt->stream);
    }
    // the evaluation's host preparation while the device completes the brood, then one wait (the
    // evaluation clears the status word, so the brood's status is read first)
    EvalSetup eval_setup;
    st = PrepareEvaluation(rt, counts->action_space_word_count, rt->e_slots, E, eo, es, &eval_setup);
    if (st != NE_OK) return st;
    st = Finish(rt);
    if (st != NE_OK) return st;
    if (screened) rt->times.assemble_ms += EventMs(rt, 13);
    const float assemble_ms = rt->times.assemble_ms, book_ms = rt->times.bookkeeping_ms;
    st = RunEvaluation(rt, eval_setup, rt->e_slots, E, rt->e_fit, rt->e_evals, rt->e_has, rt->cur_index + 1);
    if (st != NE_OK) return st;
    std::vector<float> e_fit(E);
    if (!Ok(hipMemcpy(e_fit.data(), rt->e_fit, E * 4, hipMemcpyDeviceToHost))) return Fail(rt, NE_DEVICE_FAILURE);

    // the audit's rank correlation, from predictions made before any training on these children
    std::vector<float> ap, af;
    for (const uint32_t r : audit) {
        ap.push_back(pred[r]);
        af.push_back(e_fit[size_t(entry_of[r])]);
    }
    const double rho = RankCorrelation(ap.data(), af.data(), ap.size());

    // every evaluated child joins the replay buffer (the last `capacity` when more), then training
    const uint64_t cap = rt->screen.replay_capacity;
    const uint64_t first = E > cap ? E - cap : 0;
    LaunchReplayAppend(rt->fp, rt->e_rows + first, rt->e_fit + first, uint32_t(E - first), rt->replay_x, rt->replay_f,
                       uint32_t(rt->replay_head), uint32_t(cap), rt->stream);
    for (uint64_t e = first; e < E; ++e) rt->h_replay_f[(rt->replay_head + e - first) % cap] = e_fit[e];
    rt->replay_head = (rt->replay_head + (E - first)) % cap;
    rt->replay_count = std::min<uint64_t>(rt->replay_count + (E - first), cap);
    // training starts on its own stream and overlaps the rest of this step and the next step's planning,
    // brood assembly and fingerprints; the next step's predictions wait for it
    const float loss = rt->last_loss;  // of the most recently completed training (the previous step's)
    st = StartTraining(rt);
    if (st != NE_OK) return st;

    // the new generation: the selected children; every other brood member goes back to the free stack
    std::vector<uint32_t> keep(BN, 0), sel_slots(N);
    for (uint64_t c = 0; c < N; ++c) {
        keep[rows_e[c]] = 1;
        sel_slots[c] = e_slots[c];
    }
    bool ok = Ok(hipMemcpyAsync(rt->refcounts, keep.data(), BN * 4, hipMemcpyHostToDevice, rt->stream));
    LaunchCollectReleased(rt->nxt_slots, uint32_t(BN), rt->refcounts, rt->slot_state, rt->free_stack, rt->free_count,
                          uint32_t(rt->slot_count), rt->status, rt->stream);
    LaunchFillU32(rt->nxt_slots, kInvalidSlot, BN, rt->stream);
    ok &= Ok(hipMemcpyAsync(rt->nxt_slots, sel_slots.data(), N * 4, hipMemcpyHostToDevice, rt->stream));
    ok &= Ok(hipMemcpyAsync(rt->nxt_fit, rt->e_fit, N * sizeof(float), hipMemcpyDeviceToDevice, rt->stream));
    ok &= Ok(hipMemcpyAsync(rt->nxt_evals, rt->e_evals, N * sizeof(uint32_t), hipMemcpyDeviceToDevice, rt->stream));
    ok &= Ok(hipMemcpyAsync(rt->nxt_has, rt->e_has, N, hipMemcpyDeviceToDevice, rt->stream));
    // the new generation's summary behind the copies, one wait for both
    LaunchSummary(rt->nxt_fit, rt->nxt_has, rt->nxt_slots, uint32_t(N), rt->summary, rt->status, rt->stream);
    if (!ok) return Fail(rt, NE_DEVICE_FAILURE);
    st = Finish(rt);
    if (st != NE_OK) return st;
    rt->next_size = N;
    const ne_stage_times eval_times = rt->times;
    SwapGenerations(rt);
    rt->cur_evaluated = true;
    RecordFitnessInputs(rt, counts->action_space_word_count, *offs, *secs, N);
    SummaryRecord sr{};
    if (!Ok(hipMemcpy(&sr, rt->summary, sizeof(sr), hipMemcpyDeviceToHost))) return Fail(rt, NE_DEVICE_FAILURE);
    rt->last_summary.best_fitness = sr.best_fitness;
    rt->last_summary.average_fitness = sr.average_fitness;
    rt->last_summary.best_index = sr.best_index;
    rt->last_summary.best_slot_index = sr.best_slot;
    rt->last_summary.generation_index = rt->cur_index;
    rt->last_summary.action_count = rt->action_count;
    rt->last_summary.population_size = N;
    rt->have_summary = true;

    rt->times = eval_times;
    rt->times.plan_ms = plan_ms;
    rt->times.assemble_ms = assemble_ms;
    rt->times.bookkeeping_ms = book_ms;
    rt->times.children = BN;
    rt->screen_active = rt->replay_count >= rt->screen.min_replay && std::isfinite(rho) &&
                        rho >= double(rt->screen.spearman_threshold);
    rep.screened = screened ? 1u : 0u;
    rep.next_screened = rt->screen_active ? 1u : 0u;
    rep.brood = uint32_t(BN);
    rep.evaluated = uint32_t(E);
    rep.audited = uint32_t(audit.size());
    rep.audit_spearman = float(rho);
    rep.train_loss = loss;
    rep.surrogate_ms = sg_ms;
    rep.best_fitness = sr.best_fitness;
    rep.average_fitness = sr.average_fitness;
    rep.replay_count = rt->replay_count;
    if (report) *report = rep;
    rt->last_status = NE_OK;
    return NE_OK;
}

int ne_fingerprint(ne_runtime* rt, const uint32_t* slot_indices, const float* parent_fitness, uint64_t count,
                   float* out) {
    if (!rt || !slot_indices || !parent_fitness || !out || count == 0) return Fail(rt, NE_INVALID_ARGUMENT);
    if (!rt->screen_on) return Fail(rt, NE_UNSUPPORTED);
    if (count > rt->max_gen) return Fail(rt, NE_INVALID_ARGUMENT);
    for (uint64_t i = 0; i < count; ++i)
        if (slot_indices[i] >= rt->slot_count) return Fail(rt, NE_INVALID_ARGUMENT);
    uint32_t* d_pairs = nullptr;
    float* d_pf = nullptr;
    std::vector<uint32_t> pairs(2 * count);
    for (uint64_t i = 0; i < 2 * count; ++i) pairs[i] = uint32_t(i);
    if (!Alloc(&d_pairs, 2 * count) || !Alloc(&d_pf, 2 * count)) {
        Free(d_pairs);
        return Fail(rt, NE_DEVICE_FAILURE);
    }
    bool ok = Ok(hipMemcpyAsync(rt->e_slots, slot_indices, count * 4, hipMemcpyHostToDevice, rt->stream));
    ok &= Ok(hipMemcpyAsync(d_pairs, pairs.data(), 2 * count * 4, hipMemcpyHostToDevice, rt->stream));
    ok &= Ok(hipMemcpyAsync(d_pf, parent_fitness, 2 * count * 4, hipMemcpyHostToDevice, rt->stream));
    if (ok) LaunchFingerprints(rt, rt->e_slots, uint32_t(count), d_pairs, d_pf, rt->fp);
    int st = ok ? Finish(rt) : Fail(rt, NE_DEVICE_FAILURE);
    if (st == NE_OK && !Ok(hipMemcpy(out, rt->fp, count * kSgIn * 4, hipMemcpyDeviceToHost))) st = Fail(rt, NE_DEVICE_FAILURE);
    Free(d_pairs);
    Free(d_pf);
    return st;
}

int ne_get_probe_states(ne_runtime* rt, uint32_t* dst, uint64_t capacity, uint64_t* count) {
    if (!rt || !dst) return Fail(rt, NE_INVALID_ARGUMENT);
    if (!rt->screen_on) return Fail(rt, NE_UNSUPPORTED);
    constexpr uint64_t n = uint64_t(kProbeSecrets) * kProbeDepths;
    if (count) *count = n;
    if (capacity < n) return Fail(rt, NE_INVALID_ARGUMENT);
    std::memcpy(dst, rt->h_probe_words, n * sizeof(uint32_t));
    return NE_OK;
}

int ne_surrogate_predict(ne_runtime* rt, const float* x, uint64_t count, float* y) {
    if (!rt || !x || !y || count == 0 || count > 0xFFFFFFFFull / kSgIn) return Fail(rt, NE_INVALID_ARGUMENT);
    if (!rt->screen_on) return Fail(rt, NE_UNSUPPORTED);
    float *dx = nullptr, *dy = nullptr;
    if (!Alloc(&dx, count * kSgIn) || !Alloc(&dy, count)) {
        Free(dx);
        return Fail(rt, NE_DEVICE_FAILURE);
    }
    int st = (JoinTraining(rt) == NE_OK && Ok(hipMemcpyAsync(dx, x, count * kSgIn * 4, hipMemcpyHostToDevice,
                                                                rt->stream)))
                 ? NE_OK
                 : Fail(rt, NE_DEVICE_FAILURE);
    if (st == NE_OK) {
        LaunchSurrogateForward(rt->sg_params, dx, uint32_t(count), dy, nullptr, nullptr, rt->stream);
        st = Finish(rt);
    }
    if (st == NE_OK && !Ok(hipMemcpy(y, dy, count * 4, hipMemcpyDeviceToHost))) st = Fail(rt, NE_DEVICE_FAILURE);
    Free(dx);
    Free(dy);
    return st;
}

int ne_surrogate_train_step(ne_runtime* rt, const float* x, const float* t, uint64_t count) {
    if (!rt || !x || !t || count == 0) return Fail(rt, NE_INVALID_ARGUMENT);
    if (!rt->screen_on) return Fail(rt, NE_UNSUPPORTED);
    if (count > std::max<uint64_t>(rt->screen.batch_size, 1024)) return Fail(rt, NE_INVALID_ARGUMENT);
    if (ReadTraining(rt) != NE_OK) return NE_DEVICE_FAILURE;  // its batch buffers are reused here
    float coef[2];
    AdamCoef(++rt->sg_step, coef);
    if (!Ok(hipMemcpyAsync(rt->bx, x, count * kSgIn * 4, hipMemcpyHostToDevice, rt->stream)) ||
        !Ok(hipMemcpyAsync(rt->bt, t, count * 4, hipMemcpyHostToDevice, rt->stream)) ||
        !Ok(hipMemcpyAsync(rt->sg_coef1, coef, sizeof(coef), hipMemcpyHostToDevice, rt->stream)))
        return Fail(rt, NE_DEVICE_FAILURE);
    SurrogateStep(rt, rt->bx, rt->bt, uint32_t(count), rt->sg_coef1, rt->stream);
    return Finish(rt);
}

int ne_surrogate_get_params(ne_runtime* rt, float* dst, uint64_t capacity, uint64_t* count) {
    if (!rt || !dst) return Fail(rt, NE_INVALID_ARGUMENT);
    if (!rt->screen_on) return Fail(rt, NE_UNSUPPORTED);
    if (count) *count = kSgParams;
    if (capacity < kSgParams) return Fail(rt, NE_INVALID_ARGUMENT);
    if (JoinTraining(rt) != NE_OK || Sync(rt) != NE_OK ||
        !Ok(hipMemcpy(dst, rt->sg_params, kSgParams * 4, hipMemcpyDeviceToHost)))
        return Fail(rt, NE_DEVICE_FAILURE);
    return NE_OK;
}

int ne_surrogate_set_params(ne_runtime* rt, const float* src, uint64_t count) {
    if (!rt || !src || count != kSgParams) return Fail(rt, NE_INVALID_ARGUMENT);
    if (!rt->screen_on) return Fail(rt, NE_UNSUPPORTED);
    bool ok = JoinTraining(rt) == NE_OK;
    ok &= Ok(hipMemcpyAsync(rt->sg_params, src, kSgParams * 4, hipMemcpyHostToDevice, rt->stream));
    ok &= Ok(hipMemsetAsync(rt->sg_m, 0, kSgParams * 4, rt->stream));
    ok &= Ok(hipMemsetAsync(rt->sg_v, 0, kSgParams * 4, rt->stream));
    if (!ok) return Fail(rt, NE_DEVICE_FAILURE);
    rt->sg_step = 0;
    return Finish(rt);
}

double ne_rank_correlation(const float* a, const float* b, uint64_t n) {
    if (!a || !b) return std::nan("");
    return RankCorrelation(a, b, n);
}

}  // extern "C"

// ---- checkpoint and resume (device side; the file format is in ne_checkpoint.cpp) ---------------

// Copies the genomes of `slots` to dst[n][stride] through two pinned staging buffers of up to 64 MiB:
// the DMA into one overlaps the host copy out of the other.
static int DownloadGenomes(ne_runtime* rt, const uint32_t* slots, uint64_t n, uint8_t* dst) {
    if (n == 0) return NE_OK;
    const uint64_t S = rt->stride;
    const uint64_t per = std::max<uint64_t>(1, std::min<uint64_t>(n, (64ull << 20) / S));
    uint8_t* pin[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    uint64_t first[2] = {0, 0}, cnt[2] = {0, 0};
    bool ok = true;
    for (int i = 0; i < 2; ++i) {
        ok = ok && Ok(hipHostMalloc(reinterpret_cast<void**>(&pin[i]), per * S, hipHostMallocDefault));
        ok = ok && Ok(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
    }
    auto drain = [&](int i) {
        if (!ok || cnt[i] == 0) return;
        ok = Ok(hipEventSynchronize(done[i]));
        if (ok) std::memcpy(dst + first[i] * S, pin[i], cnt[i] * S);
        cnt[i] = 0;
    };
    for (uint64_t b = 0, c = 0; ok && b < n; b += per, ++c) {
        const int i = int(c & 1);
        drain(i);
        const uint64_t k = std::min(per, n - b);
        for (uint64_t j = 0; ok && j < k; ++j)
            ok = Ok(hipMemcpyAsync(pin[i] + j * S, rt->slab + uint64_t(slots[b + j]) * S, S, hipMemcpyDeviceToHost,
                                   rt->stream));
        ok = ok && Ok(hipEventRecord(done[i], rt->stream));
        first[i] = b;
        cnt[i] = k;
    }
    drain(0);
    drain(1);
    (void)hipStreamSynchronize(rt->stream);  // nothing may still write the staging buffers
    for (int i = 0; i < 2; ++i) {
        if (pin[i]) (void)hipHostFree(pin[i]);
        if (done[i]) (void)hipEventDestroy(done[i]);
    }
    return ok ? NE_OK : Fail(rt, NE_DEVICE_FAILURE);
}

static void StoreShape(uint8_t* out, uint64_t rows, uint64_t cols) {
    ne_ckpt::StoreU64(out, rows);
    ne_ckpt::StoreU64(out + 8, cols);
    ne_ckpt::StoreU64(out + 16, rows * cols);
}

extern "C" {

int ne_create_prebreeding_checkpoint(ne_runtime* rt, uint32_t generation_seed, const ne_word_counts* counts,
                                     const ne_assembly_config* config, const ne_pending_injection* injection,
                                     const uint8_t* catalog_letters, uint64_t catalog_word_count, uint32_t verbose,
                                     const ne_callbacks* callbacks, const ne_shard_release_history* history,
                                     ne_checkpoint** out_checkpoint) {
    if (!rt || !counts || !out_checkpoint) return Fail(rt, NE_INVALID_ARGUMENT);
    struct VerboseScope {
        ne_runtime* rt;
        ~VerboseScope() { rt->verbose = false; }
    } verbose_scope{rt};
    rt->verbose = verbose != 0;
    *out_checkpoint = nullptr;
    if (history && (history->release_count > ne_ckpt::kReleaseCapacity ||
                    (history->release_count && !history->release_generations)))
        return Fail(rt, NE_INVALID_ARGUMENT);
    const ne_assembly_config cfg = config ? *config : DefaultAssembly();
    if (!ValidAssembly(cfg)) return Fail(rt, NE_INVALID_ASSEMBLY_CONFIG);
    ne_pending_injection inj{};
    if (injection) inj = *injection;
    uint64_t parent_a = 0;
    int st = PrepareBoundary(rt, generation_seed, counts, cfg, inj, callbacks, history, &parent_a);
    if (st != NE_OK) return st;

    // TryDownloadBoundaryCheckpointPayload (slab_runtime.cu:1260-1356)
    ne_checkpoint* c = new (std::nothrow) ne_checkpoint();
    if (!c) return Fail(rt, NE_DEVICE_FAILURE);
    const uint64_t n = rt->cur_size, children = rt->planned_children, cols = rt->config.grid_column_count;
    try {
        c->generation_seed = generation_seed;
        ne_ckpt::EncodeWordCounts(*counts, c->word_counts);
        if (history && history->release_count) {
            c->release_generations.assign(history->release_generations,
                                          history->release_generations + history->release_count);
            c->release_p99.assign(history->release_count, 0.0f);
            if (history->release_fitness_p99)
                c->release_p99.assign(history->release_fitness_p99,
                                      history->release_fitness_p99 + history->release_count);
        }
        c->identity_hash = ne_ckpt::IdentityHash(c->word_counts, history != nullptr, c->release_generations,
                                                 c->release_p99, catalog_letters, catalog_word_count);
        ne_ckpt::EncodeAssembly(cfg, c->assembly);
        ne_ckpt::EncodeInjection(inj, c->injection);
        const uint64_t rc[6] = {rt->slab_bytes, rt->generation_byte_budget, rt->config.host_spillover_byte_budget,
                                rt->action_count, rt->max_gen, cols};
        std::memcpy(c->runtime_config, rc, sizeof(rc));
        const uint64_t layout[4] = {rt->action_count, rt->stride, rt->slot_count, rt->slab_bytes};
        std::memcpy(c->slab_layout, layout, sizeof(layout));
        StoreShape(c->shapes[0], n / cols, cols);
        StoreShape(c->shapes[1], children / cols, cols);
        StoreShape(c->shapes[2], rt->epi_rows, cols);
        c->generation_index = rt->cur_index;
        c->slots.resize(n);
        c->fitness.resize(n);
        c->evals.resize(n);
        c->has.resize(n);
        c->pairs.resize(2 * children);
    } catch (const std::bad_alloc&) {
        delete c;
        return Fail(rt, NE_DEVICE_FAILURE);
    }
    bool ok = Ok(hipMemcpyAsync(c->slots.data(), rt->cur_slots, n * 4, hipMemcpyDeviceToHost, rt->stream));
    ok = ok && Ok(hipMemcpyAsync(c->fitness.data(), rt->cur_fit, n * 4, hipMemcpyDeviceToHost, rt->stream));
    ok = ok && Ok(hipMemcpyAsync(c->evals.data(), rt->cur_evals, n * 4, hipMemcpyDeviceToHost, rt->stream));
    ok = ok && Ok(hipMemcpyAsync(c->has.data(), rt->cur_has, n, hipMemcpyDeviceToHost, rt->stream));
    ok = ok && Ok(hipMemcpyAsync(c->pairs.data(), rt->pairs, children * 8, hipMemcpyDeviceToHost, rt->stream));
    ok = ok && Ok(hipStreamSynchronize(rt->stream));
    if (!ok) {
        delete c;
        return Fail(rt, NE_DEVICE_FAILURE);
    }
    // one record per organism holding a slot, in organism order
    std::vector<uint32_t> live_slots;
    for (uint64_t i = 0; i < n; ++i) {
        if (c->slots[i] == kInvalidSlot) continue;
        if (c->slots[i] >= rt->slot_count) {
            delete c;
            return Fail(rt, NE_INVALID_SLAB);
        }
        c->organisms.push_back(uint32_t(i));
        live_slots.push_back(c->slots[i]);
    }
    c->genome_bytes = rt->stride;
    try {
        c->genomes.resize(live_slots.size() * rt->stride);
    } catch (const std::bad_alloc&) {
        delete c;
        return Fail(rt, NE_DEVICE_FAILURE);
    }
    st = DownloadGenomes(rt, live_slots.data(), live_slots.size(), c->genomes.data());
    if (st != NE_OK) {
        delete c;
        return st;
    }
    c->checksum = ne_ckpt::Checksum(*c);
    *out_checkpoint = c;
    rt->last_status = NE_OK;
    return NE_OK;
}

int ne_restore_prebreeding_checkpoint(const ne_checkpoint* c, void* stream, ne_runtime** out_runtime) {
    if (!c || !out_runtime) return NE_INVALID_ARGUMENT;
    *out_runtime = nullptr;
    // the validation of TryRestorePrebreedingCheckpointToDevice (slab_runtime.cu:1734-1742, 1790-1813)
    if (c->schema_version != ne_ckpt::kSchemaVersion || c->genome_layout_version != ne_ckpt::kGenomeLayoutVersion ||
        c->resume_phase != ne_ckpt::kPhasePreRecombination || c->checksum != ne_ckpt::Checksum(*c))
        return NE_CHECKPOINT_INVALID;
    const uint64_t A = c->ActionCount(), S = c->Stride(), slots = c->SlotCount();
    if (A == 0 || slots == 0 || slots >= 0xFFFFFFFFull || S != SlotStrideBytes(A) || c->SlabBytes() / S != slots ||
        c->genome_bytes != S)
        return NE_CHECKPOINT_INVALID;
    const uint64_t n = c->slots.size(), children = c->ChildCount(), cols = c->RuntimeConfigField(5);
    const uint64_t epi_rows = c->ShapeField(2, 0), L = c->Records();
    if (n == 0 || children == 0 || cols == 0 || n % cols != 0 || children % cols != 0 || epi_rows == 0 ||
        L > slots || n > 0xFFFFFFFFull)
        return NE_CHECKPOINT_INVALID;
    std::vector<uint8_t> restored(n, 0);
    std::vector<uint32_t> cur_slots(n, kInvalidSlot);
    for (uint64_t k = 0; k < L; ++k) {
        const uint32_t org = c->organisms[k];
        if (org >= n || restored[org] || c->slots[org] == kInvalidSlot) return NE_CHECKPOINT_INVALID;
        restored[org] = 1;
        cur_slots[org] = uint32_t(k);  // TryAllocateSlabSlot on a fresh slab hands out slots 0, 1, 2, ...
    }
    for (uint64_t i = 0; i < 2 * children; ++i)
        if (c->pairs[i] >= n || !restored[c->pairs[i]]) return NE_CHECKPOINT_INVALID;

    ne_runtime_config rc{};
    rc.genotype_slab_byte_budget = slots * S;
    rc.generation_byte_budget = c->RuntimeConfigField(1);
    rc.host_spillover_byte_budget = c->RuntimeConfigField(2);
    rc.action_count = A;
    rc.grid_column_count = cols;
    rc.stream = stream;
    const uint64_t max_gen = std::max(n, children);
    rc.population_size_ceiling = max_gen;
    ne_runtime* rt = nullptr;
    int st = CreateCore(rc, slots, max_gen, epi_rows, &rt);
    if (st != NE_OK) return st;
    LaunchFillU32(rt->slot_state, 1u, L, rt->stream);
    LaunchPopSlots(rt->free_count, uint32_t(L), rt->stream);
    bool ok = L == 0 || Ok(hipMemcpyAsync(rt->slab, c->genomes.data(), L * S, hipMemcpyHostToDevice, rt->stream));
    ok = ok && Ok(hipMemcpyAsync(rt->cur_slots, cur_slots.data(), n * 4, hipMemcpyHostToDevice, rt->stream));
    ok = ok && Ok(hipMemcpyAsync(rt->cur_fit, c->fitness.data(), n * 4, hipMemcpyHostToDevice, rt->stream));
    ok = ok && Ok(hipMemcpyAsync(rt->cur_evals, c->evals.data(), n * 4, hipMemcpyHostToDevice, rt->stream));
    ok = ok && Ok(hipMemcpyAsync(rt->cur_has, c->has.data(), n, hipMemcpyHostToDevice, rt->stream));
    ok = ok && Ok(hipMemcpyAsync(rt->pairs, c->pairs.data(), children * 8, hipMemcpyHostToDevice, rt->stream));
    ok = ok && Sync(rt) == NE_OK;
    if (!ok) {
        ne_destroy(rt);
        return NE_DEVICE_FAILURE;
    }
    rt->cur_index = c->generation_index;
    rt->cur_size = n;
    rt->planned_children = children;
    rt->cur_evaluated = true;
    *out_runtime = rt;
    return NE_OK;
}

int ne_resume_generation_from_checkpoint(ne_runtime* rt, const ne_checkpoint* c, uint32_t verbose) {
    if (!rt || !c) return Fail(rt, NE_INVALID_ARGUMENT);
    struct VerboseScope {
        ne_runtime* rt;
        ~VerboseScope() { rt->verbose = false; }
    } verbose_scope{rt};
    rt->verbose = verbose != 0;
    Progress(rt, "Generation %llu: resuming from the pre-recombination checkpoint of generation %llu",
             (unsigned long long)(rt->cur_index + 1), (unsigned long long)rt->cur_index);
    if (c->Stride() != rt->stride || c->ActionCount() != rt->action_count || c->generation_index != rt->cur_index ||
        c->slots.size() != rt->cur_size || c->ChildCount() != rt->planned_children || rt->planned_children == 0)
        return Fail(rt, NE_INVALID_GENERATION);
    const ne_assembly_config cfg = ne_ckpt::DecodeAssembly(c->assembly);
    if (!ValidAssembly(cfg)) return Fail(rt, NE_INVALID_ASSEMBLY_CONFIG);
    const ne_pending_injection inj = ne_ckpt::DecodeInjection(c->injection);
    // SlabDeviceAssemblyConfig.parent_action_count of the resume (slab_runtime.cu:1871-1874)
    const uint64_t parent_a = inj.enabled ? inj.first_catalog_word_index : c->ActionCount();
    if (inj.enabled && (inj.injection_count == 0 || parent_a + inj.injection_count != rt->action_count ||
                        parent_a + inj.injection_count > rt->catalog_count))
        return Fail(rt, NE_OUTPUT_EMBEDDING_INJECTION_FAILED);
    return AssemblePrepared(rt, c->generation_seed, cfg, parent_a, inj);
}

}  // extern "C"

// ---- generation telemetry and the curriculum controller (host side) ----------------------------

constexpr uint64_t kSampleOrganisms = 256;  // kDefaultSampleOrganismCount (genetic_convergence.hpp:18)
constexpr uint64_t kSampleWeights = 8192;   // kDefaultSampleWeightCount
constexpr uint64_t kSamplePairs = 512;      // kDefaultPairCount

// MixSamplingSeed (genetic_convergence.hpp:48-57): a hash-combine of the generation into the run seed,
// then the SplitMix64 finaliser.
static uint64_t MixSamplingSeed(uint32_t run_seed, uint64_t generation) {
    uint64_t v = 0x9E3779B97F4A7C15ull ^ uint64_t(run_seed);
    v ^= generation + 0xBF58476D1CE4E5B9ull + (v << 6) + (v >> 2);
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}

// SampleUniqueIndices (genetic_convergence.hpp:59-83): Floyd's sampling of `requested` distinct indices
// below `universe` in draw order (all of them, in order, when requested >= universe).
static std::vector<uint64_t> SampleUnique(uint64_t universe, uint64_t requested, std::mt19937_64& rng) {
    const uint64_t k = std::min(requested, universe);
    std::vector<uint64_t> out;
    out.reserve(k);
    if (k == universe) {
        out.resize(k);
        std::iota(out.begin(), out.end(), uint64_t{0});
        return out;
    }
    std::unordered_set<uint64_t> taken;
    taken.reserve(2 * k);
    for (uint64_t cursor = universe - k; cursor < universe; ++cursor) {
        std::uniform_int_distribution<uint64_t> draw(0, cursor);
        const uint64_t c = draw(rng);
        const uint64_t v = taken.count(c) ? cursor : c;
        taken.insert(v);
        out.push_back(v);
    }
    return out;
}

struct SamplePlan {
    std::vector<uint64_t> organisms, weights;
    std::vector<uint32_t> pairs;  // [pair][2] indices into `organisms`
};

// MakeSamplePlan (genetic_convergence.hpp:87-117); the draws come from std::mt19937_64 and
// std::uniform_int_distribution, so the plan is the reference's under the same C++ standard library.
static SamplePlan MakeSamplePlan(uint64_t population, uint64_t scalars, uint32_t run_seed, uint64_t generation) {
    SamplePlan p;
    if (population == 0 || scalars == 0) return p;
    std::mt19937_64 rng(MixSamplingSeed(run_seed, generation));
    p.organisms = SampleUnique(population, kSampleOrganisms, rng);
    p.weights = SampleUnique(scalars, kSampleWeights, rng);
    if (p.organisms.size() < 2) return p;
    std::uniform_int_distribution<uint64_t> first(0, p.organisms.size() - 1), second(0, p.organisms.size() - 2);
    p.pairs.reserve(2 * kSamplePairs);
    for (uint64_t i = 0; i < kSamplePairs; ++i) {
        const uint64_t a = first(rng);
        uint64_t b = second(rng);
        if (b >= a) ++b;
        p.pairs.push_back(uint32_t(a));
        p.pairs.push_back(uint32_t(b));
    }
    return p;
}

extern "C" {

int ne_convergence_sample_plan(uint64_t population, uint64_t trainable_scalars, uint32_t run_seed,
                               uint64_t generation, uint32_t* organisms, uint32_t* weights, uint32_t* pairs,
                               uint64_t counts[3]) {
    if (!counts || trainable_scalars > 0xFFFFFFFFull || population > 0xFFFFFFFFull) return NE_INVALID_ARGUMENT;
    const SamplePlan p = MakeSamplePlan(population, trainable_scalars, run_seed, generation);
    counts[0] = p.organisms.size();
    counts[1] = p.weights.size();
    counts[2] = p.pairs.size() / 2;
    if (organisms)
        for (size_t i = 0; i < p.organisms.size(); ++i) organisms[i] = uint32_t(p.organisms[i]);
    if (weights)
        for (size_t i = 0; i < p.weights.size(); ++i) weights[i] = uint32_t(p.weights[i]);
    if (pairs) std::copy(p.pairs.begin(), p.pairs.end(), pairs);
    return NE_OK;
}

int ne_collect_fitness_telemetry(ne_runtime* rt, const ne_word_counts* counts, uint32_t collect_breeding_metrics,
                                 ne_fitness_telemetry* out) {
    if (!rt || !counts || !out) return Fail(rt, NE_INVALID_ARGUMENT);
    const uint64_t n = rt->cur_size;
    if (n == 0) return Fail(rt, NE_INVALID_GENERATION);
    if (n > kFitnessStatsMax) return Fail(rt, NE_UNSUPPORTED);
    if (collect_breeding_metrics && rt->planned_children == 0) return Fail(rt, NE_INVALID_ASSEMBLY_PLAN);
    if (!rt->d_fstats && !Alloc(&rt->d_fstats, 1)) return Fail(rt, NE_DEVICE_FAILURE);
    uint32_t pow2 = 2;
    while (pow2 < n) pow2 <<= 1;
    // nearest-rank percentile (run_genetic_algorithm.cu:1127-1132)
    auto rank_index = [n](double percentile) {
        const uint64_t rank = uint64_t(std::ceil((percentile / 100.0) * double(n)));
        const uint64_t index = rank == 0 ? 0 : rank - 1;
        return uint32_t(index < n ? index : n - 1);
    };
    const uint32_t* refcounts = nullptr;
    if (collect_breeding_metrics) {
        if (!Ok(hipMemsetAsync(rt->refcounts, 0, rt->max_gen * sizeof(uint32_t), rt->stream)))
            return Fail(rt, NE_DEVICE_FAILURE);
        LaunchRefcount(rt->pairs, uint32_t(rt->planned_children), uint32_t(n), rt->refcounts, rt->status, rt->stream);
        refcounts = rt->refcounts;
    }
    LaunchFitnessStats(rt->cur_fit, rt->cur_has, uint32_t(n), pow2, rank_index(90.0), rank_index(99.0), refcounts,
                       rt->d_fstats, rt->stream);
    FitnessStatsRecord r{};
    if (!Ok(hipMemcpyAsync(&r, rt->d_fstats, sizeof(r), hipMemcpyDeviceToHost, rt->stream)))
        return Fail(rt, NE_DEVICE_FAILURE);
    const int st = Finish(rt);
    if (st != NE_OK) return st;
    if (r.missing != 0) return Fail(rt, NE_POPULATION_NOT_EVALUATED);
    ne_fitness_telemetry t{};
    t.generation = rt->cur_index;
    t.population_size = n;
    t.training_word_count = counts->training_word_count;
    t.fitness_min = r.min;
    t.fitness_mean = float(r.sum / double(n));
    t.fitness_median = n % 2 == 0 ? (r.median_low + r.median_high) * 0.5f : r.median_high;
    t.fitness_p90 = r.p90;
    t.fitness_p99 = r.p99;
    t.fitness_max = r.max;
    t.fitness_stddev = float(std::sqrt(r.squared_deltas / double(n)));
    t.distinct_fitness_count = r.distinct;
    if (collect_breeding_metrics) {
        t.has_breeding_metrics = 1;
        t.parent_childless_count = r.childless;
        t.parent_one_child_count = r.one_child;
        t.parent_multiple_children_count = r.multiple_children;
    }
    *out = t;
    return NE_OK;
}

int ne_collect_convergence_telemetry(ne_runtime* rt, uint32_t run_seed, ne_convergence_telemetry* out) {
    if (!rt || !out) return Fail(rt, NE_INVALID_ARGUMENT);
    const auto start = std::chrono::steady_clock::now();
    const uint64_t n = rt->cur_size;
    if (n == 0) return Fail(rt, NE_INVALID_GENERATION);
    const uint64_t scalars = GenomeValueCount(rt->action_count);
    const SamplePlan plan = MakeSamplePlan(n, scalars, run_seed, rt->cur_index);
    if (plan.organisms.empty() || plan.weights.empty()) return Fail(rt, NE_INVALID_GENERATION);
    const uint32_t n_org = uint32_t(plan.organisms.size()), n_w = uint32_t(plan.weights.size());
    const uint32_t n_pairs = uint32_t(plan.pairs.size() / 2);
    std::vector<uint32_t> slots(n);
    if (!Ok(hipMemcpyAsync(slots.data(), rt->cur_slots, n * 4, hipMemcpyDeviceToHost, rt->stream)) ||
        !Ok(hipStreamSynchronize(rt->stream)))
        return Fail(rt, NE_DEVICE_FAILURE);
    // one host block: sampled slots, scalar indices, pairs
    std::vector<uint32_t> index(kSampleOrganisms + kSampleWeights + 2 * kSamplePairs, 0u);
    for (uint32_t i = 0; i < n_org; ++i) {
        const uint32_t slot = slots[plan.organisms[i]];
        if (slot == kInvalidSlot || slot >= rt->slot_count) return Fail(rt, NE_INVALID_SLAB);
        index[i] = slot;
    }
    for (uint32_t i = 0; i < n_w; ++i) index[kSampleOrganisms + i] = uint32_t(plan.weights[i]);
    std::copy(plan.pairs.begin(), plan.pairs.end(), index.begin() + kSampleOrganisms + kSampleWeights);
    if (!rt->cv_index) {
        bool ok = Alloc(&rt->cv_index, index.size()) && Alloc(&rt->cv_by_org, kSampleOrganisms * kSampleWeights) &&
                  Alloc(&rt->cv_by_weight, kSampleOrganisms * kSampleWeights) &&
                  Alloc(&rt->cv_centroid, kSampleWeights) && Alloc(&rt->cv_squared, kSampleOrganisms + kSamplePairs);
        if (!ok) return Fail(rt, NE_DEVICE_FAILURE);
    }
    if (!Ok(hipMemcpyAsync(rt->cv_index, index.data(), index.size() * 4, hipMemcpyHostToDevice, rt->stream)))
        return Fail(rt, NE_DEVICE_FAILURE);
    const float inverse = 1.0f / float(n_org);
    LaunchConvergence(rt->slab, rt->stride, rt->cv_index, n_org, rt->cv_index + kSampleOrganisms, n_w, scalars,
                      rt->cv_index + kSampleOrganisms + kSampleWeights, n_pairs, inverse, rt->cv_by_org,
                      rt->cv_by_weight, rt->cv_centroid, rt->cv_squared, rt->stream);
    std::vector<float> squared(n_org + n_pairs);
    if (!Ok(hipMemcpyAsync(squared.data(), rt->cv_squared, squared.size() * 4, hipMemcpyDeviceToHost, rt->stream)))
        return Fail(rt, NE_DEVICE_FAILURE);
    const int st = Finish(rt);
    if (st != NE_OK) return st;
    // the rest of TryComputeMetrics (genetic_convergence.hpp:146-191): square roots, float sums in order
    ne_convergence_telemetry t{};
    t.generation = rt->cur_index;
    t.sample_organisms = n_org;
    t.sample_weights = n_w;
    t.pair_count = n_pairs;
    float sum = 0.0f;
    t.centroid_distance_min = FLT_MAX;
    for (uint32_t o = 0; o < n_org; ++o) {
        const float d = std::sqrt(squared[o]);
        sum += d;
        t.centroid_distance_min = std::min(t.centroid_distance_min, d);
        t.centroid_distance_max = std::max(t.centroid_distance_max, d);
    }
    t.centroid_distance_mean = sum * inverse;
    if (n_pairs == 0) {
        t.pairwise_distance_min = 0.0f;
    } else {
        sum = 0.0f;
        t.pairwise_distance_min = FLT_MAX;
        for (uint32_t p = 0; p < n_pairs; ++p) {
            const float d = std::sqrt(squared[n_org + p]);
            sum += d;
            t.pairwise_distance_min = std::min(t.pairwise_distance_min, d);
            t.pairwise_distance_max = std::max(t.pairwise_distance_max, d);
        }
        t.pairwise_distance_mean = sum / float(n_pairs);
    }
    t.elapsed_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - start).count();
    *out = t;
    return NE_OK;
}

int ne_maybe_release_training_shard(ne_runtime* rt, const ne_word_counts* counts,
                                    const ne_shard_release_policy* policy, uint64_t catalog_word_count,
                                    uint64_t* release_generations, float* release_fitness_p99,
                                    uint64_t* release_count, uint64_t release_capacity,
                                    ne_pending_injection* injection, uint64_t* next_word_count,
                                    uint32_t* released) {
    if (!rt || !counts || !policy || !release_count || !injection || !next_word_count || !released)
        return Fail(rt, NE_INVALID_ARGUMENT);
    *injection = ne_pending_injection{};
    *next_word_count = counts->action_space_word_count;
    *released = 0;
    if (*release_count == 0 || !release_generations || !release_fitness_p99)
        return Fail(rt, NE_INVALID_TRAINING_SHARD);  // no foundation release record
    if (counts->action_space_word_count >= catalog_word_count || counts->training_word_schedule.word_count_step == 0)
        return NE_OK;
    const uint64_t latest = *release_count - 1;
    const uint64_t release_generation = rt->cur_index + 1;
    const uint64_t previous = release_generations[latest];
    const uint64_t since = release_generation >= previous ? release_generation - previous : 0;
    if (*release_count == 1 && release_generation < policy->first_new_release_generation) return NE_OK;
    if (since < policy->min_gap_generations) return NE_OK;
    ne_fitness_telemetry fitness{};
    int st = ne_collect_fitness_telemetry(rt, counts, 0, &fitness);
    if (st != NE_OK) return st;
    ne_convergence_telemetry convergence{};
    st = ne_collect_convergence_telemetry(rt, policy->run_seed, &convergence);
    if (st != NE_OK) return st;
    const bool centroid_triggered = convergence.centroid_distance_mean <= policy->centroid_distance_threshold;
    const bool fitness_triggered = fitness.fitness_p99 > policy->fitness_p99_threshold;
    if (!centroid_triggered || !fitness_triggered) return NE_OK;
    const uint64_t first = counts->action_space_word_count;
    const uint64_t count = std::min(catalog_word_count - first, counts->training_word_schedule.word_count_step);
    injection->enabled = 1;
    injection->first_catalog_word_index = first;
    injection->injection_count = count;
    *next_word_count = first + count;
    // TryRecordTrainingDataShardRelease (training_data.hpp:119-133)
    if (*release_count >= std::min<uint64_t>(release_capacity, ne_ckpt::kReleaseCapacity) ||
        release_generation <= release_generations[latest])
        return Fail(rt, NE_INVALID_TRAINING_SHARD);
    release_generations[*release_count] = release_generation;
    release_fitness_p99[*release_count] = fitness.fitness_p99;
    ++*release_count;
    *released = 1;
    return NE_OK;
}

}  // extern "C"
