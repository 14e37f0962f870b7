// ne_model.hip -- single-model inference: one genome answering next-guess requests for played boards,
// one workgroup per board, many boards per launch.
//
// Every neuron is one lane's sequential sum in the reference's order (bias first, inputs in index order,
// separate multiply and add), and each action's score is summed fixed features first, then tail values,
// so the guesses and scores are those of the reference's fp32 policy bit for bit
// (src/inference/dynamic_policy.hpp:62-150).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <new>

#include "../../include/ne_runtime.h"
#include "ne_device.hpp"

using namespace ne;

namespace {

constexpr int kInferThreads = 256;

__device__ inline bool Better(float s, uint32_t i, float best, uint32_t bi) {
    return s > best || (s == best && i < bi);
}

// Parameter offsets (elements) of a genome layout in the reference's order (policy_model.hpp:9-12,
// shared_encoder.hpp:16-30, dense_trunk.hpp:25-40): the encoder (145 -> 64, or 145 -> kHidden -> 64), the
// trunk 321 -> 256 -> 128 -> 64, each layer's weights [out][in] then its biases, then the tail rows.
template <int kHidden>
struct ModelLayout {
    static constexpr uint64_t kEncOut1 = kHidden == 0 ? uint64_t(kEncOut) : uint64_t(kHidden);
    static constexpr uint64_t kEncW1 = 0, kEncB1 = kEncOut1 * kTurnFeatures;
    static constexpr uint64_t kEncW2 = kEncB1 + kEncOut1;  // hidden layouts only
    static constexpr uint64_t kEncB2 = kEncW2 + (kHidden == 0 ? 0 : uint64_t(kEncOut) * kHidden);
    static constexpr uint64_t kW0 = kHidden == 0 ? kEncW2 : kEncB2 + kEncOut;
    static constexpr uint64_t kB0 = kW0 + uint64_t(kH0) * kModelIn, kW1 = kB0 + kH0;
    static constexpr uint64_t kB1 = kW1 + uint64_t(kH1) * kH0, kW2 = kB1 + kH1;
    static constexpr uint64_t kB2 = kW2 + uint64_t(kPol) * kH1, kP = kB2 + kPol;
};
static_assert(ModelLayout<0>::kP == kPolicyParams && ModelLayout<0>::kW0 == kOffW0, "145->64 layout");
static_assert(ModelLayout<128>::kP == 150528, "the reference's 145->128->64 layout");

// A parameter as fp32: stored fp32, or stored IEEE binary16 (common/float16.hpp:9-13), widened exactly.
template <bool kFp16>
__device__ __forceinline__ float Param(const void* g, uint64_t i) {
    if constexpr (kFp16)
        return __half2float(static_cast<const __half*>(g)[i]);
    else
        return static_cast<const float*>(g)[i];
}

// One board per workgroup: validity, encoder (per occupied turn), trunk 321 -> 256 -> 128 -> 64, then the
// masked argmax with the runner-up score over the first A actions.  kHidden: the encoder's hidden width (0:
// one 145 -> 64 layer); kFp16: parameters stored as binary16.
template <int kHidden, bool kFp16>
__global__ __launch_bounds__(kInferThreads) void SelectGuessKernel(const void* genome, const uint8_t* words,
                                                                   uint32_t A, const ne_board* boards,
                                                                   ne_selected_action* out) {
    using L = ModelLayout<kHidden>;
    __shared__ float s_x[kModelIn];
    __shared__ float s_hid[kHidden > 0 ? kMaxTurns * kHidden : 1];
    __shared__ float s_h0[kH0];
    __shared__ float s_h1[kH1];
    __shared__ float s_pol[kPol];
    __shared__ uint8_t s_letters[kMaxTurns][5];
    __shared__ uint8_t s_tiles[kMaxTurns][5];
    __shared__ int s_state;  // 0 ok, else the board's status
    __shared__ float s_best[kInferThreads], s_second[kInferThreads];
    __shared__ uint32_t s_index[kInferThreads];
    __shared__ int s_have[kInferThreads], s_have2[kInferThreads];
    const uint32_t tid = threadIdx.x;
    const ne_board& b = boards[blockIdx.x];
    const uint32_t tc = b.turn_count;
    if (tid == 0) {
        int st = tc > uint32_t(kMaxTurns) ? NE_POLICY_FORWARD_FAILED : NE_OK;
        for (uint32_t t = 0; st == NE_OK && t < tc; ++t) {
            bool green = true;
            for (int p = 0; p < 5; ++p) {
                if (b.guesses[t][p] >= kAlpha || b.tiles[t][p] > 2) st = NE_POLICY_FORWARD_FAILED;
                green &= b.tiles[t][p] == 0;
            }
            if (green) st = NE_POLICY_FORWARD_FAILED;  // a won board is finished
        }
        if (tc >= uint32_t(kMaxTurns)) st = NE_POLICY_FORWARD_FAILED;  // six turns: finished
        s_state = st;
    }
    if (tid < uint32_t(kMaxTurns) * 5) {
        const uint32_t t = tid / 5, p = tid % 5;
        s_letters[t][p] = t < tc ? b.guesses[t][p] : 0;
        s_tiles[t][p] = t < tc ? b.tiles[t][p] : 0;
    }
    for (uint32_t k = tid; k < uint32_t(kModelIn); k += kInferThreads) s_x[k] = k == 0 && tc == 0 ? 1.0f : 0.0f;
    __syncthreads();
    ne_selected_action* o = out + blockIdx.x;
    if (s_state != NE_OK) {
        if (tid == 0) {
            ne_selected_action r{};
            r.status = uint32_t(s_state);
            *o = r;
        }
        return;
    }
    // the shared encoder on each occupied turn (shared_encoder.hpp:94-101): a one-hot of 5 letters and 5
    // tiles through one linear layer (145 -> 64), or through kHidden ReLU units and a linear layer to 64
    constexpr uint32_t kFirst = uint32_t(L::kEncOut1);
    for (uint32_t idx = tid; idx < tc * kFirst; idx += kInferThreads) {
        const uint32_t t = idx / kFirst, u = idx % kFirst;
        const uint64_t w = L::kEncW1 + uint64_t(u) * kTurnFeatures;
        float sum = Param<kFp16>(genome, L::kEncB1 + u);
        for (int k = 0; k < kTurnFeatures; ++k) {
            const bool on = k < 130 ? s_letters[t][k / kAlpha] == k % kAlpha : s_tiles[t][(k - 130) / 3] == (k - 130) % 3;
            sum += Param<kFp16>(genome, w + k) * (on ? 1.0f : 0.0f);
        }
        if constexpr (kHidden == 0)
            s_x[1 + t * kEncOut + u] = sum;
        else
            s_hid[t * kHidden + u] = sum < 0.0f ? 0.0f : sum;
    }
    __syncthreads();
    if constexpr (kHidden > 0) {
        for (uint32_t idx = tid; idx < tc * uint32_t(kEncOut); idx += kInferThreads) {
            const uint32_t t = idx / kEncOut, u = idx % kEncOut;
            const uint64_t w = L::kEncW2 + uint64_t(u) * kHidden;
            float sum = Param<kFp16>(genome, L::kEncB2 + u);
            for (int h = 0; h < kHidden; ++h) sum += Param<kFp16>(genome, w + h) * s_hid[t * kHidden + h];
            s_x[1 + t * kEncOut + u] = sum;
        }
        __syncthreads();
    }
    {  // 321 -> 256, ReLU
        const uint64_t w = L::kW0 + uint64_t(tid) * kModelIn;
        float sum = Param<kFp16>(genome, L::kB0 + tid);
        for (int k = 0; k < kModelIn; ++k) sum += Param<kFp16>(genome, w + k) * s_x[k];
        s_h0[tid] = sum < 0.0f ? 0.0f : sum;
    }
    __syncthreads();
    if (tid < uint32_t(kH1)) {  // 256 -> 128, ReLU
        const uint64_t w = L::kW1 + uint64_t(tid) * kH0;
        float sum = Param<kFp16>(genome, L::kB1 + tid);
        for (int k = 0; k < kH0; ++k) sum += Param<kFp16>(genome, w + k) * s_h0[k];
        s_h1[tid] = sum < 0.0f ? 0.0f : sum;
    }
    __syncthreads();
    if (tid < uint32_t(kPol)) {  // 128 -> 64, the policy vector
        const uint64_t w = L::kW2 + uint64_t(tid) * kH1;
        float sum = Param<kFp16>(genome, L::kB2 + tid);
        for (int k = 0; k < kH1; ++k) sum += Param<kFp16>(genome, w + k) * s_h1[k];
        s_pol[tid] = sum;
    }
    __syncthreads();
    // scoring and the per-lane top two, actions in increasing index per lane
    bool have = false, have2 = false;
    float best = 0.0f, second = 0.0f;
    uint32_t bi = 0;
    for (uint32_t a = tid; a < A; a += kInferThreads) {
        const uint8_t* w = words + 5 * uint64_t(a);
        const uint8_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4];
        bool guessed = false;
        for (uint32_t t = 0; t < tc; ++t)
            guessed |= s_letters[t][0] == w0 && s_letters[t][1] == w1 && s_letters[t][2] == w2 &&
                       s_letters[t][3] == w3 && s_letters[t][4] == w4;
        if (guessed) continue;
        float s = 0.0f;
        for (int l = 0; l < kFixed; ++l) {  // the fixed features: letter count, or -1 when absent
            const int c = (w0 == l) + (w1 == l) + (w2 == l) + (w3 == l) + (w4 == l);
            s += s_pol[l] * (c == 0 ? -1.0f : float(c));
        }
        const uint64_t tail = L::kP + uint64_t(a) * kTail;
        for (int j = 0; j < kTail; ++j) s += s_pol[kFixed + j] * Param<kFp16>(genome, tail + j);
        if (!have || Better(s, a, best, bi)) {
            if (have && (!have2 || best > second)) second = best;
            have2 = have2 || have;
            best = s;
            bi = a;
            have = true;
        } else if (!have2 || s > second) {
            second = s;
            have2 = true;
        }
    }
    s_best[tid] = best;
    s_second[tid] = second;
    s_index[tid] = bi;
    s_have[tid] = have;
    s_have2[tid] = have2;
    __syncthreads();
    // the winner: highest score, lowest index; the runner-up: the best of every other candidate
    for (uint32_t half = kInferThreads / 2; half > 0; half >>= 1) {
        if (tid < half) {
            const uint32_t p = tid + half;
            const bool mine = s_have[tid] != 0, theirs = s_have[p] != 0;
            if (theirs) {
                if (!mine || Better(s_best[p], s_index[p], s_best[tid], s_index[tid])) {
                    // their best wins; my best becomes a runner-up candidate
                    float r = s_have2[p] ? s_second[p] : 0.0f;
                    bool hr = s_have2[p] != 0;
                    if (mine && (!hr || s_best[tid] > r)) r = s_best[tid];
                    hr = hr || mine;
                    if (s_have2[tid] && (!hr || s_second[tid] > r)) r = s_second[tid];
                    hr = hr || s_have2[tid];
                    s_best[tid] = s_best[p];
                    s_index[tid] = s_index[p];
                    s_second[tid] = r;
                    s_have2[tid] = hr;
                    s_have[tid] = 1;
                } else {
                    float r = s_have2[tid] ? s_second[tid] : 0.0f;
                    bool hr = s_have2[tid] != 0;
                    if (!hr || s_best[p] > r) r = s_best[p];
                    hr = true;
                    if (s_have2[p] && s_second[p] > r) r = s_second[p];
                    s_second[tid] = r;
                    s_have2[tid] = hr;
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        ne_selected_action r{};
        if (!s_have[0]) {
            r.status = NE_ACTION_SELECTION_FAILED;
        } else {
            r.status = NE_OK;
            r.action_index = s_index[0];
            r.score = s_best[0];
            r.second_score = s_have2[0] ? s_second[0] : -__builtin_inff();
            for (int p = 0; p < 5; ++p) r.word[p] = words[5 * uint64_t(s_index[0]) + p];
        }
        *o = r;
    }
}

// ---- staging of ne_model_select_guesses_noalloc ------------------------------------------------
// One array per GPU for the life of the process: kNoAllocMaxBoards boards, then their results.  On the
// host the name `scratch` is a stand-in, not a device address: host code uses only the address
// hipGetSymbolAddress returns for each device (looked up once per device, with that device current).
constexpr uint64_t kNoAllocMaxBoards = NE_NOALLOC_MAX_BOARDS;
constexpr uint64_t kScratchResultsOffset = kNoAllocMaxBoards * sizeof(ne_board);
constexpr uint64_t kScratchBytes = kScratchResultsOffset + kNoAllocMaxBoards * sizeof(ne_selected_action);
static_assert(kScratchResultsOffset % 256 == 0, "results start on a 256-byte boundary");
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// This device's `scratch` address; the caller has made `device` current.
int ScratchAddress(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxScratchDevices) return NE_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return NE_DEVICE_FAILURE;
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return NE_OK;
}

// The kernel of a layout.
void LaunchSelectGuess(uint32_t enc_hidden, bool fp16, uint32_t count, hipStream_t stream, const void* genome,
                       const uint8_t* words, uint32_t A, const ne_board* boards, ne_selected_action* out) {
    const dim3 grid(count), block(kInferThreads);
    if (enc_hidden == 0 && !fp16)
        hipLaunchKernelGGL((SelectGuessKernel<0, false>), grid, block, 0, stream, genome, words, A, boards, out);
    else if (enc_hidden == 0)
        hipLaunchKernelGGL((SelectGuessKernel<0, true>), grid, block, 0, stream, genome, words, A, boards, out);
    else if (!fp16)
        hipLaunchKernelGGL((SelectGuessKernel<128, false>), grid, block, 0, stream, genome, words, A, boards, out);
    else
        hipLaunchKernelGGL((SelectGuessKernel<128, true>), grid, block, 0, stream, genome, words, A, boards, out);
}

uint64_t LayoutParams(uint32_t enc_hidden) { return enc_hidden == 0 ? ModelLayout<0>::kP : ModelLayout<128>::kP; }

}  // namespace

extern "C" uint64_t ne_model_genome_bytes(const ne_model_layout* layout, uint64_t action_count) {
    if (!layout || (layout->encoder_hidden != 0 && layout->encoder_hidden != 128) ||
        (layout->storage != NE_STORAGE_FP32 && layout->storage != NE_STORAGE_FP16) || action_count == 0)
        return 0;
    const uint64_t values = LayoutParams(layout->encoder_hidden) + uint64_t(kTail) * action_count;
    if (layout->storage == NE_STORAGE_FP16) return 2 * values;  // ComputeDynamicGenomeStrideBytes (2-byte alignment)
    return (4 * values + 255) / 256 * 256;                       // this build's fp32 slot stride
}

struct ne_model {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    uint32_t enc_hidden = 0;
    bool fp16 = false;
    void* genome = nullptr;
    uint8_t* words = nullptr;
    uint64_t action_count = 0;
    ne_board* boards = nullptr;
    ne_selected_action* results = nullptr;
    uint64_t capacity = 0;
    void* scratch = nullptr;  // this model's GPU's `scratch` (ne_model_select_guesses_noalloc)
};

extern "C" {

int ne_model_create_with_layout(const void* genome_bytes, uint64_t genome_byte_count, const ne_model_layout* layout,
                                const uint8_t* action_letters, uint64_t action_count, void* stream,
                                ne_model** out_model) {
    if (!genome_bytes || !action_letters || !out_model || action_count == 0 || action_count > 0xFFFFFFFFull)
        return NE_INVALID_ARGUMENT;
    *out_model = nullptr;
    // IsValidCreateInputs (single_model_device_runtime.cu:24-29): the genome is exactly one genome of the
    // layout for action_count actions
    const uint64_t expect = ne_model_genome_bytes(layout, action_count);
    if (expect == 0 || genome_byte_count != expect) return NE_INVALID_ARGUMENT;
    for (uint64_t i = 0; i < 5 * action_count; ++i)
        if (action_letters[i] >= kAlpha) return NE_INVALID_ARGUMENT;
    ne_model* m = new (std::nothrow) ne_model();
    if (!m) return NE_DEVICE_FAILURE;
    m->enc_hidden = layout->encoder_hidden;
    m->fp16 = layout->storage == NE_STORAGE_FP16;
    bool ok = true;
    if (stream) {
        m->stream = reinterpret_cast<hipStream_t>(stream);
    } else {
        ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess;
        m->own_stream = ok;
    }
    m->action_count = action_count;
    int device = -1;
    ok = ok && hipGetDevice(&device) == hipSuccess && ScratchAddress(device, &m->scratch) == NE_OK;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m->genome), genome_byte_count) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m->words), 5 * action_count) == hipSuccess;
    ok = ok && hipMemcpyAsync(m->genome, genome_bytes, genome_byte_count, hipMemcpyHostToDevice, m->stream) ==
                   hipSuccess;
    ok = ok && hipMemcpyAsync(m->words, action_letters, 5 * action_count, hipMemcpyHostToDevice, m->stream) ==
                   hipSuccess;
    ok = ok && hipStreamSynchronize(m->stream) == hipSuccess;
    if (!ok) {
        ne_model_destroy(m);
        return NE_DEVICE_FAILURE;
    }
    *out_model = m;
    return NE_OK;
}

int ne_model_create(const void* genome_bytes, uint64_t genome_byte_count, const uint8_t* action_letters,
                    uint64_t action_count, void* stream, ne_model** out_model) {
    const ne_model_layout fp32_layout{0, NE_STORAGE_FP32, {0, 0}};  // this build's slot: 145 -> 64, fp32
    return ne_model_create_with_layout(genome_bytes, genome_byte_count, &fp32_layout, action_letters, action_count,
                                       stream, out_model);
}

void ne_model_destroy(ne_model* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    if (m->genome) (void)hipFree(m->genome);
    if (m->words) (void)hipFree(m->words);
    if (m->boards) (void)hipFree(m->boards);
    if (m->results) (void)hipFree(m->results);
    if (m->own_stream && m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int ne_model_select_guesses(ne_model* m, const ne_board* boards, uint64_t count, ne_selected_action* out) {
    if (!m || !boards || !out || count == 0 || count > 0x7FFFFFFFull) return NE_INVALID_ARGUMENT;
    if (count > m->capacity) {
        if (m->boards) (void)hipFree(m->boards);
        if (m->results) (void)hipFree(m->results);
        m->boards = nullptr;
        m->results = nullptr;
        m->capacity = 0;
        if (hipMalloc(reinterpret_cast<void**>(&m->boards), count * sizeof(ne_board)) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&m->results), count * sizeof(ne_selected_action)) != hipSuccess)
            return NE_DEVICE_FAILURE;
        m->capacity = count;
    }
    bool ok = hipMemcpyAsync(m->boards, boards, count * sizeof(ne_board), hipMemcpyHostToDevice, m->stream) ==
              hipSuccess;
    if (ok) {
        LaunchSelectGuess(m->enc_hidden, m->fp16, uint32_t(count), m->stream, m->genome, m->words,
                          uint32_t(m->action_count), m->boards, m->results);
        ok = hipGetLastError() == hipSuccess;
    }
    ok = ok && hipMemcpyAsync(out, m->results, count * sizeof(ne_selected_action), hipMemcpyDeviceToHost,
                              m->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(m->stream) == hipSuccess;
    return ok ? NE_OK : NE_DEVICE_FAILURE;
}

// ne_model_select_guesses through the GPU's static staging array: no allocation of any kind.  Calls that
// overlap on one GPU (other threads or streams) share the array and race; every byte a call reads back
// is written by that call first (its boards by the upload, every result by the kernel), so nothing
// needs clearing between calls.
int ne_model_select_guesses_noalloc(ne_model* m, const ne_board* boards, uint64_t count, ne_selected_action* out) {
    if (!m || !boards || !out || count == 0 || count > kNoAllocMaxBoards) return NE_INVALID_ARGUMENT;
    if (!m->scratch) return NE_DEVICE_FAILURE;
    ne_board* d_boards = static_cast<ne_board*>(m->scratch);
    ne_selected_action* d_results =
        reinterpret_cast<ne_selected_action*>(static_cast<uint8_t*>(m->scratch) + kScratchResultsOffset);
    bool ok = hipMemcpyAsync(d_boards, boards, count * sizeof(ne_board), hipMemcpyHostToDevice, m->stream) ==
              hipSuccess;
    if (ok) {
        LaunchSelectGuess(m->enc_hidden, m->fp16, uint32_t(count), m->stream, m->genome, m->words,
                          uint32_t(m->action_count), d_boards, d_results);
        ok = hipGetLastError() == hipSuccess;
    }
    ok = ok && hipMemcpyAsync(out, d_results, count * sizeof(ne_selected_action), hipMemcpyDeviceToHost,
                              m->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(m->stream) == hipSuccess;
    return ok ? NE_OK : NE_DEVICE_FAILURE;
}

int ne_model_select_guess(ne_model* m, const ne_board* board, ne_selected_action* out) {
    const int st = ne_model_select_guesses(m, board, 1, out);
    return st != NE_OK ? st : int(out->status);
}

}  // extern "C"
