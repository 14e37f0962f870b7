// ne_device.hpp -- genome layout, draw protocol and small device helpers shared by the kernels and
// the host orchestration of the MI355X GA runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ne {

// ---- model sizes (reference src/model/input_encoder/encoder_spec.hpp:10-20,
// model_input/model_input_spec.hpp:11-15, dense_trunk/dense_trunk.hpp:15-18,
// output_embedding/output_embedding.hpp:19-21)
constexpr int kAlpha = 26;
constexpr int kTurnFeatures = 145;
constexpr int kEncOut = 64;
constexpr int kModelIn = 321;
constexpr int kH0 = 256;
constexpr int kH1 = 128;
constexpr int kPol = 64;
constexpr int kFixed = 26;
constexpr int kTail = 38;
constexpr int kMaxTurns = 6;
constexpr uint32_t kInvalidSlot = 0xFFFFFFFFu;
constexpr float kFitnessFloor = 1.40129846e-45f;  // reference src/spatial/grid.hpp:14

// Genome = the reference PolicyModelParameters order (policy_model.hpp:9-12) with each dense
// layer row-major [out][in] followed by its biases, then A tail rows of 38 values
// (genome/dynamic_layout.hpp:15-71).  Values are fp32.  Offsets in floats, 145->64 encoder.
constexpr uint64_t kOffEncW = 0;                                   // [64][145]
constexpr uint64_t kOffEncB = kOffEncW + uint64_t(kEncOut) * kTurnFeatures;  // [64]
constexpr uint64_t kOffW0 = kOffEncB + kEncOut;                    // [256][321]
constexpr uint64_t kOffB0 = kOffW0 + uint64_t(kH0) * kModelIn;     // [256]
constexpr uint64_t kOffW1 = kOffB0 + kH0;                          // [128][256]
constexpr uint64_t kOffB1 = kOffW1 + uint64_t(kH1) * kH0;          // [128]
constexpr uint64_t kOffW2 = kOffB1 + kH1;                          // [64][128]
constexpr uint64_t kOffB2 = kOffW2 + uint64_t(kPol) * kH1;         // [64]
constexpr uint64_t kPolicyParams = kOffB2 + kPol;                  // 132,928
static_assert(kPolicyParams == 132928, "policy parameter count");
static_assert(kOffW1 % 4 == 0 && kOffW2 % 4 == 0 && kOffB0 % 4 == 0 && kOffB1 % 4 == 0 && kOffB2 % 4 == 0,
              "float4-aligned layer offsets");
static_assert((kPolicyParams * 4) % 256 == 0, "tail rows start on a 256-byte boundary");

// Slot stride: the genome bytes rounded up to 256 bytes so every slot starts on a 256-byte line.
__host__ __device__ constexpr uint64_t GenomeValueCount(uint64_t action_count) {
    return kPolicyParams + action_count * kTail;
}
__host__ __device__ constexpr uint64_t SlotStrideBytes(uint64_t action_count) {
    return action_count == 0 ? 0 : ((GenomeValueCount(action_count) * 4 + 255) / 256) * 256;
}

// ---- scoring layout ----------------------------------------------------------------------------
// An action's score is the 64-term dot product of its embedding row [26 fixed | 38 tail] with the
// policy vector (output_embedding.hpp:97-119).  On v_mfma_f32_16x16x4_f32 it takes 16 k-steps per
// tile of 16 actions, with no pad: lane group g supplies, at step s, policy dim ScoreDim(s, g) --
// first its ScoreFixedCount(g) fixed dims (groups 0, 1: 6; groups 2, 3: 7), then its tail dims
// (10; 9).  Steps 0..5 read fixed features only and steps 7..15 tail values only; step 6 reads the
// tail in groups 0, 1 and the fixed features in groups 2, 3.
constexpr int kScoreSteps = 16;
__host__ __device__ constexpr int ScoreFixedCount(int g) { return g < 2 ? 6 : 7; }
__host__ __device__ constexpr int ScoreFixedBase(int g) { return 6 * g + (g == 3 ? 1 : 0); }  // 0, 6, 12, 19
__host__ __device__ constexpr int ScoreTailBase(int g) { return 10 * g - (g == 3 ? 1 : 0); }  // 0, 10, 20, 29
// Policy dim of step s for lane group g in the tail part (s >= ScoreFixedCount(g)) is ScoreTailRow(g) + s.
__host__ __device__ constexpr int ScoreTailRow(int g) { return kFixed + ScoreTailBase(g) - ScoreFixedCount(g); }
// First of the 10 tail values a lane group reads for steps 6..15 (groups 2, 3 read one value before
// their first tail dim, in place of step 6, which they take from the fixed features).
__host__ __device__ constexpr int ScoreTailRead(int g) { return ScoreTailRow(g) - kFixed + 6; }  // 0, 10, 19, 28
__host__ __device__ inline int ScoreDim(int s, int g) {
    return s < ScoreFixedCount(g) ? ScoreFixedBase(g) + s : ScoreTailRow(g) + s;
}
// The lane group holding fixed dim d (at step d - ScoreFixedBase(group)).
__host__ __device__ inline int ScoreFixedGroup(int d) { return d < 12 ? d / 6 : (d < 19 ? 2 : 3); }
// Per tile of 16 actions, the catalog-wide image read by the scoring loop (identical in HBM and
// LDS): two planes of [64 lanes][4] fixed-feature values (lane i + 16g holds steps 4h..4h+3 of
// action row i for lane group g; steps past the group's fixed dims are 0), then the 16 actions'
// canonical indices (the lowest catalog index holding the same word; a guessed word masks every copy
// of itself).
constexpr int kImgTileWords = 2 * 64 * 4 + 16;  // 528
constexpr int kImgTileBytes = 4 * kImgTileWords;

// BF16 screening operands.  Per action tile, the two A operands of v_mfma_f32_16x16x32_bf16, 16 bytes
// (8 bf16) per lane: lane i + 16g holds, for action row i, k-slots 8g..8g+7 of instruction m.
// Instruction 0 carries the 26 fixed dims then tail dims 32..37, instruction 1 tail dims 0..31, so the
// 64 dims fill the 64 k-slots exactly.  Lanes 0..47 of instruction 0 hold fixed dims only and come
// from one catalog-wide image; the rest is per genome:
//   catalog tile: instruction 0, lanes 0..47                                  768 B
//   genome tile:  instruction 0, lanes 48..63 (fixed 24, 25, tail 32..37)     256 B
//                 instruction 1, lanes 0..63 (tail 0..31)                   1,024 B
constexpr uint32_t kBf16CatTileBytes = 48 * 16;
constexpr uint32_t kBf16GenTileBytes = 16 * 16 + 64 * 16;
__host__ __device__ inline int Bf16SlotDim(int m, int slot) {  // slot = 8g + e -> policy dim
    if (m == 0) return slot < kFixed ? slot : kFixed + 32 + (slot - kFixed);
    return kFixed + slot;
}
// The inverse: policy dim d -> instruction (returned) and k-slot (*slot).
__host__ __device__ inline int Bf16DimSlot(int d, int* slot) {
    if (d < kFixed) {
        *slot = d;
        return 0;
    }
    const int t = d - kFixed;
    if (t < 32) {
        *slot = t;
        return 1;
    }
    *slot = kFixed + (t - 32);
    return 0;
}

// Per-evaluation copy of a genome's encoder, transposed to [feature][out], then W0's virgin column.
constexpr uint64_t kEncTStride = uint64_t(kTurnFeatures) * kEncOut + kH0;

// Dense layers in crossover / mutation order (reference genome_ops.cuh:264-309).
struct DenseLayerDesc {
    uint32_t w, b, in, out;
};
constexpr int kDenseLayers = 4;
constexpr int kEncoderLayers = 1;
__host__ __device__ inline DenseLayerDesc DenseLayerAt(int l) {
    switch (l) {
    case 0: return {uint32_t(kOffEncW), uint32_t(kOffEncB), uint32_t(kTurnFeatures), uint32_t(kEncOut)};
    case 1: return {uint32_t(kOffW0), uint32_t(kOffB0), uint32_t(kModelIn), uint32_t(kH0)};
    case 2: return {uint32_t(kOffW1), uint32_t(kOffB1), uint32_t(kH0), uint32_t(kH1)};
    default: return {uint32_t(kOffW2), uint32_t(kOffB2), uint32_t(kH1), uint32_t(kPol)};
    }
}

// ---- draw protocol ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Key = (seed, purpose tag); the counter names
// the draw (element / child / generation), so every draw is a pure function and the kernels can
// take them in any order.
constexpr uint32_t kTagBoot = 0x424F4F54u;
constexpr uint32_t kTagPlan = 0x504C414Eu;
constexpr uint32_t kTagXovr = 0x584F5652u;
constexpr uint32_t kTagMuta = 0x4D555441u;
constexpr uint32_t kTagNorm = 0x4E4F524Du;
constexpr uint32_t kTagRscl = 0x5253434Cu;
constexpr uint32_t kPlanSeedXor = 0xA341316Cu;  // planning seed = generation seed ^ this

struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline U4 Philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                         uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t m0 = uint64_t(0xD2511F53u) * uint64_t(c0);
        const uint64_t m1 = uint64_t(0xCD9E8D57u) * uint64_t(c2);
        const uint32_t n0 = uint32_t(m1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = uint32_t(m0 >> 32) ^ c3 ^ k1;
        c1 = uint32_t(m1);
        c3 = uint32_t(m0);
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{c0, c1, c2, c3};
}

__host__ __device__ inline float UniformOpenClosed(uint32_t x) {  // (0,1], exact in fp32
    return float((x >> 8) + 1u) * 0x1p-24f;
}

__host__ __device__ inline bool BernoulliDraw(float p, uint32_t word) {
    if (p <= 0.0f) return false;
    if (p >= 1.0f) return true;
    return UniformOpenClosed(word) < p;
}

__host__ __device__ inline uint32_t Lane(const U4& u, uint32_t i) {
    return i == 0 ? u.x : (i == 1 ? u.y : (i == 2 ? u.z : u.w));
}

// Logarithm on (0,1] and square root built from IEEE add/mul/div only, so the device result is the
// same bit pattern as any IEEE host.
__host__ __device__ inline double IeeeLog(double x) {
#pragma clang fp contract(off)
    uint64_t bits = __builtin_bit_cast(uint64_t, x);
    int e = int((bits >> 52) & 0x7FF) - 1023;
    double m = __builtin_bit_cast(double, (bits & 0x000FFFFFFFFFFFFFull) | (uint64_t(1023) << 52));
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double t = 1.0 / 23.0;
    t = t * s2 + 1.0 / 21.0;
    t = t * s2 + 1.0 / 19.0;
    t = t * s2 + 1.0 / 17.0;
    t = t * s2 + 1.0 / 15.0;
    t = t * s2 + 1.0 / 13.0;
    t = t * s2 + 1.0 / 11.0;
    t = t * s2 + 1.0 / 9.0;
    t = t * s2 + 1.0 / 7.0;
    t = t * s2 + 1.0 / 5.0;
    t = t * s2 + 1.0 / 3.0;
    t = t * s2 + 1.0;
    return double(e) * 0.6931471805599453 + 2.0 * s * t;
}

__host__ __device__ inline double IeeeSqrt(double x) {
#pragma clang fp contract(off)
    uint64_t bits = __builtin_bit_cast(uint64_t, x);
    double y = __builtin_bit_cast(double, (bits >> 1) + (uint64_t(1023) << 51));
    for (int i = 0; i < 6; ++i) y = 0.5 * (y + x / y);
    return y;
}

// N(0,1) from two words: a 52-bit uniform in (0,1) mapped by Acklam's rational inverse normal CDF.
__host__ __device__ inline float NormalFromWords(uint32_t a, uint32_t b) {
#pragma clang fp contract(off)
    const uint64_t k = (uint64_t(a >> 6) << 26) | uint64_t(b >> 6);
    const double u = (double(k) + 0.5) * 0x1p-52;
    const double plow = 0.02425;
    double z;
    if (u < plow || u > 1.0 - plow) {
        const double tu = u < plow ? u : 1.0 - u;
        const double q = IeeeSqrt(-2.0 * IeeeLog(tu));
        const double num =
            ((((-7.784894002430293e-03 * q + -3.223964580411365e-01) * q + -2.400758277161838e+00) * q +
              -2.549732539343734e+00) * q + 4.374664141464968e+00) * q + 2.938163982698783e+00;
        const double den = (((7.784695709041462e-03 * q + 3.224671290700398e-01) * q + 2.445134137142996e+00) * q +
                            3.754408661907416e+00) * q + 1.0;
        z = num / den;
        if (u > 1.0 - plow) z = -z;
    } else {
        const double q = u - 0.5;
        const double r = q * q;
        const double num = (((((-3.969683028665376e+01 * r + 2.209460984245205e+02) * r + -2.759285104469687e+02) * r +
                              1.383577518672690e+02) * r + -3.066479806614716e+01) * r + 2.506628277459239e+00) * q;
        const double den = ((((-5.447609879822406e+01 * r + 1.615858368580409e+02) * r + -1.556989798598866e+02) * r +
                             6.680131188771972e+01) * r + -1.328068155211847e+01) * r + 1.0;
        z = num / den;
    }
    return float(z);
}

// ---- words ------------------------------------------------------------------------------------
// A word is packed as 5 letters of 5 bits: l0 | l1<<5 | ... | l4<<20.
__host__ __device__ inline uint32_t WordLetter(uint32_t code, int p) { return (code >> (5 * p)) & 31u; }

// Tile feedback green=0, yellow=1, grey=2 (reference src/wordle/feedback.hpp:80-84,170-203),
// packed 2 bits per position.
__host__ __device__ inline uint32_t FeedbackPacked(uint32_t guess, uint32_t sol) {
    uint32_t counts[kAlpha];
    for (int l = 0; l < kAlpha; ++l) counts[l] = 0;
    uint32_t fb = 0;
    for (int p = 0; p < 5; ++p) {
        const uint32_t g = WordLetter(guess, p), s = WordLetter(sol, p);
        if (g == s) continue;  // green (0)
        fb |= 2u << (2 * p);
        ++counts[s];
    }
    for (int p = 0; p < 5; ++p) {
        if (((fb >> (2 * p)) & 3u) == 0) continue;
        const uint32_t g = WordLetter(guess, p);
        if (counts[g] == 0) continue;
        fb = (fb & ~(3u << (2 * p))) | (1u << (2 * p));
        --counts[g];
    }
    return fb;
}

// Crossover decision record of one child (16 words; the exported "crossover mask").
struct XoverRecord {
    uint32_t layer_first[6];
    uint32_t neuron[6];
    uint32_t tail_first;
    uint32_t row2;
    uint32_t splice_row;
    uint32_t splice_point;
};
static_assert(sizeof(XoverRecord) == 64, "crossover record is 16 words");

// Game trace record (matches ne_game_trace in include/ne_runtime.h).
struct GameTrace {
    uint32_t guesses[6];
    float best[6];
    float second[6];
    uint32_t decisions;
    uint32_t turn_count;
    float score;
    uint32_t won;
};

}  // namespace ne
