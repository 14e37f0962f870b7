// ne_oracle.cpp -- CPU restatement of the reference GA generation step.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this file; only the
// tests, the smoke check and the CPU-baseline leg of bench.py do.  It restates, in plain scalar C++,
// the algorithms of the reference files cited next to each function, so the HIP path can be compared
// against it on the same inputs and the same counter-based random draws.
//
// Numerics: compiled with -ffp-contract=off and no fast-math, so every float/double operation is a
// single IEEE operation; the only fused multiply-adds are the explicit std::fma/std::fmaf calls that
// the draw protocol specifies.
//
// Pinning (what makes each part trustworthy rather than a retyped formula) is listed in the tests:
// Philox known-answer vectors from Salmon et al. (SC'11, Random123 kat_vectors), the reference's own
// golden fixtures (feedback, hint grids, encoder/trunk/policy/scoring goldens, masking), brute-force
// checks against the reference headers compiled under oracle/_ref, and closed forms.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

// ------------------------------------------------------------------------------------------------
// Sizes (reference: src/model/input_encoder/encoder_spec.hpp:10-20,
// src/model/model_input/model_input_spec.hpp:11-15, src/model/dense_trunk/dense_trunk.hpp:15-18,
// src/model/output_embedding/output_embedding.hpp:19-21)
constexpr int kWordLen = 5;
constexpr int kAlpha = 26;
constexpr int kTurnFeatures = 145;  // 5*26 letters + 5*3 tiles
constexpr int kEncOut = 64;
constexpr int kModelIn = 321;       // 1 + 5*64
constexpr int kH0 = 256, kH1 = 128, kPol = 64;
constexpr int kFixed = 26, kTail = 38;
constexpr int kMaxTurns = 6;

enum Status : int {
    kOk = 0, kCudaFailure = 1, kInvalidRuntimeConfig = 2, kInvalidSlab = 3, kInvalidGeneration = 4,
    kInvalidTrainingShard = 5, kGuessAppendFailed = 6, kPolicyForwardFailed = 7, kActionSelectionFailed = 8,
    kPopulationNotEvaluated = 9, kInvalidAssemblyConfig = 10, kParentSelectionFailed = 11,
    kInvalidAssemblyPlan = 12, kInvalidParentIndex = 13, kSlabFull = 14, kOutputEmbeddingInjectionFailed = 15,
    kSlabRepackFailed = 16,
};

// Genome layout in float elements: the reference order of PolicyModelParameters
// (src/model/policy_model/policy_model.hpp:9-12, shared_encoder.hpp:108-122, dense_trunk.hpp:25-40),
// each layer row-major [out][in] then biases, followed by A rows of 38 tail values
// (src/genetic_algorithm/genome/dynamic_layout.hpp:15-71).  Stored as fp32 here.
struct Layout {
    int enc_hidden = 0;                      // 0: 145->64 encoder ; 128: 145->128->64 encoder
    size_t enc_w1 = 0, enc_b1 = 0, enc_w2 = 0, enc_b2 = 0;
    size_t w0 = 0, b0 = 0, w1 = 0, b1 = 0, w2 = 0, b2 = 0, P = 0;
    int enc1_out = 64;
};

Layout MakeLayout(int enc_hidden) {
    Layout L{};
    L.enc_hidden = enc_hidden;
    size_t o = 0;
    if (enc_hidden == 0) {
        L.enc1_out = kEncOut;
        L.enc_w1 = o; o += size_t(kEncOut) * kTurnFeatures;
        L.enc_b1 = o; o += kEncOut;
    } else {
        L.enc1_out = enc_hidden;
        L.enc_w1 = o; o += size_t(enc_hidden) * kTurnFeatures;
        L.enc_b1 = o; o += enc_hidden;
        L.enc_w2 = o; o += size_t(kEncOut) * enc_hidden;
        L.enc_b2 = o; o += kEncOut;
    }
    L.w0 = o; o += size_t(kH0) * kModelIn;
    L.b0 = o; o += kH0;
    L.w1 = o; o += size_t(kH1) * kH0;
    L.b1 = o; o += kH1;
    L.w2 = o; o += size_t(kPol) * kH1;
    L.b2 = o; o += kPol;
    L.P = o;
    return L;
}

// Dense layers in crossover/mutation order (reference genome_ops.cuh:264-309).
struct DenseLayer { size_t w, b; int in, out; bool encoder; };

int DenseLayers(const Layout& L, DenseLayer* out) {
    int n = 0;
    if (L.enc_hidden == 0) {
        out[n++] = {L.enc_w1, L.enc_b1, kTurnFeatures, kEncOut, true};
    } else {
        out[n++] = {L.enc_w1, L.enc_b1, kTurnFeatures, L.enc_hidden, true};
        out[n++] = {L.enc_w2, L.enc_b2, L.enc_hidden, kEncOut, true};
    }
    out[n++] = {L.w0, L.b0, kModelIn, kH0, false};
    out[n++] = {L.w1, L.b1, kH0, kH1, false};
    out[n++] = {L.w2, L.b2, kH1, kPol, false};
    return n;
}

// ------------------------------------------------------------------------------------------------
// Counter-based RNG: Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11.  Round: (hi0,lo0)=mulhilo(M0,c0); (hi1,lo1)=mulhilo(M1,c2);
// c' = {hi1^c1^k0, lo1, hi0^c3^k1, lo0}; key bumped by the Weyl constants between rounds.
void Philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void Draw(uint32_t key0, uint32_t tag, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t out[4]) {
    const uint32_t ctr[4] = {c0, c1, c2, c3};
    const uint32_t key[2] = {key0, tag};
    Philox(ctr, key, out);
}

// Purpose tags (second key word) of the draw protocol.
constexpr uint32_t kTagBoot = 0x424F4F54u;  // "BOOT" generation-0 weights
constexpr uint32_t kTagPlan = 0x504C414Eu;  // "PLAN" parent selection
constexpr uint32_t kTagXovr = 0x584F5652u;  // "XOVR" crossover decisions
constexpr uint32_t kTagMuta = 0x4D555441u;  // "MUTA" mutation Bernoulli words
constexpr uint32_t kTagNorm = 0x4E4F524Du;  // "NORM" mutation normals
constexpr uint32_t kTagRscl = 0x5253434Cu;  // "RSCL" tail row scale
constexpr uint32_t kPlanSeedXor = 0xA341316Cu;  // reference slab_runtime.cu:1443

// Uniform in (0,1]: the top 24 bits plus one, times 2^-24 (exact in fp32).
float U01(uint32_t x) { return float((x >> 8) + 1u) * 0x1p-24f; }

bool Bernoulli(float p, uint32_t word) {
    if (p <= 0.0f) return false;
    if (p >= 1.0f) return true;
    return U01(word) < p;
}

// Natural log on (0,1] from IEEE +,-,*,/ only (atanh series), so host and device agree bit-for-bit.
double DetLog(double x) {
    uint64_t bits;
    std::memcpy(&bits, &x, 8);
    int e = int((bits >> 52) & 0x7FF) - 1023;
    uint64_t mb = (bits & 0x000FFFFFFFFFFFFFull) | (uint64_t(1023) << 52);
    double m;
    std::memcpy(&m, &mb, 8);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
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
    const double logm = 2.0 * s * t;
    return double(e) * 0.6931471805599453 + logm;
}

// Square root by Newton iteration from a bit-level first guess; IEEE ops only.
double DetSqrt(double x) {
    uint64_t bits;
    std::memcpy(&bits, &x, 8);
    bits = (bits >> 1) + (uint64_t(1023) << 51);
    double y;
    std::memcpy(&y, &bits, 8);
    for (int i = 0; i < 6; ++i) y = 0.5 * (y + x / y);
    return y;
}

// Standard normal from two words: 52-bit uniform in (0,1), then Acklam's rational approximation of
// the inverse normal CDF (P. J. Acklam, 2003; relative error < 1.15e-9).
float NormalFromWords(uint32_t a, uint32_t b) {
    const uint64_t k = (uint64_t(a >> 6) << 26) | uint64_t(b >> 6);
    const double u = (double(k) + 0.5) * 0x1p-52;
    static const double A[6] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                                1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
    static const double B[5] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                                6.680131188771972e+01, -1.328068155211847e+01};
    static const double C[6] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                                -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
    static const double D[4] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                                3.754408661907416e+00};
    const double plow = 0.02425;
    double z;
    if (u < plow || u > 1.0 - plow) {
        const double tail_u = (u < plow) ? u : (1.0 - u);
        const double q = DetSqrt(-2.0 * DetLog(tail_u));
        const double num = ((((C[0] * q + C[1]) * q + C[2]) * q + C[3]) * q + C[4]) * q + C[5];
        const double den = (((D[0] * q + D[1]) * q + D[2]) * q + D[3]) * q + 1.0;
        z = num / den;
        if (u > 1.0 - plow) z = -z;
    } else {
        const double q = u - 0.5;
        const double r = q * q;
        const double num = (((((A[0] * r + A[1]) * r + A[2]) * r + A[3]) * r + A[4]) * r + A[5]) * q;
        const double den = ((((B[0] * r + B[1]) * r + B[2]) * r + B[3]) * r + B[4]) * r + 1.0;
        z = num / den;
    }
    return float(z);
}

// ------------------------------------------------------------------------------------------------
// Wordle rules (reference src/wordle/feedback.hpp:103-136, wordle_grid.hpp:13-77).
struct Grid {
    uint8_t sol[5];
    uint8_t guess[kMaxTurns][5];
    uint8_t tile[kMaxTurns][5];
    int tc = 0;
};

void Feedback(const uint8_t* guess, const uint8_t* sol, uint8_t* fb) {
    uint8_t counts[kAlpha] = {0};
    for (int p = 0; p < 5; ++p) {
        if (guess[p] == sol[p]) { fb[p] = 0; continue; }
        fb[p] = 2;
        ++counts[sol[p]];
    }
    for (int p = 0; p < 5; ++p) {
        if (fb[p] == 0) continue;
        const uint8_t g = guess[p];
        if (counts[g] == 0) continue;
        fb[p] = 1;
        --counts[g];
    }
}

bool SameWord(const uint8_t* a, const uint8_t* b) { return std::memcmp(a, b, 5) == 0; }

bool IsWon(const Grid& g) {
    for (int t = 0; t < g.tc && t < kMaxTurns; ++t)
        if (SameWord(g.guess[t], g.sol)) return true;
    return false;
}

bool IsFinished(const Grid& g) { return IsWon(g) || g.tc >= kMaxTurns; }

bool Append(Grid& g, const uint8_t* w) {
    for (int p = 0; p < 5; ++p) if (w[p] >= kAlpha) return false;
    if (IsFinished(g)) return false;
    std::memcpy(g.guess[g.tc], w, 5);
    Feedback(w, g.sol, g.tile[g.tc]);
    ++g.tc;
    return true;
}

// ------------------------------------------------------------------------------------------------
// Policy forward, sequential order of the reference (shared_encoder.hpp:38-58,94-115,
// wordle_grid_state.hpp:26-64, dense_trunk.hpp:44-114).  T = float or double.
template <typename T>
void DenseApply(const float* W, const float* b, int in, int out, const T* x, T* y, bool relu) {
    for (int o = 0; o < out; ++o) {
        T sum = T(b[o]);
        const float* row = W + size_t(o) * in;
        for (int k = 0; k < in; ++k) sum += T(row[k]) * x[k];
        y[o] = (relu && sum < T(0)) ? T(0) : sum;
    }
}

template <typename T>
void EncodeTurn(const float* g, const Layout& L, const uint8_t* letters, const uint8_t* tiles, T* enc) {
    T x[kTurnFeatures];
    for (int k = 0; k < kTurnFeatures; ++k) x[k] = T(0);
    for (int p = 0; p < 5; ++p) {
        x[p * kAlpha + letters[p]] = T(1);
        x[130 + p * 3 + tiles[p]] = T(1);
    }
    if (L.enc_hidden == 0) {
        DenseApply<T>(g + L.enc_w1, g + L.enc_b1, kTurnFeatures, kEncOut, x, enc, false);
    } else {
        T h[256];
        DenseApply<T>(g + L.enc_w1, g + L.enc_b1, kTurnFeatures, L.enc_hidden, x, h, true);
        DenseApply<T>(g + L.enc_w2, g + L.enc_b2, L.enc_hidden, kEncOut, h, enc, false);
    }
}

template <typename T>
bool PolicyForward(const float* g, const Layout& L, const Grid& grid, T* policy) {
    if (grid.tc > 5 || IsFinished(grid)) return false;
    T x[kModelIn];
    for (int k = 0; k < kModelIn; ++k) x[k] = T(0);
    x[0] = grid.tc == 0 ? T(1) : T(0);
    for (int t = 0; t < grid.tc; ++t) EncodeTurn<T>(g, L, grid.guess[t], grid.tile[t], x + 1 + 64 * t);
    T h0[kH0], h1[kH1];
    DenseApply<T>(g + L.w0, g + L.b0, kModelIn, kH0, x, h0, true);
    DenseApply<T>(g + L.w1, g + L.b1, kH0, kH1, h0, h1, true);
    DenseApply<T>(g + L.w2, g + L.b2, kH1, kPol, h1, policy, false);
    return true;
}

// BF16 trunk (the NE_EVAL_BF16_TRUNK form; no reference counterpart -- the reference's trunk is fp32,
// dense_trunk.hpp:44-114): layer 0 exactly as the fp32 forward; layers 1 and 2 on split bfloat16 operands
// (three-term "bf16x3" products): each weight w and each ReLU'd input x is split into hi = bf16(v) and
// lo = bf16(v - hi), both round to nearest even, and w.x is taken as hi_w.hi_x + hi_w.lo_x + lo_w.hi_x
// (the lo.lo term, ~2^-16 of the product, is dropped).  The exact products are summed in fp64 from the fp32
// bias and each output is rounded to fp32 -- an fp32-accumulating MFMA differs from this by its own
// accumulation rounding only.
float Bf16Round(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return x;  // inf / NaN pass through
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

void DenseBf16(const float* W, const float* b, int in, int out, const float* x, float* y, bool relu) {
    for (int o = 0; o < out; ++o) {
        double sum = double(b[o]);
        const float* row = W + size_t(o) * in;
        for (int k = 0; k < in; ++k) {
            const float wh = Bf16Round(row[k]), xh = Bf16Round(x[k]);
            const float wl = Bf16Round(row[k] - wh), xl = Bf16Round(x[k] - xh);
            sum += double(wh) * double(xh) + double(wh) * double(xl) + double(wl) * double(xh);
        }
        const float v = float(sum);
        y[o] = (relu && v < 0.0f) ? 0.0f : v;
    }
}

bool PolicyForwardT16(const float* g, const Layout& L, const Grid& grid, float* policy) {
    if (grid.tc > 5 || IsFinished(grid)) return false;
    float x[kModelIn];
    for (int k = 0; k < kModelIn; ++k) x[k] = 0.0f;
    x[0] = grid.tc == 0 ? 1.0f : 0.0f;
    for (int t = 0; t < grid.tc; ++t) EncodeTurn<float>(g, L, grid.guess[t], grid.tile[t], x + 1 + 64 * t);
    float h0[kH0], h1[kH1];
    DenseApply<float>(g + L.w0, g + L.b0, kModelIn, kH0, x, h0, true);
    DenseBf16(g + L.w1, g + L.b1, kH0, kH1, h0, h1, true);
    DenseBf16(g + L.w2, g + L.b2, kH1, kPol, h1, policy, false);
    return true;
}

// Fixed word features: letter count, or -1 when absent (output_embedding.hpp:59-73).
void FixedFeatures(const uint8_t* w, float* f) {
    int counts[kAlpha] = {0};
    for (int p = 0; p < 5; ++p) ++counts[w[p]];
    for (int l = 0; l < kAlpha; ++l) f[l] = counts[l] == 0 ? -1.0f : float(counts[l]);
}

template <typename T>
T ScoreAction(const T* policy, const float* fixed, const float* tail) {
    T s = T(0);
    for (int l = 0; l < kFixed; ++l) s += policy[l] * T(fixed[l]);
    for (int j = 0; j < kTail; ++j) s += policy[kFixed + j] * T(tail[j]);
    return s;
}

bool AlreadyGuessed(const Grid& g, const uint8_t* w) {
    for (int t = 0; t < g.tc; ++t) if (SameWord(g.guess[t], w)) return true;
    return false;
}

// Masked argmax, ties to the lowest index (inference/dynamic_policy.hpp:98-147).  Also reports the
// runner-up score over unmasked actions (the top-2 margin is best - second).
template <typename T>
bool SelectAction(const T* policy, const float* tails, const uint8_t* catalog, const float* fixed_table,
                  uint64_t A, const Grid& grid, uint64_t* best_index, T* best_score, T* second_score) {
    bool have = false, have2 = false;
    T best = T(0), second = T(0);
    uint64_t bi = 0;
    for (uint64_t a = 0; a < A; ++a) {
        const uint8_t* w = catalog + 5 * a;
        if (AlreadyGuessed(grid, w)) continue;
        const T s = ScoreAction<T>(policy, fixed_table + kFixed * a, tails + kTail * a);
        if (!have || s > best || (s == best && a < bi)) {
            if (have) { if (!have2 || best > second) second = best; have2 = true; }
            best = s; bi = a; have = true;
        } else if (!have2 || s > second) {
            second = s; have2 = true;
        }
    }
    if (!have) return false;
    *best_index = bi;
    *best_score = best;
    *second_score = have2 ? second : T(-INFINITY);
    return true;
}

// Episode initialisation (evaluation_ops.cuh:171-237): variant 0 starts empty; variants 1 and 2 start with
// the cell's next two secrets (entry+1, entry+2) or the two after them (entry+3, entry+4) already played.
// A genome's 3W games are the W fresh ones (slot = entry), then the prefilled pairs (slot W + 2 entry + v-1).
int InitEpisode(const uint8_t* catalog, const uint32_t* local, uint64_t W, uint64_t entry, int variant, Grid& g) {
    std::memset(&g, 0, sizeof(g));
    std::memcpy(g.sol, catalog + 5 * size_t(local[entry]), 5);
    if (variant == 0) return kOk;
    const uint64_t first = variant == 1 ? entry + 1 : entry + 3;
    const uint64_t second = variant == 1 ? entry + 2 : entry + 4;
    if (!Append(g, catalog + 5 * size_t(local[first % W])) || !Append(g, catalog + 5 * size_t(local[second % W])))
        return kGuessAppendFailed;
    return kOk;
}

float EpisodeScore(const Grid& g) {
    return IsWon(g) ? (10.0f + float(kMaxTurns - g.tc)) : 0.0f;  // evaluation_ops.cuh:239-242
}

// Normalised fitness (evaluation_ops.cuh:87-111).
float NormalizeFitness(float raw, uint64_t W) {
    const float maximum = 3.0f * 15.0f * float(W);
    if (maximum <= 0.0f) return 1.40129846e-45f;
    const float f = raw / maximum;
    if (f >= 1.0f) return 1.0f;
    return f > 1.40129846e-45f ? f : 1.40129846e-45f;
}

template <typename F>
void ParallelFor(uint64_t n, int threads, F&& fn) {
    if (threads <= 1 || n <= 1) {
        for (uint64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            for (uint64_t i = next.fetch_add(1); i < n; i = next.fetch_add(1)) fn(i);
        });
    for (auto& th : pool) th.join();
}

// ------------------------------------------------------------------------------------------------
// Toroidal grid (src/spatial/grid.hpp:125-218).
uint64_t Wrap(int64_t c, uint64_t side) {
    const int64_t m = int64_t(side);
    int64_t w = c % m;
    if (w < 0) w += m;
    return uint64_t(w);
}

uint64_t AxisDistance(uint64_t a, uint64_t b, uint64_t side) {
    const uint64_t fwd = a >= b ? a - b : b - a;
    const uint64_t wrapped = side - fwd;
    return fwd < wrapped ? fwd : wrapped;
}

uint64_t Chebyshev(uint64_t rows, uint64_t cols, uint64_t a, uint64_t b) {
    const uint64_t dr = AxisDistance(a / cols, b / cols, rows);
    const uint64_t dc = AxisDistance(a % cols, b % cols, cols);
    return dr > dc ? dr : dc;
}

int MooreNeighbors(uint64_t rows, uint64_t cols, uint64_t focal, uint64_t radius, uint64_t* out) {
    int count = 0;
    const int64_t fr = int64_t(focal / cols), fc = int64_t(focal % cols);
    const int64_t r = int64_t(radius);
    for (int64_t dr = -r; dr <= r; ++dr)
        for (int64_t dc = -r; dc <= r; ++dc) {
            const uint64_t idx = Wrap(fr + dr, rows) * cols + Wrap(fc + dc, cols);
            bool dup = false;
            for (int i = 0; i < count; ++i) if (out[i] == idx) { dup = true; break; }
            if (!dup) out[count++] = idx;
        }
    return count;
}

// Rank-roulette selection (device/selection_ops.cuh:26-263).
uint64_t SplitMix(uint64_t v) {
    v += 0x9E3779B97F4A7C15ull;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}

bool RanksBefore(float lf, uint64_t li, float rf, uint64_t ri, uint64_t seed) {
    if (lf < rf) return true;
    if (lf > rf) return false;
    const uint64_t lk = SplitMix(seed ^ li), rk = SplitMix(seed ^ ri);
    return lk < rk || (lk == rk && li < ri);
}

constexpr uint64_t kNone = ~uint64_t(0);

bool SelectOne(const float* fit, const uint8_t* has, const uint64_t* cand, uint64_t ncand, const float* wtable,
               uint64_t seed, float u, uint64_t excluded, uint64_t* chosen) {
    auto rank_of = [&](uint64_t i) {
        uint64_t rank = 1;
        for (uint64_t c = 0; c < ncand; ++c) {
            const uint64_t j = cand[c];
            if (j == i || j == excluded || has[j] == 0) continue;
            if (RanksBefore(fit[j], j, fit[i], i, seed)) ++rank;
        }
        return rank;
    };
    float total = 0.0f;
    bool found = false;
    uint64_t fallback = 0;
    for (uint64_t c = 0; c < ncand; ++c) {
        const uint64_t i = cand[c];
        if (i == excluded || has[i] == 0) continue;
        total += wtable[rank_of(i)];
        fallback = i;
        found = true;
    }
    if (!found || total <= 0.0f) return false;
    const float threshold = u * total;
    float cumulative = 0.0f;
    for (uint64_t c = 0; c < ncand; ++c) {
        const uint64_t i = cand[c];
        if (i == excluded || has[i] == 0) continue;
        cumulative += wtable[rank_of(i)];
        if (threshold <= cumulative) { *chosen = i; return true; }
    }
    *chosen = fallback;
    return true;
}

// ------------------------------------------------------------------------------------------------
// Crossover decisions per child (genome_ops.cuh:210-221,264-299,344-376), drawn from the
// counter-based stream: site 0 = subsystem flips, site 1+L = dense layer L, site 8/9 = tail rows.
struct XoverRecord {
    uint32_t layer_first[6];   // donor of each dense layer: 1 = first parent
    uint32_t neuron[6];        // neuron copied from the other parent, or 0xFFFFFFFF
    uint32_t tail_first;       // donor of all tail rows
    uint32_t row2;             // level-2 swapped row, or 0xFFFFFFFF
    uint32_t splice_row;       // level-3 spliced row, or 0xFFFFFFFF
    uint32_t splice_point;     // features below the point come from the splice row's source
};

void XoverDecisions(const Layout& L, uint64_t A_parent, uint32_t seed, uint32_t gen, uint32_t child, float t1,
                    float t2, float t3, XoverRecord* rec) {
    DenseLayer layers[6];
    const int nl = DenseLayers(L, layers);
    uint32_t w[4];
    Draw(seed, kTagXovr, child, gen, 0, 0, w);
    const bool enc_first = !Bernoulli(t1, w[0]);
    const bool trunk_first = !Bernoulli(t1, w[1]);
    const bool tail_first = !Bernoulli(t1, w[2]);
    for (int i = 0; i < 6; ++i) { rec->layer_first[i] = 0; rec->neuron[i] = 0xFFFFFFFFu; }
    for (int l = 0; l < nl; ++l) {
        uint32_t d[4];
        Draw(seed, kTagXovr, child, gen, 1 + uint32_t(l), 0, d);
        const bool base = layers[l].encoder ? enc_first : trunk_first;
        const bool lf = Bernoulli(t2, d[0]) ? !base : base;
        rec->layer_first[l] = lf ? 1u : 0u;
        if (Bernoulli(t3, d[1])) rec->neuron[l] = d[2] % uint32_t(layers[l].out);
    }
    rec->tail_first = tail_first ? 1u : 0u;
    rec->row2 = rec->splice_row = 0xFFFFFFFFu;
    rec->splice_point = 0;
    uint32_t d[4], e[4];
    Draw(seed, kTagXovr, child, gen, 8, 0, d);
    Draw(seed, kTagXovr, child, gen, 9, 0, e);
    if (A_parent > 0 && Bernoulli(t2, d[0])) rec->row2 = uint32_t(d[1] % A_parent);
    if (A_parent > 0 && Bernoulli(t3, d[2])) {
        rec->splice_row = uint32_t(d[3] % A_parent);
        rec->splice_point = 1u + (e[0] % uint32_t(kTail - 1));
    }
}

// Donor of element e (true = first parent), given the decisions.
bool DonorIsFirst(const Layout& L, const XoverRecord& r, uint64_t e) {
    DenseLayer layers[6];
    const int nl = DenseLayers(L, layers);
    if (e < L.P) {
        for (int l = 0; l < nl; ++l) {
            const DenseLayer& d = layers[l];
            uint64_t o;
            if (e >= d.w && e < d.w + uint64_t(d.in) * d.out) o = (e - d.w) / uint64_t(d.in);
            else if (e >= d.b && e < d.b + uint64_t(d.out)) o = e - d.b;
            else continue;
            const bool lf = r.layer_first[l] != 0;
            return (r.neuron[l] != 0xFFFFFFFFu && o == r.neuron[l]) ? !lf : lf;
        }
        return true;
    }
    const uint64_t t = e - L.P;
    const uint64_t a = t / kTail, j = t % kTail;
    const bool base = r.tail_first != 0;
    bool src = base;
    if (r.row2 != 0xFFFFFFFFu && a == r.row2) src = !base;
    if (r.splice_row != 0xFFFFFFFFu && a == r.splice_row) {
        const bool row_src = (r.row2 != 0xFFFFFFFFu && a == r.row2) ? !base : base;
        src = j < r.splice_point ? row_src : !row_src;
    }
    return src;
}

struct MutationCfg { float p, sigma, p_scale; };

// BreedAndMutateGenome (genome_ops.cuh:386-407) on the counter-based draws.  The child receives
// P + A_parent*38 values from the parents; tail rows [A_parent, A_child) stay zero.
void BreedChild(const Layout& L, const float* p1, const float* p2, float* child, uint64_t A_parent,
                uint64_t A_child, uint32_t seed, uint32_t gen, uint32_t ci, float t1, float t2, float t3,
                const MutationCfg& m, XoverRecord* rec_out) {
    XoverRecord r{};
    XoverDecisions(L, A_parent, seed, gen, ci, t1, t2, t3, &r);
    if (rec_out) *rec_out = r;
    const uint64_t n = L.P + A_parent * kTail;
    const bool mutate = m.p > 0.0f && m.sigma > 0.0f;
    for (uint64_t e = 0; e < n; ++e) {
        float v = DonorIsFirst(L, r, e) ? p1[e] : p2[e];
        if (mutate) {
            uint32_t w[4];
            Draw(seed, kTagMuta, uint32_t(e >> 2), ci, gen, 0, w);
            if (Bernoulli(m.p, w[e & 3])) {
                uint32_t z[4];
                Draw(seed, kTagNorm, uint32_t(e), ci, gen, 0, z);
                v = std::fmaf(m.sigma, NormalFromWords(z[0], z[1]), v);
            }
        }
        child[e] = v;
    }
    if (m.p_scale > 0.0f) {
        for (uint64_t a = 0; a < A_parent; ++a) {
            uint32_t w[4];
            Draw(seed, kTagRscl, uint32_t(a), ci, gen, 0, w);
            if (!Bernoulli(m.p_scale, w[0])) continue;
            const float scale = Bernoulli(0.5f, w[1]) ? 1.02f : 0.98f;  // output_tail_ops.hpp:7-12
            float* row = child + L.P + a * kTail;
            for (int j = 0; j < kTail; ++j) row[j] = row[j] * scale;
        }
    }
    for (uint64_t e = n; e < L.P + A_child * kTail; ++e) child[e] = 0.0f;
}

// ------------------------------------------------------------------------------------------------
// Hint grids (src/wordle/hint_grid.hpp:19-164): grid 0 appends every distinct left rotation of the
// word whose feedback is all yellow; grids 1 and 2 each hold one swap of two differing positions
// whose feedback is 3 green + 2 yellow, the first two distinct such swaps in (i<j) order.
bool BuildHintGroup(const uint8_t* w, Grid grids[3]) {
    for (int i = 0; i < 3; ++i) std::memset(&grids[i], 0, sizeof(Grid));
    std::memcpy(grids[0].sol, w, 5);
    uint8_t seen[4][5];
    int nseen = 0;
    for (int rot = 1; rot < 5; ++rot) {
        uint8_t g[5];
        for (int p = 0; p < 5; ++p) g[p] = w[(p + rot) % 5];
        if (SameWord(g, w)) continue;
        bool dup = false;
        for (int k = 0; k < nseen; ++k) dup |= SameWord(seen[k], g);
        if (dup) continue;
        uint8_t fb[5];
        Feedback(g, w, fb);
        bool all_yellow = true;
        for (int p = 0; p < 5; ++p) all_yellow &= fb[p] == 1;
        if (!all_yellow) continue;
        if (!Append(grids[0], g)) return false;
        std::memcpy(seen[nseen++], g, 5);
    }
    uint8_t chosen[2][5];
    int nchosen = 0;
    for (int a = 0; a < 5; ++a)
        for (int b = a + 1; b < 5; ++b) {
            if (w[a] == w[b]) continue;
            uint8_t g[5];
            std::memcpy(g, w, 5);
            std::swap(g[a], g[b]);
            uint8_t fb[5];
            Feedback(g, w, fb);
            int greens = 0, yellows = 0;
            for (int p = 0; p < 5; ++p) { greens += fb[p] == 0; yellows += fb[p] == 1; }
            if (greens != 3 || yellows != 2) continue;
            bool dup = false;
            for (int k = 0; k < nchosen; ++k) dup |= SameWord(chosen[k], g);
            if (dup) continue;
            Grid& grid = grids[1 + nchosen];
            std::memcpy(grid.sol, w, 5);
            if (!Append(grid, g)) return false;
            std::memcpy(chosen[nchosen++], g, 5);
            if (nchosen == 2) return true;
        }
    return false;
}

// TrainableTailNorm (output_embedding_injection.hpp:70-79): sequential sum of squares, sqrt.
float TailNorm(const float* t) {
    float s = 0.0f;
    for (int j = 0; j < kTail; ++j) s += t[j] * t[j];
    return std::sqrt(s);
}

// TryComputeMedianTailNorm (output_embedding_injection.hpp:80-104): the upper-middle order
// statistic, averaged with the lower-middle one for an even count.
float MedianTailNorm(const float* tails, uint64_t count) {
    std::vector<float> n(count);
    for (uint64_t a = 0; a < count; ++a) n[a] = TailNorm(tails + a * kTail);
    std::sort(n.begin(), n.end());
    const uint64_t up = count / 2;
    if (count % 2) return n[up];
    return 0.5f * (n[up - 1] + n[up]);
}

}  // namespace

// ================================================================================================
// C ABI of the oracle (ctypes from tests/ and bench.py's CPU-baseline leg only).
extern "C" {

void oro_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { Philox(ctr, key, out); }
float oro_u01(uint32_t x) { return U01(x); }
float oro_normal_from_words(uint32_t a, uint32_t b) { return NormalFromWords(a, b); }
double oro_det_log(double x) { return DetLog(x); }
double oro_det_sqrt(double x) { return DetSqrt(x); }

uint64_t oro_policy_param_count(int enc_hidden) { return MakeLayout(enc_hidden).P; }

void oro_feedback(const uint8_t* guess, const uint8_t* sol, uint8_t* fb) { Feedback(guess, sol, fb); }

// Forward one grid state.  letters/tiles: tc*5 bytes.  Writes fp32 sequential policy to out32
// and/or fp64 to out64 (either may be null).  Returns 0 or a status code.
// A grid from a played board: the solution is the guess of an all-green turn when there is one, else a
// word none of the guesses equals (so the board counts as won exactly when a turn is all green).
Grid BoardGrid(const uint8_t* letters, const uint8_t* tiles, int tc) {
    Grid g{};
    g.tc = tc;
    for (int t = 0; t < tc; ++t) {
        std::memcpy(g.guess[t], letters + 5 * t, 5);
        std::memcpy(g.tile[t], tiles + 5 * t, 5);
    }
    for (uint8_t l = 25;; --l) {  // 26 candidate words, at most 6 guesses: one is free
        std::memset(g.sol, l, 5);
        bool used = false;
        for (int t = 0; t < tc; ++t) used |= SameWord(g.guess[t], g.sol);
        if (!used) break;
    }
    for (int t = 0; t < tc; ++t) {
        bool all_green = true;
        for (int p = 0; p < 5; ++p) all_green &= tiles[5 * t + p] == 0;
        if (all_green) std::memcpy(g.sol, letters + 5 * t, 5);
    }
    return g;
}

int oro_policy_forward(const float* genome, int enc_hidden, const uint8_t* letters, const uint8_t* tiles, int tc,
                       float* out32, double* out64) {
    const Layout L = MakeLayout(enc_hidden);
    const Grid g = BoardGrid(letters, tiles, tc);
    if (out32 && !PolicyForward<float>(genome, L, g, out32)) return kPolicyForwardFailed;
    if (out64 && !PolicyForward<double>(genome, L, g, out64)) return kPolicyForwardFailed;
    return kOk;
}

void oro_fixed_features(const uint8_t* word, float* out) { FixedFeatures(word, out); }

// The BF16-trunk forward of one board (PolicyForwardT16) and the rounding it uses.
int oro_policy_forward_t16(const float* genome, int enc_hidden, const uint8_t* letters, const uint8_t* tiles, int tc,
                           float* out32) {
    const Layout L = MakeLayout(enc_hidden);
    const Grid g = BoardGrid(letters, tiles, tc);
    return PolicyForwardT16(genome, L, g, out32) ? kOk : kPolicyForwardFailed;
}
void oro_bf16_round(const float* in, float* out, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) out[i] = Bf16Round(in[i]);
}

// The next guess for a played board (TrySelectNextGuessWithSingleModelDeviceRuntime,
// single_model_device_runtime.cu:122-181; dynamic_policy.hpp:62-150): the policy forward and the masked
// argmax over the first A catalog words, in fp32 and in fp64.  Returns kPolicyForwardFailed for an
// invalid or finished board, kActionSelectionFailed when every action was guessed.
int oro_select_guess_layout(const float* genome, int enc_hidden, uint64_t A, const uint8_t* catalog,
                            const uint8_t* letters, const uint8_t* tiles, int tc, uint64_t* index32, float* best32,
                            float* second32, uint64_t* index64, double* best64, double* second64) {
    if (tc < 0 || tc > kMaxTurns) return kPolicyForwardFailed;
    for (int i = 0; i < 5 * tc; ++i)
        if (letters[i] >= kAlpha || tiles[i] > 2) return kPolicyForwardFailed;
    const Layout L = MakeLayout(enc_hidden);
    const Grid g = BoardGrid(letters, tiles, tc);
    float pol32[kPol];
    double pol64[kPol];
    if (!PolicyForward<float>(genome, L, g, pol32) || !PolicyForward<double>(genome, L, g, pol64))
        return kPolicyForwardFailed;
    std::vector<float> fixed(kFixed * A);
    for (uint64_t a = 0; a < A; ++a) FixedFeatures(catalog + 5 * a, fixed.data() + kFixed * a);
    const float* tails = genome + L.P;
    if (!SelectAction<float>(pol32, tails, catalog, fixed.data(), A, g, index32, best32, second32) ||
        !SelectAction<double>(pol64, tails, catalog, fixed.data(), A, g, index64, best64, second64))
        return kActionSelectionFailed;
    return kOk;
}

int oro_select_guess(const float* genome, uint64_t A, const uint8_t* catalog, const uint8_t* letters,
                     const uint8_t* tiles, int tc, uint64_t* index32, float* best32, float* second32,
                     uint64_t* index64, double* best64, double* second64) {
    return oro_select_guess_layout(genome, 0, A, catalog, letters, tiles, tc, index32, best32, second32, index64,
                                   best64, second64);
}

// One turn through the shared encoder (shared_encoder.hpp:94-115).
void oro_encode_turn(const float* genome, int enc_hidden, const uint8_t* letters, const uint8_t* tiles, float* out32,
                     double* out64) {
    const Layout L = MakeLayout(enc_hidden);
    if (out32) EncodeTurn<float>(genome, L, letters, tiles, out32);
    if (out64) EncodeTurn<double>(genome, L, letters, tiles, out64);
}

float oro_score_action(const float* policy, const float* fixed, const float* tail) {
    return ScoreAction<float>(policy, fixed, tail);
}

// Trace of one game (slot order per genome: W fresh games by entry, then 2W prefilled games,
// slot W+q -> entry q/2, variant 1+(q%2)).
struct OroGameTrace {
    uint32_t guesses[6];
    float best[6];
    float second[6];
    uint32_t decisions;
    uint32_t turn_count;
    float score;
    uint32_t won;
};

// Evaluate a population (TryEvaluateCurrentGenerationFitnessOnDevice semantics,
// slab_runtime.cu:1086-1175 with the per-cell CSR of secrets instead of the shard walk).
// genomes: pointers to each individual's fp32 genome (stride-agnostic).  mode 0 = fp32 sequential,
// 1 = fp64, 2 = BF16 trunk (PolicyForwardT16) with fp32 sequential scoring.  traces (optional): per game, indexed by game_offsets[n] + slot.
int oro_evaluate(const float* const* genomes, uint64_t N, int enc_hidden, uint64_t A_genome, uint64_t A_runtime,
                 const uint8_t* catalog, uint64_t n_catalog, const uint32_t* cell_offsets, const uint32_t* secret_ids,
                 int mode, int threads, float* fitness, float* raw, uint32_t* local_w, uint64_t* decisions,
                 const uint64_t* game_offsets, OroGameTrace* traces) {
    const Layout L = MakeLayout(enc_hidden);
    const uint64_t A = A_genome < A_runtime ? A_genome : A_runtime;
    if (A == 0 || A > n_catalog) return kInvalidRuntimeConfig;
    std::vector<float> fixed(size_t(n_catalog) * kFixed);
    for (uint64_t a = 0; a < n_catalog; ++a) FixedFeatures(catalog + 5 * a, fixed.data() + kFixed * a);
    std::atomic<int> status{kOk};
    std::atomic<uint64_t> total_decisions{0};
    ParallelFor(N, threads, [&](uint64_t n) {
        const uint64_t W = cell_offsets[n + 1] - cell_offsets[n];
        if (W == 0) { int z = kOk; status.compare_exchange_strong(z, kInvalidTrainingShard); return; }
        const uint32_t* local = secret_ids + cell_offsets[n];
        const float* g = genomes[n];
        const float* tails = g + L.P;
        float sum = 0.0f;
        uint64_t dec = 0;
        for (uint64_t slot = 0; slot < 3 * W; ++slot) {
            const uint64_t entry = slot < W ? slot : (slot - W) / 2;
            const int variant = slot < W ? 0 : 1 + int((slot - W) % 2);
            Grid grid;
            int st = InitEpisode(catalog, local, W, entry, variant, grid);
            OroGameTrace tr{};
            for (int i = 0; i < 6; ++i) { tr.guesses[i] = 0xFFFFFFFFu; tr.best[i] = tr.second[i] = 0.0f; }
            while (st == kOk && !IsFinished(grid)) {
                uint64_t bi = 0;
                if (mode == 0 || mode == 2) {
                    float pol[kPol], b, s2;
                    if (!(mode == 0 ? PolicyForward<float>(g, L, grid, pol) : PolicyForwardT16(g, L, grid, pol))) {
                        st = kPolicyForwardFailed; break;
                    }
                    if (!SelectAction<float>(pol, tails, catalog, fixed.data(), A, grid, &bi, &b, &s2)) {
                        st = kActionSelectionFailed; break;
                    }
                    tr.best[tr.decisions] = b; tr.second[tr.decisions] = s2;
                } else {
                    double pol[kPol], b, s2;
                    if (!PolicyForward<double>(g, L, grid, pol)) { st = kPolicyForwardFailed; break; }
                    if (!SelectAction<double>(pol, tails, catalog, fixed.data(), A, grid, &bi, &b, &s2)) {
                        st = kActionSelectionFailed; break;
                    }
                    tr.best[tr.decisions] = float(b); tr.second[tr.decisions] = float(s2);
                }
                tr.guesses[tr.decisions] = uint32_t(bi);
                ++tr.decisions;
                ++dec;
                if (!Append(grid, catalog + 5 * bi)) st = kGuessAppendFailed;
            }
            if (st != kOk) { int z = kOk; status.compare_exchange_strong(z, st); return; }
            const float sc = EpisodeScore(grid);
            tr.turn_count = uint32_t(grid.tc);
            tr.score = sc;
            tr.won = IsWon(grid) ? 1u : 0u;
            if (traces) traces[game_offsets[n] + slot] = tr;
            sum += sc;
        }
        if (raw) raw[n] = sum;
        if (local_w) local_w[n] = uint32_t(W);
        fitness[n] = NormalizeFitness(sum, W);
        total_decisions += dec;
    });
    if (decisions) *decisions = total_decisions.load();
    return status.load();
}

// Replay one game with given guesses (teacher forcing) and report, per decision, the fp64 policy,
// the fp64 score of every unmasked action's argmax and runner-up, and the fp64 score of the given
// guess.  Used to measure the GPU forward error on exactly the states the GPU visited.
int oro_replay_game(const float* genome, int enc_hidden, uint64_t A, const uint8_t* catalog, const uint32_t* local,
                    uint64_t W, uint64_t slot, const uint32_t* guesses, uint32_t n_guesses, double* policy64,
                    double* best64, double* second64, uint32_t* argmax64, double* chosen64) {
    const Layout L = MakeLayout(enc_hidden);
    std::vector<float> fixed(size_t(A) * kFixed);
    for (uint64_t a = 0; a < A; ++a) FixedFeatures(catalog + 5 * a, fixed.data() + kFixed * a);
    const uint64_t entry = slot < W ? slot : (slot - W) / 2;
    const int variant = slot < W ? 0 : 1 + int((slot - W) % 2);
    Grid grid;
    int st = InitEpisode(catalog, local, W, entry, variant, grid);
    if (st != kOk) return st;
    for (uint32_t d = 0; d < n_guesses; ++d) {
        if (IsFinished(grid)) return kGuessAppendFailed;
        double pol[kPol];
        if (!PolicyForward<double>(genome, L, grid, pol)) return kPolicyForwardFailed;
        uint64_t bi = 0;
        double b = 0, s2 = 0;
        if (!SelectAction<double>(pol, genome + L.P, catalog, fixed.data(), A, grid, &bi, &b, &s2))
            return kActionSelectionFailed;
        for (int k = 0; k < kPol; ++k) policy64[d * kPol + k] = pol[k];
        best64[d] = b;
        second64[d] = s2;
        argmax64[d] = uint32_t(bi);
        chosen64[d] = ScoreAction<double>(pol, fixed.data() + kFixed * guesses[d], genome + L.P + kTail * guesses[d]);
        if (!Append(grid, catalog + 5 * size_t(guesses[d]))) return kGuessAppendFailed;
    }
    return kOk;
}

// Population summary (slab_runtime.cu:588-631).
void oro_summary(const float* fitness, uint64_t N, float* best, float* mean, uint64_t* best_index) {
    float b = 0.0f, sum = 0.0f;
    uint64_t bi = 0;
    for (uint64_t i = 0; i < N; ++i) {
        sum += fitness[i];
        if (i == 0 || fitness[i] > b || (fitness[i] == b && i < bi)) { b = fitness[i]; bi = i; }
    }
    *best = b;
    *mean = sum / float(N);
    *best_index = bi;
}

float oro_normalize_fitness(float raw, uint64_t W) { return NormalizeFitness(raw, W); }

// Rank weight table w[r] = rank^exponent (selection_ops.cuh:59-61), tabulated in double and rounded.
void oro_rank_weights(float exponent, uint64_t n, float* out) {
    out[0] = 0.0f;
    for (uint64_t r = 1; r <= n; ++r) out[r] = float(std::pow(double(r), double(exponent)));
}

// One rank-roulette draw (TrySelectRankWeightedIndexFromCandidates, selection_ops.cuh:115-161) with the
// tie-break seed and the uniform given directly, so tests can place the threshold exactly.
// excluded = UINT64_MAX for none.  Returns 1 and the chosen cell, or 0 when no candidate is eligible.
int oro_select_one(const float* fit, const uint8_t* has, const uint64_t* cand, uint64_t ncand, const float* wtable,
                   uint64_t tie_seed, float u, uint64_t excluded, uint64_t* chosen) {
    return SelectOne(fit, has, cand, ncand, wtable, tie_seed, u, excluded, chosen) ? 1 : 0;
}

// SampleBernoulli (genome_ops.cuh:35-44) on one word of the counter-based stream.
int oro_bernoulli(float p, uint32_t word) { return Bernoulli(p, word) ? 1 : 0; }

int oro_moore(uint64_t rows, uint64_t cols, uint64_t focal, uint64_t radius, uint64_t* out) {
    return MooreNeighbors(rows, cols, focal, radius, out);
}

uint64_t oro_chebyshev(uint64_t rows, uint64_t cols, uint64_t a, uint64_t b) { return Chebyshev(rows, cols, a, b); }

// Parent plan (BuildSlabAssemblyPlanKernel, slab_runtime.cu:633-675, and
// TrySelectCellularParentPairDevice, selection_ops.cuh:218-263).  wtable[r] = weight of rank r.
static int PlanChildren(const float* fit, const uint8_t* has, uint64_t cur_size, uint64_t child_count, uint64_t cols,
                        uint64_t gen_index, uint64_t radius, const float* wtable, uint32_t generation_seed,
                        uint32_t* pairs) {
    const uint64_t rows = cur_size / cols;
    const uint32_t seed = generation_seed ^ kPlanSeedXor;
    std::vector<uint64_t> cand(cur_size > 25 ? cur_size : 25);
    for (uint64_t c = 0; c < child_count; ++c) {
        const uint64_t focal = c % cur_size;
        uint64_t ncand = 0;
        if (radius <= 2) {
            ncand = uint64_t(MooreNeighbors(rows, cols, focal, radius, cand.data()));
        } else {
            for (uint64_t i = 0; i < cur_size; ++i)
                if (Chebyshev(rows, cols, focal, i) <= radius) cand[ncand++] = i;
        }
        uint64_t chosen[2] = {0, 0};
        for (int k = 0; k < 2; ++k) {
            uint32_t w[4];
            Draw(seed, kTagPlan, uint32_t(c), uint32_t(gen_index), uint32_t(k), 0, w);
            const uint64_t tie = (uint64_t(w[0]) << 32) ^ uint64_t(w[1]);
            if (!SelectOne(fit, has, cand.data(), ncand, wtable, tie, U01(w[2]), k == 0 ? kNone : chosen[0],
                           &chosen[k]))
                return kParentSelectionFailed;
        }
        pairs[2 * c] = uint32_t(chosen[0]);
        pairs[2 * c + 1] = uint32_t(chosen[1]);
    }
    return kOk;
}

int oro_plan(const float* fit, const uint8_t* has, uint64_t cur_size, uint64_t child_count, uint64_t cols,
             uint64_t gen_index, uint64_t radius, const float* wtable, uint32_t generation_seed, uint32_t* pairs) {
    if (cols == 0 || cur_size % cols || child_count % cols || child_count > cur_size || radius == 0)
        return kInvalidAssemblyConfig;
    return PlanChildren(fit, has, cur_size, child_count, cols, gen_index, radius, wtable, generation_seed, pairs);
}

// The brood of screened evolution (this build's extension, no reference counterpart): brood x cur_size
// children, child c belonging to cell c % cur_size and drawing with child index c -- for c < cur_size
// exactly oro_plan's child c.
int oro_plan_brood(const float* fit, const uint8_t* has, uint64_t cur_size, uint64_t brood, uint64_t cols,
                   uint64_t gen_index, uint64_t radius, const float* wtable, uint32_t generation_seed,
                   uint32_t* pairs) {
    if (cols == 0 || cur_size == 0 || cur_size % cols || brood == 0 || radius == 0) return kInvalidAssemblyConfig;
    return PlanChildren(fit, has, cur_size, brood * cur_size, cols, gen_index, radius, wtable, generation_seed,
                        pairs);
}

// Slot bookkeeping of one assembly (device_runtime.cu:1434-1557 with TryAllocateSlabSlot /
// TryReleaseSlabSlot / TryReleaseParentReference, slab_allocator.hpp:271-354,
// reference_counter.hpp:135-220) in the build's deterministic order: parents with no planned child
// are freed in index order; batch b gives child i the slot free_stack[free_count-1-i]; after the
// batch every parent whose count reached zero is freed in index order.  Returns a status and the
// number of batches.
int oro_assemble_slots(uint32_t slot_count, uint32_t* free_stack, uint32_t* free_count, uint32_t* slot_state,
                       uint32_t* cur_slots, uint32_t cur_size, const uint32_t* pairs, uint32_t child_count,
                       uint32_t* child_slots, uint32_t* batches_out, uint32_t spill_capacity, uint32_t* spilled_out) {
    std::vector<uint32_t> ref(cur_size, 0);
    for (uint32_t c = 0; c < child_count; ++c) {
        if (pairs[2 * c] >= cur_size || pairs[2 * c + 1] >= cur_size) return kInvalidParentIndex;
        ++ref[pairs[2 * c]];
        ++ref[pairs[2 * c + 1]];
    }
    auto collect = [&]() -> int {
        for (uint32_t i = 0; i < cur_size; ++i) {
            if (ref[i] != 0 || cur_slots[i] == 0xFFFFFFFFu) continue;
            const uint32_t sl = cur_slots[i];
            if (sl >= slot_count || slot_state[sl] == 0 || *free_count >= slot_count) return kInvalidSlab;
            slot_state[sl] = 0;
            free_stack[(*free_count)++] = sl;
            cur_slots[i] = 0xFFFFFFFFu;
        }
        return kOk;
    };
    int st = collect();
    if (st != kOk) return st;
    uint32_t offset = 0, batches = 0, spilled = 0;
    while (offset < child_count) {
        const uint32_t fc = *free_count;
        if (fc == 0) {
            // one host spill of the assembled prefix (slab_runtime.cu:1554-1640): its slots go back onto
            // the free stack in child order
            if (spilled != 0 || offset == 0 || spill_capacity == 0) return kSlabFull;
            if (offset > spill_capacity) return 1;  // kCudaFailure
            for (uint32_t i = 0; i < offset; ++i) {
                slot_state[child_slots[i]] = 0;
                free_stack[(*free_count)++] = child_slots[i];
                child_slots[i] = 0xFFFFFFFFu;
            }
            spilled = offset;
            continue;
        }
        const uint32_t batch = std::min(child_count - offset, fc);
        for (uint32_t i = 0; i < batch; ++i) {
            const uint32_t sl = free_stack[fc - 1 - i];
            child_slots[offset + i] = sl;
            slot_state[sl] = 1;
        }
        *free_count = fc - batch;
        for (uint32_t i = 0; i < batch; ++i) {
            --ref[pairs[2 * (offset + i)]];
            --ref[pairs[2 * (offset + i) + 1]];
        }
        st = collect();
        if (st != kOk) return st;
        offset += batch;
        ++batches;
    }
    if (spilled != 0) {  // the spilled prefix takes slots again, child i from the top of the stack down
        const uint32_t fc = *free_count;
        if (fc < spilled) return 1;
        for (uint32_t i = 0; i < spilled; ++i) {
            child_slots[i] = free_stack[fc - 1 - i];
            slot_state[child_slots[i]] = 1;
        }
        *free_count = fc - spilled;
    }
    if (batches_out) *batches_out = batches;
    if (spilled_out) *spilled_out = spilled;
    return kOk;
}

// Widening repack of the slot bookkeeping (repacking.hpp:104-287; device_runtime.cu:978-1033):
// parents without a planned child are freed, the referenced ones keep their slot order and move to
// slots [S'-k, S') of the wider layout, the free stack becomes [S'-k-1 ... 0].  moves[r] receives the
// old slot of survivor r.  Returns a status.
int oro_repack_slots(uint32_t new_slot_count, uint32_t* cur_slots, uint32_t cur_size, const uint32_t* pairs,
                     uint32_t child_count, uint32_t* slot_state, uint32_t* free_stack, uint32_t* free_count,
                     uint32_t* moves, uint32_t* survivors_out) {
    std::vector<uint32_t> ref(cur_size, 0);
    for (uint32_t c = 0; c < child_count; ++c) {
        ++ref[pairs[2 * c]];
        ++ref[pairs[2 * c + 1]];
    }
    std::vector<std::pair<uint32_t, uint32_t>> surv;  // (old slot, parent)
    for (uint32_t i = 0; i < cur_size; ++i) {
        if (ref[i] == 0) { cur_slots[i] = 0xFFFFFFFFu; continue; }
        surv.emplace_back(cur_slots[i], i);
    }
    std::sort(surv.begin(), surv.end());
    const uint32_t k = uint32_t(surv.size());
    if (k == 0 || k > new_slot_count) return kSlabRepackFailed;
    const uint32_t base = new_slot_count - k;
    for (uint32_t r = 0; r < k; ++r) {
        moves[r] = surv[r].first;
        cur_slots[surv[r].second] = base + r;
    }
    for (uint32_t sl = 0; sl < new_slot_count; ++sl) slot_state[sl] = sl >= base ? 1u : 0u;
    for (uint32_t f = 0; f < base; ++f) free_stack[f] = base - 1 - f;
    *free_count = base;
    *survivors_out = k;
    return kOk;
}

void oro_xover_decisions(int enc_hidden, uint64_t A_parent, uint32_t seed, uint32_t gen, uint32_t child, float t1,
                         float t2, float t3, uint32_t* rec_words) {
    XoverRecord r{};
    XoverDecisions(MakeLayout(enc_hidden), A_parent, seed, gen, child, t1, t2, t3, &r);
    std::memcpy(rec_words, &r, sizeof(r));
}

uint64_t oro_xover_record_words() { return sizeof(XoverRecord) / 4; }

void oro_breed_child(int enc_hidden, const float* p1, const float* p2, float* child, uint64_t A_parent,
                     uint64_t A_child, uint32_t seed, uint32_t gen, uint32_t ci, float t1, float t2, float t3, float p,
                     float sigma, float p_scale, uint32_t* rec_words) {
    XoverRecord r{};
    BreedChild(MakeLayout(enc_hidden), p1, p2, child, A_parent, A_child, seed, gen, ci, t1, t2, t3,
               MutationCfg{p, sigma, p_scale}, &r);
    if (rec_words) std::memcpy(rec_words, &r, sizeof(r));
}

int oro_hint_group(const uint8_t* word, uint32_t* turn_counts, uint8_t* letters, uint8_t* tiles) {
    Grid g[3];
    if (!BuildHintGroup(word, g)) return 0;
    for (int i = 0; i < 3; ++i) {
        turn_counts[i] = uint32_t(g[i].tc);
        for (int t = 0; t < g[i].tc && t < 5; ++t)
            for (int p = 0; p < 5; ++p) {
                letters[(i * 5 + t) * 5 + p] = g[i].guess[t][p];
                tiles[(i * 5 + t) * 5 + p] = g[i].tile[t][p];
            }
    }
    return 1;
}

float oro_median_tail_norm(const float* tails, uint64_t count) { return MedianTailNorm(tails, count); }

// Output-embedding injection into one child (TryInjectExpandedOutputEmbeddingTailsConcurrently,
// injection_ops.cuh:21-64; TrySeedOutputEmbeddingTailFromHintGridsConcurrently and
// TryScaleTrainableTailToNormConcurrently, output_embedding_injection.hpp:107-205): for each new word,
// the mean over its 3 hint grids of policy dims 26..63, rescaled to the median norm of the child's
// existing rows.  mode 0 = fp32 sequential forward, 1 = fp64 forward.  Returns a status.
int oro_inject(const float* genome_in, float* genome_out, int enc_hidden, uint64_t parent_a,
               const uint8_t* catalog, uint64_t first_word, uint64_t count, int mode) {
    const Layout L = MakeLayout(enc_hidden);
    if (parent_a == 0 || count == 0 || first_word != parent_a) return kOutputEmbeddingInjectionFailed;
    const float target = MedianTailNorm(genome_in + L.P, parent_a);
    for (uint64_t k = 0; k < count; ++k) {
        Grid grids[3];
        if (!BuildHintGroup(catalog + 5 * (first_word + k), grids)) return kOutputEmbeddingInjectionFailed;
        float sums[kTail];
        for (int f = 0; f < kTail; ++f) sums[f] = 0.0f;
        for (int gi = 0; gi < 3; ++gi) {
            float pol[kPol];
            if (mode == 0) {
                if (!PolicyForward<float>(genome_in, L, grids[gi], pol)) return kOutputEmbeddingInjectionFailed;
            } else {
                double p64[kPol];
                if (!PolicyForward<double>(genome_in, L, grids[gi], p64)) return kOutputEmbeddingInjectionFailed;
                for (int d = 0; d < kPol; ++d) pol[d] = float(p64[d]);
            }
            for (int f = 0; f < kTail; ++f) sums[f] += pol[kFixed + f];
        }
        float* row = genome_out + L.P + (parent_a + k) * kTail;
        const float third = 1.0f / 3.0f;
        for (int f = 0; f < kTail; ++f) row[f] = sums[f] * third;
        float scale = 0.0f;
        if (target > 1.0e-6f) {
            const float cur = TailNorm(row);
            if (!(cur > 1.0e-6f)) return kOutputEmbeddingInjectionFailed;
            scale = target / cur;
        }
        for (int f = 0; f < kTail; ++f) row[f] = row[f] * scale;
    }
    return kOk;
}

// Generation-0 genome (InitializeRandomGenome, genome_ops.cuh:58-143) on the counter-based draws:
// weights = std_layer * z, biases 0, tails = tail_std * z.  weight_std[l] per dense layer.
void oro_bootstrap_genome(int enc_hidden, float* genome, uint64_t A, uint32_t seed, uint32_t gen, uint32_t organism,
                          const float* weight_std, float tail_std) {
    const Layout L = MakeLayout(enc_hidden);
    DenseLayer layers[6];
    const int nl = DenseLayers(L, layers);
    const uint64_t n = L.P + A * kTail;
    for (uint64_t e = 0; e < n; ++e) {
        float std = tail_std;
        bool bias = false;
        if (e < L.P) {
            for (int l = 0; l < nl; ++l) {
                if (e >= layers[l].w && e < layers[l].w + uint64_t(layers[l].in) * layers[l].out) { std = weight_std[l]; break; }
                if (e >= layers[l].b && e < layers[l].b + uint64_t(layers[l].out)) { bias = true; break; }
            }
        }
        if (bias) { genome[e] = 0.0f; continue; }
        uint32_t w[4];
        Draw(seed, kTagBoot, uint32_t(e >> 1), organism, gen, 0, w);
        const float z = (e & 1) ? NormalFromWords(w[2], w[3]) : NormalFromWords(w[0], w[1]);
        genome[e] = std * z;
    }
}

// The n-th output (1-based) of std::mt19937_64 seeded with `seed` (the engine MakeSamplePlan draws from,
// genetic_convergence.hpp:93).
uint64_t oro_mt19937_64_nth(uint64_t seed, uint64_t n) {
    std::mt19937_64 e(seed);
    uint64_t v = 0;
    for (uint64_t i = 0; i < n; ++i) v = e();
    return v;
}

// MakeSamplePlan after its seed mixing (genetic_convergence.hpp:59-117): the engine seeded with
// engine_seed, Floyd's sampling of the organism and the scalar indices (all of them, in order, when the
// request covers the universe), then `pair_count` ordered pairs of distinct sampled organisms.
int oro_sample_plan(uint64_t population, uint64_t scalars, uint64_t engine_seed, uint64_t organism_count,
                    uint64_t weight_count, uint64_t pair_count, uint32_t* organisms, uint32_t* weights,
                    uint32_t* pairs, uint64_t* counts) {
    std::mt19937_64 e(engine_seed);
    auto sample = [&e](uint64_t universe, uint64_t requested) {
        const uint64_t k = requested < universe ? requested : universe;
        std::vector<uint64_t> out(k);
        if (k == universe) {
            for (uint64_t i = 0; i < k; ++i) out[i] = i;
            return out;
        }
        std::unordered_set<uint64_t> seen;
        for (uint64_t j = 0, cursor = universe - k; cursor < universe; ++cursor, ++j) {
            const uint64_t c = std::uniform_int_distribution<uint64_t>(0, cursor)(e);
            out[j] = seen.count(c) != 0 ? cursor : c;
            seen.insert(out[j]);
        }
        return out;
    };
    const std::vector<uint64_t> org = sample(population, organism_count);
    const std::vector<uint64_t> w = sample(scalars, weight_count);
    counts[0] = org.size();
    counts[1] = w.size();
    counts[2] = 0;
    for (size_t i = 0; i < org.size(); ++i) organisms[i] = uint32_t(org[i]);
    for (size_t i = 0; i < w.size(); ++i) weights[i] = uint32_t(w[i]);
    if (org.size() < 2) return 0;
    std::uniform_int_distribution<uint64_t> d1(0, org.size() - 1), d2(0, org.size() - 2);
    for (uint64_t p = 0; p < pair_count; ++p) {
        const uint64_t a = d1(e);
        const uint64_t b0 = d2(e);
        pairs[2 * p] = uint32_t(a);
        pairs[2 * p + 1] = uint32_t(b0 >= a ? b0 + 1 : b0);
    }
    counts[2] = pair_count;
    return 0;
}

}  // extern "C"
