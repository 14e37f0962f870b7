// ne_kernels.hip -- HIP kernels of the GA generation step for gfx950 (CDNA4, wave64).
//
// Evaluation: persistent workgroups take (genome, split) items from per-XCD work queues; one wave plays
// one tile of 16 games (two in the paired kernels).  The dense contractions run in the transposed form
// (weights = A operand, games = B operand), so each layer's accumulator layout is directly the next
// layer's B operand: on v_mfma_f32_16x16x4_f32 in the FP32 form; the BF16 forms screen the actions on
// v_mfma_f32_16x16x32_bf16 and rescore candidates on the FP32 contraction, and the BF16-trunk form runs
// the trunk's layers 1 and 2 on split operands with v_mfma_f32_16x16x16_bf16.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ne_device.hpp"
#include "ne_kernels.hpp"

namespace ne {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ inline void SetStatus(int* status, int code) { atomicCAS(status, 0, code); }

__device__ inline v4f Mfma(float a, float b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// BF16 trunk (layers 1 and 2 of the NE_EVAL_BF16_TRUNK form) on split operands: every weight w and every
// ReLU'd input x is split into hi = bf16(v) and lo = bf16(v - hi) (round to nearest even; v - hi is exact in
// fp32), and w.x is taken as three products lo_w.hi_x + hi_w.lo_x + hi_w.hi_x on v_mfma_f32_16x16x16_bf16
// (FP32 accumulate; the dropped lo.lo term is ~2^-16 of the product).  The instruction takes, per lane, the
// 4 k-values that the four v_mfma_f32_16x16x4_f32 instructions r = 0..3 of the FP32 trunk take one by one
// (lane group g holds k-slots 4g..4g+3 of both operands), so the FP32 trunk's operand reads and accumulator
// layout serve it unchanged.
typedef short v4s __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
struct Split16 {
    v4s hi, lo;
};
__device__ __forceinline__ Split16 SplitBf16(float x0, float x1, float x2, float x3) {
    const bf16x4 h = {__bf16(x0), __bf16(x1), __bf16(x2), __bf16(x3)};
    const bf16x4 l = {__bf16(x0 - float(h[0])), __bf16(x1 - float(h[1])), __bf16(x2 - float(h[2])),
                      __bf16(x3 - float(h[3]))};
    return {__builtin_bit_cast(v4s, h), __builtin_bit_cast(v4s, l)};
}
__device__ __forceinline__ v4f MfmaT16(const v4f& w, const Split16& x, v4f c) {
    const Split16 ws = SplitBf16(w[0], w[1], w[2], w[3]);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ws.lo, x.hi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ws.hi, x.lo, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ws.hi, x.hi, c, 0, 0, 0);
}

__device__ inline v4f Load4(const float* p) { return *reinterpret_cast<const v4f*>(p); }

// Raw buffer loads: one 128-bit descriptor (SGPRs) per genome, a 32-bit per-lane byte offset and a
// uniform byte offset, so unrolled loops do not each hold a 64-bit per-lane address.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int kRsrcFlags = 0x00020000;

__device__ inline Rsrc MakeRsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, int(bytes), kRsrcFlags);
}
__device__ inline float BufF(Rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ inline v4f Buf4(Rsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

// LDS DMA (buffer_load ... lds): an LDS destination pointer, and the wait for a wave's transfers.
typedef __attribute__((address_space(3))) void* LdsPtr;
// vmcnt(0), leaving the other counters alone: this wave's chunk transfers have landed in LDS.
__device__ __forceinline__ void WaitChunk() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Register-only Wordle feedback (no per-letter count array): a non-green position p of letter L is
// yellow iff the number of earlier non-green positions of letter L is below the number of
// non-green solution positions holding L -- the same result as the reference's two passes
// (src/wordle/feedback.hpp:103-136).
__device__ inline uint32_t FeedbackRegs(uint32_t guess, uint32_t sol) {
    uint32_t g[5], s[5];
    bool green[5];
#pragma unroll
    for (int p = 0; p < 5; ++p) {
        g[p] = WordLetter(guess, p);
        s[p] = WordLetter(sol, p);
        green[p] = g[p] == s[p];
    }
    uint32_t fb = 0;
#pragma unroll
    for (int p = 0; p < 5; ++p) {
        if (green[p]) continue;
        int avail = 0, before = 0;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            avail += (!green[q] && s[q] == g[p]) ? 1 : 0;
            before += (q < p && !green[q] && g[q] == g[p]) ? 1 : 0;
        }
        fb |= (before < avail ? 1u : 2u) << (2 * p);
    }
    return fb;
}

// ---- trunk pieces shared by game evaluation and output-embedding injection -----------------
// Lane (j, gq) of a wave: j = lane%16 is the A-operand row / B-operand column, gq = lane/16 the
// k group.  Layer-0 accumulators acc[t] hold neuron rows 16t+4gq..16t+4gq+3 of game column j.
struct TrunkLane {
    Rsrc gr;
    uint32_t j, gq;
    uint32_t off_b0, off_virgin, off_fold, off_w1, off_b1;
};

__device__ __forceinline__ TrunkLane MakeTrunkLane(const float* genome, uint32_t j, uint32_t gq) {
    TrunkLane t;
    t.gr = MakeRsrc(genome, uint32_t(kPolicyParams * 4));
    t.j = j;
    t.gq = gq;
    t.off_b0 = uint32_t(kOffB0 + 4 * gq) * 4;
    t.off_virgin = uint32_t(kOffW0 + 4 * gq * kModelIn) * 4;
    t.off_fold = uint32_t(kOffW0 + j * kModelIn + 1 + 16 * gq) * 4;
    t.off_w1 = uint32_t(kOffW1 + j * kH0 + 4 * gq) * 4;
    t.off_b1 = uint32_t(kOffB1 + 4 * gq) * 4;
    return t;
}

__device__ __forceinline__ void InitLayer0(v4f (&acc)[16], const TrunkLane& tl) {
#pragma unroll
    for (int t = 0; t < 16; ++t) acc[t] = Buf4(tl.gr, tl.off_b0, 64 * t);
}

// Virgin flag (model input x[0] = 1 on an empty grid): adds column 0 of W0 where `lane_virgin`.
__device__ __forceinline__ void AddVirgin(v4f (&acc)[16], const TrunkLane& tl, bool lane_virgin) {
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float w = BufF(tl.gr, tl.off_virgin, uint32_t(16 * t + r) * kModelIn * 4);
            if (lane_virgin) acc[t][r] += w;
        }
}

// Encoder 145->64 on one turn from the LDS copy (transposed [feature][out]): dims 16gq..16gq+15.
// Features in increasing index order, 5 letters then 5 tiles (turn_features.hpp:28-49), added to
// the bias one by one -- the reference's sequential sum over a one-hot input.
__device__ __forceinline__ void EncodeTurnLds(const float* s_enc, const float* s_encb, uint32_t code, uint32_t fb,
                                              uint32_t gq, float (&e)[16]) {
#pragma unroll
    for (int s = 0; s < 16; ++s) e[s] = s_encb[16 * gq + s];
#pragma unroll
    for (int f = 0; f < 10; ++f) {
        const uint32_t k = f < 5 ? uint32_t(f) * 26u + WordLetter(code, f)
                                 : 130u + uint32_t(f - 5) * 3u + ((fb >> (2 * (f - 5))) & 3u);
        const float* col = s_enc + k * kEncOut + 16 * gq;
#pragma unroll
        for (int s = 0; s < 16; s += 4) {
            const v4f c = *reinterpret_cast<const v4f*>(col + s);
            e[s] += c[0];
            e[s + 1] += c[1];
            e[s + 2] += c[2];
            e[s + 3] += c[3];
        }
    }
}

// The same encoder from the transposed HBM copy of the genome's encoder ([feature][out], written
// by TransposeEncoderKernel once per evaluation): the bias from the genome, then the 10 active
// columns in increasing feature order, 128-bit loads through L2.
__device__ __forceinline__ void EncodeTurnGlobal(Rsrc enc_t, const TrunkLane& tl, uint32_t code, uint32_t fb,
                                                 float (&e)[16]) {
    uint32_t koff[10];
#pragma unroll
    for (int f = 0; f < 10; ++f) {
        const uint32_t k = f < 5 ? uint32_t(f) * 26u + WordLetter(code, f)
                                 : 130u + uint32_t(f - 5) * 3u + ((fb >> (2 * (f - 5))) & 3u);
        koff[f] = (k * kEncOut + 16 * tl.gq) * 4;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v4f acc = Buf4(tl.gr, uint32_t(kOffEncB + 16 * tl.gq) * 4, uint32_t(16 * q));
        v4f col[10];
#pragma unroll
        for (int f = 0; f < 10; ++f) col[f] = Buf4(enc_t, koff[f], uint32_t(16 * q));
#pragma unroll
        for (int f = 0; f < 10; ++f) {
            acc[0] += col[f][0];
            acc[1] += col[f][1];
            acc[2] += col[f][2];
            acc[3] += col[f][3];
        }
        e[4 * q] = acc[0];
        e[4 * q + 1] = acc[1];
        e[4 * q + 2] = acc[2];
        e[4 * q + 3] = acc[3];
    }
}

// acc += W0[:, 1+64*slot .. 1+64*slot+63] . e.  Lane (j, gq)'s 16 weights of neuron tile t are
// contiguous (row 16t + j, inputs 16gq..16gq+15 of the slot): four 128-bit loads (buffer loads need
// only dword alignment, and rows of 321 weights start at every dword phase), issued one tile ahead.
__device__ __forceinline__ void FoldSlot(v4f (&acc)[16], const TrunkLane& tl, int slot, const float (&e)[16]) {
    const uint32_t soff_slot = uint32_t(64 * slot) * 4;
    v4f a[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) a[q] = Buf4(tl.gr, tl.off_fold, soff_slot + 16 * q);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        v4f nx[4];
        if (t + 1 < 16) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                nx[q] = Buf4(tl.gr, tl.off_fold, soff_slot + uint32_t(16 * (t + 1) * kModelIn) * 4 + 16 * q);
        }
#pragma unroll
        for (int s = 0; s < 16; ++s) acc[t] = Mfma(a[s / 4][s % 4], e[s], acc[t]);
        __builtin_amdgcn_sched_barrier(0);
        if (t + 1 < 16) {
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = nx[q];
        }
    }
}

// Layers 1 and 2: pol[s] = policy dim ScoreDim(s, gq) of game column j.
// Layer 2 runs 4 row tiles; row i of tile u computes policy dim ScoreDim(4u + i%4, i/4), so the
// accumulator register r of tile u in lane group gq is exactly scoring step 4u + r.
__device__ __forceinline__ void TrunkToPolicy(const v4f (&acc)[16], const TrunkLane& tl, float (&pol)[kScoreSteps]) {
    v4f h1[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) h1[u] = Buf4(tl.gr, tl.off_b1, 64 * u);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        float hv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) hv[r] = fmaxf(acc[t][r], 0.0f);
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            v4f w[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                w[u] = Buf4(tl.gr, tl.off_w1, uint32_t(16 * (4 * half + u) * kH0 + 16 * t) * 4);
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int r = 0; r < 4; ++r) h1[4 * half + u] = Mfma(w[u][r], hv[r], h1[4 * half + u]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // Policy dim of step s in lane group g: ScoreFixedBase(g) + s for its fixed steps, else
    // ScoreTailRow(g) + s.
    constexpr int kRowTiles = kScoreSteps / 4;
    const uint32_t jr = tl.j & 3u, jg = tl.j >> 2;
    const uint32_t nf_j = uint32_t(ScoreFixedCount(int(jg)));
    const uint32_t row_f = uint32_t(ScoreFixedBase(int(jg))) + jr, row_t = uint32_t(ScoreTailRow(int(jg))) + jr;
    uint32_t roff[kRowTiles];
#pragma unroll
    for (int u = 0; u < kRowTiles; ++u) {
        const uint32_t d = 4u * u + jr < nf_j ? row_f + 4u * u : row_t + 4u * u;
        roff[u] = uint32_t(kOffW2 + d * kH1 + 4 * tl.gq) * 4;
    }
    const int nf = ScoreFixedCount(int(tl.gq));
    const uint32_t bias_f = uint32_t(kOffB2 + ScoreFixedBase(int(tl.gq))) * 4;
    const uint32_t bias_t = uint32_t(kOffB2 + ScoreTailRow(int(tl.gq))) * 4;
    v4f pv[kRowTiles];
#pragma unroll
    for (int u = 0; u < kRowTiles; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int st = 4 * u + r;
            pv[u][r] = BufF(tl.gr, st < nf ? bias_f : bias_t, uint32_t(4 * st));
        }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        float hv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) hv[r] = fmaxf(h1[t][r], 0.0f);
#pragma unroll
        for (int u = 0; u < kRowTiles; ++u) {
            const v4f w = Buf4(tl.gr, roff[u], uint32_t(16 * t) * 4);
#pragma unroll
            for (int r = 0; r < 4; ++r) pv[u] = Mfma(w[r], hv[r], pv[u]);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int s = 0; s < kScoreSteps; ++s) pol[s] = pv[s / 4][s % 4];
}

// ---- layers 1 and 2 with their weights streamed through LDS ----------------------------------
// The weights of layers 1 and 2 are the same for every wave of a workgroup, so the workgroup moves
// them once per decision by LDS DMA, chunk i + 1 in flight while chunk i is used, instead of every
// wave loading them into registers.  Chunk t < 16 is W1[0..128)[16t..16t+16) as [128][16] (8 KB);
// chunk 16 + c is W2[0..64)[32c..32c+32) as [64][32] (8 KB).  A transfer (one instruction, 64 lanes
// x 16 bytes) covers 16 rows of a W1 chunk or 8 rows of a W2 chunk; its per-lane byte offset is
// fixed and the chunk / transfer part is uniform (an SGPR), so issuing needs no per-lane arithmetic.
// The MFMA order is exactly that of TrunkToPolicy, so the results are identical (kT16 = false; with kT16
// layers 1 and 2 take the split-BF16 products of MfmaT16, in the same order in every kernel).
constexpr uint32_t kTrunkDma = 8;                       // transfers per chunk
constexpr uint32_t kTrunkW1Chunks = kH0 / 16;           // 16
constexpr uint32_t kTrunkW2Chunks = kH1 / 32;           // 4
constexpr uint32_t kTrunkChunkBytes = kTrunkDma * 1024;
static_assert(kH1 * 16 * 4 == kTrunkChunkBytes && kPol * 32 * 4 == kTrunkChunkBytes, "8 KB chunks");

// Per-lane byte offsets of the two transfer shapes (lane l: W1 row l/4, inputs 4(l%4)..+3; W2 row
// l/8, inputs 4(l%8)..+3, relative to the transfer's first row and the chunk's first input).
__device__ __forceinline__ uint32_t TrunkLaneW1(uint32_t lane) { return ((lane >> 2) * kH0 + 4 * (lane & 3u)) * 4; }
__device__ __forceinline__ uint32_t TrunkLaneW2(uint32_t lane) { return ((lane >> 3) * kH1 + 4 * (lane & 7u)) * 4; }

// Starts the transfers of trunk chunk c (0 .. kTrunkW1Chunks + kTrunkW2Chunks - 1) into buf; the
// waves of the workgroup take transfers round robin.
__device__ __forceinline__ void IssueTrunkChunk(uint8_t* buf, Rsrc gr, uint32_t c, uint32_t wave, uint32_t waves,
                                                uint32_t lw1, uint32_t lw2) {
    const bool w1 = c < kTrunkW1Chunks;
    const uint32_t voff = w1 ? lw1 : lw2;
    const uint32_t base = w1 ? uint32_t(kOffW1 + 16 * c) * 4 : uint32_t(kOffW2 + 32 * (c - kTrunkW1Chunks)) * 4;
    const uint32_t row_bytes = w1 ? 16u * kH0 * 4 : 8u * kH1 * 4;  // rows per transfer x row length
    for (uint32_t i = wave; i < kTrunkDma; i += waves)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(gr, (LdsPtr)(buf + i * 1024), 16, voff, base + i * row_bytes, 0, 0);
}

// Buffers handed to TrunkStream hold its largest layer-2 read: row 63's last 16 bytes, the chunk's end.
constexpr uint32_t kTrunkReadBytes = 63 * 32 * 4 + 64 + 48 + 16;
static_assert(kTrunkReadBytes == kTrunkChunkBytes, "layer-2 reads stay inside the chunk");

// Layers 1 and 2 for the whole workgroup (every wave calls it; `compute` = this wave has games).
// Chunk 0 must be in flight into b0; chunk i uses b0 when i is even, b1 when odd.  During the last
// chunk (in b1) next_stage(b0) starts whatever follows (the scoring stream's first chunk).
template <bool kT16 = false, typename FN>
__device__ __forceinline__ void TrunkStream(const v4f (&acc)[16], const TrunkLane& tl, float (&pol)[kScoreSteps],
                                            bool compute, uint8_t* b0, uint8_t* b1, uint32_t wave, uint32_t waves,
                                            uint32_t lane, FN&& next_stage) {
    static_assert(kTrunkW1Chunks % 2 == 0 && kTrunkW2Chunks % 2 == 0, "the last chunk sits in b1");
    const uint32_t lw1 = TrunkLaneW1(lane), lw2 = TrunkLaneW2(lane);
    v4f h1[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) h1[u] = Buf4(tl.gr, tl.off_b1, 64 * u);
    constexpr int kRowTiles = kScoreSteps / 4;
    v4f pv[kRowTiles];
    WaitChunk();
    __syncthreads();
    // layer 1 (the two loops are each unrolled whole, so every accumulator index is static)
    const uint32_t rd1 = 64 * tl.j + 16 * tl.gq;  // row 16u + j, inputs 4gq..4gq+3 of a W1 chunk
#pragma unroll
    for (uint32_t t = 0; t < kTrunkW1Chunks; ++t) {
        uint8_t* cur = (t & 1u) ? b1 : b0;
        uint8_t* nxt = (t & 1u) ? b0 : b1;
        IssueTrunkChunk(nxt, tl.gr, t + 1, wave, waves, lw1, lw2);
        if (t == kTrunkW1Chunks - 1) {  // layer-2 biases, landing during the last layer-1 chunk
            const int nf = ScoreFixedCount(int(tl.gq));
            const uint32_t bias_f = uint32_t(kOffB2 + ScoreFixedBase(int(tl.gq))) * 4;
            const uint32_t bias_t = uint32_t(kOffB2 + ScoreTailRow(int(tl.gq))) * 4;
#pragma unroll
            for (int u = 0; u < kRowTiles; ++u)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int st = 4 * u + r;
                    pv[u][r] = BufF(tl.gr, st < nf ? bias_f : bias_t, uint32_t(4 * st));
                }
        }
        if (compute) {
            float hv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) hv[r] = fmaxf(acc[t][r], 0.0f);
            const Split16 hb = SplitBf16(hv[0], hv[1], hv[2], hv[3]);
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) {
                const v4f w = *reinterpret_cast<const v4f*>(cur + u * 1024 + rd1);
                if constexpr (kT16) {
                    h1[u] = MfmaT16(w, hb, h1[u]);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) h1[u] = Mfma(w[r], hv[r], h1[u]);
                }
            }
        }
        WaitChunk();
        __syncthreads();
    }
    // layer 2: chunk c holds inputs 32c..32c+31, i.e. input tiles t = 2c, 2c + 1.  Lane (j, gq) of row
    // tile u reads row ScoreDim(4u + j%4, j/4): ScoreFixedBase(jg) + jr + 4u while that is a fixed dim
    // (u = 0; u = 1 for jr < 2, and jr = 2 in groups 2, 3), else ScoreTailRow(jg) + jr + 4u -- three base
    // addresses plus immediates.
    const uint32_t jr = tl.j & 3u, jg = tl.j >> 2;
    const uint32_t rd_f = (uint32_t(ScoreFixedBase(int(jg))) + jr) * (32 * 4) + 16 * tl.gq;
    const uint32_t rd_t = (uint32_t(ScoreTailRow(int(jg))) + jr) * (32 * 4) + 16 * tl.gq;
    const uint32_t rd_1 = jr + 4 < uint32_t(ScoreFixedCount(int(jg))) ? rd_f + 512 : rd_t + 512;
#pragma unroll
    for (uint32_t c = 0; c < kTrunkW2Chunks; ++c) {
        uint8_t* cur = (c & 1u) ? b1 : b0;
        uint8_t* nxt = (c & 1u) ? b0 : b1;
        if (c + 1 < kTrunkW2Chunks)
            IssueTrunkChunk(nxt, tl.gr, kTrunkW1Chunks + c + 1, wave, waves, lw1, lw2);
        else
            next_stage(nxt);
        if (compute) {
#pragma unroll
            for (uint32_t h = 0; h < 2; ++h) {
                const uint32_t t = 2 * c + h;
                float hv[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) hv[r] = fmaxf(h1[t][r], 0.0f);
                const Split16 hb = SplitBf16(hv[0], hv[1], hv[2], hv[3]);
#pragma unroll
                for (int u = 0; u < kRowTiles; ++u) {
                    const uint32_t rd = u == 0 ? rd_f : (u == 1 ? rd_1 : rd_t + 512 * uint32_t(u));
                    const v4f w = *reinterpret_cast<const v4f*>(cur + rd + 64 * h);
                    if constexpr (kT16) {
                        pv[u] = MfmaT16(w, hb, pv[u]);
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r) pv[u] = Mfma(w[r], hv[r], pv[u]);
                    }
                }
            }
        }
        WaitChunk();
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < kScoreSteps; ++s) pol[s] = pv[s / 4][s % 4];
}

// Stages a genome's 145->64 encoder into LDS, transposed to [feature][out].
__device__ __forceinline__ void StageEncoder(const float* genome, float* s_enc, float* s_encb) {
    for (uint32_t idx = threadIdx.x; idx < uint32_t(kEncOut * kTurnFeatures); idx += blockDim.x) {
        const uint32_t o = idx / kTurnFeatures, k = idx - o * kTurnFeatures;
        s_enc[k * kEncOut + o] = genome[kOffEncW + idx];
    }
    if (threadIdx.x < kEncOut) s_encb[threadIdx.x] = genome[kOffEncB + threadIdx.x];
    __syncthreads();
}

// Writes each genome's 145->64 encoder transposed to [feature][out], followed by the virgin-flag
// column of W0 (W0[:, 0], 256 values) made contiguous (one workgroup per genome).
__global__ void TransposeEncoderKernel(const uint8_t* slab, uint64_t stride, const uint32_t* slot_indices,
                                       uint32_t slot_count, float* enc_t) {
    const uint32_t slot = slot_indices[blockIdx.x];
    if (slot >= slot_count) return;  // reported by the evaluation kernel
    __shared__ __align__(16) float s_w[kEncOut * kTurnFeatures];  // [out][feature], as stored in the genome
    const float* genome = reinterpret_cast<const float*>(slab + uint64_t(slot) * stride);
    float* dst = enc_t + uint64_t(blockIdx.x) * kEncTStride;
    static_assert(kOffEncW % 4 == 0 && (kEncOut * kTurnFeatures) % 4 == 0 && kEncTStride % 4 == 0 && kEncOut % 4 == 0,
                  "16-byte moves");
    for (uint32_t o = threadIdx.x; o < uint32_t(kH0); o += blockDim.x)
        dst[kTurnFeatures * kEncOut + o] = genome[kOffW0 + o * kModelIn];
    // 16-byte loads into LDS, then 16-byte stores of four consecutive outputs of one feature
    constexpr uint32_t kVec = uint32_t(kEncOut * kTurnFeatures) / 4;
    for (uint32_t v = threadIdx.x; v < kVec; v += blockDim.x)
        *reinterpret_cast<v4f*>(s_w + 4 * v) = Load4(genome + kOffEncW + 4 * v);
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < kVec; v += blockDim.x) {
        const uint32_t k = (4 * v) / kEncOut, o = 4 * v - k * kEncOut;
        v4f t;
#pragma unroll
        for (int i = 0; i < 4; ++i) t[i] = s_w[(o + uint32_t(i)) * kTurnFeatures + k];
        *reinterpret_cast<v4f*>(dst + 4 * v) = t;
    }
}

// Packs the per-genome part of each genome's BF16 screening operands (kBf16GenTileBytes per action
// tile, round-to-nearest-even) and the max |tail value| per tail dim over its first A rows (the
// screening bound).  Grid: (blocks, genomes), 320 threads = 4 tiles at a time, 80 threads per tile:
// thread r of a tile writes bytes 16r..16r+15 of it, i.e. row i = r % 16 of operand class c = r / 16
// (0: instruction 0 lane group 3, 1..4: instruction 1 lane group c - 1), so its 8 dims are fixed and
// their maxima stay in registers until one reduction per workgroup.
// Workgroups per genome: measured at 64x64 / 4,739 actions (encoder transpose + pack): 16 -> 1.61 ms,
// 8 -> 1.50, 4 -> 1.42, 2 -> 1.40 without the register prefetch; 4 -> 1.05 ms with it (8 -> 1.10)
#ifndef NE_PACK_BLOCKS
#define NE_PACK_BLOCKS 4
#endif
constexpr uint32_t kPackTiles = 4;  // = blockDim.x / 80: 5 classes x 16 rows per tile
__global__ __launch_bounds__(320) void PackBf16Kernel(const uint8_t* slab, uint64_t stride, const uint32_t* slot_indices,
                                                      uint32_t slot_count, const uint32_t* img, uint32_t A,
                                                      uint32_t a_tiles, uint8_t* out, uint64_t out_stride,
                                                      uint32_t* tail_absmax) {
    __shared__ float s_mx[kPackTiles][5][8];
    // the kPackTiles tiles' tail rows (16 x 38 contiguous floats each, 16-byte aligned), staged with
    // coalesced 128-bit loads
    __shared__ __align__(16) float s_rows[kPackTiles * 16 * kTail];
    const uint32_t n = blockIdx.y;
    const uint32_t slot = slot_indices[n];
    if (slot >= slot_count) return;  // reported by the evaluation kernel
    const float* tails = reinterpret_cast<const float*>(slab + uint64_t(slot) * stride) + kPolicyParams;
    const uint32_t sub = threadIdx.x / 80, r = threadIdx.x % 80;
    const uint32_t c = r / 16, i = r % 16;
    const int m = c == 0 ? 0 : 1, g = c == 0 ? 3 : int(c) - 1;
    const uint32_t n_tail = A * kTail;  // valid tail floats
    float mx[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // the next group's rows are loaded into registers while the current group is packed
    constexpr uint32_t kQuads = kPackTiles * 16 * kTail / 4;  // float4s per group (2,432)
    constexpr uint32_t kPer = (kQuads + 319) / 320;            // per thread (blockDim.x = 320)
    v4f pre[kPer];
    auto load_group = [&](uint32_t at0) __device__ {
        const uint32_t f0 = at0 * 16 * kTail;
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t q = threadIdx.x + k * 320;
            const uint32_t f = f0 + 4 * q;
            v4f v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (q < kQuads && at0 < a_tiles) {
                if (f + 4 <= n_tail) {
                    v = *reinterpret_cast<const v4f*>(tails + f);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[e] = f + e < n_tail ? tails[f + e] : 0.0f;
                }
            }
            pre[k] = v;
        }
    };
    load_group(blockIdx.x * kPackTiles);
    for (uint32_t at0 = blockIdx.x * kPackTiles; at0 < a_tiles; at0 += gridDim.x * kPackTiles) {
        __syncthreads();  // the previous tiles are done with s_rows
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t q = threadIdx.x + k * 320;
            if (q < kQuads) *reinterpret_cast<v4f*>(s_rows + 4 * q) = pre[k];
        }
        __syncthreads();
        load_group(at0 + gridDim.x * kPackTiles);

        const uint32_t at = at0 + sub;
        if (at >= a_tiles) continue;
        const uint32_t a = at * 16 + i;
        const float* row = s_rows + (sub * 16 + i) * kTail;
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int d = Bf16SlotDim(m, 8 * g + e);
            float x = 0.0f;
            if (a < A) {
                if (d < kFixed) {  // fixed dim d: step d - ScoreFixedBase(gg) of its lane group in the FP32 image
                    const int gg = ScoreFixedGroup(d), st = d - ScoreFixedBase(gg);
                    x = __builtin_bit_cast(float, img[uint64_t(at) * kImgTileWords + (st / 4) * 256 +
                                                      4 * (i + 16 * uint32_t(gg)) + (st % 4)]);
                } else {
                    x = row[d - kFixed];
                    mx[e] = fmaxf(mx[e], fabsf(x));
                }
            }
            v[e] = __bf16(x);
        }
        *reinterpret_cast<bf16x8*>(out + uint64_t(n) * out_stride + uint64_t(at) * kBf16GenTileBytes + 16 * r) = v;
    }
    // the 16 threads of one (tile, class) are 16 aligned lanes of one wave (80 and 16 divide evenly)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) mx[e] = fmaxf(mx[e], __shfl_xor(mx[e], off));
        if (i == 0) s_mx[sub][c][e] = mx[e];
    }
    __syncthreads();
    if (threadIdx.x < uint32_t(kTail)) {
        const uint32_t t = threadIdx.x;
        const uint32_t cc = t < 32 ? 1 + t / 8 : 0, ee = t < 32 ? t % 8 : t - 32 + (kFixed - 24);
        float v = 0.0f;
#pragma unroll
        for (uint32_t q = 0; q < kPackTiles; ++q) v = fmaxf(v, s_mx[q][cc][ee]);
        // non-negative floats order like their bit patterns
        atomicMax(tail_absmax + uint64_t(n) * 40 + t, __builtin_bit_cast(uint32_t, v));
    }
}

// ================================================================================================
// Evaluation kernel.
//
// One workgroup per genome (blockIdx.y splits long game lists), one wave per tile of 16 games; the
// waves of a workgroup take their decisions in lockstep.  Per decision each wave runs its own trunk,
// then all waves score their games against every action together: the output embedding streams
// through LDS in chunks of kChunkTiles action tiles (double-buffered, loaded by the whole workgroup
// one chunk ahead), so a genome's embedding is read from HBM/L2 once per workgroup and decision.
#ifndef NE_CHUNK_TILES
#define NE_CHUNK_TILES 4  // action tiles per LDS chunk of the FP32 scan (build-time tunable)
#endif
#ifndef NE_EVAL_WAVES_PER_EU
#define NE_EVAL_WAVES_PER_EU 3  // register budget: 3 waves/SIMD = four 3-wave workgroups per CU
#endif
#ifndef NE_EVAL_MAX_WAVES
#define NE_EVAL_MAX_WAVES 4  // waves per evaluation workgroup the kernel's LDS tables are sized for
#endif
constexpr uint32_t kEvalMaxWaves = NE_EVAL_MAX_WAVES;
constexpr int kChunkTiles = NE_CHUNK_TILES;
constexpr int kChunkTilesBf16 = 2;  // the BF16 mode keeps candidate lists in LDS as well
// Chunks move by LDS DMA (buffer_load_dwordx4 ... lds): one instruction moves 64 lanes x 16 bytes to
// consecutive LDS, so each region is a whole number of 1 KB transfers.
constexpr uint32_t kDmaBytes = 1024;
template <int kCT>
struct ChunkGeom {
    static constexpr uint32_t kImgBytes = uint32_t(kCT * kImgTileBytes);
    static constexpr uint32_t kTailBytes = uint32_t(kCT * 16 * kTail * 4);
    static constexpr uint32_t kImgDma = (kImgBytes + kDmaBytes - 1) / kDmaBytes;
    static constexpr uint32_t kTailDma = (kTailBytes + kDmaBytes - 1) / kDmaBytes;
    static constexpr uint32_t kBufBytes = (kImgDma + kTailDma) * kDmaBytes;
    static constexpr uint32_t kTailRegion = kImgDma * kDmaBytes;  // byte offset of the tail rows
};
static_assert(kImgTileBytes % 16 == 0 && (16 * kTail * 4) % 16 == 0, "16-byte transfer units");

typedef uint32_t v4u __attribute__((ext_vector_type(4)));

// Issues the DMA transfers of chunk c into one LDS buffer, spread over the workgroup's waves.  Bytes
// past a descriptor's size arrive as 0.  The caller waits (vmcnt) and barriers before reading.
__device__ __forceinline__ void IssueChunk(uint8_t* buf, Rsrc img_r, Rsrc tail_r, uint32_t c, uint32_t wave,
                                           uint32_t waves, uint32_t lane, uint32_t img_bytes, uint32_t tail_bytes,
                                           uint32_t img_dma, uint32_t tail_dma) {
    for (uint32_t i = wave; i < img_dma + tail_dma; i += waves) {
        if (i < img_dma) {
            __builtin_amdgcn_raw_ptr_buffer_load_lds(img_r, (LdsPtr)(buf + i * kDmaBytes), 16,
                                                     c * img_bytes + i * kDmaBytes + 16 * lane, 0, 0, 0);
        } else {
            const uint32_t k = i - img_dma;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(tail_r, (LdsPtr)(buf + img_dma * kDmaBytes + k * kDmaBytes), 16,
                                                     c * tail_bytes + k * kDmaBytes + 16 * lane, 0, 0, 0);
        }
    }
}


// Streams every chunk of the scoring image and tail rows through two LDS buffers (b0, b1: distinct
// objects), chunk c + 1 in flight while chunk c is used.  issue(buf, c) starts the transfers of
// chunk c; per_tile(buf, k, at) runs for the kCT tiles of each chunk (tile k of the buffer is action
// tile at) when `compute`.
// The whole workgroup calls this together; chunk 0 must already be in flight into b0.
template <int kCT, typename FI, typename FT>
__device__ __forceinline__ void StreamChunks(uint8_t* b0, uint8_t* b1, uint32_t chunks, bool compute, FI&& issue,
                                             FT&& per_tile) {
    WaitChunk();
    __syncthreads();
    auto one = [&](const uint8_t* buf, uint8_t* next, uint32_t c) __device__ {
        if (c + 1 < chunks) issue(next, c + 1);
        if (compute) {
#pragma unroll
            for (int k = 0; k < kCT; ++k) per_tile(buf, k, c * uint32_t(kCT) + uint32_t(k));
        }
        WaitChunk();
        __syncthreads();
    };
    for (uint32_t c = 0; c < chunks; c += 2) {
        one(b0, b1, c);
        if (c + 1 < chunks) one(b1, b0, c + 1);
    }
}

// Waits until at most kKeep of this wave's vector memory transfers are outstanding (they complete in issue
// order, so every older transfer has landed).
template <uint32_t kKeep>
__device__ __forceinline__ void WaitKeep() {
    static_assert(kKeep < 16, "vmcnt field");
    __builtin_amdgcn_s_waitcnt(0x0F70 | kKeep);
}

// A workgroup barrier that leaves this wave's later transfers in flight: __syncthreads() is also a
// workgroup-scope fence, which waits for every outstanding transfer (vmcnt(0)) before the barrier.  Here
// only this wave's LDS reads are completed (lgkmcnt(0): a buffer may be refilled after the barrier), the
// transfers the caller needs have been waited for by the caller (WaitKeep), and compiler and scheduler
// barriers keep every memory access on its side.
__device__ __forceinline__ void BarrierKeepTransfers() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt not waited for
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" ::: "memory");
}

// StreamChunks over three buffers (distinct objects) with two chunks in flight: chunk c + 2 is issued when
// chunk c's tiles start, and before the chunk's closing barrier each wave waits only until its own transfers
// of chunk c + 1 have landed (at most kKeep transfers outstanding: those of chunk c + 2, of which every wave
// issues exactly kKeep and nothing else in between).  A buffer is refilled only after the barrier that
// closes its last reading chunk.  Chunk 0 must already have landed in b0 (and b1 must be free) when the
// workgroup calls this.
template <uint32_t kKeep, typename FI, typename FC>
__device__ __forceinline__ void StreamChunks3Whole(uint8_t* b0, uint8_t* b1, uint8_t* b2, uint32_t chunks, bool compute,
                                                   FI&& issue, FC&& per_chunk) {
    WaitChunk();
    __syncthreads();
    if (chunks > 1) issue(b1, 1u);
    auto one = [&](const uint8_t* buf, uint8_t* next2, uint32_t c) __device__ {
        if (c + 2 < chunks) issue(next2, c + 2);
        if (compute) per_chunk(buf, c);
        if (c + 2 < chunks) {
            WaitKeep<kKeep>();  // chunk c + 1 landed, chunk c + 2 may stay in flight
            BarrierKeepTransfers();
        } else {
            WaitChunk();
            __syncthreads();
        }
    };
    for (uint32_t c = 0; c < chunks; c += 3) {
        one(b0, b2, c);
        if (c + 1 < chunks) one(b1, b0, c + 1);
        if (c + 2 < chunks) one(b2, b1, c + 2);
    }
}
// The same with per_tile(buf, k, at) for the kCT tiles of each chunk.
template <int kCT, uint32_t kKeep, typename FI, typename FT>
__device__ __forceinline__ void StreamChunks3(uint8_t* b0, uint8_t* b1, uint8_t* b2, uint32_t chunks, bool compute,
                                              FI&& issue, FT&& per_tile) {
    StreamChunks3Whole<kKeep>(b0, b1, b2, chunks, compute, issue, [&](const uint8_t* buf, uint32_t c) __device__ {
#pragma unroll
        for (int k = 0; k < kCT; ++k) per_tile(buf, k, c * uint32_t(kCT) + uint32_t(k));
    });
}

// StreamChunks with one callback per chunk: per_chunk(buf, c) for the buffer holding chunk c.
template <typename FI, typename FC>
__device__ __forceinline__ void StreamChunksWhole(uint8_t* b0, uint8_t* b1, uint32_t chunks, bool compute, FI&& issue,
                                                  FC&& per_chunk) {
    WaitChunk();
    __syncthreads();
    auto one = [&](const uint8_t* buf, uint8_t* next, uint32_t c) __device__ {
        if (c + 1 < chunks) issue(next, c + 1);
        if (compute) per_chunk(buf, c);
        WaitChunk();
        __syncthreads();
    };
    for (uint32_t c = 0; c < chunks; c += 2) {
        one(b0, b1, c);
        if (c + 1 < chunks) one(b1, b0, c + 1);
    }
}

// Masked argmax state of one lane: candidates in increasing action order, the first unmasked one
// taken whatever its score, later ones only when strictly greater (evaluation_ops.cuh:429-445).
struct ArgmaxLane {
    float best, second;
    uint32_t besta;
    float gate;  // a score some unmasked row of the same game already has (rows below it cannot win)
};

// Once per chunk: the game's best score so far over its 4 lane groups becomes every lane's gate.  A row
// strictly below a score another row of the game already holds can never be the argmax; rows equal to it
// still enter (the lower index may win the tie).  Fewer tiles then take the update path (measured at
// 64x64 / 4,739 actions: 15.9 -> 15.35 ms per FP32 evaluation, same decisions).
__device__ __forceinline__ void ShareArgmaxGate(ArgmaxLane& am) {
    const float g = fmaxf(am.best, __shfl_xor(am.best, 16));
    am.gate = fmaxf(g, __shfl_xor(g, 32));
}

// The 16 embedding values lane (j, gq) feeds into the scoring contraction for action row j of a tile:
// fixed steps from the two image planes, then 10 tail values of the row from ScoreTailRead(gq) on
// (steps 6..15; lane groups 2, 3 take step 6 from the image instead).  The tail values are read one
// dword at a time: a row is 38 floats, so a lane group's first tail value is only dword aligned.
__device__ __forceinline__ void TileOperands(const uint32_t* img_tile, const float* tail_tile, uint32_t lane,
                                             uint32_t j, uint32_t gq, float (&ev)[kScoreSteps]) {
    const float* fx = reinterpret_cast<const float*>(img_tile);
    const v4f f0 = *reinterpret_cast<const v4f*>(fx + 4 * lane);
    const v4f f1 = *reinterpret_cast<const v4f*>(fx + 256 + 4 * lane);
    const float* tr = tail_tile + j * kTail + uint32_t(ScoreTailRead(int(gq)));
    float tv[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) tv[q] = tr[q];
    ev[0] = f0[0];
    ev[1] = f0[1];
    ev[2] = f0[2];
    ev[3] = f0[3];
    ev[4] = f1[0];
    ev[5] = f1[1];
    ev[6] = gq < 2 ? tv[0] : f1[2];
#pragma unroll
    for (int q = 1; q < 10; ++q) ev[6 + q] = tv[q];
}

// The FP32 scoring contraction of one tile: 16 dependent v_mfma_f32_16x16x4_f32 steps in a fixed
// order.  Every FP32 score of this build, streamed or rescored, comes from exactly this sequence.
__device__ __forceinline__ v4f Fp32Scores(const float (&ev)[kScoreSteps], const float (&pol)[kScoreSteps]) {
    v4f sc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < kScoreSteps; ++s) sc = Mfma(ev[s], pol[s], sc);
    return sc;
}

// Masked argmax over the 4 rows (actions abase..abase+3) this lane owns in a tile; candidates arrive
// in increasing action order, the first unmasked one is taken whatever its score, later ones only
// when strictly greater (evaluation_ops.cuh:429-445).  Without traces the per-row work runs only
// when some lane of the wave can change: a lane with a candidate changes only if one of its four
// scores beats its best (a tie never displaces the earlier, lower index).
template <bool kTrace>
__device__ __forceinline__ void ArgmaxUpdate(const v4f& sc, const uint32_t* cn_ptr, uint32_t abase, uint32_t A,
                                             const uint32_t (&hist)[5], ArgmaxLane& am) {
    if (!kTrace) {
        const float m4 = fmaxf(fmaxf(sc[0], sc[1]), fmaxf(sc[2], sc[3]));
        if (!__any((am.besta == kInvalidSlot || (m4 > am.best && m4 >= am.gate)) ? 1 : 0)) return;
    }
    const v4u cn = *reinterpret_cast<const v4u*>(cn_ptr);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t a = abase + uint32_t(r);
        const float v = sc[r];
        const uint32_t c = cn[r];
        const bool masked = (c == hist[0]) | (c == hist[1]) | (c == hist[2]) | (c == hist[3]) | (c == hist[4]);
        const bool ok = (a < A) & !masked;
        const bool first = am.besta == kInvalidSlot;
        const bool better = ok & (first | (v > am.best));
        if (kTrace) {
            const float demoted = first ? am.second : fmaxf(am.second, am.best);
            am.second = better ? demoted : ((ok & (v > am.second)) ? v : am.second);
        }
        am.best = better ? v : am.best;
        am.besta = better ? a : am.besta;
    }
}

// Scores one action tile (16 actions) in FP32 and folds this lane's 4 rows into its argmax.
template <bool kTrace>
__device__ __forceinline__ void ScoreTile(const uint32_t* img_tile, const float* tail_tile, uint32_t lane, uint32_t j,
                                          uint32_t gq, const float (&pol)[kScoreSteps], uint32_t abase, uint32_t A,
                                          const uint32_t (&hist)[5], ArgmaxLane& am) {
    float ev[kScoreSteps];
    TileOperands(img_tile, tail_tile, lane, j, gq, ev);
    ArgmaxUpdate<kTrace>(Fp32Scores(ev, pol), img_tile + 512 + 4 * gq, abase, A, hist, am);
}

// Tile k of an FP32 chunk buffer: its scoring image and its 16 tail rows.
template <int kCT>
__device__ __forceinline__ const uint32_t* Fp32Img(const uint8_t* buf, int k) {
    return reinterpret_cast<const uint32_t*>(buf) + k * kImgTileWords;
}
template <int kCT>
__device__ __forceinline__ const float* Fp32Tails(const uint8_t* buf, int k) {
    return reinterpret_cast<const float*>(buf + ChunkGeom<kCT>::kTailRegion) + k * 16 * kTail;
}

// ---- BF16 screening -----------------------------------------------------------------------------
// The same contraction on v_mfma_f32_16x16x32_bf16 (FP32 accumulate) over the packed operands of
// PackBf16Kernel: two instructions per tile, operands straight from LDS.
struct PolicyBf16 {
    bf16x8 k[2];
};
// A screening chunk of kB16Tiles tiles in LDS: the catalog parts (3 transfers of 1 KB), the genome
// parts (5 transfers), then the tiles' canonical indices (one 4-byte transfer).
// Tiles per screening chunk, a multiple of 4.  Measured at 64x64 / 4,739 actions: 4 tiles 11.0 ms per
// evaluation; 8 tiles 14.2 ms (50.7 KB of LDS per workgroup: three workgroups per CU instead of four).
#ifndef NE_B16_TILES
#define NE_B16_TILES 4
#endif


constexpr uint32_t kB16Tiles = NE_B16_TILES;
constexpr uint32_t kB16CatBytes = kB16Tiles * kBf16CatTileBytes;
constexpr uint32_t kB16GenBytes = kB16Tiles * kBf16GenTileBytes;
constexpr uint32_t kB16OpsDma = (kB16CatBytes + kB16GenBytes) / kDmaBytes;
constexpr uint32_t kB16CanonDma = kB16Tiles / 4;  // 4-byte transfers: 16 indices x 4 tiles each
constexpr uint32_t kB16BufBytes = kB16CatBytes + kB16GenBytes + kB16Tiles * 64;
static_assert(kB16Tiles % 4 == 0, "whole canonical-index transfers");
static_assert(kB16CatBytes % kDmaBytes == 0 && kB16GenBytes % kDmaBytes == 0, "whole transfers");

__device__ __forceinline__ v4f Bf16Screen(const uint8_t* buf, int k, uint32_t lane, const PolicyBf16& pb) {
    // one base pointer and an integer select keep the read attributable to its buffer (a pointer
    // select would make every read wait for the transfers into the other buffer)
    const uint32_t off0 = lane < 48 ? k * kBf16CatTileBytes + 16 * lane
                                    : kB16CatBytes + k * kBf16GenTileBytes + 16 * (lane - 48);
    const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(buf + off0);
    const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(buf + kB16CatBytes + k * kBf16GenTileBytes + 256 + 16 * lane);
    v4f sc = {0.0f, 0.0f, 0.0f, 0.0f};
    sc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, pb.k[0], sc, 0, 0, 0);
    sc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, pb.k[1], sc, 0, 0, 0);
    return sc;
}
__device__ __forceinline__ const uint32_t* Bf16Canon(const uint8_t* buf, int k, uint32_t gq) {
    return reinterpret_cast<const uint32_t*>(buf + kB16CatBytes + kB16GenBytes) + k * 16 + 4 * gq;
}
// Bf16Screen in two parts, so that one read of a tile's operands serves several game tiles: the lane's two
// A operands of tile k, then the two instructions against one tile's policy.
__device__ __forceinline__ void Bf16Operands(const uint8_t* buf, int k, uint32_t lane, bf16x8& a0, bf16x8& a1) {
    const uint32_t off0 = lane < 48 ? k * kBf16CatTileBytes + 16 * lane
                                    : kB16CatBytes + k * kBf16GenTileBytes + 16 * (lane - 48);
    a0 = *reinterpret_cast<const bf16x8*>(buf + off0);
    a1 = *reinterpret_cast<const bf16x8*>(buf + kB16CatBytes + k * kBf16GenTileBytes + 256 + 16 * lane);
}
__device__ __forceinline__ v4f Bf16Mfma(const bf16x8& a0, const bf16x8& a1, const PolicyBf16& pb) {
    v4f sc = {0.0f, 0.0f, 0.0f, 0.0f};
    sc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, pb.k[0], sc, 0, 0, 0);
    sc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, pb.k[1], sc, 0, 0, 0);
    return sc;
}

// Reduces a lane's masked-argmax state over the four lane groups sharing game column j (lower index
// wins ties); every lane of the column ends with the game's result.
__device__ __forceinline__ void ReduceArgmax(float& best, float& second, uint32_t& besta) {
#pragma unroll
    for (int off = 16; off <= 32; off <<= 1) {
        const float ob = __shfl_xor(best, off);
        const uint32_t oa = __shfl_xor(besta, off);
        const float os = __shfl_xor(second, off);
        const bool take = oa != kInvalidSlot && (besta == kInvalidSlot || ob > best || (ob == best && oa < besta));
        if (take) {
            second = besta == kInvalidSlot ? fmaxf(second, os) : fmaxf(fmaxf(second, best), os);
            best = ob;
            besta = oa;
        } else {
            second = oa == kInvalidSlot ? fmaxf(second, os) : fmaxf(fmaxf(second, ob), os);
        }
    }
}

__device__ inline float2 BufF2(Rsrc r, uint32_t voff) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
}

// The 16 FP32 operands of action a for lane group gq, read from HBM (rescoring one candidate): the
// same values TileOperands reads from LDS for that action's row.
__device__ __forceinline__ void ActionOperands(Rsrc img_r, Rsrc tail_r, uint32_t a, uint32_t gq,
                                               float (&ev)[kScoreSteps]) {
    const uint32_t base = ((a >> 4) * uint32_t(kImgTileWords) + 4 * ((a & 15u) + 16 * gq)) * 4;
    const v4f f0 = Buf4(img_r, base, 0);
    const v4f f1 = Buf4(img_r, base + 1024, 0);
    const uint32_t toff = (a * uint32_t(kTail) + uint32_t(ScoreTailRead(int(gq)))) * 4;
    float tv[10];
#pragma unroll
    for (int q = 0; q < 5; ++q) {  // buffer loads need only dword alignment
        const float2 v = BufF2(tail_r, toff + 8 * q);
        tv[2 * q] = v.x;
        tv[2 * q + 1] = v.y;
    }
    ev[0] = f0[0];
    ev[1] = f0[1];
    ev[2] = f0[2];
    ev[3] = f0[3];
    ev[4] = f1[0];
    ev[5] = f1[1];
    ev[6] = gq < 2 ? tv[0] : f1[2];
#pragma unroll
    for (int q = 1; q < 10; ++q) ev[6 + q] = tv[q];
}

// Layer-0 accumulators between decisions.  They are needed only by the fold and the trunk, so they wait in
// a global scratch area of the wave while the scan runs, leaving their 64 VGPRs to the scan's operand
// reads.  A wave's area is [game tile q][neuron tile t][lane] v4f (16 KB per game tile), addressed through
// a buffer resource (the wave's base) with the lane's 16-byte slot as the only per-lane offset and the
// (q, t) part as a scalar offset, so no per-(q, t) lane addresses are kept in registers.
constexpr uint32_t kAccTileFloats = 16 * 64 * 4;  // one tile's layer-0 accumulators: [neuron tile][lane] v4f
__device__ __forceinline__ v4f AccLoad(Rsrc r, uint32_t lane, uint32_t q, uint32_t t) {
    return Buf4(r, lane * 16, (q * 16 + t) * 1024);
}
__device__ __forceinline__ void AccStore(Rsrc r, uint32_t lane, uint32_t q, uint32_t t, const v4f& v) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, v), r, lane * 16, (q * 16 + t) * 1024, 0);
}

// BF16 screening bound.  With round-to-nearest-even conversions |bf16(x) - x| <= 2^-9 |x|, so each
// product is off by at most 2^-8 (1 + 2^-10) |e p|; the FP32 accumulation errors of both the BF16
// and the FP32 contraction (<= 96 and 68 roundings of 2^-24) fit in the 3% margin.  Hence
// |S_bf16 - S_fp32| <= eps = 1.03 * 2^-8 * sum_d M_d |p_d| (+ a subnormal floor), M_d = max_a |E_ad|.
constexpr float kBf16RelBound = 1.03f * 0x1p-8f;
constexpr float kBf16AbsBound = 0x1p-120f;
constexpr uint32_t kMaxCandidates = 32;  // per game; more within the window -> FP32 scan of the decision
// Per lane during the screening pass (compacted when full).  4 measured: lists overflow, 10,037 instead of
// 68 FP32 fallbacks per evaluation at 64x64 / 4,739 actions, 11.8 instead of 9.5 ms.
constexpr uint32_t kLaneCandidates = 8;

// The next (genome, split) item of a persistent evaluation workgroup (called by one thread): from the
// queue of its own XCD while that lasts, then from the other queues in turn; `q` and `tried` carry the
// queue being drained between calls.  A genome's items sit together in one queue, so the workgroups
// playing its splits run side by side on one XCD and share its L2 copy of the genome's weights and
// embedding rows.  Returns work_count once every queue is empty.
__device__ __forceinline__ uint32_t TakeItem(const EvalArgs& p, uint32_t& q, uint32_t& tried) {
    while (tried < kWorkQueues) {
        const uint32_t item = p.queue_begin[q] + atomicAdd(p.work_counter + q * kWorkCounterStride, 1u);
        if (item < p.queue_begin[q + 1]) return item;
        q = q + 1 == kWorkQueues ? 0u : q + 1;
        ++tried;
    }
    return p.work_count;
}

template <bool kTrace, bool kBf16, bool kT16 = false>
__global__ __launch_bounds__(64 * NE_EVAL_MAX_WAVES) __attribute__((amdgpu_waves_per_eu(NE_EVAL_WAVES_PER_EU, NE_EVAL_WAVES_PER_EU))) void EvaluateKernel(EvalArgs p) {
    constexpr int kCT = kBf16 ? kChunkTilesBf16 : kChunkTiles;  // FP32 stream (the BF16 form's fallback)
    using G = ChunkGeom<kCT>;
    constexpr uint32_t kBufBytes = kBf16 && kB16BufBytes > G::kBufBytes ? kB16BufBytes : G::kBufBytes;
    // two chunk buffers as distinct objects, so a read of one never waits on transfers into the other
    __shared__ __align__(16) uint8_t s_buf0[kBufBytes];
    __shared__ __align__(16) uint8_t s_buf1[kBufBytes];

    // Persistent: the workgroup takes (genome, split) items from the work queues until they are empty;
    // split s of genome n plays tiles s*waves + r*splits*waves + wave (r = 0, 1, ...) of n's game tiles.
    __shared__ uint32_t s_item;
    uint32_t queue = blockIdx.x % kWorkQueues, tried = 0;
    for (;;) {
        __syncthreads();  // the previous item is done with the LDS buffers
        if (threadIdx.x == 0) s_item = TakeItem(p, queue, tried);
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= p.work_count) break;
        const uint32_t n = p.work[2 * item], split = p.work[2 * item + 1];
        const uint32_t slot = p.slot_indices[n];
        if (slot >= p.slot_count) {
            if (threadIdx.x == 0) SetStatus(p.status, 4);  // kInvalidGeneration
            continue;
        }
        const float* genome = reinterpret_cast<const float*>(p.slab + uint64_t(slot) * p.stride);
        const uint32_t tid = threadIdx.x, bd = blockDim.x;

        const Rsrc enc_r = MakeRsrc(p.enc_t + uint64_t(n) * kEncTStride, uint32_t(kEncTStride * 4));

        const uint32_t W = p.cell_offsets[n + 1] - p.cell_offsets[n];
        if (W == 0) {
            if (tid == 0) SetStatus(p.status, 5);  // kInvalidTrainingShard
            continue;
        }
        if (tid == 0 && split == 0) p.local_w[n] = W;
        const uint32_t* local = p.secret_ids + p.cell_offsets[n];
        const uint32_t tiles_fresh = (W + 15) / 16;
        const uint32_t tiles = tiles_fresh + (2 * W + 15) / 16;

        const uint32_t lane = tid & 63u;
        const uint32_t j = lane & 15u;   // A-operand row / B-operand column within a 16-tile
        const uint32_t gq = lane >> 4;   // k group (0..3)
        const uint32_t waves = bd >> 6;
        const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);

        const TrunkLane tl = MakeTrunkLane(genome, j, gq);
        const uint32_t A = p.action_count;
        const uint32_t a_tiles = (A + 15) / 16;
        const uint32_t chunks = (a_tiles + kCT - 1) / kCT;
        const Rsrc acc_r = MakeRsrc(p.acc_scratch + (uint64_t(blockIdx.x) * kEvalMaxWaves + wave) * 2 * kAccTileFloats,
                                    kAccTileFloats * sizeof(float));
        const Rsrc img_r = MakeRsrc(p.img, p.img_bytes);
        const Rsrc tail_r = MakeRsrc(genome + kPolicyParams, uint32_t(p.stride - kPolicyParams * 4));

        auto issue = [&](uint8_t* buf, uint32_t c) __device__ {
            IssueChunk(buf, img_r, tail_r, c, wave, waves, lane, G::kImgBytes, G::kTailBytes, G::kImgDma, G::kTailDma);
        };
        // BF16 screening stream, per chunk of kB16Tiles tiles: 3 x 16-byte transfers of catalog operands,
        // 5 of the genome's packed operands, one 4-byte transfer of the 64 canonical indices
        const uint32_t chunks16 = (a_tiles + kB16Tiles - 1) / kB16Tiles;
        const Rsrc cat_r = MakeRsrc(p.bf16_cat, p.bf16_cat_bytes);
        const Rsrc ops_r = MakeRsrc(kBf16 ? p.bf16_ops + uint64_t(n) * p.bf16_stride : p.slab, a_tiles * kBf16GenTileBytes);
        const Rsrc canon_r = MakeRsrc(p.canon, p.canon_bytes);
        auto issue16 = [&](uint8_t* buf, uint32_t c) __device__ {
            constexpr uint32_t kCatDma = kB16CatBytes / kDmaBytes;
            for (uint32_t i = wave; i < kB16OpsDma + kB16CanonDma; i += waves) {
                if (i < kCatDma)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(cat_r, (LdsPtr)(buf + i * kDmaBytes), 16,
                                                             c * kB16CatBytes + i * kDmaBytes + 16 * lane, 0, 0, 0);
                else if (i < kB16OpsDma)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(ops_r, (LdsPtr)(buf + i * kDmaBytes), 16,
                                                             c * kB16GenBytes + (i - kCatDma) * kDmaBytes + 16 * lane, 0,
                                                             0, 0);
                else
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(canon_r,
                                                             (LdsPtr)(buf + kB16OpsDma * kDmaBytes + (i - kB16OpsDma) * 256),
                                                             4, c * (kB16Tiles * 64) + (i - kB16OpsDma) * 256 + 4 * lane,
                                                             0, 0, 0);
            }
        };

        const uint32_t per_round = p.splits * waves;
        const uint32_t first_tile = split * waves;
        const uint32_t rounds = tiles > first_tile ? (tiles - first_tile + per_round - 1) / per_round : 0u;
        for (uint32_t round = 0; round < rounds; ++round) {
            const uint32_t tile = first_tile + round * per_round + wave;
            const bool fresh = tile < tiles_fresh;
            const uint32_t game = fresh ? tile * 16 + j : (tile - tiles_fresh) * 16 + j;
            const bool valid = tile < tiles && (fresh ? game < W : game < 2 * W);
            const uint32_t game_slot = fresh ? game : W + game;

            // ---- episode initialisation (evaluation_ops.cuh:171-237)
            // hist[0..tc) are the canonical indices of the guessed words, most recent first (unused
            // entries never match); (last, prev) are the two most recent turns with their feedback,
            // which is all the turn-slot fold ever needs
            uint32_t hist[5] = {kInvalidSlot, kInvalidSlot, kInvalidSlot, kInvalidSlot, kInvalidSlot};
            uint32_t last_code = 0, last_fb = 0, prev_code = 0, prev_fb = 0;
            GameTrace* tr = nullptr;
            uint32_t sol = 0;
            int tc = 0;
            bool won = false;
            bool active = false;
            uint32_t decisions = 0;
            uint32_t rescored = 0, fallbacks = 0;
            if (valid) {
                const uint32_t entry = fresh ? game : (game >> 1);
                const uint32_t variant = fresh ? 0u : 1u + (game & 1u);
                sol = p.codes[local[entry]];
                bool ok = true;
                if (variant != 0) {
                    const uint32_t first = variant == 1 ? entry + 1 : entry + 3;
                    const uint32_t w1 = local[first % W], w2 = local[(first + 1) % W];
                    const uint32_t c1 = p.codes[w1];
                    const uint32_t c2 = p.codes[w2];
                    tc = 1;
                    won = c1 == sol;
                    if (won) {
                        ok = false;  // the second prefill would be appended to a finished grid
                    } else {
                        prev_code = c1;
                        prev_fb = FeedbackRegs(c1, sol);
                        last_code = c2;
                        last_fb = FeedbackRegs(c2, sol);
                        hist[0] = p.canon[w2];
                        hist[1] = p.canon[w1];
                        tc = 2;
                        won = c2 == sol;
                    }
                }
                if (!ok) {
                    if (gq == 0) SetStatus(p.status, 6);  // kGuessAppendFailed
                } else {
                    active = !won && tc < kMaxTurns;
                }
                if (kTrace && p.game_trace != nullptr && gq == 0) {
                    tr = p.game_trace + p.game_offsets[n] + game_slot;
                    for (int d = 0; d < 6; ++d) {
                        tr->guesses[d] = kInvalidSlot;
                        tr->best[d] = 0.0f;
                        tr->second[d] = 0.0f;
                    }
                }
            }

            const int base_tc = fresh ? 0 : 2;
            // layer-0 pre-activations, neuron tile t rows 4*gq+r, game column j; between decisions the FP32
            // form parks them in scratch (its scan then has their registers), the BF16 form keeps them
            v4f acc[16];
            int folded = 0;
            for (int step = 0; step < kMaxTurns; ++step) {
                if (!__syncthreads_or(active ? 1 : 0)) break;  // the whole workgroup leaves together
                const bool wave_active = __any(active ? 1 : 0) != 0;
                IssueTrunkChunk(s_buf0, tl.gr, 0, wave, waves, TrunkLaneW1(lane), TrunkLaneW2(lane));  // lands behind the fold
                float pol[kScoreSteps];
                if (wave_active) {
                    const int ttc = base_tc + step;  // turn count of every active game in this tile
                    if (!(step == 0 || (fresh && step == 1))) {
                        if constexpr (!kBf16) {
    #pragma unroll
                            for (int t = 0; t < 16; ++t) acc[t] = AccLoad(acc_r, lane, 0, uint32_t(t));
                        }
                    } else {
                        InitLayer0(acc, tl);
                        if (fresh && step == 0) {  // acc += W0[:, 0] (virgin flag), from the contiguous copy
    #pragma unroll
                            for (int t = 0; t < 16; ++t)
                                acc[t] += Buf4(enc_r, uint32_t(kTurnFeatures * kEncOut + 16 * t + 4 * gq) * 4, 0);
                        }
                        folded = 0;
                    }
                    // ---- fold newly occupied turn slots: acc += W0[:, 1+64*slot ...] . enc(turn slot)
                    for (int ts = folded; ts < ttc; ++ts) {
                        const bool is_last = ts == ttc - 1;
                        float e[16];
                        EncodeTurnGlobal(enc_r, tl, is_last ? last_code : prev_code, is_last ? last_fb : prev_fb, e);
                        FoldSlot(acc, tl, ts, e);
                    }
                    folded = ttc;
                }
                // layers 1 and 2; the scoring stream's first chunk goes into s_buf0 during the last chunk
                static_assert(kTrunkChunkBytes <= kBufBytes && kTrunkReadBytes <= kBufBytes, "trunk chunk");
                TrunkStream<kT16>(acc, tl, pol, wave_active, s_buf0, s_buf1, wave, waves, lane, [&](uint8_t* b) __device__ {
                    if constexpr (kBf16)
                        issue16(b, 0);
                    else
                        issue(b, 0);
                });
                if (!kBf16 && wave_active) {  // parked until the next decision's fold
    #pragma unroll
                    for (int t = 0; t < 16; ++t) AccStore(acc_r, lane, 0, uint32_t(t), acc[t]);
                }
                if (!wave_active) {
    #pragma unroll
                    for (int s = 0; s < kScoreSteps; ++s) pol[s] = 0.0f;
                }

                // ---- scoring: S = E . P over all actions, masked argmax (lowest index on ties)
                float best = -__builtin_inff(), second = -__builtin_inff();
                uint32_t besta = kInvalidSlot;
                if constexpr (!kBf16) {
                    ArgmaxLane am = {-__builtin_inff(), -__builtin_inff(), kInvalidSlot, -__builtin_inff()};
                    // per chunk: the operands of tile k + 1 are read from LDS before tile k's contraction, so
                    // their latency hides behind its 16 MFMAs
                    StreamChunksWhole(s_buf0, s_buf1, chunks, wave_active, issue, [&](const uint8_t* buf, uint32_t c) __device__ {
                        if (!kTrace) ShareArgmaxGate(am);
                        float ev[kScoreSteps];
                        TileOperands(Fp32Img<kCT>(buf, 0), Fp32Tails<kCT>(buf, 0), lane, j, gq, ev);
#pragma unroll
                        for (int k = 0; k < kCT; ++k) {
                            float nx[kScoreSteps];
                            if (k + 1 < kCT) TileOperands(Fp32Img<kCT>(buf, k + 1), Fp32Tails<kCT>(buf, k + 1), lane, j, gq, nx);
                            const uint32_t at = c * uint32_t(kCT) + uint32_t(k);
                            ArgmaxUpdate<kTrace>(Fp32Scores(ev, pol), Fp32Img<kCT>(buf, k) + 512 + 4 * gq, at * 16 + 4 * gq, A,
                                                 hist, am);
                            if (k + 1 < kCT) {
#pragma unroll
                                for (int s2 = 0; s2 < kScoreSteps; ++s2) ev[s2] = nx[s2];
                            }
                        }
                    });
                    best = am.best;
                    second = am.second;
                    besta = am.besta;
                    ReduceArgmax(best, second, besta);
                } else {
                    // BF16 screening with exact FP32 rescoring.  With |S_bf16 - S_fp32| <= eps for every
                    // action, the FP32 argmax (and every action tying with it) satisfies S_bf16 >= b16 - 2 eps,
                    // b16 = the BF16 masked maximum.  One screening pass keeps, per lane, every unmasked row
                    // with S_bf16 >= (a running maximum of the game's unmasked rows seen so far: the lane's own,
                    // merged with its 3 partner lanes' once per chunk) - 2 eps, a superset of that window since
                    // the running maximum never exceeds b16; the window's members are then rescored with the
                    // FP32 contraction itself, so the FP32 argmax and its score are reproduced bit for bit.
                    // Decisions the window cannot settle (overflowing lists, non-finite scores) take the
                    // FP32 scan.
                    // each wave's candidate table lives in s_buf0, free from the end of the stream until the
                    // fallback's or the next decision's first transfer
                    static_assert(sizeof(uint32_t) * kEvalMaxWaves * 16 * kMaxCandidates <= kBufBytes, "candidate table");
                    auto s_cand = reinterpret_cast<uint32_t (*)[16][kMaxCandidates]>(s_buf0);
                    __shared__ uint32_t s_ncand[kEvalMaxWaves][16];
                    __shared__ float s_lsc[kEvalMaxWaves][kLaneCandidates][64];
                    __shared__ uint32_t s_lac[kEvalMaxWaves][kLaneCandidates][64];
                    // the policy as the B operands of the two screening instructions: every lane writes its
                    // dims (bf16) at their k-slots in this wave's 2 KB of s_buf1, and reads back its own 2 x 8
                    PolicyBf16 pb;
                    {
                        // s_buf1 is free between two barriers: TrunkStream's closing barrier (every wave has
                        // consumed the last trunk chunk, which sat in s_buf1) and StreamChunks' opening
                        // barrier (after which the screening stream's second chunk is issued into s_buf1)
                        static_assert(kEvalMaxWaves * 2048 <= kBufBytes, "policy exchange");
                        __bf16* xb = reinterpret_cast<__bf16*>(s_buf1 + wave * 2048);
    #pragma unroll
                        for (int s = 0; s < kScoreSteps; ++s) {
                            const int d = ScoreDim(s, int(gq));
                            if (d < 0) continue;
                            int slot;
                            const int m = Bf16DimSlot(d, &slot);
                            xb[m * 512 + (int(j) + 16 * (slot >> 3)) * 8 + (slot & 7)] = __bf16(pol[s]);
                        }
                        __builtin_amdgcn_wave_barrier();
                        pb.k[0] = *reinterpret_cast<const bf16x8*>(xb + lane * 8);
                        pb.k[1] = *reinterpret_cast<const bf16x8*>(xb + 512 + lane * 8);
                    }
                    // the bound: sum over this lane group's steps of M_d |p_d|, then over the 4 groups
                    float esum = 0.0f, pabs = 0.0f;
    #pragma unroll
                    for (int s = 0; s < kScoreSteps; ++s) {
                        const int d = ScoreDim(s, int(gq));
                        const float m = d < 0 ? 0.0f
                                              : (d < kFixed ? p.fixed_absmax[d]
                                                            : p.tail_absmax[uint64_t(n) * 40 + uint32_t(d - kFixed)]);
                        esum += m * fabsf(pol[s]);
                        pabs += m + fabsf(pol[s]);
                    }
                    esum += __shfl_xor(esum, 16);
                    esum += __shfl_xor(esum, 32);
                    pabs += __shfl_xor(pabs, 16);
                    pabs += __shfl_xor(pabs, 32);
                    const float eps = kBf16RelBound * esum + kBf16AbsBound * (pabs + 1.0f);
                    const float eps2 = 2.0f * eps;
                    bool fail = active && (p.force_fallback != 0 || !(eps <= 1.0e37f));
                    const bool want = active && !fail;
                    // lane state: running masked maximum, its threshold, list length, overflow / non-finite
                    float lbest = -__builtin_inff(), lthr = -__builtin_inff();
                    uint32_t lcnt = 0;
                    bool lfail = false;
                    float llost = -__builtin_inff();  // best score dropped from a full list
                    // one tile's screening scores: running maximum, window threshold, candidate lists
                    auto screen = [&](const v4f& sc, uint32_t at, auto&& canon_of) __device__ {
                            const float m4 = fmaxf(fmaxf(sc[0], sc[1]), fmaxf(sc[2], sc[3]));
                            // a NaN or infinite score anywhere in the row quad (or an overflowing sum of
                            // four) sends the decision to the FP32 scan
                            lfail = lfail || !(fabsf((sc[0] + sc[1]) + (sc[2] + sc[3])) <= 3.0e38f);
                            if (!__any((want && m4 >= lthr) ? 1 : 0)) return;
                            const uint32_t ca = at * 16 + 4 * gq;
                            const v4u cn = p.canon_identity ? v4u{ca, ca + 1, ca + 2, ca + 3} : canon_of();
    #pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const uint32_t a = at * 16 + 4 * gq + uint32_t(r);
                                const uint32_t c = cn[r];
                                const bool masked = (c == hist[0]) | (c == hist[1]) | (c == hist[2]) | (c == hist[3]) |
                                                    (c == hist[4]);
                                const float v = sc[r];
                                if (!(want && a < A && !masked && v >= lthr)) continue;
                                if (v > lbest) {
                                    lbest = v;
                                    lthr = v - eps2;
                                }
                                if (lcnt == kLaneCandidates) {  // drop what the risen threshold excludes
                                    uint32_t w = 0;
                                    for (uint32_t q = 0; q < kLaneCandidates; ++q) {
                                        const float sq = s_lsc[wave][q][lane];
                                        if (sq >= lthr) {
                                            s_lsc[wave][w][lane] = sq;
                                            s_lac[wave][w][lane] = s_lac[wave][q][lane];
                                            ++w;
                                        }
                                    }
                                    lcnt = w;
                                }
                                if (lcnt < kLaneCandidates) {
                                    s_lsc[wave][lcnt][lane] = v;
                                    s_lac[wave][lcnt][lane] = a;
                                    ++lcnt;
                                } else {  // full: keep the best kLaneCandidates, remember the best score dropped
                                    uint32_t qmin = 0;
                                    float smin = s_lsc[wave][0][lane];
                                    for (uint32_t q = 1; q < kLaneCandidates; ++q) {
                                        const float sq = s_lsc[wave][q][lane];
                                        if (sq < smin) {
                                            smin = sq;
                                            qmin = q;
                                        }
                                    }
                                    if (v > smin) {
                                        s_lsc[wave][qmin][lane] = v;
                                        s_lac[wave][qmin][lane] = a;
                                        llost = fmaxf(llost, smin);
                                    } else {
                                        llost = fmaxf(llost, v);
                                    }
                                }
                            }
                    };
                    StreamChunks<kB16Tiles>(
                        s_buf0, s_buf1, chunks16, __any(want ? 1 : 0) != 0, issue16,
                        [&](const uint8_t* buf, int k, uint32_t at) __device__ {
                            if (k == 0) {
                                // once per chunk, the game's running maximum over its 4 lane groups: still at
                                // most the final BF16 maximum b16, so the window stays a superset, and every
                                // lane's threshold rises with 16 rows per tile instead of 4, so fewer tiles
                                // take the list path (measured: 11.0 -> 9.4 ms per evaluation at 64x64 /
                                // 4,739 actions; every tile instead of every chunk: 9.8 ms)
                                float gb = fmaxf(lbest, __shfl_xor(lbest, 16));
                                gb = fmaxf(gb, __shfl_xor(gb, 32));
                                if (gb > lbest) {
                                    lbest = gb;
                                    lthr = gb - eps2;
                                }
                            }
                            screen(Bf16Screen(buf, k, lane, pb), at, [&]() __device__ {
                                return *reinterpret_cast<const v4u*>(Bf16Canon(buf, k, gq));
                            });
                        });
                    // the game's window: b16 over its 4 lane groups, then every listed row within 2 eps
                    float b16 = lbest;
                    b16 = fmaxf(b16, __shfl_xor(b16, 16));
                    b16 = fmaxf(b16, __shfl_xor(b16, 32));
                    const float thr = b16 - eps2;
                    // a dropped score inside the window means the lists may miss a window member
                    int gfail = (lfail || llost >= thr) ? 1 : 0;
                    gfail |= __shfl_xor(gfail, 16);
                    gfail |= __shfl_xor(gfail, 32);
                    fail = fail || (want && (gfail != 0 || !(fabsf(b16) <= 3.0e38f)));
                    if (lane < 16) s_ncand[wave][lane] = 0u;
                    __builtin_amdgcn_wave_barrier();
                    if (want && !fail) {
                        for (uint32_t q = 0; q < lcnt; ++q) {
                            if (s_lsc[wave][q][lane] >= thr) {
                                const uint32_t slot_k = atomicAdd(&s_ncand[wave][j], 1u);
                                if (slot_k < kMaxCandidates) s_cand[wave][j][slot_k] = s_lac[wave][q][lane];
                            }
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                    const uint32_t ncand = s_ncand[wave][j];
                    fail = fail || (want && (ncand == 0 || ncand > kMaxCandidates));
                    const bool resc = active && !fail;
                    uint32_t kmax = resc ? ncand : 0u;
    #pragma unroll
                    for (int off = 1; off < 64; off <<= 1) kmax = max(kmax, uint32_t(__shfl_xor(int(kmax), off)));
                    kmax = __builtin_amdgcn_readfirstlane(kmax);
                    for (uint32_t k = 0; k < kmax; ++k) {
                        // candidate k of game i sits in MFMA row i; the diagonal D[j][j] is game j's score
                        const uint32_t cand = (resc && k < ncand) ? s_cand[wave][j][k] : 0u;
                        float ev[kScoreSteps];
                        ActionOperands(img_r, tail_r, cand, gq, ev);
                        const v4f sc = Fp32Scores(ev, pol);
                        const float mine = (j & 3u) == 0 ? sc[0] : ((j & 3u) == 1 ? sc[1] : ((j & 3u) == 2 ? sc[2] : sc[3]));
                        const float v = __shfl(mine, int(j + 16 * (j >> 2)));
                        if (resc && k < ncand) {
                            const bool better = besta == kInvalidSlot || v > best || (v == best && cand < besta);
                            if (kTrace) {
                                const float demoted = besta == kInvalidSlot ? second : fmaxf(second, best);
                                second = better ? demoted : fmaxf(second, v);
                            }
                            best = better ? v : best;
                            besta = better ? cand : besta;
                        }
                    }
                    if (resc && gq == 0) rescored += ncand;
                    const bool need = active && fail;
                    if (__syncthreads_or(need ? 1 : 0)) {
                        issue(s_buf0, 0);
                        ArgmaxLane af = {-__builtin_inff(), -__builtin_inff(), kInvalidSlot, -__builtin_inff()};
                        StreamChunks<kCT>(s_buf0, s_buf1, chunks, __any(need ? 1 : 0) != 0, issue,
                                          [&](const uint8_t* buf, int k, uint32_t at) __device__ {
                                              ScoreTile<kTrace>(Fp32Img<kCT>(buf, k), Fp32Tails<kCT>(buf, k), lane, j, gq,
                                                                pol, at * 16 + 4 * gq, A, hist, af);
                                          });
                        float fb = af.best, fs = af.second;
                        uint32_t fa = af.besta;
                        ReduceArgmax(fb, fs, fa);
                        if (need) {
                            best = fb;
                            second = fs;
                            besta = fa;
                        }
                    }
                    if (need && gq == 0) ++fallbacks;
                }

                if (kTrace && active && p.policy_trace != nullptr) {
                    float* dst = p.policy_trace + ((p.game_offsets[n] + game_slot) * uint64_t(kMaxTurns) + decisions) * kPol;
    #pragma unroll
                    for (int s = 0; s < kScoreSteps; ++s) {
                        const int d = ScoreDim(s, int(gq));
                        if (d >= 0) dst[d] = pol[s];
                    }
                }
                // ---- apply the guess (wordle_grid.hpp:69-77)
                if (active) {
                    if (besta == kInvalidSlot) {
                        if (gq == 0) SetStatus(p.status, 8);  // kActionSelectionFailed
                        active = false;
                    } else {
                        const uint32_t code = p.codes[besta];
    #pragma unroll
                        for (int t = 4; t > 0; --t) hist[t] = hist[t - 1];
                        hist[0] = p.canon[besta];
                        prev_code = last_code;
                        prev_fb = last_fb;
                        last_code = code;
                        last_fb = FeedbackRegs(code, sol);
                        if (kTrace && tr != nullptr) {  // every active game has taken `step` decisions
                            tr->guesses[step] = besta;
                            tr->best[step] = best;
                            tr->second[step] = second;
                        }
                        ++tc;
                        ++decisions;
                        won = code == sol;
                        active = !won && tc < kMaxTurns;
                    }
                }
            }

            // ---- score the finished episodes (evaluation_ops.cuh:239-242) and sum over the tile
            float score = (valid && won) ? 10.0f + float(kMaxTurns - tc) : 0.0f;
            uint32_t dec = (valid && gq == 0) ? decisions : 0u;
            // turns that went through the encoder: every turn present at the game's last decision
            uint32_t enc = (valid && gq == 0 && decisions > 0) ? uint32_t(base_tc) + decisions - 1u : 0u;
            if (gq != 0) score = 0.0f;
            if (kTrace && tr != nullptr) {
                tr->decisions = decisions;
                tr->turn_count = uint32_t(tc);
                tr->score = score;
                tr->won = won ? 1u : 0u;
            }
    #pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                score += __shfl_xor(score, off);
                dec += __shfl_xor(dec, off);
                enc += __shfl_xor(enc, off);
                if (kBf16) {
                    rescored += __shfl_xor(rescored, off);
                    fallbacks += __shfl_xor(fallbacks, off);
                }
            }
            if (lane == 0 && tile < tiles) {
                atomicAdd(p.raw_fitness + n, score);
                atomicAdd(p.counters + 0, (unsigned long long)dec);
                atomicAdd(p.counters + 1, (unsigned long long)enc);
                if (kBf16) {
                    atomicAdd(p.counters + 2, (unsigned long long)rescored);
                    atomicAdd(p.counters + 3, (unsigned long long)fallbacks);
                }
            }
        }
    }
}

// ================================================================================================
// Evaluation with two game tiles per wave (the untraced FP32 form).
//
// The same games, decisions and arithmetic as EvaluateKernel<false, false>, with every wave playing two
// tiles of 16 games: a 3-wave workgroup then holds all 96 games of a 32-secret cell, so a genome's trunk
// weights and tail rows stream through LDS once per decision instead of once per split, and every operand
// read from LDS feeds two independent MFMA chains.  The layer-0 accumulators of two tiles (2 x 64 VGPRs)
// do not fit beside the rest, so they live in a global scratch area of the workgroup (16 KB per tile,
// lane-major, so each access of a wave is one contiguous KB): the fold loads, updates and stores them, and
// the trunk reads them back one neuron tile per layer-1 chunk.  Per tile, every MFMA runs in the same
// order as in EvaluateKernel, so both kernels give the same guesses, scores and fitness.

// One lane's game: lane (j, gq) of a tile plays game column j (all four lane groups hold the same state).
struct GameLane {
    uint32_t hist[5];  // canonical indices of the guessed words, most recent first
    uint32_t last_code, last_fb, prev_code, prev_fb;  // the two most recent turns and their feedback
    uint32_t sol;
    int tc;
    uint32_t decisions;
    bool won, active, valid;
};

// Episode initialisation (evaluation_ops.cuh:171-237), as in EvaluateKernel: fresh games first, then the
// two prefilled variants of every local secret.
__device__ __forceinline__ void InitGameLane(const EvalArgs& p, const uint32_t* local, uint32_t W, uint32_t tiles_fresh,
                                             uint32_t tiles, uint32_t tile, uint32_t j, uint32_t gq, GameLane& g) {
    const bool fresh = tile < tiles_fresh;
    const uint32_t game = fresh ? tile * 16 + j : (tile - tiles_fresh) * 16 + j;
    g.valid = tile < tiles && (fresh ? game < W : game < 2 * W);
#pragma unroll
    for (int h = 0; h < 5; ++h) g.hist[h] = kInvalidSlot;
    g.last_code = g.last_fb = g.prev_code = g.prev_fb = 0;
    g.sol = 0;
    g.tc = 0;
    g.decisions = 0;
    g.won = false;
    g.active = false;
    if (!g.valid) return;
    const uint32_t entry = fresh ? game : (game >> 1);
    const uint32_t variant = fresh ? 0u : 1u + (game & 1u);
    g.sol = p.codes[local[entry]];
    bool ok = true;
    if (variant != 0) {
        const uint32_t first = variant == 1 ? entry + 1 : entry + 3;
        const uint32_t w1 = local[first % W], w2 = local[(first + 1) % W];
        const uint32_t c1 = p.codes[w1];
        const uint32_t c2 = p.codes[w2];
        g.tc = 1;
        g.won = c1 == g.sol;
        if (g.won) {
            ok = false;  // the second prefill would be appended to a finished grid
        } else {
            g.prev_code = c1;
            g.prev_fb = FeedbackRegs(c1, g.sol);
            g.last_code = c2;
            g.last_fb = FeedbackRegs(c2, g.sol);
            g.hist[0] = p.canon[w2];
            g.hist[1] = p.canon[w1];
            g.tc = 2;
            g.won = c2 == g.sol;
        }
    }
    if (!ok) {
        if (gq == 0) SetStatus(p.status, 6);  // kGuessAppendFailed
    } else {
        g.active = !g.won && g.tc < kMaxTurns;
    }
}

// One pass of the layer-0 fold over the 16 neuron tiles of a wave's two game tiles, one neuron tile at a
// time, so only 4 accumulators per game tile are live: tile q (when on_q) starts from the bias (+ W0[:, 0]
// where virgin_q) when restart_q, else from scratch; where add_q it gains W0[:, 1+64 slot ...] . e_q (the
// 16 MFMAs of FoldSlot, same order) -- one slot for both tiles, whose W0 operands are loaded once; the
// result goes back to scratch.  Everything neuron tile t + 1 starts from (stored accumulators or bias and
// virgin column, and its W0 operands) loads while tile t folds.
__device__ __forceinline__ void FoldPairPass(Rsrc acc_r, const TrunkLane& tl, Rsrc enc_r, int slot,
                                             bool on0, bool restart0, bool virgin0, bool add0, const float (&e0)[16],
                                             bool on1, bool restart1, bool virgin1, bool add1, const float (&e1)[16],
                                             uint32_t lane) {
    const uint32_t so = uint32_t(64 * slot) * 4;
    const bool any_add = add0 || add1;
    const v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};
    // the loads neuron tile t starts from: W0 operands, and per game tile its base (stored accumulators or
    // the bias) and the virgin column
    auto issue = [&](int t, v4f (&w)[4], v4f& b0, v4f& v0, v4f& b1, v4f& v1) __device__ {
        const uint32_t row = uint32_t(16 * t * kModelIn) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = any_add ? Buf4(tl.gr, tl.off_fold, so + row + 16 * q) : zero;
        const uint32_t voff = uint32_t(kTurnFeatures * kEncOut + 16 * t + 4 * tl.gq) * 4;
        b0 = !on0 ? zero : (restart0 ? Buf4(tl.gr, tl.off_b0, 64 * t) : AccLoad(acc_r, lane, 0, t));
        v0 = (on0 && virgin0) ? Buf4(enc_r, voff, 0) : zero;
        b1 = !on1 ? zero : (restart1 ? Buf4(tl.gr, tl.off_b0, 64 * t) : AccLoad(acc_r, lane, 1, t));
        v1 = (on1 && virgin1) ? Buf4(enc_r, voff, 0) : zero;
    };
    v4f w[4], b0, v0, b1, v1;
    issue(0, w, b0, v0, b1, v1);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        v4f nw[4], nb0, nv0, nb1, nv1;
        if (t + 1 < 16) issue(t + 1, nw, nb0, nv0, nb1, nv1);
        v4f a0 = virgin0 ? b0 + v0 : b0;
        v4f a1 = virgin1 ? b1 + v1 : b1;
        if (add0 && add1) {
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                a0 = Mfma(w[s / 4][s % 4], e0[s], a0);
                a1 = Mfma(w[s / 4][s % 4], e1[s], a1);
            }
        } else if (add0) {
#pragma unroll
            for (int s = 0; s < 16; ++s) a0 = Mfma(w[s / 4][s % 4], e0[s], a0);
        } else if (add1) {
#pragma unroll
            for (int s = 0; s < 16; ++s) a1 = Mfma(w[s / 4][s % 4], e1[s], a1);
        }
        if (on0) AccStore(acc_r, lane, 0, t, a0);
        if (on1) AccStore(acc_r, lane, 1, t, a1);
        __builtin_amdgcn_sched_barrier(0);
        if (t + 1 < 16) {
#pragma unroll
            for (int q = 0; q < 4; ++q) w[q] = nw[q];
            b0 = nb0;
            v0 = nv0;
            b1 = nb1;
            v1 = nv1;
        }
    }
}

// Layer-0 fold of a wave's two game tiles for decision `step` (EvaluateKernel's fold, same order per
// accumulator).  The older prefilled turn (slot 0) of a prefilled tile went in before the first decision
// (FoldPrefill), so every decision adds at most one slot per tile, that of its latest turn: a fresh tile
// restarts at its first decision (bias + virgin column, no slot) and at its second (bias + slot 0, the grid
// no longer empty); otherwise a tile adds slot ttc - 1 to its stored accumulators.  Tiles of the same kind
// add the same slot in one pass.
__device__ __forceinline__ void FoldPair(Rsrc acc_r, const TrunkLane& tl, Rsrc enc_r,
                                         const GameLane& g0, const GameLane& g1, bool act0, bool act1, bool fresh0,
                                         bool fresh1, int step, uint32_t lane) {
    const int ttc0 = (fresh0 ? 0 : 2) + step, ttc1 = (fresh1 ? 0 : 2) + step;
    const bool rs0 = fresh0 && step <= 1, rs1 = fresh1 && step <= 1;
    const bool add0 = act0 && ttc0 > 0, add1 = act1 && ttc1 > 0;
    float e0[16], e1[16];
    if (add0) EncodeTurnGlobal(enc_r, tl, g0.last_code, g0.last_fb, e0);
    if (add1) EncodeTurnGlobal(enc_r, tl, g1.last_code, g1.last_fb, e1);
    const int slot0 = ttc0 - 1, slot1 = ttc1 - 1;
    if (!(add0 && add1) || slot0 == slot1) {
        FoldPairPass(acc_r, tl, enc_r, add0 ? slot0 : slot1, act0, rs0, rs0 && step == 0, add0, e0, act1,
                     rs1, rs1 && step == 0, add1, e1, lane);
    } else {  // a fresh and a prefilled tile: one pass each
        FoldPairPass(acc_r, tl, enc_r, slot0, act0, rs0, rs0 && step == 0, add0, e0, false, false, false,
                     false, e1, lane);
        FoldPairPass(acc_r, tl, enc_r, slot1, false, false, false, false, e0, act1, rs1, rs1 && step == 0,
                     add1, e1, lane);
    }
}

// Before the first decision: the accumulators of every active prefilled tile become bias + W0[:, slot 0]
// . enc(older prefilled turn).
__device__ __forceinline__ void FoldPrefill(Rsrc acc_r, const TrunkLane& tl, Rsrc enc_r,
                                            const GameLane& g0, const GameLane& g1, bool pre0, bool pre1,
                                            uint32_t lane) {
    float e0[16], e1[16];
    if (pre0) EncodeTurnGlobal(enc_r, tl, g0.prev_code, g0.prev_fb, e0);
    if (pre1) EncodeTurnGlobal(enc_r, tl, g1.prev_code, g1.prev_fb, e1);
    FoldPairPass(acc_r, tl, enc_r, 0, pre0, true, false, pre0, e0, pre1, true, false, pre1, e1, lane);
}

// Layers 1 and 2 of two tiles for the whole workgroup, TrunkStream's chunks and MFMA order per tile;
// the layer-0 accumulators come from scratch (acc_r), one neuron tile per layer-1 chunk, loaded
// one chunk ahead.
// Layers 1 and 2 of two tiles for the whole workgroup, TrunkStream's chunks and MFMA order per tile;
// the layer-0 accumulators come from scratch (acc_r), one neuron tile per layer-1 chunk, loaded
// one chunk ahead.  The layer-2 biases travel with the first layer-2 chunk (one 256-byte transfer behind
// it), so no register holds them during layer 1.
constexpr uint32_t kTrunkBiasOff = 9216;  // byte offset of the layer-2 biases in that chunk's buffer
static_assert(kTrunkBiasOff >= kTrunkReadBytes && kTrunkBiasOff % 16 == 0, "behind every layer-2 read");
template <bool kTwo, bool kT16 = false, typename FN>
__device__ __forceinline__ void TrunkStreamPair(Rsrc acc_r, const TrunkLane& tl,
                                                float (&pol0)[kScoreSteps], float (&pol1)[kScoreSteps], bool compute,
                                                uint8_t* b0, uint8_t* b1, uint32_t wave, uint32_t waves, uint32_t lane,
                                                FN&& next_stage) {
    static_assert(kTrunkW1Chunks % 2 == 0 && kTrunkW2Chunks % 2 == 0, "the last chunk sits in b1");
    const uint32_t lw1 = TrunkLaneW1(lane), lw2 = TrunkLaneW2(lane);
    v4f h1a[8], h1b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        h1a[u] = Buf4(tl.gr, tl.off_b1, 64 * u);
        h1b[u] = h1a[u];
    }
    v4f xa = AccLoad(acc_r, lane, 0, 0);
    v4f xb = kTwo ? AccLoad(acc_r, lane, 1, 0) : xa;
    WaitChunk();
    __syncthreads();
    const uint32_t rd1 = 64 * tl.j + 16 * tl.gq;
#pragma unroll
    for (uint32_t t = 0; t < kTrunkW1Chunks; ++t) {
        uint8_t* cur = (t & 1u) ? b1 : b0;
        uint8_t* nxt = (t & 1u) ? b0 : b1;
        IssueTrunkChunk(nxt, tl.gr, t + 1, wave, waves, lw1, lw2);
        v4f na = xa, nb = xb;
        if (t + 1 < kTrunkW1Chunks) {
            na = AccLoad(acc_r, lane, 0, t + 1);
            if (kTwo) nb = AccLoad(acc_r, lane, 1, t + 1);
        } else if (wave == 0) {  // the layer-2 biases, B2[0..64), behind the first layer-2 chunk
            __builtin_amdgcn_raw_ptr_buffer_load_lds(tl.gr, (LdsPtr)(nxt + kTrunkBiasOff), 4, 4 * lane,
                                                     uint32_t(kOffB2) * 4, 0, 0);
        }
        if (compute) {
            float ha[4], hb[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                ha[r] = fmaxf(xa[r], 0.0f);
                hb[r] = fmaxf(xb[r], 0.0f);
            }
            const Split16 ta = SplitBf16(ha[0], ha[1], ha[2], ha[3]), tb = SplitBf16(hb[0], hb[1], hb[2], hb[3]);
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) {
                const v4f w = *reinterpret_cast<const v4f*>(cur + u * 1024 + rd1);
                if constexpr (kT16) {
                    h1a[u] = MfmaT16(w, ta, h1a[u]);
                    if (kTwo) h1b[u] = MfmaT16(w, tb, h1b[u]);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        h1a[u] = Mfma(w[r], ha[r], h1a[u]);
                        if (kTwo) h1b[u] = Mfma(w[r], hb[r], h1b[u]);
                    }
                }
            }
        }
        WaitChunk();
        __syncthreads();
        xa = na;
        xb = nb;
    }
    // layer 2 starts from the biases: pv[u][r] = B2[ScoreDim(4u + r, gq)], as TrunkStream loads them
    constexpr int kRowTiles = kScoreSteps / 4;
    v4f pva[kRowTiles], pvb[kRowTiles];
    {
        const float* sb = reinterpret_cast<const float*>(b0 + kTrunkBiasOff);
        const int nf = ScoreFixedCount(int(tl.gq));
        const int bf = ScoreFixedBase(int(tl.gq)), bt = ScoreTailRow(int(tl.gq));
#pragma unroll
        for (int u = 0; u < kRowTiles; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int st = 4 * u + r;
                pva[u][r] = sb[st < nf ? bf + st : bt + st];
                pvb[u][r] = pva[u][r];
            }
    }
    const uint32_t jr = tl.j & 3u, jg = tl.j >> 2;
    const uint32_t rd_f = (uint32_t(ScoreFixedBase(int(jg))) + jr) * (32 * 4) + 16 * tl.gq;
    const uint32_t rd_t = (uint32_t(ScoreTailRow(int(jg))) + jr) * (32 * 4) + 16 * tl.gq;
    const uint32_t rd_1 = jr + 4 < uint32_t(ScoreFixedCount(int(jg))) ? rd_f + 512 : rd_t + 512;
#pragma unroll
    for (uint32_t c = 0; c < kTrunkW2Chunks; ++c) {
        uint8_t* cur = (c & 1u) ? b1 : b0;
        uint8_t* nxt = (c & 1u) ? b0 : b1;
        if (c + 1 < kTrunkW2Chunks)
            IssueTrunkChunk(nxt, tl.gr, kTrunkW1Chunks + c + 1, wave, waves, lw1, lw2);
        else
            next_stage(nxt);
        if (compute) {
#pragma unroll
            for (uint32_t h = 0; h < 2; ++h) {
                const uint32_t t = 2 * c + h;
                float ha[4], hb[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    ha[r] = fmaxf(h1a[t][r], 0.0f);
                    hb[r] = fmaxf(h1b[t][r], 0.0f);
                }
                const Split16 ta = SplitBf16(ha[0], ha[1], ha[2], ha[3]), tb = SplitBf16(hb[0], hb[1], hb[2], hb[3]);
#pragma unroll
                for (int u = 0; u < kRowTiles; ++u) {
                    const uint32_t rd = u == 0 ? rd_f : (u == 1 ? rd_1 : rd_t + 512 * uint32_t(u));
                    const v4f w = *reinterpret_cast<const v4f*>(cur + rd + 64 * h);
                    if constexpr (kT16) {
                        pva[u] = MfmaT16(w, ta, pva[u]);
                        if (kTwo) pvb[u] = MfmaT16(w, tb, pvb[u]);
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            pva[u] = Mfma(w[r], ha[r], pva[u]);
                            if (kTwo) pvb[u] = Mfma(w[r], hb[r], pvb[u]);
                        }
                    }
                }
            }
        }
        WaitChunk();
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < kScoreSteps; ++s) {
        pol0[s] = pva[s / 4][s % 4];
        pol1[s] = pvb[s / 4][s % 4];
    }
}

// Applies a game's decision (wordle_grid.hpp:69-77), as EvaluateKernel does.
__device__ __forceinline__ void ApplyGuess(const EvalArgs& p, uint32_t besta, uint32_t gq, GameLane& g) {
    if (!g.active) return;
    if (besta == kInvalidSlot) {
        if (gq == 0) SetStatus(p.status, 8);  // kActionSelectionFailed
        g.active = false;
        return;
    }
    const uint32_t code = p.codes[besta];
#pragma unroll
    for (int t = 4; t > 0; --t) g.hist[t] = g.hist[t - 1];
    g.hist[0] = p.canon[besta];
    g.last_code = code;  // (the older turn is no longer needed: every later fold adds the latest turn only)
    g.last_fb = FeedbackRegs(code, g.sol);
    ++g.tc;
    ++g.decisions;
    g.won = code == g.sol;
    g.active = !g.won && g.tc < kMaxTurns;
}

// Scores the finished episodes of a tile (evaluation_ops.cuh:239-242) into the genome's raw fitness and
// the decision / encoded-turn counters (and, in the BF16 form, the rescored-candidate and FP32-fallback
// counters: the lane's counts, nonzero in lane group 0 only).
template <bool kBf16 = false>
__device__ __forceinline__ void FinishTile(const EvalArgs& p, uint32_t n, bool fresh, bool tile_exists, uint32_t lane,
                                           uint32_t gq, const GameLane& g, uint32_t rescored = 0,
                                           uint32_t fallbacks = 0) {
    float score = (g.valid && g.won) ? 10.0f + float(kMaxTurns - g.tc) : 0.0f;
    uint32_t dec = (g.valid && gq == 0) ? g.decisions : 0u;
    const uint32_t base_tc = fresh ? 0u : 2u;
    uint32_t enc = (g.valid && gq == 0 && g.decisions > 0) ? base_tc + g.decisions - 1u : 0u;
    if (gq != 0) score = 0.0f;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        score += __shfl_xor(score, off);
        dec += __shfl_xor(dec, off);
        enc += __shfl_xor(enc, off);
        if (kBf16) {
            rescored += __shfl_xor(rescored, off);
            fallbacks += __shfl_xor(fallbacks, off);
        }
    }
    if (lane == 0 && tile_exists) {
        atomicAdd(p.raw_fitness + n, score);
        atomicAdd(p.counters + 0, (unsigned long long)dec);
        atomicAdd(p.counters + 1, (unsigned long long)enc);
        if (kBf16) {
            atomicAdd(p.counters + 2, (unsigned long long)rescored);
            atomicAdd(p.counters + 3, (unsigned long long)fallbacks);
        }
    }
}

// ---- BF16 screening of one game tile (the BF16 form of EvaluateStreamed) --------------------------------
// EvaluateKernel's BF16 screening, with the lane's state of one game tile gathered in one place, so that a
// wave playing two tiles screens both from one read of each action tile's operands.
struct TileScreen {
    PolicyBf16 pb;      // the tile's policy as the B operands of the two screening instructions
    float eps2;         // twice the screening bound of the lane's game
    bool want;          // the game is active and screened
    bool fail;          // the decision goes to the FP32 scan
    float lbest, lthr;  // the lane's running masked maximum and its window threshold
    float llost;        // the best score dropped from a full list
    uint32_t lcnt;      // list length
    bool lfail;         // a non-finite screening score
};

// The tile's policy as bf16 B operands, exchanged through the wave's 2 KB area `xb` (every lane writes its
// dims at their k-slots and reads back its own 2 x 8; all 1,024 entries are written, so the area can serve
// the next tile at once), and the bound eps = 1.03 2^-8 sum_d M_d |p_d| + 2^-120 (sum (M_d + |p_d|) + 1).
__device__ __forceinline__ void SetupTileScreen(const EvalArgs& p, uint32_t n, const float (&pol)[kScoreSteps],
                                                bool active, __bf16* xb, uint32_t lane, uint32_t j, uint32_t gq,
                                                TileScreen& ts) {
#pragma unroll
    for (int s = 0; s < kScoreSteps; ++s) {
        const int d = ScoreDim(s, int(gq));
        int slot;
        const int m = Bf16DimSlot(d, &slot);
        xb[m * 512 + (int(j) + 16 * (slot >> 3)) * 8 + (slot & 7)] = __bf16(pol[s]);
    }
    __builtin_amdgcn_wave_barrier();
    ts.pb.k[0] = *reinterpret_cast<const bf16x8*>(xb + lane * 8);
    ts.pb.k[1] = *reinterpret_cast<const bf16x8*>(xb + 512 + lane * 8);
    __builtin_amdgcn_wave_barrier();  // the reads precede any later write of the area (LDS keeps a wave's order)
    float esum = 0.0f, pabs = 0.0f;
#pragma unroll
    for (int s = 0; s < kScoreSteps; ++s) {
        const int d = ScoreDim(s, int(gq));
        const float m = d < kFixed ? p.fixed_absmax[d] : p.tail_absmax[uint64_t(n) * 40 + uint32_t(d - kFixed)];
        esum += m * fabsf(pol[s]);
        pabs += m + fabsf(pol[s]);
    }
    esum += __shfl_xor(esum, 16);
    esum += __shfl_xor(esum, 32);
    pabs += __shfl_xor(pabs, 16);
    pabs += __shfl_xor(pabs, 32);
    const float eps = kBf16RelBound * esum + kBf16AbsBound * (pabs + 1.0f);
    ts.eps2 = 2.0f * eps;
    ts.fail = active && (p.force_fallback != 0 || !(eps <= 1.0e37f));
    ts.want = active && !ts.fail;
    ts.lbest = -__builtin_inff();
    ts.lthr = -__builtin_inff();
    ts.llost = -__builtin_inff();
    ts.lcnt = 0;
    ts.lfail = false;
}

// Once per chunk: the game's running maximum over its 4 lane groups (still at most the final BF16 maximum,
// so the window stays a superset) raises every lane's threshold.
__device__ __forceinline__ void ShareRunningMax(TileScreen& ts) {
    float gb = fmaxf(ts.lbest, __shfl_xor(ts.lbest, 16));
    gb = fmaxf(gb, __shfl_xor(gb, 32));
    if (gb > ts.lbest) {
        ts.lbest = gb;
        ts.lthr = gb - ts.eps2;
    }
}

// One action tile's screening scores of the lane's 4 rows into its list (lsc / lac: the lane's kLC slots
// of this game tile, [slot][lane]); EvaluateKernel's screen step.
// With canon_identity (no word occurs twice: the canonical index of action a is a) the masks compare action
// indices and the rows' canonical indices are not read.
template <uint32_t kLC>
__device__ __forceinline__ void ScreenTileRows(const v4f& sc, uint32_t at, const uint32_t* canon_ptr,
                                               bool canon_identity, uint32_t gq, uint32_t A,
                                               const uint32_t (&hist)[5], TileScreen& ts, float (*lsc)[64],
                                               uint32_t (*lac)[64], uint32_t lane) {
    const float m4 = fmaxf(fmaxf(sc[0], sc[1]), fmaxf(sc[2], sc[3]));
    // a NaN or infinite score anywhere in the row quad (or an overflowing sum of four) sends the decision to
    // the FP32 scan
    ts.lfail = ts.lfail || !(fabsf((sc[0] + sc[1]) + (sc[2] + sc[3])) <= 3.0e38f);
    if (!__any((ts.want && m4 >= ts.lthr) ? 1 : 0)) return;
    const uint32_t a0 = at * 16 + 4 * gq;
    const v4u cn = canon_identity ? v4u{a0, a0 + 1, a0 + 2, a0 + 3} : *reinterpret_cast<const v4u*>(canon_ptr);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t a = a0 + uint32_t(r);
        const uint32_t c = cn[r];
        const bool masked = (c == hist[0]) | (c == hist[1]) | (c == hist[2]) | (c == hist[3]) | (c == hist[4]);
        const float v = sc[r];
        if (!(ts.want && a < A && !masked && v >= ts.lthr)) continue;
        if (v > ts.lbest) {
            ts.lbest = v;
            ts.lthr = v - ts.eps2;
        }
        if (ts.lcnt == kLC) {  // drop what the risen threshold excludes
            uint32_t w = 0;
            for (uint32_t q = 0; q < kLC; ++q) {
                const float sq = lsc[q][lane];
                if (sq >= ts.lthr) {
                    lsc[w][lane] = sq;
                    lac[w][lane] = lac[q][lane];
                    ++w;
                }
            }
            ts.lcnt = w;
        }
        if (ts.lcnt < kLC) {
            lsc[ts.lcnt][lane] = v;
            lac[ts.lcnt][lane] = a;
            ++ts.lcnt;
        } else {  // full: keep the best kLC, remember the best score dropped
            uint32_t qmin = 0;
            float smin = lsc[0][lane];
            for (uint32_t q = 1; q < kLC; ++q) {
                const float sq = lsc[q][lane];
                if (sq < smin) {
                    smin = sq;
                    qmin = q;
                }
            }
            if (v > smin) {
                lsc[qmin][lane] = v;
                lac[qmin][lane] = a;
                ts.llost = fmaxf(ts.llost, smin);
            } else {
                ts.llost = fmaxf(ts.llost, v);
            }
        }
    }
}

// The game's window after the screening pass: b16 over its 4 lane groups, then every listed row with
// S16 >= b16 - 2 eps goes into cand[j][0..) (count in *ncount, this wave's counter of the tile); the decision
// fails (FP32 scan) on a non-finite score, a dropped score inside the window, no candidate or more than
// kMaxCandidates.  Returns the game's candidate count.
template <uint32_t kLC>
__device__ __forceinline__ uint32_t CollectWindow(TileScreen& ts, float (*lsc)[64], const uint32_t (*lac)[64],
                                                  uint32_t (*cand)[kMaxCandidates], uint32_t* ncount, uint32_t lane,
                                                  uint32_t j) {
    float b16 = ts.lbest;
    b16 = fmaxf(b16, __shfl_xor(b16, 16));
    b16 = fmaxf(b16, __shfl_xor(b16, 32));
    const float thr = b16 - ts.eps2;
    int gfail = (ts.lfail || ts.llost >= thr) ? 1 : 0;
    gfail |= __shfl_xor(gfail, 16);
    gfail |= __shfl_xor(gfail, 32);
    ts.fail = ts.fail || (ts.want && (gfail != 0 || !(fabsf(b16) <= 3.0e38f)));
    if (lane < 16) ncount[lane] = 0u;
    __builtin_amdgcn_wave_barrier();
    if (ts.want && !ts.fail) {
        for (uint32_t q = 0; q < ts.lcnt; ++q) {
            if (lsc[q][lane] >= thr) {
                const uint32_t k = atomicAdd(&ncount[j], 1u);
                if (k < kMaxCandidates) cand[j][k] = lac[q][lane];
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    const uint32_t nc = ncount[j];
    ts.fail = ts.fail || (ts.want && (nc == 0 || nc > kMaxCandidates));
    return nc;
}

// Rescoring of candidate k of the game (when k < nc and `resc`) by the FP32 contraction itself: candidate
// of game i in MFMA row i, game i's policy in column i, the diagonal D[j][j] is game j's score; the FP32
// argmax over the candidates (lowest index on ties) reproduces the FP32 form's decision and score.
__device__ __forceinline__ void RescoreCandidate(Rsrc img_r, Rsrc tail_r, const uint32_t (*cand)[kMaxCandidates],
                                                 uint32_t k, uint32_t nc, bool resc, const float (&pol)[kScoreSteps],
                                                 uint32_t j, uint32_t gq, float& best, uint32_t& besta) {
    const bool take = resc && k < nc;
    const uint32_t c = take ? cand[j][k] : 0u;
    float ev[kScoreSteps];
    ActionOperands(img_r, tail_r, c, gq, ev);
    const v4f sc = Fp32Scores(ev, pol);
    const float mine = (j & 3u) == 0 ? sc[0] : ((j & 3u) == 1 ? sc[1] : ((j & 3u) == 2 ? sc[2] : sc[3]));
    const float v = __shfl(mine, int(j + 16 * (j >> 2)));
    if (take) {
        const bool better = besta == kInvalidSlot || v > best || (v == best && c < besta);
        best = better ? v : best;
        besta = better ? c : besta;
    }
}

// kBf16: the BF16-screened form (EvaluateKernel<false, true>'s decisions, per tile), with LDS lists of kLC
// entries per lane and game tile for at most kListWaves waves per workgroup.  kRing (BF16 form, workgroups of
// exactly kListWaves = 3 waves): the screening stream runs through three buffers with two chunks in flight
// (StreamChunks3), every wave issuing exactly 3 of a chunk's 9 transfers.
template <int kTiles, int kCT, bool kBf16 = false, int kListWaves = 1, uint32_t kLC = kLaneCandidates,
          bool kRing = false, bool kT16 = false>
__device__ __forceinline__ void EvaluateStreamed(const EvalArgs& p) {
    constexpr bool kTwo = kTiles == 2;
    using G = ChunkGeom<kCT>;
    // chunk buffers: FP32 scan chunks, trunk chunks (with the layer-2 biases behind) and screening chunks
    constexpr uint32_t kTrunkNeed = kTrunkBiasOff + 4 * 68;
    constexpr uint32_t kBufA = G::kBufBytes > kTrunkNeed ? G::kBufBytes : kTrunkNeed;
    constexpr uint32_t kBufBytes = (kBf16 && kB16BufBytes > kBufA) ? kB16BufBytes : kBufA;
    static_assert(kTrunkChunkBytes <= kBufBytes && kTrunkNeed <= kBufBytes && kBufBytes % 16 == 0, "trunk chunk");
    static_assert(!kBf16 || (kListWaves * 2048 <= kBufBytes &&
                             sizeof(uint32_t) * kListWaves * 16 * kMaxCandidates <= kBufBytes),
                  "policy exchange and candidate tables fit a chunk buffer");
    static_assert(!kRing || (kBf16 && kListWaves == 3 && kB16OpsDma + kB16CanonDma == 9 && kTiles == 1),
                  "the ring's waits count 3 transfers per wave and chunk (3-wave workgroups, one tile per wave)");
    __shared__ __align__(16) uint8_t s_buf0[kBufBytes];
    __shared__ __align__(16) uint8_t s_buf1[kBufBytes];
    __shared__ uint32_t s_item;
    const uint32_t tid = threadIdx.x, bd = blockDim.x;
    const uint32_t lane = tid & 63u;
    const uint32_t j = lane & 15u;
    const uint32_t gq = lane >> 4;
    const uint32_t waves = bd >> 6;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const Rsrc acc_r = MakeRsrc(p.acc_scratch + (uint64_t(blockIdx.x) * kEvalMaxWaves + wave) * 2 * kAccTileFloats,
                                2 * kAccTileFloats * sizeof(float));
    // Persistent: (genome, split) items from the work queues; split s of genome n plays tiles
    // kTiles (s waves + r splits waves + wave) + q, q < kTiles (r = 0, 1, ...).
    uint32_t queue = blockIdx.x % kWorkQueues, tried = 0;
    for (;;) {
        __syncthreads();  // the previous item is done with the LDS buffers
        if (tid == 0) s_item = TakeItem(p, queue, tried);
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= p.work_count) break;
        const uint32_t n = p.work[2 * item], split = p.work[2 * item + 1];
        const uint32_t slot = p.slot_indices[n];
        if (slot >= p.slot_count) {
            if (tid == 0) SetStatus(p.status, 4);  // kInvalidGeneration
            continue;
        }
        const float* genome = reinterpret_cast<const float*>(p.slab + uint64_t(slot) * p.stride);
        const Rsrc enc_r = MakeRsrc(p.enc_t + uint64_t(n) * kEncTStride, uint32_t(kEncTStride * 4));
        const uint32_t W = p.cell_offsets[n + 1] - p.cell_offsets[n];
        if (W == 0) {
            if (tid == 0) SetStatus(p.status, 5);  // kInvalidTrainingShard
            continue;
        }
        if (tid == 0 && split == 0) p.local_w[n] = W;
        const uint32_t* local = p.secret_ids + p.cell_offsets[n];
        const uint32_t tiles_fresh = (W + 15) / 16;
        const uint32_t tiles = tiles_fresh + (2 * W + 15) / 16;
        const TrunkLane tl = MakeTrunkLane(genome, j, gq);
        const uint32_t A = p.action_count;
        const uint32_t a_tiles = (A + 15) / 16;
        const uint32_t chunks = (a_tiles + kCT - 1) / kCT;
        const Rsrc img_r = MakeRsrc(p.img, p.img_bytes);
        const Rsrc tail_r = MakeRsrc(genome + kPolicyParams, uint32_t(p.stride - kPolicyParams * 4));
        auto issue = [&](uint8_t* buf, uint32_t c) __device__ {
            IssueChunk(buf, img_r, tail_r, c, wave, waves, lane, G::kImgBytes, G::kTailBytes, G::kImgDma, G::kTailDma);
        };
        // BF16 form: the screening stream, per chunk of kB16Tiles tiles 3 x 16-byte transfers of catalog
        // operands, 5 of the genome's packed operands, one 4-byte transfer of the 64 canonical indices
        const uint32_t chunks16 = (a_tiles + kB16Tiles - 1) / kB16Tiles;
        const Rsrc cat_r = MakeRsrc(p.bf16_cat, p.bf16_cat_bytes);
        const Rsrc ops_r = MakeRsrc(kBf16 ? p.bf16_ops + uint64_t(n) * p.bf16_stride : p.slab, a_tiles * kBf16GenTileBytes);
        const Rsrc canon_r = MakeRsrc(p.canon, p.canon_bytes);
        auto issue16 = [&](uint8_t* buf, uint32_t c) __device__ {
            constexpr uint32_t kCatDma = kB16CatBytes / kDmaBytes;
            for (uint32_t i = wave; i < kB16OpsDma + kB16CanonDma; i += waves) {
                if (i < kCatDma)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(cat_r, (LdsPtr)(buf + i * kDmaBytes), 16,
                                                             c * kB16CatBytes + i * kDmaBytes + 16 * lane, 0, 0, 0);
                else if (i < kB16OpsDma)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(ops_r, (LdsPtr)(buf + i * kDmaBytes), 16,
                                                             c * kB16GenBytes + (i - kCatDma) * kDmaBytes + 16 * lane, 0,
                                                             0, 0);
                else
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(canon_r,
                                                             (LdsPtr)(buf + kB16OpsDma * kDmaBytes + (i - kB16OpsDma) * 256),
                                                             4, c * (kB16Tiles * 64) + (i - kB16OpsDma) * 256 + 4 * lane,
                                                             0, 0, 0);
            }
        };
        const uint32_t per_round = p.splits * waves * kTiles;
        const uint32_t first_tile = split * waves * kTiles;
        const uint32_t rounds = tiles > first_tile ? (tiles - first_tile + per_round - 1) / per_round : 0u;
        for (uint32_t round = 0; round < rounds; ++round) {
            const uint32_t tile0 = first_tile + round * per_round + kTiles * wave, tile1 = kTwo ? tile0 + 1 : tiles;
            const bool fresh0 = tile0 < tiles_fresh, fresh1 = kTwo && tile1 < tiles_fresh;
            uint32_t rescored0 = 0, rescored1 = 0, fallbacks0 = 0, fallbacks1 = 0;  // BF16 form, lane group 0
            GameLane g0, g1;
            InitGameLane(p, local, W, tiles_fresh, tiles, tile0, j, gq, g0);
            InitGameLane(p, local, W, tiles_fresh, tiles, tile1, j, gq, g1);
            {
                const bool pre0 = !fresh0 && __any(g0.active ? 1 : 0) != 0;
                const bool pre1 = kTwo && !fresh1 && __any(g1.active ? 1 : 0) != 0;
                if (pre0 || pre1) FoldPrefill(acc_r, tl, enc_r, g0, g1, pre0, pre1, lane);
            }
            for (int step = 0; step < kMaxTurns; ++step) {
                if (!__syncthreads_or((g0.active || (kTwo && g1.active)) ? 1 : 0)) break;  // the workgroup leaves together
                const bool act0 = __any(g0.active ? 1 : 0) != 0;
                const bool act1 = kTwo && __any(g1.active ? 1 : 0) != 0;
                IssueTrunkChunk(s_buf0, tl.gr, 0, wave, waves, TrunkLaneW1(lane), TrunkLaneW2(lane));
                FoldPair(acc_r, tl, enc_r, g0, g1, act0, act1, fresh0, fresh1, step, lane);
                const bool compute = act0 || act1;
                float pol0[kScoreSteps], pol1[kScoreSteps];
                TrunkStreamPair<kTwo, kT16>(acc_r, tl, pol0, pol1, compute, s_buf0, s_buf1, wave, waves, lane,
                                      [&](uint8_t* b) __device__ {
                                          if constexpr (kBf16)
                                              issue16(b, 0);
                                          else
                                              issue(b, 0);
                                      });
                if (!act0) {
#pragma unroll
                    for (int s = 0; s < kScoreSteps; ++s) pol0[s] = 0.0f;
                }
                if (!act1) {
#pragma unroll
                    for (int s = 0; s < kScoreSteps; ++s) pol1[s] = 0.0f;
                }
                // ---- FP32 scoring: S = E . P over all actions for both tiles, masked argmax (lowest index on
                // ties); the whole workgroup streams the FP32 chunks (chunk 0 already in flight into s_buf0)
                auto fp32_scan = [&](ArgmaxLane& am0, ArgmaxLane& am1, bool scan) __device__ {
                    StreamChunksWhole(s_buf0, s_buf1, chunks, scan, issue, [&](const uint8_t* buf, uint32_t c) __device__ {
                        ShareArgmaxGate(am0);
                        if (kTwo) ShareArgmaxGate(am1);
                        float ev[kScoreSteps];
                        TileOperands(Fp32Img<kCT>(buf, 0), Fp32Tails<kCT>(buf, 0), lane, j, gq, ev);
#pragma unroll
                        for (int k = 0; k < kCT; ++k) {
                            float nx[kScoreSteps];  // the next tile's operands, read during this tile's MFMAs
                            if (k + 1 < kCT) TileOperands(Fp32Img<kCT>(buf, k + 1), Fp32Tails<kCT>(buf, k + 1), lane, j, gq, nx);
                            v4f s0 = {0.0f, 0.0f, 0.0f, 0.0f}, s1 = s0;
#pragma unroll
                            for (int s = 0; s < kScoreSteps; ++s) {
                                s0 = Mfma(ev[s], pol0[s], s0);
                                if (kTwo) s1 = Mfma(ev[s], pol1[s], s1);
                            }
                            const uint32_t at = c * uint32_t(kCT) + uint32_t(k);
                            const uint32_t* cn = Fp32Img<kCT>(buf, k) + 512 + 4 * gq;
                            ArgmaxUpdate<false>(s0, cn, at * 16 + 4 * gq, A, g0.hist, am0);
                            if (kTwo) ArgmaxUpdate<false>(s1, cn, at * 16 + 4 * gq, A, g1.hist, am1);
                            if (k + 1 < kCT) {
#pragma unroll
                                for (int s = 0; s < kScoreSteps; ++s) ev[s] = nx[s];
                            }
                        }
                    });
                    ReduceArgmax(am0.best, am0.second, am0.besta);
                    if (kTwo) ReduceArgmax(am1.best, am1.second, am1.besta);
                };
                if constexpr (!kBf16) {
                    ArgmaxLane am0 = {-__builtin_inff(), -__builtin_inff(), kInvalidSlot, -__builtin_inff()};
                    ArgmaxLane am1 = am0;
                    fp32_scan(am0, am1, compute);
                    ApplyGuess(p, am0.besta, gq, g0);
                    if (kTwo) ApplyGuess(p, am1.besta, gq, g1);
                } else {
                    // ---- BF16 screening with exact FP32 rescoring, per game tile as in EvaluateKernel's BF16
                    // form: one screening stream serves the wave's tiles (each LDS read of an action tile's
                    // operands feeds one pair of instructions per game tile); every tile keeps its own running
                    // maximum, lane lists and window, and its candidates are rescored by the FP32 contraction.
                    __shared__ float s_lsc[kListWaves][kTiles][kLC][64];
                    __shared__ uint32_t s_lac[kListWaves][kTiles][kLC][64];
                    __shared__ uint32_t s_ncand[kListWaves][kTiles][16];
                    TileScreen ts0, ts1;
                    {
                        // the policy exchange in this wave's 2 KB of s_buf1, free between TrunkStreamPair's closing
                        // barrier (every wave has consumed the last trunk chunk, which sat in s_buf1) and
                        // StreamChunks' opening barrier (after which the screening stream's second chunk is
                        // issued into s_buf1)
                        __bf16* xb = reinterpret_cast<__bf16*>(s_buf1 + wave * 2048);
                        SetupTileScreen(p, n, pol0, g0.active, xb, lane, j, gq, ts0);
                        if (kTwo) SetupTileScreen(p, n, pol1, g1.active, xb, lane, j, gq, ts1);
                    }
                    const bool any_want = __any((ts0.want || (kTwo && ts1.want)) ? 1 : 0) != 0;
                    auto screen_tile = [&](const uint8_t* buf, int k, uint32_t at) __device__ {
                        if (k == 0) {
                            ShareRunningMax(ts0);
                            if (kTwo) ShareRunningMax(ts1);
                        }
                        bf16x8 a0, a1;
                        Bf16Operands(buf, k, lane, a0, a1);
                        const uint32_t* cp = Bf16Canon(buf, k, gq);
                        const bool ci = p.canon_identity != 0;
                        const v4f sc0 = Bf16Mfma(a0, a1, ts0.pb);
                        if (kTwo) {
                            const v4f sc1 = Bf16Mfma(a0, a1, ts1.pb);
                            ScreenTileRows<kLC>(sc0, at, cp, ci, gq, A, g0.hist, ts0, s_lsc[wave][0], s_lac[wave][0], lane);
                            ScreenTileRows<kLC>(sc1, at, cp, ci, gq, A, g1.hist, ts1, s_lsc[wave][kTiles - 1],
                                                s_lac[wave][kTiles - 1], lane);
                        } else {
                            ScreenTileRows<kLC>(sc0, at, cp, ci, gq, A, g0.hist, ts0, s_lsc[wave][0], s_lac[wave][0], lane);
                        }
                    };
                    if constexpr (kRing) {
                        __shared__ __align__(16) uint8_t s_buf2[kB16BufBytes];
                        // exactly 3 of the chunk's 9 transfers per wave (3 catalog, 5 genome, 1 canonical-index)
                        auto issue16_3 = [&](uint8_t* buf, uint32_t c) __device__ {
                            constexpr uint32_t kCatDma = kB16CatBytes / kDmaBytes;
#pragma unroll
                            for (uint32_t r = 0; r < 3; ++r) {
                                const uint32_t i = wave + 3 * r;
                                if (i < kCatDma)
                                    __builtin_amdgcn_raw_ptr_buffer_load_lds(cat_r, (LdsPtr)(buf + i * kDmaBytes), 16,
                                                                             c * kB16CatBytes + i * kDmaBytes + 16 * lane,
                                                                             0, 0, 0);
                                else if (i < kB16OpsDma)
                                    __builtin_amdgcn_raw_ptr_buffer_load_lds(
                                        ops_r, (LdsPtr)(buf + i * kDmaBytes), 16,
                                        c * kB16GenBytes + (i - kCatDma) * kDmaBytes + 16 * lane, 0, 0, 0);
                                else
                                    __builtin_amdgcn_raw_ptr_buffer_load_lds(canon_r, (LdsPtr)(buf + kB16OpsDma * kDmaBytes),
                                                                             4, c * (kB16Tiles * 64) + 4 * lane, 0, 0, 0);
                            }
                        };
                        // software-pipelined: tile k + 1's operand reads and instructions are issued before tile k's
                        // rows are screened, so their latency overlaps that screening (the same scores, screened
                        // in the same order)
                        static_assert(!kRing || !kTwo, "one game tile per wave");
                        auto screen_chunk = [&](const uint8_t* buf, uint32_t c) __device__ {
                            ShareRunningMax(ts0);
                            const bool ci = p.canon_identity != 0;
                            bf16x8 a0, a1;
                            Bf16Operands(buf, 0, lane, a0, a1);
                            v4f sc = Bf16Mfma(a0, a1, ts0.pb);
#pragma unroll
                            for (int k = 0; k < int(kB16Tiles); ++k) {
                                v4f scn = sc;
                                if (k + 1 < int(kB16Tiles)) {
                                    Bf16Operands(buf, k + 1, lane, a0, a1);
                                    scn = Bf16Mfma(a0, a1, ts0.pb);
                                }
                                ScreenTileRows<kLC>(sc, c * kB16Tiles + uint32_t(k), Bf16Canon(buf, k, gq), ci, gq, A,
                                                    g0.hist, ts0, s_lsc[wave][0], s_lac[wave][0], lane);
                                sc = scn;
                            }
                        };
                        StreamChunks3Whole<3>(s_buf0, s_buf1, s_buf2, chunks16, any_want, issue16_3, screen_chunk);
                    } else {
                        StreamChunks<kB16Tiles>(s_buf0, s_buf1, chunks16, any_want, issue16, screen_tile);
                    }
                    // the windows; tile 0's candidate table in s_buf0, tile 1's in s_buf1 (both free from the
                    // end of the stream until the fallback's or the next decision's first transfer)
                    auto cand0 = reinterpret_cast<uint32_t (*)[16][kMaxCandidates]>(s_buf0) + wave;
                    auto cand1 = reinterpret_cast<uint32_t (*)[16][kMaxCandidates]>(s_buf1) + wave;
                    const uint32_t nc0 = CollectWindow<kLC>(ts0, s_lsc[wave][0], s_lac[wave][0], *cand0, s_ncand[wave][0], lane, j);
                    uint32_t nc1 = 0;
                    if (kTwo)
                        nc1 = CollectWindow<kLC>(ts1, s_lsc[wave][kTiles - 1], s_lac[wave][kTiles - 1], *cand1,
                                            s_ncand[wave][kTiles - 1], lane, j);
                    const bool resc0 = g0.active && !ts0.fail, resc1 = kTwo && g1.active && !ts1.fail;
                    uint32_t kmax = max(resc0 ? nc0 : 0u, resc1 ? nc1 : 0u);
#pragma unroll
                    for (int off = 1; off < 64; off <<= 1) kmax = max(kmax, uint32_t(__shfl_xor(int(kmax), off)));
                    kmax = __builtin_amdgcn_readfirstlane(kmax);
                    ArgmaxLane r0 = {-__builtin_inff(), -__builtin_inff(), kInvalidSlot, -__builtin_inff()};
                    ArgmaxLane r1 = r0;
                    for (uint32_t k = 0; k < kmax; ++k) {
                        if (__any((resc0 && k < nc0) ? 1 : 0))
                            RescoreCandidate(img_r, tail_r, *cand0, k, nc0, resc0, pol0, j, gq, r0.best, r0.besta);
                        if (kTwo && __any((resc1 && k < nc1) ? 1 : 0))
                            RescoreCandidate(img_r, tail_r, *cand1, k, nc1, resc1, pol1, j, gq, r1.best, r1.besta);
                    }
                    if (resc0 && gq == 0) rescored0 += nc0;
                    if (resc1 && gq == 0) rescored1 += nc1;
                    const bool need0 = g0.active && ts0.fail, need1 = kTwo && g1.active && ts1.fail;
                    if (__syncthreads_or((need0 || need1) ? 1 : 0)) {
                        issue(s_buf0, 0);
                        ArgmaxLane f0 = {-__builtin_inff(), -__builtin_inff(), kInvalidSlot, -__builtin_inff()};
                        ArgmaxLane f1 = f0;
                        fp32_scan(f0, f1, __any((need0 || need1) ? 1 : 0) != 0);
                        if (need0) r0 = f0;
                        if (need1) r1 = f1;
                    }
                    if (need0 && gq == 0) ++fallbacks0;
                    if (need1 && gq == 0) ++fallbacks1;
                    ApplyGuess(p, r0.besta, gq, g0);
                    if (kTwo) ApplyGuess(p, r1.besta, gq, g1);
                }
            }
            FinishTile<kBf16>(p, n, fresh0, tile0 < tiles, lane, gq, g0, rescored0, fallbacks0);
            if (kTwo) FinishTile<kBf16>(p, n, fresh1, tile1 < tiles, lane, gq, g1, rescored1, fallbacks1);
        }
    }
}

__global__ __launch_bounds__(64 * NE_EVAL_MAX_WAVES) __attribute__((amdgpu_waves_per_eu(NE_EVAL_WAVES_PER_EU, NE_EVAL_WAVES_PER_EU))) void EvaluatePairKernel(EvalArgs p) {
    EvaluateStreamed<2, kChunkTiles>(p);
}

// One game tile per wave with the streamed accumulators: 128 VGPRs (4 waves per SIMD) and chunks of 3 action
// tiles (31 KB of LDS), so five 3-wave workgroups share a CU instead of four.  Measured at 64x64 / 4,739
// actions against EvaluateKernel: 14.66 vs 14.87 ms; ragged 64x64: 201 vs 213.5 ms.  Tried and slower: 2 or
// 4 tiles per chunk, 5 waves per SIMD.
#ifndef NE_STREAMED_WAVES_PER_EU
#define NE_STREAMED_WAVES_PER_EU 4
#endif
#ifndef NE_STREAMED_CHUNK_TILES
#define NE_STREAMED_CHUNK_TILES 3
#endif
__global__ __launch_bounds__(64 * NE_EVAL_MAX_WAVES) __attribute__((amdgpu_waves_per_eu(NE_STREAMED_WAVES_PER_EU, NE_STREAMED_WAVES_PER_EU))) void EvaluateStreamedKernel(EvalArgs p) {
    EvaluateStreamed<1, NE_STREAMED_CHUNK_TILES>(p);
}

// The BF16-screened form on the same two kernels.  Two tiles per wave: a 3-wave workgroup plays the 96 games
// of a 32-secret cell, so each decision streams the genome's packed screening operands and trunk weights once
// instead of once per split, and every read of an action tile's screening operands feeds both tiles' pairs
// of instructions.  FP32 chunks of 2 tiles (the fallback scan's; the trunk chunks and screening chunks fit
// the same 10 KB buffers) and LDS lists for at most 3 waves keep a workgroup at 45 KB of LDS.
#ifndef NE_PAIR_BF16_WAVES
#define NE_PAIR_BF16_WAVES 3
#endif
#ifndef NE_PAIR_BF16_LIST
#define NE_PAIR_BF16_LIST 6
#endif
template <bool kT16>
__global__ __launch_bounds__(64 * NE_PAIR_BF16_WAVES) __attribute__((amdgpu_waves_per_eu(NE_EVAL_WAVES_PER_EU, NE_EVAL_WAVES_PER_EU))) void EvaluatePairBf16Kernel(EvalArgs p) {
    EvaluateStreamed<2, kChunkTilesBf16, true, NE_PAIR_BF16_WAVES, NE_PAIR_BF16_LIST, false, kT16>(p);
}
// One tile per wave in the BF16 form, at exactly 3 waves per workgroup, with the screening stream through a
// 3-buffer ring (two chunks in flight per workgroup instead of one).  FP32 fallback chunks of one tile and LDS
// lists for 3 waves keep the workgroup under 40 KB of LDS, so four still share a CU.  (Measured at 64x64 /
// 4,739 actions: 8.56 ms against 9.25 ms for EvaluateKernel<false, true>, same decisions.  The same code at
// 1-2 waves with 8-entry lists for 4 waves and 128 VGPRs gave run-to-run differences in the count of
// candidates rescored -- never in a decision -- that 6-entry lists did not; it is not built, and the BF16
// form's streamed kernel runs only at 3 waves.)
template <bool kT16>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(NE_EVAL_WAVES_PER_EU, NE_EVAL_WAVES_PER_EU))) void EvaluateStreamedBf16RingKernel(EvalArgs p) {
    EvaluateStreamed<1, 1, true, 3, kLaneCandidates, true, kT16>(p);
}

// ================================================================================================
// Child fingerprint for surrogate screening (kT16: the trunk's layers 1 and 2 on split-BF16 MFMA, as the
// BF16-trunk evaluation form runs them).  One workgroup of 4 waves per child; wave d plays the
// 16 probe boards of depth d (lane column k = probe secret k): the turns are folded into layer 0
// exactly as in a game, layers 1 and 2 stream through LDS, all actions are scored in FP32 with the
// masked argmax of a decision.  Per probe p = 16 d + k the child's guess g yields 3 inputs:
//   [3p]     quality: (2 greens + yellows) / 10 of g against probe secret k;
//   [3p + 1] depth >= 1: 1 if g is consistent with every turn on the board (it would have produced the
//            same feedback as the secret), else 0; depth 0 (no hints): the fitness of parent k % 2,
//            as (f - shift) * scale;
//   [3p + 2] 1 if g is the secret.
template <bool kT16>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NE_EVAL_WAVES_PER_EU, NE_EVAL_WAVES_PER_EU))) void FingerprintKernel(FingerprintArgs p) {
    constexpr int kCT = kChunkTiles;
    using G = ChunkGeom<kCT>;
    __shared__ __align__(16) uint8_t s_buf0[G::kBufBytes];
    __shared__ __align__(16) uint8_t s_buf1[G::kBufBytes];
    static_assert(kTrunkChunkBytes <= G::kBufBytes && kTrunkReadBytes <= G::kBufBytes, "trunk chunk");
    const uint32_t n = blockIdx.x;
    const uint32_t slot = p.slots[n];
    if (slot >= p.slot_count) {
        if (threadIdx.x == 0) SetStatus(p.status, 4);  // kInvalidGeneration
        return;
    }
    const float* genome = reinterpret_cast<const float*>(p.slab + uint64_t(slot) * p.stride);
    const uint32_t tid = threadIdx.x, lane = tid & 63u, j = lane & 15u, gq = lane >> 4;
    const uint32_t waves = blockDim.x >> 6;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t d = wave;  // turns on the board
    const TrunkLane tl = MakeTrunkLane(genome, j, gq);
    const Rsrc enc_r = MakeRsrc(p.enc_t + uint64_t(n) * kEncTStride, uint32_t(kEncTStride * 4));
    const uint32_t A = p.action_count;
    const uint32_t a_tiles = (A + 15) / 16;
    const uint32_t chunks = (a_tiles + kCT - 1) / kCT;
    const Rsrc img_r = MakeRsrc(p.img, p.img_bytes);
    const Rsrc tail_r = MakeRsrc(genome + kPolicyParams, uint32_t(p.stride - kPolicyParams * 4));
    auto issue = [&](uint8_t* buf, uint32_t c) __device__ {
        IssueChunk(buf, img_r, tail_r, c, wave, waves, lane, G::kImgBytes, G::kTailBytes, G::kImgDma, G::kTailDma);
    };
    const ProbeSet& ps = *p.probes;
    uint32_t hist[5] = {kInvalidSlot, kInvalidSlot, kInvalidSlot, kInvalidSlot, kInvalidSlot};
    for (uint32_t t = 0; t < d; ++t) hist[t] = ps.turn_canon[j][t];

    IssueTrunkChunk(s_buf0, tl.gr, 0, wave, waves, TrunkLaneW1(lane), TrunkLaneW2(lane));  // lands behind the fold
    v4f acc[16];
    InitLayer0(acc, tl);
    if (d == 0) {  // virgin flag: the board is empty
#pragma unroll
        for (int t = 0; t < 16; ++t) acc[t] += Buf4(enc_r, uint32_t(kTurnFeatures * kEncOut + 16 * t + 4 * gq) * 4, 0);
    }
    for (uint32_t ts = 0; ts < d; ++ts) {
        float e[16];
        EncodeTurnGlobal(enc_r, tl, ps.turn_code[j][ts], ps.turn_fb[j][ts], e);
        FoldSlot(acc, tl, int(ts), e);
    }
    float pol[kScoreSteps];
    TrunkStream<kT16>(acc, tl, pol, true, s_buf0, s_buf1, wave, waves, lane, [&](uint8_t* b) __device__ { issue(b, 0); });
    ArgmaxLane am = {-__builtin_inff(), -__builtin_inff(), kInvalidSlot, -__builtin_inff()};
    StreamChunks<kCT>(s_buf0, s_buf1, chunks, true, issue, [&](const uint8_t* buf, int k, uint32_t at) __device__ {
        if (k == 0) ShareArgmaxGate(am);
        ScoreTile<false>(Fp32Img<kCT>(buf, k), Fp32Tails<kCT>(buf, k), lane, j, gq, pol, at * 16 + 4 * gq, A, hist, am);
    });
    float best = am.best, second = am.second;
    uint32_t besta = am.besta;
    ReduceArgmax(best, second, besta);
    if (gq != 0) return;
    float* out = p.out + uint64_t(n) * kFingerprintDims + 3 * (d * kProbeSecrets + j);
    if (besta == kInvalidSlot) {  // nothing selectable (every action masked)
        out[0] = 0.0f;
        out[1] = 0.0f;
        out[2] = 0.0f;
        return;
    }
    const uint32_t g = p.codes[besta], sc = ps.secret_code[j];
    const uint32_t fb = FeedbackRegs(g, sc);
    int greens = 0, yellows = 0;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const uint32_t tv = (fb >> (2 * q)) & 3u;
        greens += tv == 0u ? 1 : 0;
        yellows += tv == 1u ? 1 : 0;
    }
    float c;
    if (d == 0) {
        const uint32_t parent = p.pairs[2 * n + (j & 1u)];
        c = (p.parent_fitness[parent] - p.parent_shift) * p.parent_scale;
    } else {
        bool consistent = true;
        for (uint32_t t = 0; t < d; ++t) consistent = consistent && FeedbackRegs(ps.turn_code[j][t], g) == ps.turn_fb[j][t];
        c = consistent ? 1.0f : 0.0f;
    }
    out[0] = float(2 * greens + yellows) * 0.1f;
    out[1] = c;
    out[2] = g == sc ? 1.0f : 0.0f;
}

// Fitness normalisation (FinalizeSlabGenerationFitnessKernel, slab_runtime.cu:551-586 and
// NormalizeFitnessForSelection, evaluation_ops.cuh:97-111).
__global__ void FinalizeKernel(const float* raw, const uint32_t* local_w, uint32_t n_genomes, float* fitness,
                               uint32_t* eval_counts, uint8_t* has_fitness, int* status) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_genomes) return;
    const uint32_t W = local_w[i];
    if (W == 0) {
        SetStatus(status, 5);
        return;
    }
    const float maximum = 3.0f * 15.0f * float(W);
    const float f = raw[i] / maximum;
    fitness[i] = f >= 1.0f ? 1.0f : (f > kFitnessFloor ? f : kFitnessFloor);
    eval_counts[i] += 1u;
    has_fitness[i] = 1;
}

// Population summary (SummarizeSlabGenerationKernel, slab_runtime.cu:588-631): best = max with the
// lowest index on ties (a parallel reduction, exact), mean = the sequential fp32 sum in index order
// (thread 0 over an LDS copy, the reference's summation order).
constexpr uint32_t kSummaryChunk = 8192;
__global__ __launch_bounds__(1024) void SummaryKernel(const float* fitness, const uint8_t* has_fitness,
                                                      const uint32_t* slots, uint32_t n_genomes, SummaryRecord* out,
                                                      int* status) {
    __shared__ float s_fit[kSummaryChunk];
    __shared__ float s_best[16];
    __shared__ uint32_t s_bi[16];
    __shared__ int s_missing;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid == 0) s_missing = 0;
    float best = -__builtin_inff();
    uint32_t bi = kInvalidSlot;
    float sum = 0.0f;
    for (uint32_t base = 0; base < n_genomes; base += kSummaryChunk) {
        const uint32_t len = min(kSummaryChunk, n_genomes - base);
        __syncthreads();
        for (uint32_t k = tid; k < len; k += blockDim.x) {
            const float f = fitness[base + k];
            if (has_fitness[base + k] == 0) s_missing = 1;
            s_fit[k] = f;
            if (bi == kInvalidSlot || f > best) {  // k increases per thread: ties keep the lower index
                best = f;
                bi = base + k;
            }
        }
        __syncthreads();
        if (tid == 0)
            for (uint32_t k = 0; k < len; ++k) sum += s_fit[k];
    }
    for (int off = 32; off > 0; off >>= 1) {
        const float ob = __shfl_xor(best, off);
        const uint32_t oi = __shfl_xor(bi, off);
        if (oi != kInvalidSlot && (bi == kInvalidSlot || ob > best || (ob == best && oi < bi))) {
            best = ob;
            bi = oi;
        }
    }
    if (lane == 0) {
        s_best[wave] = best;
        s_bi[wave] = bi;
    }
    __syncthreads();
    if (tid != 0) return;
    if (s_missing) {
        SetStatus(status, 9);
        return;
    }
    for (uint32_t w = 1; w < (blockDim.x >> 6); ++w)
        if (s_bi[w] != kInvalidSlot && (bi == kInvalidSlot || s_best[w] > best || (s_best[w] == best && s_bi[w] < bi))) {
            best = s_best[w];
            bi = s_bi[w];
        }
    if (bi == kInvalidSlot) {  // empty generation
        best = 0.0f;
        bi = 0;
    }
    out->best_fitness = best;
    out->average_fitness = sum / float(n_genomes);
    out->best_index = bi;
    out->best_slot = slots[bi];
}

// ================================================================================================
// Parent selection (BuildSlabAssemblyPlanKernel, slab_runtime.cu:633-675;
// TrySelectCellularParentPairDevice and helpers, selection_ops.cuh:26-263).
__device__ inline uint64_t SplitMix64(uint64_t v) {
    v += 0x9E3779B97F4A7C15ull;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}

__device__ inline bool RanksBefore(float lf, uint32_t li, float rf, uint32_t ri, uint64_t seed) {
    if (lf < rf) return true;
    if (lf > rf) return false;
    const uint64_t lk = SplitMix64(seed ^ uint64_t(li)), rk = SplitMix64(seed ^ uint64_t(ri));
    return lk < rk || (lk == rk && li < ri);
}

__device__ inline uint32_t AxisDist(uint32_t a, uint32_t b, uint32_t side) {
    const uint32_t f = a >= b ? a - b : b - a;
    const uint32_t w = side - f;
    return f < w ? f : w;
}

__device__ inline uint32_t ChebyshevDist(uint32_t rows, uint32_t cols, uint32_t a, uint32_t b) {
    const uint32_t dr = AxisDist(a / cols, b / cols, rows);
    const uint32_t dc = AxisDist(a % cols, b % cols, cols);
    return dr > dc ? dr : dc;
}

constexpr int kPlanBlock = 64;
constexpr int kMaxMoore = 25;

__global__ __launch_bounds__(kPlanBlock) void PlanKernel(PlanArgs p) {
    __shared__ uint32_t s_cand[kPlanBlock][kMaxMoore];
    if (p.step != nullptr) {
        p.gen_index = p.step[0];
        p.generation_seed = p.step[1];
    }
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= p.child_count) return;
    const uint32_t cols = p.cols, rows = p.cur_size / cols;
    const uint32_t focal = (c / cols) * cols + (c % cols);
    const bool listed = p.radius <= 2;
    uint32_t* cand = s_cand[threadIdx.x];
    uint32_t ncand = 0;
    if (listed) {
        const int r = int(p.radius);
        const int fr = int(focal / cols), fc = int(focal % cols);
        for (int dr = -r; dr <= r; ++dr)
            for (int dc = -r; dc <= r; ++dc) {
                int rr = (fr + dr) % int(rows);
                if (rr < 0) rr += int(rows);
                int cc = (fc + dc) % int(cols);
                if (cc < 0) cc += int(cols);
                const uint32_t idx = uint32_t(rr) * cols + uint32_t(cc);
                bool dup = false;
                for (uint32_t i = 0; i < ncand; ++i) dup |= cand[i] == idx;
                if (!dup) cand[ncand++] = idx;
            }
    }
    const uint32_t total = listed ? ncand : p.cur_size;
    auto cand_at = [&](uint32_t i) -> uint32_t { return listed ? cand[i] : i; };
    auto eligible = [&](uint32_t i, uint32_t excluded) -> bool {
        if (i == excluded || p.has_fitness[i] == 0) return false;
        return listed || ChebyshevDist(rows, cols, focal, i) <= p.radius;
    };
    const uint32_t seed = p.generation_seed ^ kPlanSeedXor;
    uint32_t chosen[2] = {0, 0};
    for (int k = 0; k < 2; ++k) {
        const U4 w = Philox4x32(c, p.gen_index, uint32_t(k), 0, seed, kTagPlan);
        const uint64_t tie = (uint64_t(w.x) << 32) ^ uint64_t(w.y);
        const uint32_t excluded = k == 0 ? kInvalidSlot : chosen[0];
        auto rank_of = [&](uint32_t i) {
            uint32_t rank = 1;
            for (uint32_t q = 0; q < total; ++q) {
                const uint32_t jdx = cand_at(q);
                if (jdx == i || !eligible(jdx, excluded)) continue;
                if (RanksBefore(p.fitness[jdx], jdx, p.fitness[i], i, tie)) ++rank;
            }
            return rank;
        };
        float sum = 0.0f;
        bool found = false;
        uint32_t fallback = 0;
        for (uint32_t q = 0; q < total; ++q) {
            const uint32_t i = cand_at(q);
            if (!eligible(i, excluded)) continue;
            sum += p.rank_weights[rank_of(i)];
            fallback = i;
            found = true;
        }
        if (!found || sum <= 0.0f) {
            SetStatus(p.status, 11);  // kParentSelectionFailed
            return;
        }
        const float threshold = UniformOpenClosed(w.z) * sum;
        float cum = 0.0f;
        uint32_t pick = fallback;
        for (uint32_t q = 0; q < total; ++q) {
            const uint32_t i = cand_at(q);
            if (!eligible(i, excluded)) continue;
            cum += p.rank_weights[rank_of(i)];
            if (threshold <= cum) {
                pick = i;
                break;
            }
        }
        chosen[k] = pick;
    }
    p.pairs[2 * c] = chosen[0];
    p.pairs[2 * c + 1] = chosen[1];
}

// Radius <= 2 (the Moore list of <= 25 distinct cells): one wave per child.  Lane q owns candidate q:
// it builds the (dr, dc)-ordered, de-duplicated list, hashes its tie key once and counts the eligible
// candidates ranked before it; lane 0 then sums the rank weights and scans them in candidate order,
// exactly the sequential fp32 arithmetic of the one-thread form above.
constexpr int kPlanWaves = 4;
__global__ __launch_bounds__(64 * kPlanWaves) void PlanListedKernel(PlanArgs p) {
    __shared__ uint32_t s_raw[kPlanWaves][32];
    __shared__ uint32_t s_idx[kPlanWaves][32];
    __shared__ float s_fit[kPlanWaves][32];
    __shared__ uint64_t s_key[kPlanWaves][32];
    __shared__ uint32_t s_elig[kPlanWaves][32];
    __shared__ float s_wt[kPlanWaves][32];
    if (p.step != nullptr) {
        p.gen_index = p.step[0];
        p.generation_seed = p.step[1];
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t c = blockIdx.x * kPlanWaves + wave;
    if (c >= p.child_count) return;  // whole wave
    const uint32_t cols = p.cols, rows = p.cur_size / cols;
    const int r = int(p.radius), side = 2 * r + 1, cells = side * side;
    const int fr = int(c / cols), fc = int(c % cols);
    const bool in_list = int(lane) < cells;
    uint32_t idx = 0;
    if (in_list) {
        int rr = (fr + int(lane) / side - r) % int(rows);
        if (rr < 0) rr += int(rows);
        int cc = (fc + int(lane) % side - r) % int(cols);
        if (cc < 0) cc += int(cols);
        idx = uint32_t(rr) * cols + uint32_t(cc);
        s_raw[wave][lane] = idx;
    }
    __builtin_amdgcn_wave_barrier();
    bool keep = in_list;
    for (uint32_t q = 0; q < lane && keep; ++q) keep = s_raw[wave][q] != idx;
    const uint64_t keep_mask = __ballot(keep ? 1 : 0);
    const uint32_t ncand = uint32_t(__popcll(keep_mask));
    const uint32_t pos = __builtin_amdgcn_mbcnt_hi(uint32_t(keep_mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(keep_mask), 0u));
    if (keep) s_idx[wave][pos] = idx;
    __builtin_amdgcn_wave_barrier();
    const bool mine = lane < ncand;
    const uint32_t my = mine ? s_idx[wave][lane] : 0u;
    const float my_f = mine ? p.fitness[my] : 0.0f;
    const bool my_has = mine && p.has_fitness[my] != 0;
    const uint32_t seed = p.generation_seed ^ kPlanSeedXor;
    uint32_t chosen[2] = {0, 0};
    for (int k = 0; k < 2; ++k) {
        const U4 w = Philox4x32(c, p.gen_index, uint32_t(k), 0, seed, kTagPlan);
        const uint64_t tie = (uint64_t(w.x) << 32) ^ uint64_t(w.y);
        const uint32_t excluded = k == 0 ? kInvalidSlot : chosen[0];
        const bool elig = my_has && my != excluded;
        const uint64_t key = SplitMix64(tie ^ uint64_t(my));
        if (mine) {
            s_fit[wave][lane] = my_f;
            s_key[wave][lane] = key;
            s_elig[wave][lane] = elig ? 1u : 0u;
        }
        __builtin_amdgcn_wave_barrier();
        uint32_t rank = 1;
        if (elig)
            for (uint32_t q = 0; q < ncand; ++q) {
                if (q == lane || s_elig[wave][q] == 0) continue;
                const float f = s_fit[wave][q];
                const uint64_t kq = s_key[wave][q];
                const uint32_t iq = s_idx[wave][q];
                if (f < my_f || (f == my_f && (kq < key || (kq == key && iq < my)))) ++rank;
            }
        if (mine) s_wt[wave][lane] = elig ? p.rank_weights[rank] : 0.0f;
        __builtin_amdgcn_wave_barrier();
        uint32_t pick = 0;
        int fail = 0;
        if (lane == 0) {
            float sum = 0.0f;
            bool found = false;
            uint32_t fallback = 0;
            for (uint32_t q = 0; q < ncand; ++q) {
                if (s_elig[wave][q] == 0) continue;
                sum += s_wt[wave][q];
                fallback = s_idx[wave][q];
                found = true;
            }
            if (!found || sum <= 0.0f) {
                fail = 1;
            } else {
                const float threshold = UniformOpenClosed(w.z) * sum;
                float cum = 0.0f;
                pick = fallback;
                for (uint32_t q = 0; q < ncand; ++q) {
                    if (s_elig[wave][q] == 0) continue;
                    cum += s_wt[wave][q];
                    if (threshold <= cum) {
                        pick = s_idx[wave][q];
                        break;
                    }
                }
            }
        }
        fail = __builtin_amdgcn_readfirstlane(fail);
        if (fail) {
            if (lane == 0) SetStatus(p.status, 11);  // kParentSelectionFailed
            return;
        }
        chosen[k] = __builtin_amdgcn_readfirstlane(pick);
        __builtin_amdgcn_wave_barrier();  // the LDS rows are rewritten for the second parent
    }
    if (lane == 0) {
        p.pairs[2 * c] = chosen[0];
        p.pairs[2 * c + 1] = chosen[1];
    }
}

// ================================================================================================
// Crossover decisions (genome_ops.cuh:210-221,264-299,344-376), one thread per child.
__global__ void XoverDecisionKernel(XoverArgs p) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= p.child_count) return;
    if (p.step != nullptr) {
        p.gen_index = p.step[0];
        p.seed = p.step[1];
    }
    XoverRecord r;
    const U4 s0 = Philox4x32(c, p.gen_index, 0, 0, p.seed, kTagXovr);
    const bool enc_first = !BernoulliDraw(p.t1, s0.x);
    const bool trunk_first = !BernoulliDraw(p.t1, s0.y);
    const bool tail_first = !BernoulliDraw(p.t1, s0.z);
    for (int l = 0; l < 6; ++l) {
        r.layer_first[l] = 0;
        r.neuron[l] = kInvalidSlot;
    }
    for (int l = 0; l < kDenseLayers; ++l) {
        const U4 d = Philox4x32(c, p.gen_index, 1 + uint32_t(l), 0, p.seed, kTagXovr);
        const bool base = l < kEncoderLayers ? enc_first : trunk_first;
        const bool lf = BernoulliDraw(p.t2, d.x) ? !base : base;
        r.layer_first[l] = lf ? 1u : 0u;
        if (BernoulliDraw(p.t3, d.y)) r.neuron[l] = d.z % DenseLayerAt(l).out;
    }
    r.tail_first = tail_first ? 1u : 0u;
    r.row2 = kInvalidSlot;
    r.splice_row = kInvalidSlot;
    r.splice_point = 0;
    const U4 d8 = Philox4x32(c, p.gen_index, 8, 0, p.seed, kTagXovr);
    const U4 d9 = Philox4x32(c, p.gen_index, 9, 0, p.seed, kTagXovr);
    if (p.parent_action_count > 0 && BernoulliDraw(p.t2, d8.x)) r.row2 = d8.y % p.parent_action_count;
    if (p.parent_action_count > 0 && BernoulliDraw(p.t3, d8.z)) {
        r.splice_row = d8.w % p.parent_action_count;
        r.splice_point = 1u + d9.x % uint32_t(kTail - 1);
    }
    p.records[c] = r;
    // the donor-change positions, ascending (a breeding wave whose 256-value window holds none of them
    // takes one donor for the whole window)
    uint32_t* out = p.breaks + uint64_t(c) * kBreakStride;
    static_assert(7 * kDenseLayers + 6 <= int(kBreakStride) - 1, "break list");
    uint32_t nb = 0;
    uint32_t tmp[kBreakStride - 1];
    auto push = [&](uint32_t b) {
        uint32_t k = nb++;
        for (; k > 0 && tmp[k - 1] > b; --k) tmp[k] = tmp[k - 1];
        tmp[k] = b;
    };
    for (int l = 0; l < kDenseLayers; ++l) {
        const DenseLayerDesc d = DenseLayerAt(l);
        push(d.w);
        push(d.b);
        push(d.b + d.out);
        if (r.neuron[l] != kInvalidSlot) {
            push(d.w + r.neuron[l] * d.in);
            push(d.w + (r.neuron[l] + 1) * d.in);
            push(d.b + r.neuron[l]);
            push(d.b + r.neuron[l] + 1);
        }
    }
    const uint32_t P = uint32_t(kPolicyParams);
    if (r.row2 != kInvalidSlot) {
        push(P + r.row2 * kTail);
        push(P + (r.row2 + 1) * kTail);
    }
    if (r.splice_row != kInvalidSlot) {
        push(P + r.splice_row * kTail);
        push(P + r.splice_row * kTail + r.splice_point);
        push(P + (r.splice_row + 1) * kTail);
    }
    push(P + p.parent_action_count * kTail);
    out[0] = nb;
    for (uint32_t i = 0; i < nb; ++i) out[1 + i] = tmp[i];
}

__device__ inline bool ElementDonorFirst(const XoverRecord& r, uint32_t e) {
    if (e < uint32_t(kPolicyParams)) {
#pragma unroll
        for (int l = 0; l < kDenseLayers; ++l) {
            const DenseLayerDesc d = DenseLayerAt(l);
            const bool in_w = e >= d.w && e < d.w + d.in * d.out;
            const bool in_b = e >= d.b && e < d.b + d.out;
            if (!in_w && !in_b) continue;
            const bool lf = r.layer_first[l] != 0;
            if (r.neuron[l] == kInvalidSlot) return lf;
            const uint32_t nb = r.neuron[l];
            const bool hit = in_w ? (e >= d.w + nb * d.in && e < d.w + (nb + 1) * d.in) : (e == d.b + nb);
            return hit ? !lf : lf;
        }
        return true;
    }
    const bool base = r.tail_first != 0;
    const uint32_t t = e - uint32_t(kPolicyParams);
    bool src = base;
    const bool in_row2 = r.row2 != kInvalidSlot && t >= r.row2 * kTail && t < (r.row2 + 1) * kTail;
    if (in_row2) src = !base;
    if (r.splice_row != kInvalidSlot && t >= r.splice_row * kTail && t < (r.splice_row + 1) * kTail) {
        const bool row_src = in_row2 ? !base : base;
        src = (t - r.splice_row * kTail) < r.splice_point ? row_src : !row_src;
    }
    return src;
}

// The mutated value of element e (taken for about mutation_p of the elements): a normal draw of its own.
__device__ __forceinline__ float MutatedValue(float v, uint32_t e, uint32_t child, uint32_t gen,
                                                          uint32_t seed, float sigma) {
    const U4 z = Philox4x32(e, child, gen, 0, seed, kTagNorm);
    return __builtin_fmaf(sigma, NormalFromWords(z.x, z.y), v);
}

#ifndef NE_BREED_BATCH
#define NE_BREED_BATCH 4
#endif
// quads per lane per batch (at most 8: 32 pending bits); measured on 4,096 children of 313,010 elements at
// mutation p = 0.01: 1 quad 2.12 ms, 4 quads 1.87 ms, 8 quads 1.95 ms, without mutation 1.79 ms
constexpr int kBreedBatch = NE_BREED_BATCH;
static_assert(kBreedBatch >= 1 && kBreedBatch <= 8, "breed batch");

// Crossover + mutation, element-parallel: each thread builds 4 consecutive values of one child
// (BreedAndMutateGenome, genome_ops.cuh:386-407, on counter-based draws).  Writes the full slot.
__global__ __launch_bounds__(256) void BreedKernel(BreedArgs p) {
    const uint32_t bc = blockIdx.y;  // entry of the batch
    const uint32_t brood = p.interleave_cells != 0 ? gridDim.y / p.interleave_cells : 1u;
    const uint32_t entry = p.interleave_cells != 0 ? (bc % brood) * p.interleave_cells + bc / brood : bc;
    const uint32_t child = p.child_list != nullptr ? p.child_list[bc] : p.child_offset + entry;
    __shared__ XoverRecord s_rec;
    __shared__ const float* s_p1;
    __shared__ const float* s_p2;
    __shared__ float* s_child;
    __shared__ int s_bad;
    // the child's donor-change positions (count, then ascending), from the crossover-decision kernel
    __shared__ uint32_t s_brk[kBreakStride];
    if (p.step != nullptr) {
        p.gen_index = p.step[0];
        p.seed = p.step[1];
    }
    if (threadIdx.x == 0) {
        s_rec = p.records[child];
        const uint32_t pa = p.pairs[2 * child], pb = p.pairs[2 * child + 1];
        const uint32_t sa = p.parent_slots[pa], sb = p.parent_slots[pb];
        const uint32_t fcb = p.free_count_dev != nullptr ? *p.free_count_dev : p.free_count_before;
        const bool room = p.given_slots != nullptr || fcb >= gridDim.y;
        const uint32_t cs = p.given_slots != nullptr ? p.given_slots[child]
                                                     : (room ? p.free_stack[fcb - 1 - entry] : kInvalidSlot);
        s_bad = (!room || sa >= p.slot_count || sb >= p.slot_count || cs >= p.slot_count) ? 1 : 0;
        if (!room)
            SetStatus(p.status, 14);  // kSlabFull
        else if (s_bad)
            SetStatus(p.status, 13);  // kInvalidParentIndex
        s_p1 = reinterpret_cast<const float*>(p.slab + uint64_t(sa) * p.parent_stride);
        s_p2 = reinterpret_cast<const float*>(p.slab + uint64_t(sb) * p.parent_stride);
        s_child = reinterpret_cast<float*>(p.slab + uint64_t(cs) * p.child_stride);
        if (blockIdx.x == 0 && !s_bad && p.given_slots == nullptr) {
            p.child_slots[child] = cs;
            p.slot_state[cs] = 1u;
        }
    }
    const uint32_t n_parent = uint32_t(kPolicyParams) + p.parent_action_count * kTail;
    if (threadIdx.x < kBreakStride) s_brk[threadIdx.x] = p.breaks[uint64_t(child) * kBreakStride + threadIdx.x];
    __syncthreads();
    if (s_bad) return;
    const XoverRecord r = s_rec;
    const float* p1 = s_p1;
    const float* p2 = s_p2;
    float* out = s_child;
    const uint32_t nbreaks = s_brk[0];
    const uint32_t* s_breaks = s_brk + 1;
    const bool mutate = p.mutation_p > 0.0f && p.mutation_sigma > 0.0f;
    const uint32_t quads = p.elem_end / 4;  // the loop covers quads [elem_begin / 4, elem_end / 4)
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t kb = 0;  // first break above the previous window's start
    // crossover: the donor values of quad q (called in increasing q, as the cursor requires)
    auto load_quad = [&](uint32_t q) __device__ -> v4f {
        const uint32_t e0 = 4 * q;
        v4f v = {0.0f, 0.0f, 0.0f, 0.0f};
        if (e0 >= n_parent) return v;
        // the wave's window [w0, w0 + 256): uniform donor unless a break lies strictly inside, i.e. unless
        // the first break above w0 is below w0 + 256
        const uint32_t w0 = __builtin_amdgcn_readfirstlane(e0 - 4 * lane);
        while (kb < nbreaks && s_breaks[kb] <= w0) ++kb;
        const bool uniform = w0 + 256 <= n_parent && (kb == nbreaks || s_breaks[kb] >= w0 + 256);
        if (uniform) return Load4((ElementDonorFirst(r, w0) ? p1 : p2) + e0);
        const bool d0 = ElementDonorFirst(r, e0), d1 = ElementDonorFirst(r, e0 + 1);
        const bool d2 = ElementDonorFirst(r, e0 + 2), d3 = ElementDonorFirst(r, e0 + 3);
        if (d0 == d1 && d1 == d2 && d2 == d3 && e0 + 4 <= n_parent) return Load4((d0 ? p1 : p2) + e0);
        const bool dd[4] = {d0, d1, d2, d3};
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (e0 + i < n_parent) v[i] = (dd[i] ? p1 : p2)[e0 + i];
        return v;
    };
    // tail-row scaling of quad q (element draws per row), then the store
    auto store_quad = [&](uint32_t q, v4f v) __device__ {
        const uint32_t e0 = 4 * q;
        if (e0 < n_parent && p.row_scale_p > 0.0f) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t e = e0 + i;
                if (e < uint32_t(kPolicyParams) || e >= n_parent) continue;
                const uint32_t row = (e - uint32_t(kPolicyParams)) / kTail;
                const U4 w = Philox4x32(row, child, p.gen_index, 0, p.seed, kTagRscl);
                if (BernoulliDraw(p.row_scale_p, w.x)) v[i] = v[i] * (BernoulliDraw(0.5f, w.y) ? 1.02f : 0.98f);
            }
        }
        *reinterpret_cast<v4f*>(out + e0) = v;
    };
    // Batches of kBreedBatch quads per lane: the batch's loads are in flight while its Bernoulli draws are
    // computed, and the elements drawn for mutation are then mutated one per pass across the whole batch,
    // so the wave runs the normal draw as often as its busiest lane needs over kBreedBatch quads (about
    // once per 4 quads at p = 0.01) instead of once for every element index some lane drew.
    constexpr int kB = kBreedBatch;
    const v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t base = p.elem_begin / 4 + blockIdx.x * blockDim.x + threadIdx.x; base < quads;
         base += uint32_t(kB) * stride) {
        v4f v[kB];
#pragma unroll
        for (int k = 0; k < kB; ++k) {
            const uint32_t qk = base + uint32_t(k) * stride;
            v[k] = qk < quads ? load_quad(qk) : zero;
        }
        if (mutate) {
            uint32_t pending = 0;  // bit 4k+i: element i of quad k drawn for mutation
#pragma unroll
            for (int k = 0; k < kB; ++k) {
                const uint32_t qk = base + uint32_t(k) * stride;
                if (qk >= quads || 4 * qk >= n_parent) continue;
                const U4 w = Philox4x32(qk, child, p.gen_index, 0, p.seed, kTagMuta);
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (4 * qk + i < n_parent && BernoulliDraw(p.mutation_p, Lane(w, i))) pending |= 1u << (4 * k + i);
            }
            while (pending != 0) {
                const uint32_t b = uint32_t(__builtin_ctz(pending));
                pending &= pending - 1u;
                float x = 0.0f;
#pragma unroll
                for (int k = 0; k < kB; ++k)
#pragma unroll
                    for (int i = 0; i < 4; ++i) x = b == uint32_t(4 * k + i) ? v[k][i] : x;
                const uint32_t e = 4 * (base + (b >> 2) * stride) + (b & 3u);
                const float m = MutatedValue(x, e, child, p.gen_index, p.seed, p.mutation_sigma);
#pragma unroll
                for (int k = 0; k < kB; ++k)
#pragma unroll
                    for (int i = 0; i < 4; ++i) v[k][i] = b == uint32_t(4 * k + i) ? m : v[k][i];
            }
        }
#pragma unroll
        for (int k = 0; k < kB; ++k) {
            const uint32_t qk = base + uint32_t(k) * stride;
            if (qk < quads) store_quad(qk, v[k]);
        }
    }
}

// ================================================================================================
// Generation 0 (InitializeRandomGenome, genome_ops.cuh:58-143; BootstrapRandomGenerationKernel,
// device_runtime.cu:258-314): weights std_l * z, biases 0, tails tail_std * z.
__global__ __launch_bounds__(256) void BootstrapKernel(BootArgs p) {
    const uint32_t o = blockIdx.y;
    __shared__ uint32_t s_slot;
    if (threadIdx.x == 0) {
        s_slot = p.free_stack[p.free_count_before - 1 - o];
        if (blockIdx.x == 0) {
            p.slots[o] = s_slot;
            p.slot_state[s_slot] = 1u;
        }
    }
    __syncthreads();
    float* g = reinterpret_cast<float*>(p.slab + uint64_t(s_slot) * p.stride);
    const uint32_t n_values = uint32_t(kPolicyParams) + p.action_count * kTail;
    const uint32_t n_slot = uint32_t(p.stride / 4);
    for (uint32_t e2 = (blockIdx.x * blockDim.x + threadIdx.x) * 2; e2 < n_slot; e2 += gridDim.x * blockDim.x * 2) {
        const U4 w = Philox4x32(e2 >> 1, o, p.gen_index, 0, p.seed, kTagBoot);
        float v[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const uint32_t e = e2 + i;
            if (e >= n_values) {
                v[i] = 0.0f;
                continue;
            }
            float sd = p.tail_std;
            bool bias = false;
            if (e < uint32_t(kPolicyParams)) {
#pragma unroll
                for (int l = 0; l < kDenseLayers; ++l) {
                    const DenseLayerDesc d = DenseLayerAt(l);
                    if (e >= d.w && e < d.w + d.in * d.out) sd = p.layer_std[l];
                    if (e >= d.b && e < d.b + d.out) bias = true;
                }
            }
            const float z = i == 0 ? NormalFromWords(w.x, w.y) : NormalFromWords(w.z, w.w);
            v[i] = bias ? 0.0f : sd * z;
        }
        *reinterpret_cast<float2*>(g + e2) = make_float2(v[0], v[1]);
    }
}

// ================================================================================================
// Slab bookkeeping (reference_counter.hpp:135-220, device_runtime.cu:372-446,619-663).  Slot
// pushes and pops happen in index order on one thread, so slot assignment is deterministic.
__global__ void RefcountKernel(const uint32_t* pairs, uint32_t child_count, uint32_t cur_size, uint32_t* refcounts,
                               int* status) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= child_count) return;
    const uint32_t a = pairs[2 * c], b = pairs[2 * c + 1];
    if (a >= cur_size || b >= cur_size) {
        SetStatus(status, 13);  // kInvalidParentIndex
        return;
    }
    atomicAdd(refcounts + a, 1u);
    atomicAdd(refcounts + b, 1u);
}

// Frees every parent whose reference count is zero and whose slot is still held, pushing the slots in
// generation-index order (TryCollectZeroReferenceParents, reference_counter.hpp:135-220).  One
// workgroup of 1024 threads: a ballot prefix sum gives each released parent its stack position; the
// slot-state exchange catches a slot released twice.
__global__ __launch_bounds__(1024) void CollectReleasedParentsKernel(uint32_t* cur_slots, uint32_t cur_size,
                                                                     const uint32_t* refcounts, uint32_t* slot_state,
                                                                     uint32_t* free_stack, uint32_t* free_count,
                                                                     uint32_t slot_count, int* status) {
    __shared__ uint32_t s_wave_total[16];
    __shared__ int s_bad;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid == 0) s_bad = 0;
    uint32_t fc = *free_count;
    __syncthreads();
    for (uint32_t base = 0; base < cur_size; base += blockDim.x) {
        const uint32_t i = base + tid;
        uint32_t s = kInvalidSlot;
        bool release = false;
        if (i < cur_size && refcounts[i] == 0) {
            s = cur_slots[i];
            release = s != kInvalidSlot;
        }
        const uint64_t mask = __ballot(release ? 1 : 0);
        const uint32_t rank_in_wave = __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
        if (lane == 0) s_wave_total[wave] = uint32_t(__popcll(mask));
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) {
            before += w < wave ? s_wave_total[w] : 0u;
            total += s_wave_total[w];
        }
        if (release) {
            const uint32_t pos = fc + before + rank_in_wave;
            if (s >= slot_count || pos >= slot_count || atomicExch(slot_state + s, 0u) == 0u) {
                atomicExch(&s_bad, 1);
            } else {
                free_stack[pos] = s;
                cur_slots[i] = kInvalidSlot;
            }
        }
        fc += total;
        __syncthreads();  // s_wave_total is rewritten next round
    }
    if (tid == 0) {
        if (s_bad)
            SetStatus(status, 3);  // kInvalidSlab
        else
            *free_count = fc;
    }
}

__global__ void ReleaseBatchKernel(const uint32_t* pairs, uint32_t child_offset, uint32_t batch, uint32_t* refcounts,
                                   int* status) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= batch) return;
    const uint32_t i = child_offset + c;
    const uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
    if (atomicSub(refcounts + a, 1u) == 0u || atomicSub(refcounts + b, 1u) == 0u) SetStatus(status, 3);
}

// Host spill of an assembled prefix (ReleaseNextGenerationPrefixKernel / AllocateRestoredChildPrefixKernel,
// slab_runtime.cu:733-807), in a fixed order: the slots of children 0..k-1 are pushed onto the free stack
// at positions fc..fc+k-1; on the way back child i takes free_stack[fc-1-i], as the breeding batches do.
__global__ void SpillReleaseKernel(uint32_t* child_slots, uint32_t k, uint32_t fc, uint32_t* slot_state,
                                   uint32_t* free_stack, uint32_t* free_count, uint32_t slot_count, int* status) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *free_count = fc + k;
    if (i >= k) return;
    const uint32_t s = child_slots[i];
    if (s >= slot_count || slot_state[s] == 0 || fc + i >= slot_count) {
        SetStatus(status, 3);  // kInvalidSlab
        return;
    }
    slot_state[s] = 0;
    free_stack[fc + i] = s;
    child_slots[i] = kInvalidSlot;
}

__global__ void SpillRestoreKernel(uint32_t* child_slots, uint32_t k, uint32_t fc, uint32_t* slot_state,
                                   const uint32_t* free_stack, uint32_t* free_count, int* status) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *free_count = fc - k;
    if (i >= k) return;
    const uint32_t s = free_stack[fc - 1 - i];
    if (child_slots[i] != kInvalidSlot || slot_state[s] != 0) {
        SetStatus(status, 3);
        return;
    }
    slot_state[s] = 1;
    child_slots[i] = s;
}

__global__ void PopSlotsKernel(uint32_t* free_count, uint32_t count) {
    // a batch larger than the free count was refused by the assembly kernel (slab full): nothing popped
    if (threadIdx.x == 0 && blockIdx.x == 0 && *free_count >= count) *free_count -= count;
}

__global__ void InitSlabMetadataKernel(uint32_t* slot_state, uint32_t* free_stack, uint32_t* free_count,
                                       uint32_t slot_count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < slot_count) {
        slot_state[i] = 0;
        free_stack[i] = slot_count - 1 - i;  // device_runtime.cu:230-256: stack [S-1 ... 0]
    }
    if (i == 0) *free_count = slot_count;
}

// ================================================================================================
// Widening repack (repacking.hpp:104-287; PrepareSlabForExpandedActionCountKernel,
// device_runtime.cu:978-1033).  The plan is computed on one thread; the copies run across the chip
// through a staging buffer so sources and destinations never overlap.
__global__ void RepackPlanKernel(RepackArgs p) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t survivors = 0;
    for (uint32_t i = 0; i < p.cur_size; ++i) {
        if (p.refcounts[i] == 0) continue;
        const uint32_t s = p.cur_slots[i];
        if (s == kInvalidSlot || s >= p.old_slot_count || p.slot_state[s] == 0 || p.owner[s] != kInvalidSlot) {
            SetStatus(p.status, 16);
            return;
        }
        p.owner[s] = i;
        ++survivors;
    }
    if (survivors == 0 || survivors > p.new_slot_count) {
        SetStatus(p.status, 16);
        return;
    }
    const uint32_t base = p.new_slot_count - survivors;
    uint32_t rank = 0;
    for (uint32_t s = 0; s < p.old_slot_count; ++s) {
        const uint32_t i = p.owner[s];
        if (i == kInvalidSlot) continue;
        p.move_src[rank] = s;
        p.cur_slots[i] = base + rank;
        ++rank;
    }
    for (uint32_t i = 0; i < p.cur_size; ++i)
        if (p.refcounts[i] == 0) p.cur_slots[i] = kInvalidSlot;
    for (uint32_t s = 0; s < p.new_slot_count; ++s) p.slot_state[s] = s >= base ? 1u : 0u;
    for (uint32_t f = 0; f < base; ++f) p.free_stack[f] = base - 1 - f;
    *p.free_count = base;
    *p.survivors = survivors;
}

__global__ void RepackGatherKernel(const uint8_t* slab, uint64_t old_stride, const uint32_t* move_src,
                                   uint32_t survivors, uint8_t* staging) {
    const uint32_t r = blockIdx.y;
    if (r >= survivors) return;
    const v4f* src = reinterpret_cast<const v4f*>(slab + uint64_t(move_src[r]) * old_stride);
    v4f* dst = reinterpret_cast<v4f*>(staging + uint64_t(r) * old_stride);
    const uint64_t n = old_stride / 16;
    for (uint64_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

__global__ void RepackScatterKernel(uint8_t* slab, uint64_t new_stride, uint64_t old_stride, uint32_t base,
                                    uint32_t survivors, const uint8_t* staging) {
    const uint32_t r = blockIdx.y;
    if (r >= survivors) return;
    const v4f* src = reinterpret_cast<const v4f*>(staging + uint64_t(r) * old_stride);
    v4f* dst = reinterpret_cast<v4f*>(slab + uint64_t(base + r) * new_stride);
    const uint64_t n_old = old_stride / 16, n_new = new_stride / 16;
    const v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint64_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_new; i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = i < n_old ? src[i] : zero;
}

__global__ void FillU32Kernel(uint32_t* dst, uint32_t value, uint64_t count) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = value;
}

__global__ void FillU8Kernel(uint8_t* dst, uint8_t value, uint64_t count) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = value;
}

__global__ void StepReadbackKernel(const int* status, const uint32_t* free_count, const SummaryRecord* summary,
                                   const unsigned long long* counters, StepReadback* out) {
    if (threadIdx.x != 0) return;
    out->status = *status;
    out->free_count = *free_count;
    out->summary = *summary;
    for (int i = 0; i < 4; ++i) out->counters[i] = counters[i];
}

// ================================================================================================
// Output-embedding injection on growth generations (TryInjectExpandedOutputEmbeddingTailsConcurrently,
// injection_ops.cuh:21-64; output_embedding_injection.hpp:25-205).

// TrainableTailNorm: sequential sum of squares over the 38 values, then sqrt.
__device__ inline float TailNorm(const float* row) {
    float s = 0.0f;
    for (int f = 0; f < kTail; ++f) s += row[f] * row[f];
    return sqrtf(s);
}

// Median of a child's existing tail-row norms (TryComputeMedianTailNorm): bitonic sort in LDS.
__global__ __launch_bounds__(256) void MedianNormKernel(InjectArgs p) {
    __shared__ float s[kMaxSortedNorms];
    const uint32_t c = blockIdx.x;
    const uint32_t slot = p.child_slots[c];
    const float* tails = reinterpret_cast<const float*>(p.slab + uint64_t(slot) * p.stride) + kPolicyParams;
    const uint32_t A = p.parent_action_count;
    uint32_t n2 = 1;
    while (n2 < A) n2 <<= 1;
    for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) s[i] = i < A ? TailNorm(tails + uint64_t(i) * kTail) : __builtin_inff();
    __syncthreads();
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
            for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
                const uint32_t ixj = i ^ jj;
                if (ixj > i) {
                    const bool ascending = (i & k) == 0;
                    const float a = s[i], b = s[ixj];
                    if ((a > b) == ascending) {
                        s[i] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        const uint32_t up = A / 2;
        p.medians[c] = (A % 2) ? s[up] : 0.5f * (s[up - 1] + s[up]);
    }
}

// Seeds the new tail rows of one child from its policy on the 3 hint grids of each new word, then
// rescales each row to the child's median tail norm.  One workgroup per child; each wave runs tiles
// of 5 words x 3 grids through the trunk (lane column j = 3*word + grid, column 15 idle).
__global__ __launch_bounds__(512) void InjectKernel(InjectArgs p) {
    __shared__ float s_enc[kTurnFeatures * kEncOut];
    __shared__ float s_encb[kEncOut];
    const uint32_t c = blockIdx.x;
    const uint32_t slot = p.child_slots[c];
    float* genome = reinterpret_cast<float*>(p.slab + uint64_t(slot) * p.stride);
    StageEncoder(genome, s_enc, s_encb);
    const uint32_t lane = threadIdx.x & 63u, j = lane & 15u, gq = lane >> 4;
    const uint32_t waves = blockDim.x >> 6;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const TrunkLane tl = MakeTrunkLane(genome, j, gq);
    const uint32_t tiles = (p.count + 4) / 5;
    for (uint32_t tile = wave; tile < tiles; tile += waves) {
        const uint32_t word = tile * 5 + j / 3, grid = j % 3;
        const bool valid = j < 15 && word < p.count;
        HintState st = {0, {0, 0, 0, 0}, {0, 0, 0, 0}};
        if (valid) st = p.states[word * 3 + grid];
        const int tc = valid ? int(st.tc) : 0;
        int max_tc = tc;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) max_tc = max(max_tc, __shfl_xor(max_tc, off));
        max_tc = __builtin_amdgcn_readfirstlane(max_tc);
        v4f acc[16];
        InitLayer0(acc, tl);
        AddVirgin(acc, tl, valid && tc == 0);
        for (int ts = 0; ts < max_tc; ++ts) {
            const uint32_t code = ts == 0 ? st.code[0] : (ts == 1 ? st.code[1] : (ts == 2 ? st.code[2] : st.code[3]));
            const uint32_t fb = ts == 0 ? st.fb[0] : (ts == 1 ? st.fb[1] : (ts == 2 ? st.fb[2] : st.fb[3]));
            float e[16];
            EncodeTurnLds(s_enc, s_encb, code, fb, gq, e);
            if (ts >= tc) {
#pragma unroll
                for (int q = 0; q < 16; ++q) e[q] = 0.0f;
            }
            FoldSlot(acc, tl, ts, e);
        }
        float pol[kScoreSteps];
        TrunkToPolicy(acc, tl, pol);
        // mean over the word's 3 grids in grid order: ((0 + p0) + p1) + p2, times 1/3
        float* row = genome + kPolicyParams + uint64_t(p.parent_action_count + word) * kTail;
#pragma unroll
        for (int q = 6; q < kScoreSteps; ++q) {  // steps 0..5 hold fixed dims in every lane group
            const float p1 = __shfl_down(pol[q], 1);
            const float p2 = __shfl_down(pol[q], 2);
            const float mean = (((0.0f + pol[q]) + p1) + p2) * (1.0f / 3.0f);
            const int dim = ScoreDim(q, int(gq));
            if (valid && grid == 0 && dim >= kFixed) row[dim - kFixed] = mean;
        }
    }
    __syncthreads();
    const float target = p.medians[c];
    for (uint32_t k = threadIdx.x; k < p.count; k += blockDim.x) {
        float* row = genome + kPolicyParams + uint64_t(p.parent_action_count + k) * kTail;
        float scale = 0.0f;
        if (target > 1.0e-6f) {
            const float cur = TailNorm(row);
            if (!(cur > 1.0e-6f)) {
                SetStatus(p.status, 15);  // kOutputEmbeddingInjectionFailed
                continue;
            }
            scale = target / cur;
        }
        for (int f = 0; f < kTail; ++f) row[f] = row[f] * scale;
    }
}

// ================================================================================================
// Launch wrappers (host side).
void LaunchPackBf16(const uint8_t* slab, uint64_t stride, const uint32_t* slots, uint32_t n, uint32_t slot_count,
                    const uint32_t* img, uint32_t A, uint8_t* out, uint64_t out_stride, uint32_t* tail_absmax,
                    hipStream_t s) {
    const uint32_t a_tiles = (A + 15) / 16;
    const uint32_t groups = (a_tiles + kPackTiles - 1) / kPackTiles;
    const uint32_t blocks = groups < NE_PACK_BLOCKS ? groups : NE_PACK_BLOCKS;
    hipLaunchKernelGGL(PackBf16Kernel, dim3(blocks, n), dim3(320), 0, s, slab, stride, slots, slot_count, img, A, a_tiles,
                       out, out_stride, tail_absmax);
}

void LaunchFingerprint(const FingerprintArgs& a, uint32_t n, hipStream_t s) {
    if (a.trunk_bf16)
        hipLaunchKernelGGL(FingerprintKernel<true>, dim3(n), dim3(64 * kProbeDepths), 0, s, a);
    else
        hipLaunchKernelGGL(FingerprintKernel<false>, dim3(n), dim3(64 * kProbeDepths), 0, s, a);
}

void LaunchTransposeEncoder(const uint8_t* slab, uint64_t stride, const uint32_t* slots, uint32_t n,
                            uint32_t slot_count, float* enc_t, hipStream_t s) {
    hipLaunchKernelGGL(TransposeEncoderKernel, dim3(n), dim3(256), 0, s, slab, stride, slots, slot_count, enc_t);
}

void LaunchEvaluate(const EvalArgs& a, uint32_t blocks, uint32_t waves, uint32_t kernel, bool trace, bool bf16,
                    hipStream_t s) {
    const dim3 grid(blocks);
    const dim3 block(64 * waves);
    // the BF16 form's streamed kernel is the ring, which counts exactly 3 transfers per wave and chunk
    kernel = EvaluateKernelFor(kernel, waves, trace, bf16);
    const bool t16 = bf16 && a.trunk_bf16 != 0;  // the BF16-trunk form: the BF16 kernels with kT16
    if (!trace && kernel == kEvalKernelPaired) {
        if (t16)
            hipLaunchKernelGGL(EvaluatePairBf16Kernel<true>, grid, block, 0, s, a);
        else if (bf16)
            hipLaunchKernelGGL(EvaluatePairBf16Kernel<false>, grid, block, 0, s, a);
        else
            hipLaunchKernelGGL(EvaluatePairKernel, grid, block, 0, s, a);
        return;
    }
    if (!trace && kernel == kEvalKernelStreamed) {
        if (t16)
            hipLaunchKernelGGL(EvaluateStreamedBf16RingKernel<true>, grid, block, 0, s, a);
        else if (bf16)
            hipLaunchKernelGGL(EvaluateStreamedBf16RingKernel<false>, grid, block, 0, s, a);
        else
            hipLaunchKernelGGL(EvaluateStreamedKernel, grid, block, 0, s, a);
        return;
    }
    if (t16) {
        if (trace)
            hipLaunchKernelGGL((EvaluateKernel<true, true, true>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((EvaluateKernel<false, true, true>), grid, block, 0, s, a);
    } else if (bf16) {
        if (trace)
            hipLaunchKernelGGL((EvaluateKernel<true, true>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((EvaluateKernel<false, true>), grid, block, 0, s, a);
    } else {
        if (trace)
            hipLaunchKernelGGL((EvaluateKernel<true, false>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((EvaluateKernel<false, false>), grid, block, 0, s, a);
    }
}

uint32_t EvaluateMaxWaves() { return kEvalMaxWaves; }

uint32_t EvaluateMaxWavesFor(uint32_t kernel, bool trace, bool bf16) {
    if (!trace && bf16 && kernel == kEvalKernelPaired) return uint32_t(NE_PAIR_BF16_WAVES);
    if (!trace && bf16 && kernel == kEvalKernelStreamed) return 3u;  // the ring kernel: exactly 3
    return kEvalMaxWaves;
}

uint32_t EvaluateKernelFor(uint32_t kernel, uint32_t waves, bool trace, bool bf16) {
    if (trace) return kEvalKernelRegisters;
    if (bf16 && kernel == kEvalKernelStreamed && waves != 3) return kEvalKernelRegisters;
    return kernel;
}

uint64_t EvaluateAccScratchBytes() { return uint64_t(kEvalMaxWaves) * 2 * kAccTileFloats * sizeof(float); }

uint32_t EvaluateResidentPerCu(uint32_t waves, uint32_t kernel, bool trace, bool bf16, bool t16) {
    // what the registers and LDS of the kernel built here allow, as the runtime reports it
    if (waves > EvaluateMaxWavesFor(kernel, trace, bf16)) return 0;
    const bool plain = !trace && !bf16;
    t16 = t16 && bf16;
    const void* k = (plain && kernel == kEvalKernelPaired) ? reinterpret_cast<const void*>(&EvaluatePairKernel)
                  : (plain && kernel == kEvalKernelStreamed) ? reinterpret_cast<const void*>(&EvaluateStreamedKernel)
                  : (!trace && kernel == kEvalKernelPaired)
                      ? (t16 ? reinterpret_cast<const void*>(&EvaluatePairBf16Kernel<true>)
                             : reinterpret_cast<const void*>(&EvaluatePairBf16Kernel<false>))
                  : (!trace && kernel == kEvalKernelStreamed)
                      ? (t16 ? reinterpret_cast<const void*>(&EvaluateStreamedBf16RingKernel<true>)
                             : reinterpret_cast<const void*>(&EvaluateStreamedBf16RingKernel<false>))
                  : t16 ? (trace ? reinterpret_cast<const void*>(&EvaluateKernel<true, true, true>)
                                 : reinterpret_cast<const void*>(&EvaluateKernel<false, true, true>))
                  : bf16 ? (trace ? reinterpret_cast<const void*>(&EvaluateKernel<true, true>)
                                  : reinterpret_cast<const void*>(&EvaluateKernel<false, true>))
                         : (trace ? reinterpret_cast<const void*>(&EvaluateKernel<true, false>)
                                  : reinterpret_cast<const void*>(&EvaluateKernel<false, false>));
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, int(64 * waves), 0) != hipSuccess || blocks <= 0)
        return 0;
    return uint32_t(blocks);
}

void LaunchFinalize(const float* raw, const uint32_t* local_w, uint32_t n, float* fitness, uint32_t* evals,
                    uint8_t* has, int* status, hipStream_t s) {
    hipLaunchKernelGGL(FinalizeKernel, dim3((n + 255) / 256), dim3(256), 0, s, raw, local_w, n, fitness, evals, has,
                       status);
}

void LaunchSummary(const float* fitness, const uint8_t* has, const uint32_t* slots, uint32_t n, SummaryRecord* out,
                   int* status, hipStream_t s) {
    hipLaunchKernelGGL(SummaryKernel, dim3(1), dim3(1024), 0, s, fitness, has, slots, n, out, status);
}

void LaunchPlan(const PlanArgs& a, hipStream_t s) {
    if (a.radius <= 2)
        hipLaunchKernelGGL(PlanListedKernel, dim3((a.child_count + kPlanWaves - 1) / kPlanWaves), dim3(64 * kPlanWaves),
                           0, s, a);
    else
        hipLaunchKernelGGL(PlanKernel, dim3((a.child_count + kPlanBlock - 1) / kPlanBlock), dim3(kPlanBlock), 0, s, a);
}

void LaunchXover(const XoverArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(XoverDecisionKernel, dim3((a.child_count + 127) / 128), dim3(128), 0, s, a);
}

void LaunchBreed(const BreedArgs& a, uint32_t batch, hipStream_t s) {
    const uint32_t quads = (a.elem_end - a.elem_begin) / 4;
    // about 16k workgroups in all (measured at 4,096 children: 4 per child 2.66 ms, 16 2.78 ms, 64 3.14 ms
    // -- each workgroup's long forward walk through its share of the child beats many short ones), at
    // least 4 and at most 64 per child, so that small generations still fill the GPU
    uint32_t blocks_x = (16384 + batch - 1) / batch;
    blocks_x = blocks_x < 4 ? 4 : (blocks_x > 64 ? 64 : blocks_x);
    if (blocks_x > (quads + 255) / 256) blocks_x = (quads + 255) / 256;
    hipLaunchKernelGGL(BreedKernel, dim3(blocks_x, batch), dim3(256), 0, s, a);
}

void LaunchBootstrap(const BootArgs& a, uint32_t organisms, hipStream_t s) {
    const uint32_t pairs = uint32_t(a.stride / 8);
    uint32_t blocks_x = (pairs + 255) / 256;
    if (blocks_x > 64) blocks_x = 64;
    hipLaunchKernelGGL(BootstrapKernel, dim3(blocks_x, organisms), dim3(256), 0, s, a);
}

void LaunchRefcount(const uint32_t* pairs, uint32_t child_count, uint32_t cur_size, uint32_t* refcounts, int* status,
                    hipStream_t s) {
    hipLaunchKernelGGL(RefcountKernel, dim3((child_count + 255) / 256), dim3(256), 0, s, pairs, child_count, cur_size,
                       refcounts, status);
}

void LaunchCollectReleased(uint32_t* cur_slots, uint32_t cur_size, const uint32_t* refcounts, uint32_t* slot_state,
                           uint32_t* free_stack, uint32_t* free_count, uint32_t slot_count, int* status,
                           hipStream_t s) {
    hipLaunchKernelGGL(CollectReleasedParentsKernel, dim3(1), dim3(1024), 0, s, cur_slots, cur_size, refcounts,
                       slot_state, free_stack, free_count, slot_count, status);
}

void LaunchReleaseBatch(const uint32_t* pairs, uint32_t child_offset, uint32_t batch, uint32_t* refcounts,
                        int* status, hipStream_t s) {
    hipLaunchKernelGGL(ReleaseBatchKernel, dim3((batch + 255) / 256), dim3(256), 0, s, pairs, child_offset, batch,
                       refcounts, status);
}

void LaunchSpillRelease(uint32_t* child_slots, uint32_t k, uint32_t fc, uint32_t* slot_state, uint32_t* free_stack,
                        uint32_t* free_count, uint32_t slot_count, int* status, hipStream_t s) {
    hipLaunchKernelGGL(SpillReleaseKernel, dim3((k + 255) / 256), dim3(256), 0, s, child_slots, k, fc, slot_state,
                       free_stack, free_count, slot_count, status);
}

void LaunchSpillRestore(uint32_t* child_slots, uint32_t k, uint32_t fc, uint32_t* slot_state,
                        const uint32_t* free_stack, uint32_t* free_count, int* status, hipStream_t s) {
    hipLaunchKernelGGL(SpillRestoreKernel, dim3((k + 255) / 256), dim3(256), 0, s, child_slots, k, fc, slot_state,
                       free_stack, free_count, status);
}

void LaunchPopSlots(uint32_t* free_count, uint32_t count, hipStream_t s) {
    hipLaunchKernelGGL(PopSlotsKernel, dim3(1), dim3(1), 0, s, free_count, count);
}

void LaunchInitSlabMetadata(uint32_t* slot_state, uint32_t* free_stack, uint32_t* free_count, uint32_t slot_count,
                            hipStream_t s) {
    hipLaunchKernelGGL(InitSlabMetadataKernel, dim3((slot_count + 255) / 256), dim3(256), 0, s, slot_state, free_stack,
                       free_count, slot_count);
}

void LaunchRepackPlan(const RepackArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(RepackPlanKernel, dim3(1), dim3(1), 0, s, a);
}

void LaunchRepackCopies(uint8_t* slab, uint64_t old_stride, uint64_t new_stride, const uint32_t* move_src,
                        uint32_t survivors, uint32_t base, uint8_t* staging, hipStream_t s) {
    const uint32_t bx = 16;
    hipLaunchKernelGGL(RepackGatherKernel, dim3(bx, survivors), dim3(256), 0, s, slab, old_stride, move_src, survivors,
                       staging);
    hipLaunchKernelGGL(RepackScatterKernel, dim3(bx, survivors), dim3(256), 0, s, slab, new_stride, old_stride, base,
                       survivors, staging);
}

void LaunchFillU32(uint32_t* dst, uint32_t value, uint64_t count, hipStream_t s) {
    if (count == 0) return;
    uint64_t blocks = (count + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(FillU32Kernel, dim3(uint32_t(blocks)), dim3(256), 0, s, dst, value, count);
}

void LaunchFillU8(uint8_t* dst, uint8_t value, uint64_t count, hipStream_t s) {
    if (count == 0) return;
    uint64_t blocks = (count + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(FillU8Kernel, dim3(uint32_t(blocks)), dim3(256), 0, s, dst, value, count);
}

void LaunchStepReadback(const int* status, const uint32_t* free_count, const SummaryRecord* summary,
                        const unsigned long long* counters, StepReadback* out, hipStream_t s) {
    hipLaunchKernelGGL(StepReadbackKernel, dim3(1), dim3(64), 0, s, status, free_count, summary, counters, out);
}

}  // namespace ne

namespace ne {
void LaunchInject(const InjectArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(MedianNormKernel, dim3(a.children), dim3(256), 0, s, a);
    const uint32_t tiles = (a.count + 4) / 5;
    const uint32_t waves = tiles < 8 ? (tiles == 0 ? 1 : tiles) : 8;
    hipLaunchKernelGGL(InjectKernel, dim3(a.children), dim3(64 * waves), 0, s, a);
}
}  // namespace ne
