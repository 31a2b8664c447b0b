// Path-guiding MLP on the matrix cores of gfx950 (BF16 inputs, FP32 accumulation).
//
// Kernels:
//   guideInferKernel     encoding + 5-layer forward, register resident: every layer is
//                        computed transposed (H^T = W X^T, 16x16x32 MFMA with the weights as
//                        the A operand), so a layer's accumulators are already the next
//                        layer's B operand; weights are read from LDS in fragment order (one
//                        16-byte read per lane per fragment, shared by the wave's 2 tiles of 16
//                        queries); one 1,024-thread workgroup per CU shares one LDS copy of the
//                        weights, so 4 waves per SIMD fit (register file and LDS both); the
//                        last layer's rows are permuted so that each lane holds the outputs of
//                        the two lobes it decodes (8 vMF lobes per query, FP16 record)
//   guideFwdBwdKernel    training, 64 samples per wave: the same register-resident transposed
//                        forward, the loss -c*log p(dir) and its gradient w.r.t. the 40 outputs
//                        (through a small LDS stage), and the backward pass in the same
//                        transposed form with a second, transposed fragment copy of the weights
//                        (dA^T = W^T dZ^T; the ReLU masks are kept as bits).  Every hidden
//                        layer input A_l (l >= 1) and output gradient dZ_l is written
//                        feature-major to HBM; the register operands (lane = sample) are turned
//                        into rows (lane = feature) by a product with an identity fragment on
//                        the matrix cores, which is exact for BF16 values.  For A_0 only the 9
//                        input scalars are written (36 of its 448 bytes per sample)
//   guideWeightGradKernel  dW_l = dZ_l^T A_l and db_l = column sums of dZ_l: one 64 x 64 tile of
//                        one layer per wave over a chunk of 512 or 1024 samples (16x16x32 MFMA with
//                        the samples as the reduction index), into a per-chunk partial
//                        gradient; A_0 is recomputed from the scalars bit for bit
//   guideGradReduceKernel  sums the partial gradients in chunk order (deterministic)
//   guideAdamKernel      Adam on the FP32 master weights, refreshed BF16 copies
// All matrix-core work uses v_mfma_f32_16x16x32_bf16: lane l holds the A rows / B columns
// l % 16 and D[4*(l/16)+i][l%16], i = 0..3.  Only the M/N placement is fixed by the hardware;
// which K index a (lane group, slot) stands for is the caller's choice as long as A and B
// agree, and guideFragPos / guideTFragPos / guideEncodeBlock fix it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.h"
#include "guide.h"
#include "guide_dev.h"

namespace rbr {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kWaves = 16;                    // waves per inference workgroup (one weight copy in LDS)
constexpr int kTile = 16;                     // queries / samples per wave tile

// Orders a wave's LDS writes before its later LDS reads by other lanes (no block barrier: the
// waves of a workgroup run different numbers of tiles).
__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace

// ---------------------------------------------------------------- inference

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kIT = 2;                         // 16-query tiles per wave
// Floats per query row of a training wave's LDS output stage: 52, not kGOutPad = 48, so that the 16 rows
// of a tile start in 16 different banks (with 48 a column of the tile fell in only 4 of the
// 64 banks, and every lane-parallel read of one output across the tile's queries conflicted).
constexpr int kStageRow = 52;
constexpr int kInferStage = kTile * kStageRow;  // floats of LDS output stage per wave

// Inference computes layer 4 with its rows permuted so that a lane's four outputs of each
// 16-row block (rows 16 mb + 4 g + i of query r in lane (r, g)) are the ten outputs of lobes
// 2g, 2g + 1 and the selection logit: row 16 mb + 4 g + i carries output outOfRow(...) of
// slot 4 mb + i.  The lobe decode then reads its inputs from registers, with no LDS stage.
__device__ __forceinline__ int outOfRow(int row) {
    const int g = (row & 15) >> 2, slot = 4 * (row >> 4) + (row & 3);
    return slot < 10 ? 10 * g + slot : slot == 10 ? kGSel : -1;
}
constexpr int kInferW4Frags = 6;   // layer-4 fragments: 3 row blocks x 2 k-blocks

__device__ __forceinline__ v4f mfma32(u32x4 a, u32x4 b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// ReLU + BF16 of output blocks 2kb, 2kb+1 -> the next layer's B operand for k-block kb.
__device__ __forceinline__ u32x4 reluPack(const v4f& lo, const v4f& hi) {
    u32x4 r;
    r[0] = reluBf2(packBf2(lo[0], lo[1]));
    r[1] = reluBf2(packBf2(lo[2], lo[3]));
    r[2] = reluBf2(packBf2(hi[0], hi[1]));
    r[3] = reluBf2(packBf2(hi[2], hi[3]));
    return r;
}

__device__ __forceinline__ v4f biasInit(const float* lb, int off, int mb, int g) {
    const float4 b = *reinterpret_cast<const float4*>(lb + off + mb * 16 + 4 * g);
    return v4f{b.x, b.y, b.z, b.w};
}

// One hidden layer (64 -> 64) on all tiles: acc holds the previous layer's pre-activations.
template <int L>
__device__ __forceinline__ void hiddenLayer(v4f (&acc)[kIT][4], const u32x4* lwf, const float* lb, int lane, int g) {
    u32x4 h[kIT][2];
#pragma unroll
    for (int t = 0; t < kIT; ++t) {
        h[t][0] = reluPack(acc[t][0], acc[t][1]);
        h[t][1] = reluPack(acc[t][2], acc[t][3]);
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) acc[t][mb] = biasInit(lb, 64 * L, mb, g);
    }
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const u32x4 af = lwf[(kGFragOff[L] + mb * 2 + kb) * 64 + lane];
#pragma unroll
            for (int t = 0; t < kIT; ++t) acc[t][mb] = mfma32(af, h[t][kb], acc[t][mb]);
        }
}

}  // namespace

// Queries come either from an explicit array (9 floats: position, normal, wo) or from the
// render's extension queue (hit point, outward normal, -ray direction); each wave takes
// kIT * 16 queries at a time.  Queries that are not guided are skipped (no output).
__global__ __launch_bounds__(64 * kWaves, 1) void guideInferKernel(GuideInferArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    u32x4* lwf = reinterpret_cast<u32x4*>(smem);
    float* lb = reinterpret_cast<float*>(smem + kGWeights * 2);
    u32x4* lw4 = reinterpret_cast<u32x4*>(smem + kGWeights * 2 + kGBias * 4);                       // permuted layer 4
    float* lb4 = reinterpret_cast<float*>(smem + kGWeights * 2 + kGBias * 4 + kInferW4Frags * 64 * 16);   // its biases
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int i = threadIdx.x; i < kGWeights / 8; i += blockDim.x) lwf[i] = reinterpret_cast<const u32x4*>(a.g.wf)[i];
    for (int i = threadIdx.x; i < kGBias; i += blockDim.x) lb[i] = a.g.b[i];
    __syncthreads();
    // the permuted layer 4: fragment (mb, kb), lane (row, k-group) <- output outOfRow(16 mb + row)
    for (int i = threadIdx.x; i < kInferW4Frags * 64; i += blockDim.x) {
        const int f = i >> 6, l = i & 63, o = outOfRow(16 * (f >> 1) + (l & 15));
        lw4[i] = o < 0 ? u32x4{0u, 0u, 0u, 0u} : lwf[(kGFragOff[4] + (o >> 4) * 2 + (f & 1)) * 64 + (o & 15) + 16 * (l >> 4)];
    }
    for (int i = threadIdx.x; i < kGOutPad; i += blockDim.x) {
        const int o = outOfRow(i);
        lb4[i] = o < 0 ? 0.0f : lb[4 * kGHid + o];
    }
    __syncthreads();
    const uint32_t n = a.nDev ? *a.nDev : a.n;
    const uint32_t groups = (n + kTile * kIT - 1) / (kTile * kIT);
    for (uint32_t grp = blockIdx.x * kWaves + wave; grp < groups; grp += gridDim.x * kWaves) {
        float s9[kIT][9];
        bool valid[kIT];
        bool any = false;
#pragma unroll
        for (int t = 0; t < kIT; ++t) {
            const uint32_t qi = (grp * kIT + t) * kTile + r;
            valid[t] = qi < n && guideQuery(a, qi, s9[t]);
            if (!valid[t])
                for (int j = 0; j < 9; ++j) s9[t][j] = 0.0f;
            any |= valid[t];
        }
        if (__ballot(any ? 1 : 0) == 0ull) continue;
        // layer 0: features are produced one k-block at a time
        v4f acc[kIT][4];
        float2 st[kIT];
#pragma unroll
        for (int t = 0; t < kIT; ++t) {
            st[t] = make_float2(0.0f, 1.0f);
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) acc[t][mb] = biasInit(lb, 0, mb, g);
        }
#pragma unroll
        for (int kb = 0; kb < 7; ++kb) {
            u32x4 bx[kIT];
#pragma unroll
            for (int t = 0; t < kIT; ++t) bx[t] = guideEncodeBlock(s9[t], g, kb, st[t]);
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                const u32x4 af = lwf[(kGFragOff[0] + mb * 7 + kb) * 64 + lane];
#pragma unroll
                for (int t = 0; t < kIT; ++t) acc[t][mb] = mfma32(af, bx[t], acc[t][mb]);
            }
        }
        hiddenLayer<1>(acc, lwf, lb, lane, g);
        hiddenLayer<2>(acc, lwf, lb, lane, g);
        hiddenLayer<3>(acc, lwf, lb, lane, g);
        // layer 4: 64 -> 48
        v4f o[kIT][3];
        {
            u32x4 h[kIT][2];
#pragma unroll
            for (int t = 0; t < kIT; ++t) {
                h[t][0] = reluPack(acc[t][0], acc[t][1]);
                h[t][1] = reluPack(acc[t][2], acc[t][3]);
#pragma unroll
                for (int mb = 0; mb < 3; ++mb) o[t][mb] = biasInit(lb4, 0, mb, g);
            }
#pragma unroll
            for (int kb = 0; kb < 2; ++kb)
#pragma unroll
                for (int mb = 0; mb < 3; ++mb) {
                    const u32x4 af = lw4[(mb * 2 + kb) * 64 + lane];
#pragma unroll
                    for (int t = 0; t < kIT; ++t) o[t][mb] = mfma32(af, h[t][kb], o[t][mb]);
                }
        }
        // outputs -> lobes: lane (r, g) holds lobes 2g, 2g+1 of query r (ten outputs, slots 0-9 of
        // the permuted layer 4) and the selection logit (slot 10).
#pragma unroll
        for (int t = 0; t < kIT; ++t) {
            {
                // The softmax max and sum over the query's 8 weights are taken across its four
                // lanes (r, r + 16, r + 32, r + 48) with two exchanges each, so each weight is
                // exponentiated once.  Every lane takes part: the four lanes of a query agree on
                // valid[t].  Hardware exp / rsq / rcp: the lobes are stored in FP16.
                float v[10];
#pragma unroll
                for (int q = 0; q < 10; ++q) v[q] = o[t][q >> 2][q & 3];
                const float selLogit = o[t][2][2];
                float mx = fmaxf(v[4], v[9]);
                mx = fmaxf(mx, __shfl_xor(mx, 16));
                mx = fmaxf(mx, __shfl_xor(mx, 32));
                const float e[2] = {__expf(v[4] - mx), __expf(v[9] - mx)};
                float se = e[0] + e[1];
                se += __shfl_xor(se, 16);
                se += __shfl_xor(se, 32);
                // The stored weights carry the selection probability s (kGSel): they sum to s,
                // which the renderer reads back as the guide's share of the samples.
                const float rse = __builtin_amdgcn_rcpf(se) *
                                  (kGSelMin + (kGSelMax - kGSelMin) * __builtin_amdgcn_rcpf(1.0f + __expf(-selLogit)));
                float vals[2][5];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const float m0 = v[5 * j], m1 = v[5 * j + 1], m2 = v[5 * j + 2];
                    const float il = __builtin_amdgcn_rsqf(fmaxf(m0 * m0 + m1 * m1 + m2 * m2, 1e-12f));
                    vals[j][0] = m0 * il;
                    vals[j][1] = m1 * il;
                    vals[j][2] = m2 * il;
                    vals[j][3] = __expf(fminf(fmaxf(v[5 * j + 3], kGLogKappaMin), kGLogKappaMax));
                    vals[j][4] = e[j] * rse;
                }
                if (valid[t]) {
                    const uint32_t qi = (grp * kIT + t) * kTile + r;
                    uint32_t* dst = reinterpret_cast<uint32_t*>(a.out + size_t(qi) * 5);
#pragma unroll
                    for (int a5 = 0; a5 < 5; ++a5) dst[4 * a5 + g] = packH2(vals[0][a5], vals[1][a5]);
                }
            }
        }
    }
}

// ---------------------------------------------------------------- training

namespace {

constexpr int kTT = 2;                 // 16-sample tiles per training wave group: 32 samples
constexpr int kTrainWaves = 8;         // waves per training workgroup
// Samples per partial weight gradient: 1,024, or 512 for batches of at most 2^17 samples, whose
// 1,024-sample chunks leave too few waves to fill the chip (measured optimum of each size).
__host__ __device__ constexpr uint32_t gradChunk(size_t n) { return n <= (size_t(1) << 17) ? 512u : 1024u; }
constexpr int kGradL0Tiles = kGInPad / 32;  // weight-gradient tiles of layer 0: 64 x 32 column blocks
constexpr int kGradTiles = kGradL0Tiles + 4;  // and one 64 x 64 (48 x 64) tile for each of layers 1-4
                                       // (the last has 32 columns), layers 1-4 one each

// Identity B operand that turns a forward operand (lane = sample, slots = 32 features of a
// k-block) into rows: mfma(op, E, 0) = D[sample][feature], lane (n, g) holding feature
// 16 hi + n of the block for samples 4g..4g+3.  Lane (n, g) carries one BF16 1.0 in slot
// 4 hi + (n & 3) when g = n / 4 (slot i of lane group g stands for feature 4g + (i & 3) + 16 (i >> 2)).
__device__ __forceinline__ u32x4 identityOp(int lane, int hi) {
    const int n = lane & 15, g = lane >> 4;
    const int i = 4 * hi + (n & 3);
    const uint32_t one = g == (n >> 2) ? (0x3F80u << (16 * (i & 1))) : 0u;
    u32x4 r;
#pragma unroll
    for (int w = 0; w < 4; ++w) r[w] = w == (i >> 1) ? one : 0u;
    return r;
}

// The training arrays A_l and dZ_l are blocked by group: group q holds F rows (features) of
// its 32 samples, row f at (q F + f) 32.  Writes one 16-feature half of a forward operand of
// tile t: row f0 + n gets samples 16 t + 4 g .. +3 (8 bytes per lane; the lane part of the
// address is fixed, the rest is uniform).  The operand holds BF16 values, so the FP32 product
// with the identity is exact and its upper halves are those values.
__device__ __forceinline__ void storeRows(uint16_t* grpBase, int f0, int t, u32x4 op, u32x4 id, int lane) {
    const v4f z = {0.0f, 0.0f, 0.0f, 0.0f};
    const v4f d = mfma32(op, id, z);
    uint2 w;
    w.x = packBf2(d[0], d[1]);
    w.y = packBf2(d[2], d[3]);
    *reinterpret_cast<uint2*>(grpBase + (f0 + (lane & 15)) * 32 + 16 * t + 4 * (lane >> 4)) = w;
}

// Both halves of k-block kb of a group's operands (F rows per group in dst).
__device__ __forceinline__ void storeBlock(uint16_t* dst, int F, uint32_t grp, int kb, const u32x4 (&op)[kTT], u32x4 idLo,
                                           u32x4 idHi, int lane) {
    uint16_t* base = dst + size_t(grp) * F * 32;
#pragma unroll
    for (int t = 0; t < kTT; ++t) storeRows(base, 32 * kb, t, op[t], idLo, lane);
#pragma unroll
    for (int t = 0; t < kTT; ++t) storeRows(base, 32 * kb + 16, t, op[t], idHi, lane);
}

// ReLU masks of the layer inputs, kept as two words per tile: for layer L and k-block kb, bit
// o + w (o = 4 (2 (L - 1) + kb)) of `lo` / `hi` is set when the low / high BF16 half of operand
// word w is not zero (reluPack's slot order: word w holds slots 2w, 2w+1).
struct ReluMask {
    uint32_t lo, hi;
};

// Adds the flags of a ReLU output operand at offset o: min(bits, 1) per half is the flag
// (ReLU outputs are +0 or positive, so nonzero bits are nonzero numbers), at bits w and 16 + w.
// The v_pk_min_u16 is written as an instruction of its own, with (1, 1) in a register (an
// inline constant of a packed instruction feeds its high half 0): left to the compiler,
// min(max(x, 0), 1) and the shifts after it became compares whose masks the gradient test
// against the FP64 oracle found wrong.
__device__ __forceinline__ uint32_t minOne16(uint32_t x, uint32_t ones) {
    uint32_t y;
    asm("v_pk_min_u16 %0, %1, %2" : "=v"(y) : "v"(x), "v"(ones));
    return y;
}

__device__ __forceinline__ void addMask(ReluMask& m, u32x4 h, int o) {
    uint32_t f = 0u;
#pragma unroll
    for (int w = 0; w < 4; ++w) f |= minOne16(h[w], 0x00010001u) << w;
    m.lo |= (f & 0xFu) << o;
    m.hi |= (f >> 16) << o;
}

// Masked BF16 pack of output blocks 2kb, 2kb+1 into a k-block operand (reluPack's slot order):
// each word is packed and ANDed with the halves its flags keep (a one-bit field sign-extended
// to 0 or ~0 per half, merged with v_bfi).
__device__ __forceinline__ u32x4 maskPack(const v4f& lo, const v4f& hi, const ReluMask& m, int o) {
    const float v[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    u32x4 r;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint32_t keepLo = uint32_t(int32_t(m.lo << (31 - (o + w))) >> 31);
        const uint32_t keepHi = uint32_t(int32_t(m.hi << (31 - (o + w))) >> 31);
        r[w] = packBf2(v[2 * w], v[2 * w + 1]) & ((keepLo & 0xFFFFu) | (keepHi & 0xFFFF0000u));
    }
    return r;
}

// Layer L (1..4) of the training forward: acc holds layer L-1's pre-activations on entry and
// layer L's on return.  Stores A_L, records its ReLU mask.
template <int L, int MB>
__device__ __forceinline__ void trainLayer(v4f (&acc)[kTT][4], ReluMask (&mask)[kTT], const u32x4* lwf, const float* lb,
                                           const GuideTrainArgs& a, uint32_t grp, u32x4 idLo, u32x4 idHi, int lane) {
    const int g = lane >> 4;
    u32x4 h[2][kTT];
#pragma unroll
    for (int t = 0; t < kTT; ++t) {
        h[0][t] = reluPack(acc[t][0], acc[t][1]);
        h[1][t] = reluPack(acc[t][2], acc[t][3]);
        addMask(mask[t], h[0][t], 8 * (L - 1));
        addMask(mask[t], h[1][t], 8 * (L - 1) + 4);
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) acc[t][mb] = biasInit(lb, 64 * L, mb, g);
    }
    storeBlock(a.act[L], kGHid, grp, 0, h[0], idLo, idHi, lane);
    storeBlock(a.act[L], kGHid, grp, 1, h[1], idLo, idHi, lane);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const u32x4 af = lwf[(kGFragOff[L] + mb * 2 + kb) * 64 + lane];
#pragma unroll
            for (int t = 0; t < kTT; ++t) acc[t][mb] = mfma32(af, h[kb][t], acc[t][mb]);
        }
}

// Backward through layer L (4..1): dz holds dZ_L as k-block operands; on return it holds
// dZ_{L-1} = BF16(W_L^T dZ_L masked by A_L > 0), which is also stored.
template <int L>
__device__ __forceinline__ void trainBack(u32x4 (&dz)[2][kTT], v4f (&acc)[kTT][4], const ReluMask (&mask)[kTT],
                                          const u32x4* lwt, const GuideTrainArgs& a, uint32_t grp, u32x4 idLo,
                                          u32x4 idHi, int lane) {
#pragma unroll
    for (int t = 0; t < kTT; ++t)
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) acc[t][mb] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const u32x4 at = lwt[(8 * (L - 1) + mb * 2 + kb) * 64 + lane];
#pragma unroll
            for (int t = 0; t < kTT; ++t) acc[t][mb] = mfma32(at, dz[kb][t], acc[t][mb]);
        }
#pragma unroll
    for (int t = 0; t < kTT; ++t) {
        dz[0][t] = maskPack(acc[t][0], acc[t][1], mask[t], 8 * (L - 1));
        dz[1][t] = maskPack(acc[t][2], acc[t][3], mask[t], 8 * (L - 1) + 4);
    }
    storeBlock(a.dz[L - 1], kGHid, grp, 0, dz[0], idLo, idHi, lane);
    storeBlock(a.dz[L - 1], kGHid, grp, 1, dz[1], idLo, idHi, lane);
}

}  // namespace

__global__ __launch_bounds__(512) void guideFwdBwdKernel(GuideTrainArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    u32x4* lwf = reinterpret_cast<u32x4*>(smem);
    u32x4* lwt = reinterpret_cast<u32x4*>(smem + kGWeights * 2);
    float* lb = reinterpret_cast<float*>(smem + kGWeights * 2 + kGTWeights * 2);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    float* stage = reinterpret_cast<float*>(smem + kGWeights * 2 + kGTWeights * 2 + kGBias * 4) + wave * kInferStage;
    for (int i = threadIdx.x; i < kGWeights / 8; i += blockDim.x) lwf[i] = reinterpret_cast<const u32x4*>(a.g.wf)[i];
    for (int i = threadIdx.x; i < kGTWeights / 8; i += blockDim.x) lwt[i] = reinterpret_cast<const u32x4*>(a.g.wt)[i];
    for (int i = threadIdx.x; i < kGBias; i += blockDim.x) lb[i] = a.g.b[i];
    __syncthreads();
    const u32x4 idLo = identityOp(lane, 0), idHi = identityOp(lane, 1);
    const uint32_t n = a.n;
    const uint32_t groups = (n + kTile * kTT - 1) / (kTile * kTT);
    float lossAcc = 0.0f;
    for (uint32_t grp = blockIdx.x * kTrainWaves + wave; grp < groups; grp += gridDim.x * kTrainWaves) {
        const uint32_t s0 = grp * kTile * kTT;
        float s9[kTT][9], dir[kTT][3], c[kTT], pb[kTT], pd[kTT];
        bool valid[kTT];
#pragma unroll
        for (int t = 0; t < kTT; ++t) {
            const uint32_t si = s0 + kTile * t + r;
            valid[t] = si < n;
            c[t] = 0.0f;
            pb[t] = 0.0f;
            pd[t] = 0.0f;
            dir[t][0] = 0.0f;
            dir[t][1] = 0.0f;
            dir[t][2] = 1.0f;
            for (int j = 0; j < 9; ++j) s9[t][j] = 0.0f;
            if (valid[t]) guideSample(a, si, s9[t], dir[t], c[t], pb[t], pd[t]);
        }
        // layer 0: features are produced one k-block at a time
        v4f acc[kTT][4];
        float2 st[kTT];
        ReluMask mask[kTT];
#pragma unroll
        for (int t = 0; t < kTT; ++t) {
            st[t] = make_float2(0.0f, 1.0f);
            mask[t] = ReluMask{0u, 0u};
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) acc[t][mb] = biasInit(lb, 0, mb, g);
        }
#pragma unroll
        for (int kb = 0; kb < 7; ++kb) {
            u32x4 bx[kTT];
#pragma unroll
            for (int t = 0; t < kTT; ++t) bx[t] = guideEncodeBlock(s9[t], g, kb, st[t]);
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                const u32x4 af = lwf[(kGFragOff[0] + mb * 7 + kb) * 64 + lane];
#pragma unroll
                for (int t = 0; t < kTT; ++t) acc[t][mb] = mfma32(af, bx[t], acc[t][mb]);
            }
        }
        // the weight gradient re-derives A_0 from the input scalars (36 B per sample instead of
        // the 448 B of the BF16 features): lane group g writes scalars g, g + 4, g + 8
#pragma unroll
        for (int t = 0; t < kTT; ++t) {
            float* sb = a.scal + size_t(grp) * 9 * 32 + kTile * t + r;
#pragma unroll
            for (int j = 0; j < 9; ++j)
                if ((j & 3) == g) sb[j * 32] = s9[t][j];
        }
        trainLayer<1, 4>(acc, mask, lwf, lb, a, grp, idLo, idHi, lane);
        trainLayer<2, 4>(acc, mask, lwf, lb, a, grp, idLo, idHi, lane);
        trainLayer<3, 4>(acc, mask, lwf, lb, a, grp, idLo, idHi, lane);
        trainLayer<4, 3>(acc, mask, lwf, lb, a, grp, idLo, idHi, lane);
        // Output gradient, one tile at a time through the wave's LDS stage: lane (r, g) takes
        // lobes 2g, 2g+1 of sample r; the gradient returns as the k-block operands of dZ_4.
        u32x4 dz[2][kTT];
#pragma unroll
        for (int t = 0; t < kTT; ++t) {
#pragma unroll
            for (int mb = 0; mb < 3; ++mb)
                *reinterpret_cast<float4*>(stage + r * kStageRow + mb * 16 + 4 * g) =
                    make_float4(acc[t][mb][0], acc[t][mb][1], acc[t][mb][2], acc[t][mb][3]);
            waveSync();
            float grad[10], gsel;
            const float l = outputGrad(stage + r * kStageRow, dir[t], c[t] * a.scale, c[t], pb[t], pd[t], g, grad, gsel);
            if (g == 0 && valid[t]) lossAcc += l;
            waveSync();
#pragma unroll
            for (int q = 0; q < 10; ++q) stage[r * kStageRow + 10 * g + q] = grad[q];
            if (g == 0)
                for (int q = kGOut; q < kGOutPad; ++q) stage[r * kStageRow + q] = q == kGSel ? gsel : 0.0f;
            waveSync();
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                float v[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int idx = 32 * kb + 4 * g + (i & 3) + 16 * (i >> 2);
                    v[i] = idx < kGOutPad ? stage[r * kStageRow + idx] : 0.0f;
                }
                dz[kb][t] = u32x4{packBf2(v[0], v[1]), packBf2(v[2], v[3]), packBf2(v[4], v[5]), packBf2(v[6], v[7])};
            }
            waveSync();
        }
        // dZ_4 rows 0..47 (k-block 0 both halves, k-block 1 the lower half), then the layers down
        {
            uint16_t* base = a.dz[4] + size_t(grp) * kGOutPad * 32;
#pragma unroll
            for (int t = 0; t < kTT; ++t) {
                storeRows(base, 0, t, dz[0][t], idLo, lane);
                storeRows(base, 16, t, dz[0][t], idHi, lane);
                storeRows(base, 32, t, dz[1][t], idLo, lane);
            }
        }
        trainBack<4>(dz, acc, mask, lwt, a, grp, idLo, idHi, lane);
        trainBack<3>(dz, acc, mask, lwt, a, grp, idLo, idHi, lane);
        trainBack<2>(dz, acc, mask, lwt, a, grp, idLo, idHi, lane);
        trainBack<1>(dz, acc, mask, lwt, a, grp, idLo, idHi, lane);
    }
    // Loss: one atomic per workgroup (atomics on one address are serialized, and one per wave
    // made the end of the launch queue behind them).
    __shared__ float lossSh[kTrainWaves];
    for (int off = 32; off > 0; off >>= 1) lossAcc += __shfl_down(lossAcc, off);
    if (lane == 0) lossSh[wave] = lossAcc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kTrainWaves; ++w) t += double(lossSh[w]);
        atomicAdd(a.loss, t);
    }
}

// One (layer, 64-column) tile of dW_l[n][k] = sum_s dZ_l[n][s] A_l[k][s] over one chunk of
// samples, and for the first column tile of a layer db_l[n] = sum_s dZ_l[n][s]; written to the
// chunk's partial gradient (every parameter of a chunk is written by exactly one wave).  Per
// group of 32 samples, lane (r, g) loads samples 8g..8g+7 of row r of each 16-row block (one
// 16-byte load; a block is 1 KB contiguous), so the sample index is the MFMA's reduction index.  Workgroups are dealt to the 8 XCDs in turn; the
// mapping keeps the 8 tiles of a chunk on one XCD, whose L2 then serves the dZ_0 rows that the
// four layer-0 tiles share.
__global__ __launch_bounds__(64) void guideWeightGradKernel(GuideTrainArgs a, uint32_t chunks, uint32_t chunkGroups, float* __restrict__ partial) {
    const uint32_t w = blockIdx.x, xcd = w % 8u, slot = w / 8u;
    const int tile = int(slot % kGradTiles);
    const uint32_t ch = (slot / kGradTiles) * 8u + xcd;
    if (ch >= chunks) return;
    const int layer = tile < kGradL0Tiles ? 0 : tile - (kGradL0Tiles - 1);
    const int kt = tile < kGradL0Tiles ? tile : 0;
    const int K = layer == 0 ? kGInPad : kGHid;
    const int nrb = layer == 4 ? kGOutPad / 16 : 4;
    const int ncb = layer == 0 ? 2 : 4;      // 16-column blocks of the tile
    const int c0 = layer == 0 ? 32 * kt : 0; // its first column
    const int N = layer == 4 ? kGOutPad : kGHid;
    const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
    const uint32_t groups = (a.n + 31u) / 32u;
    const uint32_t q0 = ch * chunkGroups, q1 = min(groups, q0 + chunkGroups);
    const uint16_t* dz = a.dz[layer] + r * 32 + 8 * g;
    const uint16_t* act = layer == 0 ? nullptr : a.act[layer] + r * 32 + 8 * g;
    // Layer 0: the lane's B operand of column block cb is feature c0 + 16 cb + r of samples
    // 8 g .. 8 g + 7, computed from the stored input scalars (guidePair order; the 4 padding
    // pairs are zero, as in the forward).  Lanes r and r ^ 1 hold the sine and the cosine of
    // one pair: each evaluates both for four of the eight samples (the even lane samples 0-3,
    // the odd lane 4-7) and they swap the half the other needs.
    int fj[4], fk[4];
    bool fcos[4], flive[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
        const int f = c0 + 16 * cb + r;
        int j = 0, k = 0;
        flive[cb] = layer == 0 && cb < ncb && guidePair(f >> 1, j, k);
        fj[cb] = j;
        fk[cb] = k;
        fcos[cb] = (f & 1) != 0;
    }
    v4f acc[4][4];
    float bs[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t q = q0; q < q1; ++q) {
        const uint16_t* dq = dz + size_t(q) * N * 32;
        u32x4 A[4], B[4];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
            A[rb] = rb < nrb ? *reinterpret_cast<const u32x4*>(dq + 16 * rb * 32) : u32x4{0u, 0u, 0u, 0u};
        if (layer == 0) {
            const float* sq = a.scal + size_t(q) * 9 * 32 + 8 * g + 4 * (r & 1);
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                const float4 x = *reinterpret_cast<const float4*>(sq + fj[cb] * 32);
                const float xs[4] = {x.x, x.y, x.z, x.w};
                float sn[4], cs[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float2 sc = guideFeaturePair(xs[i], fk[cb]);
                    sn[i] = sc.x;
                    cs[i] = sc.y;
                }
                const uint32_t own0 = fcos[cb] ? packBf2(cs[0], cs[1]) : packBf2(sn[0], sn[1]);
                const uint32_t own1 = fcos[cb] ? packBf2(cs[2], cs[3]) : packBf2(sn[2], sn[3]);
                const uint32_t oth0 = __shfl_xor(fcos[cb] ? packBf2(sn[0], sn[1]) : packBf2(cs[0], cs[1]), 1);
                const uint32_t oth1 = __shfl_xor(fcos[cb] ? packBf2(sn[2], sn[3]) : packBf2(cs[2], cs[3]), 1);
                B[cb] = fcos[cb] ? u32x4{oth0, oth1, own0, own1} : u32x4{own0, own1, oth0, oth1};
                if (!flive[cb]) B[cb] = u32x4{0u, 0u, 0u, 0u};
            }
            B[2] = B[3] = u32x4{0u, 0u, 0u, 0u};
        } else {
            const uint16_t* aq = act + size_t(q) * K * 32;
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
                B[cb] = cb < ncb ? *reinterpret_cast<const u32x4*>(aq + 16 * cb * 32) : u32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
                if (rb < nrb && cb < ncb) acc[rb][cb] = mfma32(A[rb], B[cb], acc[rb][cb]);
        if (kt == 0)
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                for (int q = 0; q < 4; ++q) bs[rb] += bf2f(uint16_t(A[rb][q] & 0xFFFFu)) + bf2f(uint16_t(A[rb][q] >> 16));
    }
    float* out = partial + size_t(ch) * kGParams;
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
            if (rb < nrb && cb < ncb)
#pragma unroll
                for (int j = 0; j < 4; ++j) out[kGOffW[layer] + (16 * rb + 4 * g + j) * K + c0 + 16 * cb + r] = acc[rb][cb][j];
    if (kt == 0)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            float v = bs[rb];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (g == 0 && rb < nrb) out[kGOffB[layer] + 16 * rb + r] = v;
        }
}

// grad[p] = sum over chunks of the partial gradients, in chunk order.
__global__ __launch_bounds__(256) void guideGradReduceKernel(const float* __restrict__ partial, uint32_t chunks,
                                                             float* __restrict__ grad) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kGParams) return;
    float s = 0.0f;
    uint32_t c = 0;
    for (; c + 8 <= chunks; c += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = partial[size_t(c + k) * kGParams + p];
#pragma unroll
        for (int k = 0; k < 8; ++k) s += v[k];
    }
    for (; c < chunks; ++c) s += partial[size_t(c) * kGParams + p];
    grad[p] = s;
}

// Element index of weight i (row-major parameter layout) in the fragment-ordered copy.
__host__ __device__ inline int fragIndexOf(int i) {
    const int l = i < kGOffW[1] ? 0 : i < kGOffW[2] ? 1 : i < kGOffW[3] ? 2 : i < kGOffW[4] ? 3 : 4;
    const int K = l == 0 ? kGInPad : kGHid;
    const int o = i - kGOffW[l];
    return guideFragPos(l, o / K, o % K);
}

// Element index of weight i of layers 1-4 in the transposed fragment copy (-1 for layer 0 and
// for layer 4's rows from kGOut on: the selection row's gradient stays out of the hidden
// layers, and the padding rows are zero).
__host__ __device__ inline int tfragIndexOf(int i) {
    if (i < kGOffW[1]) return -1;
    const int l = i < kGOffW[2] ? 1 : i < kGOffW[3] ? 2 : i < kGOffW[4] ? 3 : 4;
    const int o = i - kGOffW[l];
    if (l == 4 && o / kGHid >= kGOut) return -1;
    return guideTFragPos(l, o / kGHid, o % kGHid);
}

__global__ __launch_bounds__(256) void guideAdamKernel(float* __restrict__ p, float* __restrict__ m,
                                                       float* __restrict__ v, const float* __restrict__ g,
                                                       uint16_t* __restrict__ wbf, uint16_t* __restrict__ wfrag,
                                                       uint16_t* __restrict__ wtrans, float lr, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kGParams) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
    p[i] = pi;
    if (i < kGWeights) {
        const uint16_t b = f2bf(pi);
        wbf[i] = b;
        wfrag[fragIndexOf(i)] = b;
        const int ti = tfragIndexOf(i);
        if (ti >= 0) wtrans[ti] = b;
    }
}

// Direction i drawn from query i's mixture with the renderer's sampler, and its density.
__global__ __launch_bounds__(256) void guideSampleKernel(const uint4* __restrict__ rec, const float* __restrict__ u,
                                                         uint32_t n, uint32_t perQuery, float* __restrict__ dirs,
                                                         float* __restrict__ pdfs) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    uint4 r[5];
    for (int k = 0; k < 5; ++k) r[k] = rec[size_t(i / perQuery) * 5 + k];
    float w[3];
    guideSampleDir(r, u[3 * size_t(i)], u[3 * size_t(i) + 1], u[3 * size_t(i) + 2], w);
    dirs[3 * size_t(i)] = w[0];
    dirs[3 * size_t(i) + 1] = w[1];
    dirs[3 * size_t(i) + 2] = w[2];
    pdfs[i] = guideMixPdf(r, w);
}

// Synthetic resident inputs for timing: kind 0 writes n queries (9 floats: position in [0,1]^3,
// unit normal, unit wo); kind 1 writes n training samples in the packed 4 x float4 layout.
__device__ __forceinline__ float synthU(uint32_t i, uint32_t d) {
    uint64_t z = (uint64_t(i) << 8 | d) * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return float(uint32_t((z ^ (z >> 31)) >> 40)) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ void synthDir(uint32_t i, uint32_t d, float* w) {
    const float z = 1.0f - 2.0f * synthU(i, d), ph = 6.2831853f * synthU(i, d + 1);
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    w[0] = r * cosf(ph);
    w[1] = r * sinf(ph);
    w[2] = z;
}
__global__ __launch_bounds__(256) void guideSynthKernel(float* out, uint32_t n, int kind) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float p[3] = {synthU(i, 0), synthU(i, 1), synthU(i, 2)}, nn[3], wo[3], d[3];
    synthDir(i, 3, nn);
    synthDir(i, 5, wo);
    synthDir(i, 7, d);
    if (kind == 0) {
        float* q = out + size_t(i) * 9;
        for (int k = 0; k < 3; ++k) {
            q[k] = p[k];
            q[3 + k] = nn[k];
            q[6 + k] = wo[k];
        }
    } else {
        float4* s = reinterpret_cast<float4*>(out) + size_t(i) * 4;
        s[0] = make_float4(p[0], p[1], p[2], synthU(i, 9));
        s[1] = make_float4(nn[0], nn[1], nn[2], 0.0f);
        const float pbs = 0.05f + synthU(i, 10);   // a BSDF density: the selection term runs
        s[2] = make_float4(wo[0], wo[1], wo[2], pbs);
        s[3] = make_float4(d[0], d[1], d[2], pbs);
    }
}

__global__ void guideClearRecordsKernel(uint32_t* count, double* csum) {
    if (threadIdx.x == 0) {
        count[0] = 0u;
        count[1] = 0u;
        *csum = 0.0;
    }
}

__global__ void guideZeroKernel(float* p, int n, double* loss) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0f;
    if (i == 0 && loss) *loss = 0.0;
}

// ---------------------------------------------------------------- host side

namespace {

GuideState& state(rbr_ctx* ctx) {
    if (!ctx->guide) ctx->guide = new GuideState();
    return *ctx->guide;
}

size_t trainSmem() {
    return size_t(kGWeights) * 2 + size_t(kGTWeights) * 2 + kGBias * 4 + size_t(kInferStage) * 4 * kTrainWaves;
}
size_t inferSmem() { return size_t(kGWeights) * 2 + kGBias * 4 + size_t(kInferW4Frags) * 64 * 16 + kGOutPad * 4; }

int refreshBf16(rbr_ctx* ctx, GuideState& g) {
    std::vector<uint16_t> bf(kGWeights), frag(kGWeights), trans(kGTWeights, 0);
    for (int i = 0; i < kGWeights; ++i) {
        uint32_t u;
        std::memcpy(&u, &g.master[i], 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        bf[i] = uint16_t(u >> 16);
        frag[fragIndexOf(i)] = bf[i];
        const int ti = tfragIndexOf(i);
        if (ti >= 0) trans[ti] = bf[i];
    }
    int rc;
    if ((rc = uploadBuffer(ctx, g.dW, bf.data(), bf.size() * 2))) return rc;
    if ((rc = uploadBuffer(ctx, g.dWf, frag.data(), frag.size() * 2))) return rc;
    if ((rc = uploadBuffer(ctx, g.dWt, trans.data(), trans.size() * 2))) return rc;
    if ((rc = uploadBuffer(ctx, g.dMaster, g.master.data(), g.master.size() * 4))) return rc;
    return RBR_OK;
}

uint32_t gradChunks(size_t n) { return uint32_t(((n + 31) / 32 * 32 + gradChunk(n) - 1) / gradChunk(n)); }

int ensureTrain(rbr_ctx* ctx, GuideState& g, size_t n) {
    const size_t cap = (n + kTile * kTT - 1) / (kTile * kTT) * (kTile * kTT);
    if (cap > g.cap) {
        for (int l = 0; l < 5; ++l) {
            const int N = l == 4 ? kGOutPad : kGHid;
            // act[0] holds the 9 FP32 input scalars per sample, the others BF16 activations
            if (int rc = ensureBuffer(ctx, g.act[l], l == 0 ? cap * 9 * 4 : cap * kGHid * 2)) return rc;
            if (int rc = ensureBuffer(ctx, g.dz[l], cap * N * 2)) return rc;
        }
        g.cap = cap;
    }
    const size_t chunks = gradChunks(n);
    if (chunks > g.partialChunks) {
        if (int rc = ensureBuffer(ctx, g.dPartial, chunks * kGParams * 4)) return rc;
        g.partialChunks = chunks;
    }
    int rc;
    if ((rc = ensureBuffer(ctx, g.dGrad, size_t(kGParams) * 4))) return rc;
    if ((rc = ensureBuffer(ctx, g.dM, size_t(kGParams) * 4))) return rc;
    if ((rc = ensureBuffer(ctx, g.dV, size_t(kGParams) * 4))) return rc;
    if ((rc = ensureBuffer(ctx, g.dLoss, 8))) return rc;
    return RBR_OK;
}

GuideTrainArgs trainArgs(rbr_ctx* ctx, GuideState& g, const float4* samples, uint32_t stride, uint32_t n, float scale) {
    GuideTrainArgs a{};
    a.g = guideDev(g);
    a.samples = samples;
    a.stride = stride;
    a.n = n;
    a.cap = g.cap;
    a.scale = scale;
    a.cmax = INFINITY;
    a.scal = static_cast<float*>(g.act[0].ptr);
    a.act[0] = nullptr;
    for (int l = 0; l < 5; ++l) {
        if (l > 0) a.act[l] = static_cast<uint16_t*>(g.act[l].ptr);
        a.dz[l] = static_cast<uint16_t*>(g.dz[l].ptr);
    }
    a.grad = static_cast<float*>(g.dGrad.ptr);
    a.loss = static_cast<double*>(g.dLoss.ptr);
    (void)ctx;
    return a;
}

int launchLossGrad(rbr_ctx* ctx, GuideState& g, const GuideTrainArgs& a) {
    // the loss accumulator only: every gradient element is written by the reduction
    hipLaunchKernelGGL(guideZeroKernel, dim3(1), dim3(64), 0, ctx->stream, a.grad, 0, a.loss);
    const unsigned groups = (a.n + kTile * kTT - 1) / (kTile * kTT);
    const unsigned grid = std::max(1u, std::min<unsigned>((groups + kTrainWaves - 1) / kTrainWaves, unsigned(ctx->cuCount)));
    hipLaunchKernelGGL(guideFwdBwdKernel, dim3(grid), dim3(64 * kTrainWaves), trainSmem(), ctx->stream, a);
    const uint32_t chunks = gradChunks(a.n);
    float* partial = static_cast<float*>(g.dPartial.ptr);
    hipLaunchKernelGGL(guideWeightGradKernel, dim3((chunks + 7) / 8 * 8 * kGradTiles), dim3(64), 0, ctx->stream, a, chunks,
                       gradChunk(a.n) / 32, partial);
    hipLaunchKernelGGL(guideGradReduceKernel, dim3((kGParams + 255) / 256), dim3(256), 0, ctx->stream, partial, chunks,
                       a.grad);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "guide training launch");
}

}  // namespace

GuideDev guideDev(const GuideState& g) {
    GuideDev d{};
    d.w = static_cast<const uint16_t*>(g.dW.ptr);
    d.wf = static_cast<const uint16_t*>(g.dWf.ptr);
    d.wt = static_cast<const uint16_t*>(g.dWt.ptr);
    d.b = static_cast<const float*>(g.dMaster.ptr) + kGWeights;
    for (int k = 0; k < 3; ++k) {
        d.lo[k] = g.lo[k];
        d.inv[k] = g.inv[k];
    }
    return d;
}

int guideSetAttributes(rbr_ctx* ctx) {
    GuideState& g = state(ctx);
    if (g.attrSet) return RBR_OK;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(guideInferKernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(inferSmem()));
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(guideFwdBwdKernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(trainSmem()));
    if (e != hipSuccess) return failHip(ctx, e, "guide LDS attribute");
    g.attrSet = true;
    return RBR_OK;
}

int launchGuideInfer(rbr_ctx* ctx, const GuideInferArgs& a, uint32_t maxN) {
    if (int rc = guideSetAttributes(ctx)) return rc;
    const unsigned groups = (maxN + kTile * kIT - 1) / (kTile * kIT);
    const unsigned grid = std::max(1u, std::min<unsigned>((groups + kWaves - 1) / kWaves, unsigned(ctx->cuCount)));
    hipLaunchKernelGGL(guideInferKernel, dim3(grid), dim3(64 * kWaves), inferSmem(), ctx->stream, a);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "guide inference launch");
}

int guideTrainDevice(rbr_ctx* ctx, const float4* samples, uint32_t stride, uint32_t n, uint32_t total, float scale,
                     float cmax, float lr, double* lossOut) {
    GuideState& g = state(ctx);
    if (!g.ready) return fail(ctx, RBR_ERR_STATE, "guide not initialised (rbr_guide_init)");
    if (total && n > total) return fail(ctx, RBR_ERR_INVALID, "batch larger than the record set");
    if (int rc = ensureTrain(ctx, g, n)) return rc;
    if (!g.adamInit) {
        hipLaunchKernelGGL(guideZeroKernel, dim3((kGParams + 255) / 256), dim3(256), 0, ctx->stream,
                           static_cast<float*>(g.dM.ptr), kGParams, nullptr);
        hipLaunchKernelGGL(guideZeroKernel, dim3((kGParams + 255) / 256), dim3(256), 0, ctx->stream,
                           static_cast<float*>(g.dV.ptr), kGParams, nullptr);
        g.adamInit = true;
        g.step = 0;
    }
    if (int rc = guideSetAttributes(ctx)) return rc;
    GuideTrainArgs a = trainArgs(ctx, g, samples, stride, n, scale);
    a.cmax = cmax;
    if (total) {
        // xxHash's 32-bit primes; total < 2^27 is never a multiple of one
        static const uint32_t primes[5] = {2654435761u, 2246822519u, 3266489917u, 668265263u, 374761393u};
        g.trainRng = g.trainRng * 6364136223846793005ull + 1442695040888963407ull;
        a.total = total;
        a.permMul = primes[(g.trainRng >> 33) % 5] % total;
        a.invTotal = 1.0 / double(total);
        a.permAdd = uint32_t((g.trainRng >> 11) % total);
    }
    if (int rc = launchLossGrad(ctx, g, a)) return rc;
    g.step += 1;
    const float c1 = 1.0f / (1.0f - std::pow(0.9f, float(g.step)));
    const float c2 = 1.0f / (1.0f - std::pow(0.999f, float(g.step)));
    hipLaunchKernelGGL(guideAdamKernel, dim3((kGParams + 255) / 256), dim3(256), 0, ctx->stream,
                       static_cast<float*>(g.dMaster.ptr), static_cast<float*>(g.dM.ptr), static_cast<float*>(g.dV.ptr),
                       static_cast<const float*>(g.dGrad.ptr), static_cast<uint16_t*>(g.dW.ptr),
                       static_cast<uint16_t*>(g.dWf.ptr), static_cast<uint16_t*>(g.dWt.ptr), lr, c1, c2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return failHip(ctx, e, "guide Adam launch");
    g.masterDirty = true;
    if (lossOut) {
        double l = 0.0;
        e = hipMemcpyAsync(&l, a.loss, 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return failHip(ctx, e, "guide loss download");
        *lossOut = l;
    }
    return RBR_OK;
}

int guideEnsureRecords(rbr_ctx* ctx, GuideState& g) {
    if (g.recCap) return RBR_OK;
    int rc;
    if ((rc = ensureBuffer(ctx, g.dRec, size_t(g.recReserve) * kRecStride * sizeof(float4)))) return rc;
    if ((rc = ensureBuffer(ctx, g.dRecCount, 8))) return rc;
    if ((rc = ensureBuffer(ctx, g.dCsum, 8))) return rc;
    g.recCap = g.recReserve;
    hipLaunchKernelGGL(guideClearRecordsKernel, dim3(1), dim3(64), 0, ctx->stream, static_cast<uint32_t*>(g.dRecCount.ptr),
                       static_cast<double*>(g.dCsum.ptr));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "record clear");
}

void destroyGuideState(rbr_ctx* ctx) {
    GuideState* g = ctx->guide;
    if (!g) return;
    for (DevBuf* b : {&g->dW, &g->dWf, &g->dWt, &g->dPartial, &g->dMaster, &g->dM, &g->dV, &g->dGrad, &g->dLoss, &g->dQ, &g->dOut, &g->dRec, &g->dRecCount,
                      &g->dCsum})
        freeBuffer(*b);
    for (int l = 0; l < 5; ++l) {
        freeBuffer(g->act[l]);
        freeBuffer(g->dz[l]);
    }
    delete g;
    ctx->guide = nullptr;
}

}  // namespace rbr

using namespace rbr;

// IEEE binary16 -> float (host side of the guide records).
static float halfToFloat(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16, ex = (h >> 10) & 0x1Fu, man = h & 0x3FFu;
    uint32_t u;
    if (ex == 0) {
        if (man == 0) u = sign;
        else {   // subnormal: normalise
            int e = -1;
            uint32_t m = man;
            do {
                ++e;
                m <<= 1;
            } while (!(m & 0x400u));
            u = sign | uint32_t(127 - 15 - e) << 23 | (m & 0x3FFu) << 13;
        }
    } else if (ex == 31) u = sign | 0x7F800000u | man << 13;
    else u = sign | (ex + 112u) << 23 | man << 13;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

extern "C" {

int rbr_guide_init(rbr_ctx* ctx, uint64_t seed, const float lo[3], const float hi[3]) {
    if (!ctx || !lo || !hi) return RBR_ERR_INVALID;
    if (int rc = useDevice(ctx)) return rc;
    GuideState& g = state(ctx);
    g.master.assign(kGParams, 0.0f);
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
    auto u01 = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return float(z >> 40) * (1.0f / 16777216.0f);
    };
    // He-uniform initialisation of the real (unpadded) weights; biases zero.
    const int ins[5] = {kGIn, kGHid, kGHid, kGHid, kGHid};
    const int Ks[5] = {kGInPad, kGHid, kGHid, kGHid, kGHid};
    const int outs[5] = {kGHid, kGHid, kGHid, kGHid, kGOut};
    std::vector<int> real0;   // input columns that carry a feature (guidePair order)
    for (int p = 0; p < kGInPad / 2; ++p) {
        int j, k;
        if (guidePair(p, j, k)) {
            real0.push_back(2 * p);
            real0.push_back(2 * p + 1);
        }
    }
    for (int l = 0; l < 5; ++l) {
        const float lim = std::sqrt(6.0f / float(ins[l])) * (l == 4 ? 0.1f : 1.0f);
        for (int o = 0; o < outs[l]; ++o)
            for (int c = 0; c < ins[l]; ++c) {
                const int i = l == 0 ? real0[c] : c;
                g.master[kGOffW[l] + o * Ks[l] + i] = (2.0f * u01() - 1.0f) * lim;
            }
    }
    for (int k = 0; k < 3; ++k) {
        g.lo[k] = lo[k];
        const float ext = hi[k] - lo[k];
        g.inv[k] = ext > 0.0f ? 1.0f / ext : 1.0f;
    }
    g.adamInit = false;
    g.step = 0;
    g.ready = true;
    return refreshBf16(ctx, g);
}

int rbr_guide_get_params(rbr_ctx* ctx, float* params, size_t n) {
    if (!ctx || !params || n != size_t(kGParams)) return ctx ? fail(ctx, RBR_ERR_INVALID, "params size") : RBR_ERR_INVALID;
    GuideState& g = state(ctx);
    if (!g.ready) return fail(ctx, RBR_ERR_STATE, "guide not initialised");
    if (int rc = useDevice(ctx)) return rc;
    hipError_t e = hipMemcpyAsync(params, g.dMaster.ptr, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "guide params download");
    g.master.assign(params, params + n);
    return RBR_OK;
}

int rbr_guide_set_params(rbr_ctx* ctx, const float* params, size_t n) {
    if (!ctx || !params || n != size_t(kGParams)) return ctx ? fail(ctx, RBR_ERR_INVALID, "params size") : RBR_ERR_INVALID;
    GuideState& g = state(ctx);
    if (!g.ready) return fail(ctx, RBR_ERR_STATE, "guide not initialised");
    if (int rc = useDevice(ctx)) return rc;
    g.master.assign(params, params + n);
    g.adamInit = false;
    return refreshBf16(ctx, g);
}

int rbr_guide_param_count(void) { return kGParams; }

// Guide file: "RBRGUID1", uint32 version 1, parameter count, flags (bit 0: Adam moments
// present), Adam step; float scene box lo[3] and 1 / extent[3]; uint64 batch-permutation
// state; FP32 master parameters, then the Adam moments m and v when present.  Little-endian,
// as the device that wrote it.
namespace {
constexpr char kGuideMagic[8] = {'R', 'B', 'R', 'G', 'U', 'I', 'D', '1'};
struct GuideFileHead {
    char magic[8];
    uint32_t version, nparams, flags, step;
    float lo[3], inv[3];
    uint64_t rng;
};
}  // namespace

int rbr_guide_save(rbr_ctx* ctx, const char* path) {
    if (!ctx || !path) return RBR_ERR_INVALID;
    GuideState& g = state(ctx);
    if (!g.ready) return fail(ctx, RBR_ERR_STATE, "guide not initialised");
    if (int rc = useDevice(ctx)) return rc;
    std::vector<float> buf(size_t(kGParams) * 3, 0.0f);
    hipError_t e = hipMemcpyAsync(buf.data(), g.dMaster.ptr, size_t(kGParams) * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && g.adamInit) {
        e = hipMemcpyAsync(buf.data() + kGParams, g.dM.ptr, size_t(kGParams) * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(buf.data() + 2 * size_t(kGParams), g.dV.ptr, size_t(kGParams) * 4, hipMemcpyDeviceToHost,
                               ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "guide state download");
    GuideFileHead h{};
    std::memcpy(h.magic, kGuideMagic, 8);
    h.version = 1;
    h.nparams = kGParams;
    h.flags = g.adamInit ? 1u : 0u;
    h.step = uint32_t(g.step);
    for (int k = 0; k < 3; ++k) {
        h.lo[k] = g.lo[k];
        h.inv[k] = g.inv[k];
    }
    h.rng = g.trainRng;
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(ctx, RBR_ERR_IO, std::string("cannot write ") + path);
    const size_t words = size_t(kGParams) * (g.adamInit ? 3 : 1);
    const bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && std::fwrite(buf.data(), 4, words, f) == words;
    if (std::fclose(f) != 0 || !ok) return fail(ctx, RBR_ERR_IO, std::string("write failed: ") + path);
    return RBR_OK;
}

int rbr_guide_load(rbr_ctx* ctx, const char* path) {
    if (!ctx || !path) return RBR_ERR_INVALID;
    if (int rc = useDevice(ctx)) return rc;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(ctx, RBR_ERR_IO, std::string("cannot read ") + path);
    GuideFileHead h{};
    std::vector<float> buf;
    bool ok = std::fread(&h, sizeof h, 1, f) == 1 && std::memcmp(h.magic, kGuideMagic, 8) == 0 && h.version == 1 &&
              h.nparams == uint32_t(kGParams) && (h.flags & ~1u) == 0u;
    if (ok) {
        const size_t words = size_t(kGParams) * ((h.flags & 1u) ? 3 : 1);
        buf.resize(words);
        ok = std::fread(buf.data(), 4, words, f) == words && std::fgetc(f) == EOF;
    }
    std::fclose(f);
    if (!ok) return fail(ctx, RBR_ERR_IO, std::string("not a guide file of this network: ") + path);
    GuideState& g = state(ctx);
    for (int k = 0; k < 3; ++k) {
        g.lo[k] = h.lo[k];
        g.inv[k] = h.inv[k];
    }
    g.master.assign(buf.begin(), buf.begin() + kGParams);
    g.trainRng = h.rng;
    g.ready = true;
    if (int rc = refreshBf16(ctx, g)) return rc;
    g.adamInit = false;
    g.step = 0;
    if (h.flags & 1u) {
        int rc;
        if ((rc = ensureBuffer(ctx, g.dM, size_t(kGParams) * 4))) return rc;
        if ((rc = ensureBuffer(ctx, g.dV, size_t(kGParams) * 4))) return rc;
        if ((rc = uploadBuffer(ctx, g.dM, buf.data() + kGParams, size_t(kGParams) * 4))) return rc;
        if ((rc = uploadBuffer(ctx, g.dV, buf.data() + 2 * size_t(kGParams), size_t(kGParams) * 4))) return rc;
        g.adamInit = true;
        g.step = int(h.step);
    }
    return RBR_OK;
}

int rbr_guide_eval(rbr_ctx* ctx, const float* queries, size_t n, float* lobes) {
    if (!ctx || (n && (!queries || !lobes))) return RBR_ERR_INVALID;
    GuideState& g = state(ctx);
    if (!g.ready) return fail(ctx, RBR_ERR_STATE, "guide not initialised");
    if (int rc = useDevice(ctx)) return rc;
    if (n == 0) return RBR_OK;
    int rc;
    if ((rc = ensureBuffer(ctx, g.dQ, n * 9 * 4))) return rc;
    if ((rc = ensureBuffer(ctx, g.dOut, n * kGuideRecBytes))) return rc;
    hipError_t e = hipMemcpyAsync(g.dQ.ptr, queries, n * 9 * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "query upload");
    GuideInferArgs a{};
    a.g = guideDev(g);
    a.queries = static_cast<const float*>(g.dQ.ptr);
    a.n = uint32_t(n);
    a.out = static_cast<uint4*>(g.dOut.ptr);
    if ((rc = launchGuideInfer(ctx, a, uint32_t(n)))) return rc;
    std::vector<uint16_t> rec(n * kGuideRecBytes / 2);
    e = hipMemcpyAsync(rec.data(), g.dOut.ptr, n * kGuideRecBytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return failHip(ctx, e, "guide eval");
    for (size_t q = 0; q < n; ++q)
        for (int k = 0; k < kGLobes; ++k)
            for (int v = 0; v < 5; ++v) lobes[q * 40 + 5 * k + v] = halfToFloat(rec[q * 40 + 8 * v + k]);
    return RBR_OK;
}

int rbr_guide_sample(rbr_ctx* ctx, const float* queries, size_t nq, size_t per_query, const float* u, float* dirs,
                     float* pdfs) {
    if (!ctx || (nq && (!queries || !u || !dirs || !pdfs)) || per_query == 0) return RBR_ERR_INVALID;
    const size_t n = nq * per_query;
    if (n > (1u << 28)) return fail(ctx, RBR_ERR_INVALID, "too many samples");
    GuideState& g = state(ctx);
    if (!g.ready) return fail(ctx, RBR_ERR_STATE, "guide not initialised");
    if (int rc = useDevice(ctx)) return rc;
    if (n == 0) return RBR_OK;
    int rc;
    DevBuf du, dd, dp;
    if ((rc = ensureBuffer(ctx, g.dQ, nq * 9 * 4))) return rc;
    if ((rc = ensureBuffer(ctx, g.dOut, nq * kGuideRecBytes))) return rc;
    if ((rc = uploadBuffer(ctx, du, u, n * 12))) return rc;
    if ((rc = ensureBuffer(ctx, dd, n * 12)) || (rc = ensureBuffer(ctx, dp, n * 4))) {
        freeBuffer(du);
        freeBuffer(dd);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(g.dQ.ptr, queries, nq * 9 * 4, hipMemcpyHostToDevice, ctx->stream);
    GuideInferArgs a{};
    a.g = guideDev(g);
    a.queries = static_cast<const float*>(g.dQ.ptr);
    a.n = uint32_t(nq);
    a.out = static_cast<uint4*>(g.dOut.ptr);
    if (e == hipSuccess) rc = launchGuideInfer(ctx, a, uint32_t(nq));
    if (e == hipSuccess && !rc) {
        hipLaunchKernelGGL(guideSampleKernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ctx->stream,
                           static_cast<const uint4*>(g.dOut.ptr), static_cast<const float*>(du.ptr), uint32_t(n),
                           uint32_t(per_query), static_cast<float*>(dd.ptr), static_cast<float*>(dp.ptr));
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(dirs, dd.ptr, n * 12, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(pdfs, dp.ptr, n * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    freeBuffer(du);
    freeBuffer(dd);
    freeBuffer(dp);
    if (rc) return rc;
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "guide sample");
}

// samples: 13 floats each (position 3, normal 3, wo 3, direction 3, weight c).
static int uploadSamples(rbr_ctx* ctx, GuideState& g, const float* samples, size_t n) {
    std::vector<float4> packed(n * 4);
    for (size_t i = 0; i < n; ++i) {
        const float* s = samples + 14 * i;
        packed[4 * i + 0] = make_float4(s[0], s[1], s[2], s[12]);
        packed[4 * i + 1] = make_float4(s[3], s[4], s[5], 0.0f);
        packed[4 * i + 2] = make_float4(s[6], s[7], s[8], s[13]);   // drawn from the BSDF: density pb
        packed[4 * i + 3] = make_float4(s[9], s[10], s[11], s[13]);
    }
    return uploadBuffer(ctx, g.dQ, packed.data(), packed.size() * sizeof(float4));
}

int rbr_guide_loss_grad(rbr_ctx* ctx, const float* samples, size_t n, double* loss, float* grad) {
    if (!ctx || !samples || n == 0 || n > (1u << 26)) return RBR_ERR_INVALID;
    GuideState& g = state(ctx);
    if (!g.ready) return fail(ctx, RBR_ERR_STATE, "guide not initialised");
    if (int rc = useDevice(ctx)) return rc;
    int rc;
    if ((rc = uploadSamples(ctx, g, samples, n))) return rc;
    if ((rc = ensureTrain(ctx, g, n))) return rc;
    if ((rc = guideSetAttributes(ctx))) return rc;
    GuideTrainArgs a = trainArgs(ctx, g, static_cast<const float4*>(g.dQ.ptr), 4, uint32_t(n), 1.0f / float(n));
    if ((rc = launchLossGrad(ctx, g, a))) return rc;
    hipError_t e = hipSuccess;
    if (grad) e = hipMemcpyAsync(grad, g.dGrad.ptr, size_t(kGParams) * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, g.dLoss.ptr, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "guide gradient download");
}

int rbr_guide_train_step(rbr_ctx* ctx, const float* samples, size_t n, float lr, double* loss) {
    if (!ctx || !samples || n == 0 || n > (1u << 26)) return RBR_ERR_INVALID;
    GuideState& g = state(ctx);
    if (!g.ready) return fail(ctx, RBR_ERR_STATE, "guide not initialised");
    if (int rc = useDevice(ctx)) return rc;
    if (int rc = uploadSamples(ctx, g, samples, n)) return rc;
    return guideTrainDevice(ctx, static_cast<const float4*>(g.dQ.ptr), 4, uint32_t(n), 0, 1.0f / float(n), INFINITY, lr,
                            loss);
}

int rbr_guide_time(rbr_ctx* ctx, int kind, size_t n, int iters, float* mean_ms) {
    if (!ctx || !mean_ms || (kind != 0 && kind != 1) || n == 0 || n > (1u << 26) || iters < 1)
        return ctx ? fail(ctx, RBR_ERR_INVALID, "rbr_guide_time arguments") : RBR_ERR_INVALID;
    GuideState& g = state(ctx);
    if (!g.ready) return fail(ctx, RBR_ERR_STATE, "guide not initialised");
    if (int rc = useDevice(ctx)) return rc;
    int rc;
    if ((rc = ensureBuffer(ctx, g.dQ, n * (kind == 0 ? 36 : 64)))) return rc;
    if (kind == 0 && (rc = ensureBuffer(ctx, g.dOut, n * kGuideRecBytes))) return rc;
    hipLaunchKernelGGL(guideSynthKernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       static_cast<float*>(g.dQ.ptr), uint32_t(n), kind);
    auto once = [&]() -> int {
        if (kind == 0) {
            GuideInferArgs a{};
            a.g = guideDev(g);
            a.queries = static_cast<const float*>(g.dQ.ptr);
            a.n = uint32_t(n);
            a.out = static_cast<uint4*>(g.dOut.ptr);
            return launchGuideInfer(ctx, a, uint32_t(n));
        }
        // a full training step with a zero learning rate: every kernel runs, the weights stay
        return guideTrainDevice(ctx, static_cast<const float4*>(g.dQ.ptr), 4, uint32_t(n), 0, 1.0f / float(n), INFINITY,
                                0.0f, nullptr);
    };
    if ((rc = once())) return rc;
    hipError_t e = hipEventRecord(ctx->ev0, ctx->stream);
    for (int k = 0; k < iters && e == hipSuccess; ++k)
        if ((rc = once())) return rc;
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (e != hipSuccess) return failHip(ctx, e, "guide timing");
    *mean_ms = ms / float(iters);
    return RBR_OK;
}

int rbr_guide_reserve(rbr_ctx* ctx, uint32_t capacity) {
    if (!ctx || capacity == 0 || capacity > (1u << 27)) return ctx ? fail(ctx, RBR_ERR_INVALID, "record capacity") : RBR_ERR_INVALID;
    GuideState& g = state(ctx);
    if (int rc = useDevice(ctx)) return rc;
    g.recReserve = capacity;
    if (g.recCap && g.recCap != capacity) {
        freeBuffer(g.dRec);
        g.recCap = 0;
    }
    return guideEnsureRecords(ctx, g);
}

int rbr_guide_collected(rbr_ctx* ctx, uint64_t* count, double* weight_sum) {
    if (!ctx) return RBR_ERR_INVALID;
    GuideState& g = state(ctx);
    if (int rc = useDevice(ctx)) return rc;
    uint32_t c[2] = {0, 0};
    double cs = 0.0;
    if (g.recCap) {
        hipError_t e = hipMemcpyAsync(c, g.dRecCount.ptr, 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&cs, g.dCsum.ptr, 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return failHip(ctx, e, "record count download");
    }
    if (count) *count = c[1];
    if (weight_sum) *weight_sum = cs;
    return RBR_OK;
}

int rbr_guide_clear_collected(rbr_ctx* ctx) {
    if (!ctx) return RBR_ERR_INVALID;
    GuideState& g = state(ctx);
    if (int rc = useDevice(ctx)) return rc;
    if (!g.recCap) return RBR_OK;
    hipLaunchKernelGGL(guideClearRecordsKernel, dim3(1), dim3(64), 0, ctx->stream, static_cast<uint32_t*>(g.dRecCount.ptr),
                       static_cast<double*>(g.dCsum.ptr));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RBR_OK : failHip(ctx, e, "record clear");
}

int rbr_guide_train_collected(rbr_ctx* ctx, int steps, uint32_t batch, float lr, double* loss) {
    if (!ctx || steps < 0 || batch == 0 || batch > (1u << 26) || !(lr > 0.0f)) return ctx ? fail(ctx, RBR_ERR_INVALID, "train arguments") : RBR_ERR_INVALID;
    GuideState& g = state(ctx);
    if (!g.ready) return fail(ctx, RBR_ERR_STATE, "guide not initialised");
    uint64_t n = 0;
    double cs = 0.0;
    if (int rc = rbr_guide_collected(ctx, &n, &cs)) return rc;
    if (loss) *loss = 0.0;
    if (n == 0 || !(cs > 0.0) || steps == 0) return RBR_OK;
    const uint32_t B = uint32_t(std::min<uint64_t>(batch, n));
    // c is normalised by its mean over the records, so the loss scale does not depend on the
    // scene's brightness, and clamped at kTargetClamp times that mean so that a few extreme
    // path samples do not steer the fit (the target only shapes the guide; the estimator
    // stays unbiased whatever the guide is).
    const double cmean = cs / double(n);
    const float scale = float(1.0 / (cmean * double(B)));
    const float cmax = float(kTargetClamp * cmean);
    for (int s = 0; s < steps; ++s)
        if (int rc = guideTrainDevice(ctx, static_cast<const float4*>(g.dRec.ptr), kRecStride, B, uint32_t(n), scale, cmax,
                                      lr, s + 1 == steps ? loss : nullptr))
            return rc;
    return RBR_OK;
}

}  // extern "C"
