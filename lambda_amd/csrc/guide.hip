// Learned path guiding: the network's MFMA forward (inference for every shaded path of a bounce)
// and forward + backward (training on the renderer's own radiance records), Adam, and the C ABI.
//
// MFMA layout (v_mfma_f32_16x16x16_bf16, 64 lanes, g = lane / 16): A (16 x 16, M x K) lane holds
// A[lane % 16][4 g .. 4 g + 3]; B (K x N) lane holds B[4 g .. 4 g + 3][lane % 16]; D (M x N) lane
// holds D[4 g .. 4 g + 3][lane % 16]. Every layer is computed transposed, Y = W X with the
// queries as columns: then a layer's D tiles (neurons 4 g + r of query lane % 16) are exactly
// the B operand of the next layer, and activations never leave registers in the forward pass.
// The backward pass reuses that for dX = W^T dZ and transposes through LDS for dW = dZ X^T.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "guide.h"
#include "host.h"
#include "lgpu_internal.h"

namespace lgpu {

typedef short Bf4 __attribute__((ext_vector_type(4)));
typedef float F4v __attribute__((ext_vector_type(4)));

__device__ inline F4v Mfma(Bf4 a, Bf4 b, F4v c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
// gfx950's K = 32 form (twice the work per instruction). Lane group g supplies K positions
// 8 g .. 8 g + 7; the kernels below map them to 4 g .. 4 g + 3 of two 16-wide K tiles, the same
// mapping for A and B, so the product is unchanged and each layer's D tiles still feed the next
// layer as they are.
typedef __bf16 Bf8v __attribute__((ext_vector_type(8)));
typedef short S8 __attribute__((ext_vector_type(8)));
__device__ inline F4v Mfma32(S8 a, S8 b, F4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(Bf8v, a), __builtin_bit_cast(Bf8v, b), c, 0, 0, 0);
}
__device__ inline S8 Cat8(Bf4 lo, Bf4 hi) { return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7); }
__device__ inline F4v Mfma32(Bf4 a0, Bf4 a1, Bf4 b0, Bf4 b1, F4v c) { return Mfma32(Cat8(a0, a1), Cat8(b0, b1), c); }
// FP32 -> BF16 pairs on gfx950's v_cvt_pk_bf16_f32 (round to nearest even, as ToBf16).
typedef __bf16 Bf2v __attribute__((ext_vector_type(2)));
typedef float F2v __attribute__((ext_vector_type(2)));
typedef short S2 __attribute__((ext_vector_type(2)));
__device__ inline S2 CvtBf2(float a, float b) { return __builtin_bit_cast(S2, __builtin_convertvector((F2v){a, b}, Bf2v)); }
__device__ inline Bf4 Pack4(float a, float b, float c, float d) {
  return __builtin_shufflevector(CvtBf2(a, b), CvtBf2(c, d), 0, 1, 2, 3);
}
// max(x, 0) of four accumulators, as BF16. A BF16 pattern read as a 16-bit integer orders like
// its value around zero (negative values have the sign bit, so they are negative integers), so
// the ReLU is one packed integer max per pair after the conversion.
__device__ inline Bf4 PackRelu4(F4v v) {
  const S2 z = {0, 0};
  return __builtin_shufflevector(__builtin_elementwise_max(CvtBf2(v[0], v[1]), z),
                                 __builtin_elementwise_max(CvtBf2(v[2], v[3]), z), 0, 1, 2, 3);
}
__device__ inline Bf4 LoadBf4(const uint16_t *p) { return *reinterpret_cast<const Bf4 *>(p); }  // 8-byte aligned
typedef uint32_t U2v __attribute__((ext_vector_type(2)));
__device__ inline uint2 AsU2(Bf4 v) {
  const U2v u = __builtin_bit_cast(U2v, v);
  return make_uint2(u.x, u.y);
}
__device__ inline Bf4 FromU2(uint2 u) { return __builtin_bit_cast(Bf4, (U2v){u.x, u.y}); }
// Register-resident access to tile pt of a 4-tile row with a run-time pt (selects, no indexing).
__device__ inline Bf4 Get4(const Bf4 v[4], int i) { return i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3])); }
__device__ inline void Put4(Bf4 v[4], int i, Bf4 a) {
  v[0] = i == 0 ? a : v[0]; v[1] = i == 1 ? a : v[1]; v[2] = i == 2 ? a : v[2]; v[3] = i == 3 ? a : v[3];
}
__device__ inline float Lane4Sum(float v) { v += __shfl_xor(v, 16); v += __shfl_xor(v, 32); return v; }
__device__ inline float Lane4Max(float v) { v = fmaxf(v, __shfl_xor(v, 16)); v = fmaxf(v, __shfl_xor(v, 32)); return v; }

// Encoded inputs of one query as this lane's B operand of layer 0: features 16 kt + 4 g + r.
// (K tile kt holds the 16 features of input kt: octaves 2 g and 2 g + 1, sin and cos.) The angle
// of octave 2 g + 1 is twice that of octave 2 g, so its sine and cosine come from the double-angle
// identities (sin 2a = 2 sin a cos a, cos 2a = 1 - 2 sin^2 a) instead of two more transcendentals.
__device__ inline void EncodeB(const float s[kGuideScalars], int g, Bf4 x[5]) {
  const float scale = (float)(1 << (2 * g)) * .5f;  // 2^(2 g - 1): octave 2 g
#pragma unroll
  for (int kt = 0; kt < 5; ++kt) {
    const float t = Fract(s[kt] * scale);
    const float s1 = __builtin_amdgcn_sinf(t), c1 = __builtin_amdgcn_cosf(t);  // revolutions
    const float m2 = -2.f * s1;
    x[kt] = Pack4(s1, c1, -m2 * c1, fmaf(m2, s1, 1.f));
  }
}

// Hidden layer L of one wave's 64 queries (4 column tiles): x[kt][pt] in, result in x (the next
// layer's B operands). W: BF16 weights (LDS or global), B: FP32 biases. `store`: per-lane store
// of each layer's input (training), layout [slot][pt][lane] with 5 slots for layer 0 and 4 per
// hidden layer. Compile-time layer shapes keep every operand array in registers.
template <int L, bool kStore>
__device__ inline void HiddenLayer(const uint16_t *W, const float *B, Bf4 x[5][4], uint2 *store) {
  constexpr int C = GuideCols(L), nk = C / 16;
  const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
  const uint16_t *Wl = W + GuideWOff(L);
  const float *Bl = B + GuideBOff(L);
  F4v acc[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int pt = 0; pt < 4; ++pt) acc[m][pt] = F4v{0.f, 0.f, 0.f, 0.f};
  // K tiles in a rolled loop (operands picked by selects): the scheduler then keeps one K tile's
  // weights in flight instead of hoisting the whole layer's loads into registers
#pragma unroll 1
  for (int kt = 0; kt < nk; ++kt) {
    Bf4 xb[4];
#pragma unroll
    for (int pt = 0; pt < 4; ++pt)
      xb[pt] = kt == 0 ? x[0][pt] : (kt == 1 ? x[1][pt] : (kt == 2 ? x[2][pt] : (kt == 3 ? x[3][pt] : x[4][pt])));
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const Bf4 a = LoadBf4(Wl + (16 * m + c16) * C + 16 * kt + 4 * g);
#pragma unroll
      for (int pt = 0; pt < 4; ++pt) acc[m][pt] = Mfma(a, xb[pt], acc[m][pt]);
    }
  }
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const float b0 = Bl[16 * m + 4 * g], b1 = Bl[16 * m + 4 * g + 1], b2 = Bl[16 * m + 4 * g + 2],
                b3 = Bl[16 * m + 4 * g + 3];
#pragma unroll
    for (int pt = 0; pt < 4; ++pt)
      x[m][pt] = Pack4(fmaxf(acc[m][pt][0] + b0, 0.f), fmaxf(acc[m][pt][1] + b1, 0.f), fmaxf(acc[m][pt][2] + b2, 0.f),
                       fmaxf(acc[m][pt][3] + b3, 0.f));
  }
  if (kStore) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int pt = 0; pt < 4; ++pt) store[((5 + 4 * L + m) * 4 + pt) * 64 + lane] = AsU2(x[m][pt]);
  }
}
template <bool kStore>
__device__ inline void HiddenLayers(const uint16_t *W, const float *B, Bf4 x[5][4], uint2 *store) {
  if (kStore) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int kt = 0; kt < 5; ++kt)
#pragma unroll
      for (int pt = 0; pt < 4; ++pt) store[(kt * 4 + pt) * 64 + lane] = AsU2(x[kt][pt]);
  }
  HiddenLayer<0, kStore>(W, B, x, store);
  HiddenLayer<1, kStore>(W, B, x, store);
  HiddenLayer<2, kStore>(W, B, x, store);
  HiddenLayer<3, kStore>(W, B, x, store);
  HiddenLayer<4, kStore>(W, B, x, store);
}

// Inference: a wave handles kInferPT column tiles (16 kInferPT queries); the kInferBlock threads
// of a block share one LDS copy of the network.
constexpr int kInferPT = 4;  // = 64 / 16: lane l derives the inputs of query l of the wave
constexpr int kInferBlock = 512;

// The inference kernel keeps the weights in LDS in MFMA fragment order: for each layer, each
// 16-row tile m and each K pair j (32 columns), the 64 lanes' A operands of one K = 32 MFMA are
// 64 consecutive 16-byte fragments, lane (c16, g) holding row 16 m + c16, columns 32 j + 4 g ..
// + 3 and 32 j + 16 + 4 g .. + 3; a layer whose width is an odd number of 16-column tiles (layer
// 0) ends with its last tile as 8-byte K = 16 fragments. One ds_read_b128 per lane and K pair,
// consecutive lanes at consecutive addresses, instead of rows 128 bytes apart in the same banks.
// Same size as the row-major weights: fragments of layer l start at byte 2 GuideWOff(l).
LGPU_HD constexpr int GuidePairs(int l) { return GuideCols(l) / 32; }
__device__ inline void FillWeightFragments(const uint16_t *__restrict__ Wg, uint16_t *W) {
#pragma unroll 1
  for (int l = 0; l < kGuideLayers; ++l) {
    const int R = GuideRows(l), C = GuideCols(l), np = GuidePairs(l);
    const uint16_t *src = Wg + GuideWOff(l);
    uint16_t *dst = W + GuideWOff(l);
    const int n32 = (R / 16) * np * 64;  // 16-byte fragments
    for (int f = threadIdx.x; f < n32; f += blockDim.x) {
      const int lane = f & 63, t = f >> 6, m = t / np, j = t % np;
      const uint16_t *row = src + (16 * m + (lane & 15)) * C + 32 * j + 4 * (lane >> 4);
      const uint2 lo = *reinterpret_cast<const uint2 *>(row), hi = *reinterpret_cast<const uint2 *>(row + 16);
      reinterpret_cast<uint4 *>(dst)[f] = make_uint4(lo.x, lo.y, hi.x, hi.y);
    }
    if (C & 16) {  // the last 16 columns as K = 16 fragments, after the pairs
      const int n16 = (R / 16) * 64;
      for (int f = threadIdx.x; f < n16; f += blockDim.x) {
        const int lane = f & 63, m = f >> 6;
        reinterpret_cast<uint2 *>(dst + 8 * n32)[f] =
            *reinterpret_cast<const uint2 *>(src + (16 * m + (lane & 15)) * C + 32 * np + 4 * (lane >> 4));
      }
    }
  }
}
__device__ inline S8 LoadFrag32(const uint16_t *W, int l, int m, int j, int lane) {
  return reinterpret_cast<const S8 *>(W + GuideWOff(l))[(m * GuidePairs(l) + j) * 64 + lane];
}

// Activations of the inference kernel: the B operands of a 64-wide layer as two K pairs per
// column tile (xp[j][pt]: K tiles 2 j and 2 j + 1, one K = 32 MFMA operand each); layer 0's
// fifth K tile (features 64 .. 79) apart, as a K = 16 operand.
// Inference form of a hidden layer on fragment-ordered weights: the biases as the accumulators'
// start, ReLU and BF16 conversion on the packed results; same arithmetic as HiddenLayer.
template <int L>
__device__ inline void HiddenLayer32(const uint16_t *W, const float *B, S8 xp[2][kInferPT], const Bf4 *x4) {
  const int lane = threadIdx.x & 63, g = lane >> 4;
  const float *Bl = B + GuideBOff(L);
  F4v acc[4][kInferPT];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const F4v b = *reinterpret_cast<const F4v *>(Bl + 16 * m + 4 * g);
#pragma unroll
    for (int pt = 0; pt < kInferPT; ++pt) acc[m][pt] = b;
  }
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const S8 a = LoadFrag32(W, L, m, j, lane);
#pragma unroll
      for (int pt = 0; pt < kInferPT; ++pt) acc[m][pt] = Mfma32(a, xp[j][pt], acc[m][pt]);
    }
  if constexpr (L == 0) {  // K = 16 fragments of the last 16 columns, after the pairs
    const Bf4 *tail = reinterpret_cast<const Bf4 *>(W + GuideWOff(0) + 8 * 4 * GuidePairs(0) * 64);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const Bf4 a = tail[m * 64 + lane];
#pragma unroll
      for (int pt = 0; pt < kInferPT; ++pt) acc[m][pt] = Mfma(a, x4[pt], acc[m][pt]);
    }
  }
#pragma unroll
  for (int pt = 0; pt < kInferPT; ++pt) {
    xp[0][pt] = Cat8(PackRelu4(acc[0][pt]), PackRelu4(acc[1][pt]));
    xp[1][pt] = Cat8(PackRelu4(acc[2][pt]), PackRelu4(acc[3][pt]));
  }
}
// Output layer of one column tile on fragment-ordered weights: z[m] (m < 9) as OutputLayer.
__device__ inline void OutputLayer32(const uint16_t *W, const float *B, S8 x0, S8 x1, F4v z[9]) {
  constexpr int L = kGuideLayers - 1;
  const int lane = threadIdx.x & 63, g = lane >> 4;
  const float *Bl = B + GuideBOff(L);
#pragma unroll
  for (int m = 0; m < 9; ++m) {
    F4v acc = *reinterpret_cast<const F4v *>(Bl + 16 * m + 4 * g);
    acc = Mfma32(LoadFrag32(W, L, m, 0, lane), x0, acc);
    acc = Mfma32(LoadFrag32(W, L, m, 1, lane), x1, acc);
    z[m] = acc;
  }
}

// Output layer for column tile pt: z[m] (m < 9) D tiles = logits 16 m + 4 g + r of query 16 pt + lane % 16.
__device__ inline void OutputLayer(const uint16_t *W, const float *B, const Bf4 x[5][4], int pt, F4v z[9]) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
  const uint16_t *Wl = W + GuideWOff(kGuideLayers - 1);
  const float *Bl = B + GuideBOff(kGuideLayers - 1);
#pragma unroll
  for (int m = 0; m < 9; ++m) {
    F4v acc = F4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) acc = Mfma(LoadBf4(Wl + (16 * m + c16) * kGuideHidden + 16 * kt + 4 * g), Get4(x[kt], pt), acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] += Bl[16 * m + 4 * g + r];
    z[m] = acc;
  }
}

// Query inputs: from path state (render: query q is queue slot list[q]) or from an array of
// scalars (test hook: query q is slot q).
struct GuideQuerySrc {
  const float4 *ray_o, *ray_d, *hit;  // path state of the current queue
  const uint32_t *list;               // queue slots to infer (render)
  const float *scalars;               // or 5 scalars per query
  float bmin[3], binv[3];
};
// A query's raw inputs (loads only), so the next query's loads can be in flight while this one
// is evaluated: hit, ray origin and direction of its queue slot, or its 5 scalars (in h and o.x).
struct QueryRaw {
  float4 h, o, d;
};
__device__ inline QueryRaw QueryLoad(const GuideQuerySrc &src, uint32_t q) {
  QueryRaw r;
  if (src.scalars) {
    const float *p = src.scalars + kGuideScalars * (size_t)q;
    r.h = make_float4(p[0], p[1], p[2], p[3]);
    r.o = make_float4(p[4], 0.f, 0.f, 0.f);
    r.d = make_float4(0.f, 0.f, 0.f, 0.f);
    return r;
  }
  q = src.list[q];
  r.h = src.hit[q];
  r.o = src.ray_o[q];
  r.d = src.ray_d[q];
  return r;
}
__device__ inline void QueryScalars(const GuideQuerySrc &src, const QueryRaw &r, float s[kGuideScalars]) {
  if (src.scalars) {
    s[0] = r.h.x; s[1] = r.h.y; s[2] = r.h.z; s[3] = r.h.w; s[4] = r.o.x;
    return;
  }
  if (__float_as_uint(r.h.w) == kNone) {
#pragma unroll
    for (int k = 0; k < kGuideScalars; ++k) s[k] = 0.f;
    return;
  }
  const float4 o = r.o, d = r.d;
  const float t = r.h.x;
  const V3 p = v3(o.x + d.x * t, o.y + d.y * t, o.z + d.z * t);
  // local copies: the kernel argument's address would otherwise reach the callee (scratch copy)
  const float bmin[3] = {src.bmin[0], src.bmin[1], src.bmin[2]}, binv[3] = {src.binv[0], src.binv[1], src.binv[2]};
  GuideInputs(p, v3(-d.x, -d.y, -d.z), bmin, binv, s);
}

// Inference: per query, the quantised bin table (bytes max(1, round(255 exp(z - z_max))), the
// distribution is byte / sum) and the mixing probability. Persistent blocks; weights in LDS in
// fragment order. kList: queries are the queue slots src.list[q] (render), else q (test hook).
template <bool kList>
__global__ void __launch_bounds__(kInferBlock) k_guide_infer(const uint16_t *__restrict__ Wg, const float *__restrict__ Bg,
                                                               GuideQuerySrc src, uint32_t n, const uint32_t *np,
                                                               uint32_t *__restrict__ bins, float2 *__restrict__ hdr) {
  __shared__ __align__(16) uint16_t Wlds[kGuideWeights];
  __shared__ __align__(16) float Blds[kGuideBiases];
  constexpr int kPerBlock = (kInferBlock / 64) * 16 * kInferPT;
  if (np) n = *np;
  if ((uint64_t)blockIdx.x * kPerBlock >= n) return;
  FillWeightFragments(Wg, Wlds);
  for (int i = threadIdx.x; i < kGuideBiases; i += kInferBlock) Blds[i] = Bg[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15, wave = threadIdx.x >> 6;
  const float kLog2e = 1.44269504088896341f;
  const uint64_t stride = (uint64_t)gridDim.x * kPerBlock;
  // each lane derives the inputs of one query (column tile g, column c16) of its wave; the loads
  // of the next iteration's query are issued before this iteration's network
  uint64_t base = (uint64_t)blockIdx.x * kPerBlock;
  QueryRaw raw = QueryLoad(src, (uint32_t)std::min<uint64_t>(base + 16 * kInferPT * wave + lane, n - 1));
  for (; base < n; base += stride) {
    const uint64_t wbase = base + 16 * kInferPT * wave;
    const QueryRaw cur = raw;
    if (base + stride < n) raw = QueryLoad(src, (uint32_t)std::min<uint64_t>(wbase + stride + lane, n - 1));
    // an offset the compiler cannot see through: the LDS weights are loop-invariant, and hoisting
    // their loads out of this loop would hold the whole network in (spilled) registers
    int lo = 0;
    asm volatile("" : "+s"(lo));
    const uint16_t *W = Wlds + lo;
    const float *B = Blds + lo;
    // the four lanes that encode a query's features gather its inputs
    float sq[kGuideScalars];
    QueryScalars(src, cur, sq);
    S8 xp[2][kInferPT];
    Bf4 x4[kInferPT];
#pragma unroll
    for (int pt = 0; pt < kInferPT; ++pt) {
      float sp[kGuideScalars];
#pragma unroll
      for (int k = 0; k < kGuideScalars; ++k) sp[k] = __shfl(sq[k], 16 * pt + c16);
      Bf4 e[5];
      EncodeB(sp, g, e);
      xp[0][pt] = Cat8(e[0], e[1]);
      xp[1][pt] = Cat8(e[2], e[3]);
      x4[pt] = e[4];
    }
    HiddenLayer32<0>(W, B, xp, x4);
    HiddenLayer32<1>(W, B, xp, nullptr);
    HiddenLayer32<2>(W, B, xp, nullptr);
    HiddenLayer32<3>(W, B, xp, nullptr);
    HiddenLayer32<4>(W, B, xp, nullptr);
#pragma unroll
    for (int pt = 0; pt < kInferPT; ++pt) {
      F4v z[9];
      OutputLayer32(W, B, xp[0][pt], xp[1][pt], z);
      float zmax = -INFINITY;
#pragma unroll
      for (int m = 0; m < 8; ++m) zmax = fmaxf(zmax, fmaxf(fmaxf(z[m][0], z[m][1]), fmaxf(z[m][2], z[m][3])));
      zmax = Lane4Max(zmax);
      const float off = -zmax * kLog2e;
      uint32_t words[8], tot = 0u;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        uint32_t w = 0u;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          // exp(z - zmax) = 2^(z log2 e - zmax log2 e) on the hardware exp2 (inputs <= 0)
          const float e = __builtin_amdgcn_exp2f(fmaf(z[m][r], kLog2e, off));
          const uint32_t b = (uint32_t)fmaf(e, 255.f, .5f);
          w |= (b < 1u ? 1u : b) << (8 * r);
        }
        words[m] = w;
        tot = __builtin_amdgcn_udot4(w, 0x01010101u, tot, false);
      }
      const float sum = Lane4Sum((float)tot);
      const float zc = __shfl(z[8][0], c16);  // the mixing logit, row 128: lane group 0
      // Stores without a branch, so that the four column tiles' output layers and quantisation
      // form one straight-line block the scheduler can interleave. Columns past the last query
      // were given query n - 1's inputs (the clamped loads above), so they write its slot with
      // the same values; every lane group holds the same c and sum.
      const uint64_t q = std::min<uint64_t>(wbase + 16 * pt + c16, n - 1);
      const uint64_t slot = kList ? src.list[q] : q;
      // lane group g holds bytes 16 m + 4 g .. + 3: words 4 m + g
#pragma unroll
      for (int m = 0; m < 8; ++m) bins[slot * 32 + 4 * m + g] = words[m];
      hdr[slot] = make_float2(kGuideCMin + kGuideCSpan / (1.f + __expf(-zc)), sum);
    }
  }
}

// Training records: finalise the weights once the wave's paths are done.
__global__ void k_guide_finalize(GuideRecord *rec, const uint32_t *count, uint32_t cap, const float4 *L, int mode) {
  const uint32_t n = min(*count, cap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    GuideRecord r = rec[i];
    const uint32_t sid = __float_as_uint(r.b.w);
    const float lf = Luminance(S4(L[sid]), mode);
    float w = 0.f;
    if (!isnan(r.c.x) && r.c.y > 0.f) w = fmaxf(0.f, (lf - r.c.x) / r.c.y);
    if (!isfinite(w)) w = 0.f;
    rec[i].c.z = w;
  }
}

// A training batch: random record indices (with replacement) and the sum of their weights.
__global__ void k_guide_batch(const GuideRecord *rec, const uint32_t *count, uint32_t cap, uint32_t batch, uint32_t step,
                              uint32_t seed, uint32_t *idx, float *sumW) {
  const uint32_t n = min(*count, cap);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch) return;
  uint32_t c[4] = {i, step, 0x47554944u, 0u};
  Philox(c, seed, 7u);
  const uint32_t k = n ? (uint32_t)(((uint64_t)c[0] * n) >> 32) : 0u;
  idx[i] = k;
  if (n) atomicAdd(sumW, rec[k].c.z);
}

// Backward step of layer L for one wave: dW = dZ X^T and db = sum dZ added to grad (the inner
// dimension, the 64 queries, transposed through LDS), then dZ of layer L - 1 = (W^T dZ) masked by
// the ReLU of layer L's input, in place in dz (D layout, BF16).
template <int L>
__device__ inline void BackLayer(const uint16_t *W, const uint2 *st, uint16_t *dzt, uint16_t *xt, Bf4 dz[9][4],
                                 float *grad) {
  constexpr int R = GuideRows(L), C = GuideCols(L), nm = R / 16, nk = C / 16;
  constexpr int slot0 = L == 0 ? 0 : 5 + 4 * (L - 1);
  const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
#pragma unroll
  for (int kt = 0; kt < nk; ++kt)
#pragma unroll
    for (int pt = 0; pt < 4; ++pt) {
      const Bf4 v = FromU2(st[((slot0 + kt) * 4 + pt) * 64 + lane]);
#pragma unroll
      for (int r = 0; r < 4; ++r) xt[(16 * kt + 4 * g + r) * 64 + 16 * pt + c16] = (uint16_t)v[r];
    }
#pragma unroll
  for (int m = 0; m < nm; ++m)
#pragma unroll
    for (int pt = 0; pt < 4; ++pt)
#pragma unroll
      for (int r = 0; r < 4; ++r) dzt[(16 * m + 4 * g + r) * 64 + 16 * pt + c16] = (uint16_t)dz[m][pt][r];
  __syncthreads();
  float *gW = grad + GuideWOff(L);
#pragma unroll 1
  for (int m = 0; m < nm; ++m)
#pragma unroll
    for (int nt = 0; nt < nk; ++nt) {
      F4v acc = F4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kp = 0; kp < 4; ++kp)
        acc = Mfma(LoadBf4(dzt + (16 * m + c16) * 64 + 16 * kp + 4 * g), LoadBf4(xt + (16 * nt + c16) * 64 + 16 * kp + 4 * g),
                   acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) unsafeAtomicAdd(&gW[(16 * m + 4 * g + r) * C + 16 * nt + c16], acc[r]);
    }
  float *gB = grad + kGuideWeights + GuideBOff(L);
  for (int row = lane; row < R; row += 64) {
    float s = 0.f;
    for (int k = 0; k < 64; ++k) s += FromBf16(dzt[row * 64 + k]);
    unsafeAtomicAdd(&gB[row], s);
  }
  __syncthreads();
  if constexpr (L > 0) {
    const uint16_t *Wl = W + GuideWOff(L);
    Bf4 nd[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      F4v acc[4];
#pragma unroll
      for (int pt = 0; pt < 4; ++pt) acc[pt] = F4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kt = 0; kt < nm; ++kt) {
        Bf4 a;
#pragma unroll
        for (int r = 0; r < 4; ++r) a[r] = (short)Wl[(16 * kt + 4 * g + r) * C + 16 * m + c16];
#pragma unroll
        for (int pt = 0; pt < 4; ++pt) acc[pt] = Mfma(a, dz[kt][pt], acc[pt]);
      }
#pragma unroll
      for (int pt = 0; pt < 4; ++pt) {
        const Bf4 xin = FromU2(st[((slot0 + m) * 4 + pt) * 64 + lane]);
        float d[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) d[r] = FromBf16((uint16_t)xin[r]) > 0.f ? acc[pt][r] : 0.f;
        nd[m][pt] = Pack4(d[0], d[1], d[2], d[3]);
      }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int pt = 0; pt < 4; ++pt) dz[m][pt] = nd[m][pt];
  }
}

constexpr int kTrainBlock = 64;  // one wave: the LDS transposes are per wave

// Forward and backward of the guiding objective on 64 records per block:
//   loss = -sum_i (w_i / sum w) log(c q(wi_i) + (1 - c) p_bsdf,i), q = softmax_bin / (4 pi / 128).
// Gradients (FP32) are added to grad with atomics. `act`: per-block store of the layer inputs.
__global__ void __launch_bounds__(kTrainBlock) k_guide_grad(const uint16_t *__restrict__ W, const float *__restrict__ B,
                                                           const GuideRecord *__restrict__ rec, const uint32_t *__restrict__ idx,
                                                           uint32_t batch, const float *sumW, float *grad, float *loss,
                                                           uint2 *act) {
  __shared__ __align__(16) uint16_t dzt[kGuideOut * 64];  // dZ^T: [neuron][query]
  __shared__ __align__(16) uint16_t xt[kGuideIn * 64];    // X: [feature][query]
  const int lane = threadIdx.x, g = lane >> 4, c16 = lane & 15;
  const uint32_t base = blockIdx.x * 64;
  uint2 *st = act + (size_t)blockIdx.x * 25 * 4 * 64;
  const float invW = *sumW > 0.f ? 1.f / *sumW : 0.f;
  // ---- forward
  Bf4 x[5][4];
#pragma unroll 1
  for (int pt = 0; pt < 4; ++pt) {
    const uint32_t j = base + 16 * pt + c16;
    float s[kGuideScalars] = {0.f, 0.f, 0.f, 0.f, 0.f};
    if (j < batch) {
      const GuideRecord r = rec[idx[j]];
      s[0] = r.a.x; s[1] = r.a.y; s[2] = r.a.z; s[3] = r.a.w; s[4] = r.b.x;
    }
    Bf4 e[5];
    EncodeB(s, g, e);
#pragma unroll
    for (int kt = 0; kt < 5; ++kt) Put4(x[kt], pt, e[kt]);
  }
  HiddenLayers<true>(W, B, x, st);
  // ---- output layer and dZ (bf16, D layout) of the objective
  Bf4 dz[9][4];
  float lsum = 0.f;
#pragma unroll 1
  for (int pt = 0; pt < 4; ++pt) {
    const uint32_t j = base + 16 * pt + c16;
    float pbv = 1.f, wnv = 0.f;
    uint32_t kbv = 0u;
    if (j < batch) {
      const GuideRecord r = rec[idx[j]];
      kbv = __float_as_uint(r.b.y);
      pbv = r.b.z;
      wnv = r.c.z * invW;
    }
    F4v z[9];
    OutputLayer(W, B, x, pt, z);
    float zmax = -INFINITY;
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) zmax = fmaxf(zmax, z[m][r]);
    zmax = Lane4Max(zmax);
    float e[8][4], sum = 0.f, ek = 0.f;
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        e[m][r] = expf(z[m][r] - zmax);
        sum += e[m][r];
        if ((uint32_t)(16 * m + 4 * g + r) == kbv) ek = e[m][r];
      }
    sum = Lane4Sum(sum);
    ek = Lane4Sum(ek);
    const float zc = Lane4Sum(g == 0 ? z[8][0] : 0.f);
    const float sg = 1.f / (1.f + expf(-zc));
    const float c = kGuideCMin + kGuideCSpan * sg;
    const float q = ek / sum / kGuideBinSr;
    const float pmix = c * q + (1.f - c) * pbv;
    const float coef = wnv > 0.f ? -wnv / pmix : 0.f;
    if (g == 0 && wnv > 0.f) lsum += -wnv * logf(pmix);
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      float d[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float sj = e[m][r] / sum;
        const float delta = (uint32_t)(16 * m + 4 * g + r) == kbv ? 1.f : 0.f;
        d[r] = coef * c * q * (delta - sj);
      }
      Put4(dz[m], pt, Pack4(d[0], d[1], d[2], d[3]));
    }
    const float dc = coef * (q - pbv) * kGuideCSpan * sg * (1.f - sg);
    Put4(dz[8], pt, Pack4(g == 0 ? dc : 0.f, 0.f, 0.f, 0.f));
  }
  if (lsum != 0.f) atomicAdd(loss, lsum);
  // ---- backward, layer by layer
  BackLayer<5>(W, st, dzt, xt, dz, grad);
  BackLayer<4>(W, st, dzt, xt, dz, grad);
  BackLayer<3>(W, st, dzt, xt, dz, grad);
  BackLayer<2>(W, st, dzt, xt, dz, grad);
  BackLayer<1>(W, st, dzt, xt, dz, grad);
  BackLayer<0>(W, st, dzt, xt, dz, grad);
}

// Adam (Kingma and Ba 2015) on the FP32 master parameters; writes the BF16 weights and FP32
// biases the kernels read, and clears the gradient.
__global__ void k_guide_adam(float *p, float *m, float *v, float *grad, uint16_t *wbf, float *bias, float lr, float b1t,
                             float b2t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kGuideParams) return;
  const float b1 = .9f, b2 = .999f, eps = 1e-8f;
  const float gr = grad[i];
  const float mi = b1 * m[i] + (1.f - b1) * gr;
  const float vi = b2 * v[i] + (1.f - b2) * gr * gr;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / (1.f - b1t)) / (sqrtf(vi / (1.f - b2t)) + eps);
  p[i] = pi;
  grad[i] = 0.f;
  if (i < kGuideWeights) wbf[i] = ToBf16(pi);
  else bias[i - kGuideWeights] = pi;
}

__global__ void k_guide_publish(const float *p, uint16_t *wbf, float *bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kGuideParams) return;
  if (i < kGuideWeights) wbf[i] = ToBf16(p[i]);
  else bias[i - kGuideWeights] = p[i];
}

}  // namespace lgpu

using namespace lgpu;

struct lgpu_guide_s {
  lgpu_device dev = nullptr;  // null once the device was destroyed first
  lgpu_guide_desc desc{};
  DevBuf params, m, v, grad, wbf, bias, loss, sumw, idx, act;
  DevBuf rec, rec_count;
  uint32_t steps = 0;
  uint32_t last_records = 0;
  float last_loss = 0.f;
  uint64_t expected_vertices = 0;  // guided vertices of the last wave (recording rate)
  uint32_t infer_grid = 0;         // persistent inference grid: resident blocks per CU x CUs
};

namespace lgpu {

static lgpu_status GuideEnsureScratch(lgpu_guide g, uint32_t batch) {
  const uint32_t blocks = (batch + 63) / 64;
  LGPU_HIP(g->idx.alloc(4ull * std::max<uint32_t>(batch, 1)));
  LGPU_HIP(g->act.alloc((size_t)blocks * 25 * 4 * 64 * sizeof(uint2) + 256));
  return LGPU_OK;
}

// One gradient over `batch` records (indices already in g->idx, weights summed in g->sumw).
static lgpu_status GuideGradient(lgpu_guide g, uint32_t batch, const GuideRecord *rec, hipStream_t s) {
  if (!batch) return LGPU_OK;
  const uint32_t blocks = (batch + 63) / 64;
  hipLaunchKernelGGL(k_guide_grad, dim3(blocks), dim3(kTrainBlock), 0, s, g->wbf.as<uint16_t>(), g->bias.as<float>(), rec,
                     g->idx.as<uint32_t>(), batch, g->sumw.as<float>(), g->grad.as<float>(), g->loss.as<float>(),
                     g->act.as<uint2>());
  LGPU_HIP(hipGetLastError());
  return LGPU_OK;
}

static lgpu_status GuideAdam(lgpu_guide g, hipStream_t s) {
  g->steps++;
  const float b1t = powf(.9f, (float)g->steps), b2t = powf(.999f, (float)g->steps);
  hipLaunchKernelGGL(k_guide_adam, dim3((kGuideParams + 255) / 256), dim3(256), 0, s, g->params.as<float>(), g->m.as<float>(),
                     g->v.as<float>(), g->grad.as<float>(), g->wbf.as<uint16_t>(), g->bias.as<float>(), g->desc.learning_rate,
                     b1t, b2t);
  LGPU_HIP(hipGetLastError());
  return LGPU_OK;
}

// Training steps on the records of the last wave (the caller holds the device mutex).
lgpu_status GuideTrainSteps(lgpu_guide g, uint32_t steps, hipStream_t s) {
  if (!steps) return LGPU_OK;
  const uint32_t batch = g->desc.batch;
  lgpu_status st = GuideEnsureScratch(g, batch);
  if (st != LGPU_OK) return st;
  LGPU_HIP(hipMemsetAsync(g->loss.ptr, 0, 4, s));
  for (uint32_t k = 0; k < steps; ++k) {
    LGPU_HIP(hipMemsetAsync(g->sumw.ptr, 0, 4, s));
    hipLaunchKernelGGL(k_guide_batch, dim3((batch + 255) / 256), dim3(256), 0, s, g->rec.as<GuideRecord>(),
                       g->rec_count.as<uint32_t>(), g->desc.record_capacity, batch, g->steps, g->desc.seed,
                       g->idx.as<uint32_t>(), g->sumw.as<float>());
    LGPU_HIP(hipGetLastError());
    st = GuideGradient(g, batch, g->rec.as<GuideRecord>(), s);
    if (st != LGPU_OK) return st;
    st = GuideAdam(g, s);
    if (st != LGPU_OK) return st;
  }
  float l = 0.f;
  LGPU_HIP(hipMemcpyAsync(&l, g->loss.ptr, 4, hipMemcpyDeviceToHost, s));
  LGPU_HIP(hipStreamSynchronize(s));
  g->last_loss = l / (float)steps;
  return LGPU_OK;
}

// Render-loop hooks (render.hip).
lgpu_status GuidePrepareWave(lgpu_guide g, const float bounds[6], uint64_t expectedVertices, GuideRefs *refs,
                             const uint4 *bins, const float2 *hdr, hipStream_t s) {
  refs->bins = bins;
  refs->hdr = hdr;
  refs->rec = g->rec.as<GuideRecord>();
  refs->rec_count = g->rec_count.as<uint32_t>();
  refs->rec_cap = g->desc.record_capacity;
  refs->max_depth = g->desc.max_depth ? g->desc.max_depth : 0xFFFFFFFFu;
  const uint64_t ev = expectedVertices ? expectedVertices : 1;
  const double frac = std::min(1.0, (double)g->desc.record_capacity / (1.25 * (double)ev));
  refs->rec_threshold = frac >= 1.0 ? 0xFFFFFFFFu : (uint32_t)(frac * 4294967295.0);
  for (int k = 0; k < 3; ++k) {
    refs->bmin[k] = bounds[k];
    const float ext = bounds[3 + k] - bounds[k];
    refs->binv[k] = ext > 0.f ? 1.f / ext : 0.f;
  }
  LGPU_HIP(hipMemsetAsync(g->rec_count.ptr, 0, 4, s));
  return LGPU_OK;
}

// `list` holds the bounce's guided vertices (`count` of them, at most `ub`), written by the
// extension stage.
lgpu_status GuideInfer(lgpu_guide g, const GuideRefs &refs, const PathState &ps, uint32_t ub, const uint32_t *list,
                       const uint32_t *count, uint4 *bins, float2 *hdr, hipStream_t s) {
  GuideQuerySrc src{};
  src.ray_o = ps.ray_o; src.ray_d = ps.ray_d; src.hit = ps.hit; src.list = list; src.scalars = nullptr;
  for (int k = 0; k < 3; ++k) { src.bmin[k] = refs.bmin[k]; src.binv[k] = refs.binv[k]; }
  constexpr uint32_t kPerBlock = (kInferBlock / 64) * 16 * kInferPT;
  const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((ub + kPerBlock - 1) / kPerBlock, g->infer_grid));
  hipLaunchKernelGGL(k_guide_infer<true>, dim3(blocks), dim3(kInferBlock), 0, s, g->wbf.as<uint16_t>(), g->bias.as<float>(), src, ub,
                     count, reinterpret_cast<uint32_t *>(bins), hdr);
  LGPU_HIP(hipGetLastError());
  return LGPU_OK;
}

lgpu_status GuideFinishWave(lgpu_guide g, const SampleState &ss, int mode, uint32_t trainSteps, uint64_t vertices,
                            hipStream_t s, uint32_t *records) {
  hipLaunchKernelGGL(k_guide_finalize, dim3(1024), dim3(256), 0, s, g->rec.as<GuideRecord>(), g->rec_count.as<uint32_t>(),
                     g->desc.record_capacity, ss.L, mode);
  LGPU_HIP(hipGetLastError());
  uint32_t cnt = 0;
  LGPU_HIP(hipMemcpyAsync(&cnt, g->rec_count.ptr, 4, hipMemcpyDeviceToHost, s));
  LGPU_HIP(hipStreamSynchronize(s));
  g->last_records = std::min(cnt, g->desc.record_capacity);
  if (records) *records = g->last_records;
  g->expected_vertices = vertices;
  if (g->last_records && trainSteps) return GuideTrainSteps(g, trainSteps, s);
  return LGPU_OK;
}

uint64_t GuideExpectedVertices(lgpu_guide g) { return g->expected_vertices; }
bool GuideOnDevice(lgpu_guide g, lgpu_device dev) { return g->dev == dev; }

}  // namespace lgpu

// ---------------------------------------------------------------- C ABI
extern "C" uint32_t lgpu_guide_num_params(void) { return (uint32_t)kGuideParams; }

extern "C" lgpu_status lgpu_guide_create(lgpu_device dev, const lgpu_guide_desc *desc, lgpu_guide *out) {
  if (!dev || !out) return Fail(LGPU_ERR_INVALID_ARGUMENT, "guide_create: null argument");
  *out = nullptr;
  lgpu_guide_desc d{};
  if (desc) d = *desc;
  if (!d.record_capacity) d.record_capacity = 1u << 22;
  if (!d.batch) d.batch = 16384;
  if (!(d.learning_rate > 0.f)) d.learning_rate = 2e-3f;
  if (d.record_capacity > (1u << 28)) return Fail(LGPU_ERR_INVALID_ARGUMENT, "guide_create: record_capacity above 2^28");
  if (d.batch > (1u << 22)) return Fail(LGPU_ERR_INVALID_ARGUMENT, "guide_create: batch above 2^22");
  std::lock_guard<std::mutex> lk(dev->mu);
  LGPU_HIP(hipSetDevice(dev->device));
  lgpu_guide g = new lgpu_guide_s();
  g->dev = dev;
  g->desc = d;
  auto fail = [&](hipError_t e) {
    delete g;
    return Fail(LGPU_ERR_HIP, std::string("guide_create: ") + hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = g->params.alloc(4ull * kGuideParams)) != hipSuccess) return fail(e);
  if ((e = g->m.alloc(4ull * kGuideParams)) != hipSuccess) return fail(e);
  if ((e = g->v.alloc(4ull * kGuideParams)) != hipSuccess) return fail(e);
  if ((e = g->grad.alloc(4ull * kGuideParams)) != hipSuccess) return fail(e);
  if ((e = g->wbf.alloc(2ull * kGuideWeights)) != hipSuccess) return fail(e);
  if ((e = g->bias.alloc(4ull * kGuideBiases)) != hipSuccess) return fail(e);
  if ((e = g->loss.alloc(256)) != hipSuccess) return fail(e);
  if ((e = g->sumw.alloc(256)) != hipSuccess) return fail(e);
  if ((e = g->rec.alloc(sizeof(GuideRecord) * (size_t)d.record_capacity)) != hipSuccess) return fail(e);
  if ((e = g->rec_count.alloc(256)) != hipSuccess) return fail(e);
  {  // the persistent inference grid: as many blocks as are resident at once
    int perCu = 0, cus = 0;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_guide_infer<true>, kInferBlock, 0)) != hipSuccess) return fail(e);
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev->device)) != hipSuccess) return fail(e);
    g->infer_grid = (uint32_t)std::max(1, perCu * cus);
  }
  // He-uniform initial weights from the seed; zero biases except the mixing logit (c = 0.3 at start)
  std::vector<float> p(kGuideParams, 0.f);
  std::mt19937 rng(d.seed ^ 0x9E3779B9u);
  for (int l = 0; l < kGuideLayers; ++l) {
    const int R = GuideRows(l), C = GuideCols(l);
    const float lim = l == kGuideLayers - 1 ? 0.f : std::sqrt(6.f / (float)C);
    std::uniform_real_distribution<float> U(-lim, lim);
    for (int i = 0; i < R * C; ++i) p[GuideWOff(l) + i] = lim > 0.f ? U(rng) : 0.f;
  }
  // sigmoid(z) = (0.3 - cmin) / span
  const float sg = (0.3f - kGuideCMin) / kGuideCSpan;
  p[kGuideWeights + GuideBOff(kGuideLayers - 1) + kGuideBins] = std::log(sg / (1.f - sg));
  hipStream_t s = dev->stream;
  if ((e = hipMemcpyAsync(g->params.ptr, p.data(), 4ull * kGuideParams, hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
  (void)hipMemsetAsync(g->m.ptr, 0, 4ull * kGuideParams, s);
  (void)hipMemsetAsync(g->v.ptr, 0, 4ull * kGuideParams, s);
  (void)hipMemsetAsync(g->grad.ptr, 0, 4ull * kGuideParams, s);
  (void)hipMemsetAsync(g->rec_count.ptr, 0, 4, s);
  hipLaunchKernelGGL(k_guide_publish, dim3((kGuideParams + 255) / 256), dim3(256), 0, s, g->params.as<float>(),
                     g->wbf.as<uint16_t>(), g->bias.as<float>());
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e);
  dev->guides.push_back(g);
  dev->guide_orphan = [](void *p) {  // device destroyed first: free the device memory, keep the handle
    lgpu_guide gg = static_cast<lgpu_guide>(p);
    gg->params.release(); gg->m.release(); gg->v.release(); gg->grad.release(); gg->wbf.release();
    gg->bias.release(); gg->loss.release(); gg->sumw.release(); gg->idx.release(); gg->act.release();
    gg->rec.release(); gg->rec_count.release();
    gg->dev = nullptr;
  };
  *out = g;
  return LGPU_OK;
}

extern "C" lgpu_status lgpu_guide_destroy(lgpu_guide g) {
  if (!g) return LGPU_OK;
  if (lgpu_device dev = g->dev) {
    std::lock_guard<std::mutex> lk(dev->mu);
    (void)hipSetDevice(dev->device);
    (void)hipStreamSynchronize(dev->stream);
    dev->guides.erase(std::remove(dev->guides.begin(), dev->guides.end(), static_cast<void *>(g)), dev->guides.end());
  }
  delete g;
  return LGPU_OK;
}

extern "C" lgpu_status lgpu_guide_get_params(lgpu_guide g, float *params) {
  if (!g || !params) return Fail(LGPU_ERR_INVALID_ARGUMENT, "guide_get_params: null argument");
  if (!g->dev) return Fail(LGPU_ERR_STATE, "guide_get_params: the guide's device was destroyed");
  std::lock_guard<std::mutex> lk(g->dev->mu);
  LGPU_HIP(hipSetDevice(g->dev->device));
  LGPU_HIP(hipStreamSynchronize(g->dev->stream));
  LGPU_HIP(hipMemcpy(params, g->params.ptr, 4ull * kGuideParams, hipMemcpyDeviceToHost));
  return LGPU_OK;
}

extern "C" lgpu_status lgpu_guide_set_params(lgpu_guide g, const float *params) {
  if (!g || !params) return Fail(LGPU_ERR_INVALID_ARGUMENT, "guide_set_params: null argument");
  if (!g->dev) return Fail(LGPU_ERR_STATE, "guide_set_params: the guide's device was destroyed");
  for (int i = 0; i < kGuideParams; ++i)
    if (!std::isfinite(params[i])) return Fail(LGPU_ERR_INVALID_ARGUMENT, "guide_set_params: non-finite parameter");
  std::lock_guard<std::mutex> lk(g->dev->mu);
  LGPU_HIP(hipSetDevice(g->dev->device));
  hipStream_t s = g->dev->stream;
  LGPU_HIP(hipMemcpyAsync(g->params.ptr, params, 4ull * kGuideParams, hipMemcpyHostToDevice, s));
  LGPU_HIP(hipMemsetAsync(g->m.ptr, 0, 4ull * kGuideParams, s));
  LGPU_HIP(hipMemsetAsync(g->v.ptr, 0, 4ull * kGuideParams, s));
  LGPU_HIP(hipMemsetAsync(g->grad.ptr, 0, 4ull * kGuideParams, s));
  hipLaunchKernelGGL(k_guide_publish, dim3((kGuideParams + 255) / 256), dim3(256), 0, s, g->params.as<float>(),
                     g->wbf.as<uint16_t>(), g->bias.as<float>());
  LGPU_HIP(hipGetLastError());
  LGPU_HIP(hipStreamSynchronize(s));
  g->steps = 0;
  return LGPU_OK;
}

extern "C" lgpu_status lgpu_guide_forward(lgpu_guide g, uint32_t n, const float *inputs, float *out) {
  if (!g || (n && (!inputs || !out))) return Fail(LGPU_ERR_INVALID_ARGUMENT, "guide_forward: null argument");
  if (!g->dev) return Fail(LGPU_ERR_STATE, "guide_forward: the guide's device was destroyed");
  if (!n) return LGPU_OK;
  std::lock_guard<std::mutex> lk(g->dev->mu);
  LGPU_HIP(hipSetDevice(g->dev->device));
  hipStream_t s = g->dev->stream;
  DevBuf din, dbins, dhdr;
  LGPU_HIP(din.upload(inputs, 4ull * kGuideScalars * n));
  LGPU_HIP(dbins.alloc(128ull * n));
  LGPU_HIP(dhdr.alloc(8ull * n));
  GuideQuerySrc src{};
  src.list = nullptr;
  src.scalars = din.as<float>();
  constexpr uint32_t kPerBlock = (kInferBlock / 64) * 16 * kInferPT;
  const uint32_t blocks = std::min<uint32_t>((n + kPerBlock - 1) / kPerBlock, g->infer_grid);
  hipLaunchKernelGGL(k_guide_infer<false>, dim3(blocks), dim3(kInferBlock), 0, s, g->wbf.as<uint16_t>(), g->bias.as<float>(), src, n,
                     nullptr, dbins.as<uint32_t>(), dhdr.as<float2>());
  const hipError_t le = hipGetLastError();
  const hipError_t se = hipStreamSynchronize(s);
  LGPU_HIP(le);
  LGPU_HIP(se);
  std::vector<uint8_t> b(128ull * n);
  std::vector<float> h(2ull * n);
  LGPU_HIP(hipMemcpy(b.data(), dbins.ptr, b.size(), hipMemcpyDeviceToHost));
  LGPU_HIP(hipMemcpy(h.data(), dhdr.ptr, 4 * h.size(), hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < n; ++i) {
    for (int k = 0; k < kGuideBins; ++k) out[(size_t)i * (kGuideBins + 1) + k] = (float)b[(size_t)i * 128 + k] / h[2 * i + 1];
    out[(size_t)i * (kGuideBins + 1) + kGuideBins] = h[2 * i];
  }
  return LGPU_OK;
}

extern "C" lgpu_status lgpu_guide_gradient(lgpu_guide g, uint32_t n, const float *records, float *grad, float *loss) {
  if (!g || !grad || (n && !records)) return Fail(LGPU_ERR_INVALID_ARGUMENT, "guide_gradient: null argument");
  if (!g->dev) return Fail(LGPU_ERR_STATE, "guide_gradient: the guide's device was destroyed");
  std::lock_guard<std::mutex> lk(g->dev->mu);
  LGPU_HIP(hipSetDevice(g->dev->device));
  hipStream_t s = g->dev->stream;
  std::vector<GuideRecord> rv(std::max<uint32_t>(n, 1));
  std::vector<uint32_t> iv(std::max<uint32_t>(n, 1));
  float sw = 0.f;
  for (uint32_t i = 0; i < n; ++i) {
    const float *r = records + 8ull * i;
    const uint32_t bin = (uint32_t)r[5];
    if (bin >= (uint32_t)kGuideBins || !(r[6] >= 0.f) || !(r[7] >= 0.f))
      return Fail(LGPU_ERR_INVALID_ARGUMENT, "guide_gradient: record with a bad bin, pdf or weight");
    rv[i].a = make_float4(r[0], r[1], r[2], r[3]);
    rv[i].b = make_float4(r[4], AsF32(bin), r[6], 0.f);
    rv[i].c = make_float4(0.f, 1.f, r[7], 0.f);
    iv[i] = i;
    sw += r[7];
  }
  lgpu_status st = GuideEnsureScratch(g, n);
  if (st != LGPU_OK) return st;
  DevBuf drec;
  LGPU_HIP(drec.upload(rv.data(), sizeof(GuideRecord) * rv.size()));
  LGPU_HIP(hipMemcpyAsync(g->idx.ptr, iv.data(), 4ull * iv.size(), hipMemcpyHostToDevice, s));
  LGPU_HIP(hipMemcpyAsync(g->sumw.ptr, &sw, 4, hipMemcpyHostToDevice, s));
  LGPU_HIP(hipMemsetAsync(g->grad.ptr, 0, 4ull * kGuideParams, s));
  LGPU_HIP(hipMemsetAsync(g->loss.ptr, 0, 4, s));
  st = GuideGradient(g, n, drec.as<GuideRecord>(), s);
  const hipError_t se = hipStreamSynchronize(s);
  if (st != LGPU_OK) return st;
  LGPU_HIP(se);
  LGPU_HIP(hipMemcpy(grad, g->grad.ptr, 4ull * kGuideParams, hipMemcpyDeviceToHost));
  if (loss) LGPU_HIP(hipMemcpy(loss, g->loss.ptr, 4, hipMemcpyDeviceToHost));
  LGPU_HIP(hipMemsetAsync(g->grad.ptr, 0, 4ull * kGuideParams, s));
  LGPU_HIP(hipStreamSynchronize(s));
  return LGPU_OK;
}

extern "C" lgpu_status lgpu_guide_train(lgpu_guide g, uint32_t steps, float *loss) {
  if (!g) return Fail(LGPU_ERR_INVALID_ARGUMENT, "guide_train: null guide");
  if (!g->dev) return Fail(LGPU_ERR_STATE, "guide_train: the guide's device was destroyed");
  std::lock_guard<std::mutex> lk(g->dev->mu);
  LGPU_HIP(hipSetDevice(g->dev->device));
  if (!g->last_records) return Fail(LGPU_ERR_STATE, "guide_train: no training records (render with the guide first)");
  lgpu_status st = GuideTrainSteps(g, steps, g->dev->stream);
  if (st != LGPU_OK) return st;
  if (loss) *loss = g->last_loss;
  return LGPU_OK;
}

extern "C" lgpu_status lgpu_guide_stats(lgpu_guide g, uint32_t *records, uint32_t *steps, float *last_loss) {
  if (!g) return Fail(LGPU_ERR_INVALID_ARGUMENT, "guide_stats: null guide");
  if (records) *records = g->last_records;
  if (steps) *steps = g->steps;
  if (last_loss) *last_loss = g->last_loss;
  return LGPU_OK;
}
