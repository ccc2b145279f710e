// mlp.hip — learned warm start: the 16-64-64-64-3 MLP of mlp.h on CDNA MFMA.
//
// Formulation.  Every layer is evaluated transposed, Z^T = W . H^T + b, with the
// neurons on the MFMA M axis and 16 contacts on the N axis of a 16x16x16 bf16
// MFMA.  For that instruction lane l holds A[l%16][4(l/16)+i], B[4(l/16)+i][l%16]
// and D[4(l/16)+i][l%16] (i = 0..3), so the D fragment of one layer's 16-neuron
// tile is exactly the B fragment of the next layer's 16-wide K step: activations
// never leave registers.  Weights are pre-packed in fragment order (8 bytes per
// lane, one coalesced load per fragment).  The backward pass reuses the same
// trick with W^T fragments; weight gradients contract over contacts, for which
// the activations are transposed once through LDS and stored fragment-ready.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>

#include "mlp.h"

namespace jbg {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));

constexpr int kTileFrags = 26;  // per 16-sample training tile: H0 H1x4 H2x4 H3x4 dZ1x4 dZ2x4 dZ3x4 dZ4

// fp32 -> bf16, round to nearest even (Inf keeps its sign, NaN stays NaN).  gfx950 converts in hardware
// (v_cvt_pk_bf16_f32, two values per instruction); a software rounding costs ~8 VALU operations per value,
// which made the inference kernel VALU bound.
__device__ inline uint16_t toBf16(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }

typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));

// two values per conversion instruction, already in fragment order
__device__ inline bf16x4 pack4(float a, float b, float c, float d) {
  const bf16x2v lo = bf16x2v{(__bf16)a, (__bf16)b}, hi = bf16x2v{(__bf16)c, (__bf16)d};
  const uint2 u = uint2{__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi)};
  return __builtin_bit_cast(bf16x4, u);
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ inline bf16x4 frag(const uint16_t* __restrict__ P, int f, int lane) {
  return *reinterpret_cast<const bf16x4*>(P + ((size_t)f * 64 + lane) * 4);
}

// fragment f, lane l, element i of the packed operands (see file header)
__host__ __device__ inline float packedSource(const float* W, int f, int l, int i) {
  const int row = l & 15, k = 4 * (l >> 4) + i;
  if (f < 4) return W[kMlpW1 + (16 * f + row) * kMlpIn + k];  // W1 tile mt = f
  if (f < 20) {
    int mt = (f - 4) >> 2, kt = (f - 4) & 3;
    return W[kMlpW2 + (16 * mt + row) * kMlpHid + 16 * kt + k];
  }
  if (f < 36) {
    int mt = (f - 20) >> 2, kt = (f - 20) & 3;
    return W[kMlpW3 + (16 * mt + row) * kMlpHid + 16 * kt + k];
  }
  if (f < 40) {  // W4 padded to 16 rows, K step kt = f - 36
    int kt = f - 36;
    return row < kMlpOut ? W[kMlpW4 + row * kMlpHid + 16 * kt + k] : 0.0f;
  }
  if (f < 44) {  // W4^T: M = 64 inputs of layer 4, K = 16 (3 outputs used)
    int mt = f - 40;
    return k < kMlpOut ? W[kMlpW4 + k * kMlpHid + 16 * mt + row] : 0.0f;
  }
  if (f < 60) {  // W3^T
    int mt = (f - 44) >> 2, kt = (f - 44) & 3;
    return W[kMlpW3 + (16 * kt + k) * kMlpHid + 16 * mt + row];
  }
  int mt = (f - 60) >> 2, kt = (f - 60) & 3;  // W2^T
  return W[kMlpW2 + (16 * kt + k) * kMlpHid + 16 * mt + row];
}

// Inference fragment f (K = 32 step), lane l, element i.  The B operand of K step j is the concatenation of the
// D fragments of the previous layer's neuron tiles 2j and 2j+1, so lane group g = l/16 element i stands for
// neuron 32j + 4g + i (i < 4) or 32j + 16 + 4g + (i - 4) (i >= 4); the A operand (weights) uses the same
// neuron for the same (g, i).  The product pairs A and B elements by (g, i) whatever K order the hardware
// gives them, so no further layout is assumed.
__host__ __device__ inline float packedSource32(const float* W, int f, int l, int i) {
  const int row = l & 15, g = l >> 4;
  const int kk = i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4);
  if (f < 8) return W[kMlpW2 + (16 * (f >> 1) + row) * kMlpHid + 32 * (f & 1) + kk];
  if (f < 16) return W[kMlpW3 + (16 * ((f - 8) >> 1) + row) * kMlpHid + 32 * (f & 1) + kk];
  return row < kMlpOut ? W[kMlpW4 + row * kMlpHid + 32 * (f - 16) + kk] : 0.0f;  // W4 padded to 16 rows
}

__global__ void kMlpPack(const float* __restrict__ W, uint16_t* __restrict__ P) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < kMlpFrags * 256) {
    int f = t >> 8, l = (t >> 2) & 63, i = t & 3;
    P[t] = toBf16(packedSource(W, f, l, i));
  } else if (t < kMlpPackedWords) {
    int u = t - kMlpFrags * 256;
    int f = u >> 9, l = (u >> 3) & 63, i = u & 7;
    P[t] = toBf16(packedSource32(W, f, l, i));
  }
}

__device__ inline f32x4 biasInit(const float* __restrict__ b, int mt, int lane) {
  const float* p = b + 16 * mt + 4 * (lane >> 4);
  return f32x4{p[0], p[1], p[2], p[3]};
}

// max(z, 0) as a signed-integer max of the bit patterns (negative floats are negative integers): one
// instruction, where fmaxf first canonicalises its operand; -0 becomes +0
__device__ inline float relu1(float z) { return __builtin_bit_cast(float, max(__builtin_bit_cast(int, z), 0)); }
__device__ inline f32x4 relu(f32x4 z) { return f32x4{relu1(z.x), relu1(z.y), relu1(z.z), relu1(z.w)}; }
__device__ inline bf16x4 toB(f32x4 v) { return pack4(v.x, v.y, v.z, v.w); }
// bf16(max(v, 0)) as max(bf16(v), 0) on the 16-bit patterns, two per instruction: rounding keeps the sign, so
// the two agree (up to the sign of zero).  The empty asm keeps the backend from folding the 16-bit max into
// the conversions, which splits each paired conversion in two plus a permute.
__device__ inline bf16x4 toBRelu(f32x4 v) {
  uint2 u = __builtin_bit_cast(uint2, toB(v));
  asm("" : "+v"(u.x), "+v"(u.y));
  return __builtin_bit_cast(bf16x4, __builtin_elementwise_max(__builtin_bit_cast(s16x4, u), s16x4{0, 0, 0, 0}));
}

struct Acts {  // per-lane state of one 16-contact tile (D layout)
  bf16x4 h0, h1[4], h2[4], h3[4];
  f32x4 z1[4], z2[4], z3[4];
};

// forward of one tile whose input fragment is h0; returns the normalised outputs (rows 0..2 of lanes 0..15)
template <bool Keep>
__device__ inline f32x4 forward(const uint16_t* __restrict__ P, const float* __restrict__ W, bf16x4 h0, int lane,
                                Acts* a) {
  f32x4 z1[4], z2[4], z3[4];
  bf16x4 h1[4], h2[4], h3[4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    z1[mt] = mfma(frag(P, mt, lane), h0, biasInit(W + kMlpB1, mt, lane));
    h1[mt] = toB(relu(z1[mt]));
  }
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    f32x4 acc = biasInit(W + kMlpB2, mt, lane);
#pragma unroll
    for (int kt = 0; kt < 4; kt++) acc = mfma(frag(P, 4 + 4 * mt + kt, lane), h1[kt], acc);
    z2[mt] = acc;
    h2[mt] = toB(relu(acc));
  }
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    f32x4 acc = biasInit(W + kMlpB3, mt, lane);
#pragma unroll
    for (int kt = 0; kt < 4; kt++) acc = mfma(frag(P, 20 + 4 * mt + kt, lane), h2[kt], acc);
    z3[mt] = acc;
    h3[mt] = toB(relu(acc));
  }
  f32x4 y = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  if (lane < 16) y = f32x4{W[kMlpB4], W[kMlpB4 + 1], W[kMlpB4 + 2], 0.0f};
#pragma unroll
  for (int kt = 0; kt < 4; kt++) y = mfma(frag(P, 36 + kt, lane), h3[kt], y);
  if (Keep) {
    a->h0 = h0;
#pragma unroll
    for (int m = 0; m < 4; m++) {
      a->h1[m] = h1[m];
      a->h2[m] = h2[m];
      a->h3[m] = h3[m];
      a->z1[m] = z1[m];
      a->z2[m] = z2[m];
      a->z3[m] = z3[m];
    }
  }
  return y;
}

// inference operands of one lane, loaded once per wave: the layer-1 fragments (K = 16), the K = 32 fragments of
// layers 2-4 (80 VGPRs in all) and the bias words of the lane's output rows
struct InferOps {
  bf16x4 w1[4];
  bf16x8 w[kMlpInferFrags];
  f32x4 b[12];  // layers 1-3, tile mt of layer L at 4(L-1) + mt
  f32x4 b4;
};

__device__ inline void loadInferOps(const uint16_t* __restrict__ P, const float* __restrict__ W, int lane,
                                    InferOps& o) {
#pragma unroll
  for (int f = 0; f < 4; f++) o.w1[f] = frag(P, f, lane);
  const uint16_t* P32 = P + kMlpFrags * 256;
#pragma unroll
  for (int f = 0; f < kMlpInferFrags; f++) o.w[f] = *reinterpret_cast<const bf16x8*>(P32 + ((size_t)f * 64 + lane) * 8);
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    o.b[mt] = biasInit(W + kMlpB1, mt, lane);
    o.b[4 + mt] = biasInit(W + kMlpB2, mt, lane);
    o.b[8 + mt] = biasInit(W + kMlpB3, mt, lane);
  }
  o.b4 = lane < 16 ? f32x4{W[kMlpB4], W[kMlpB4 + 1], W[kMlpB4 + 2], 0.0f} : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
}

__device__ inline f32x4 mfma32(bf16x8 a, bf16x4 lo, bf16x4 hi, f32x4 c) {
  const bf16x8 b = __builtin_shufflevector(__builtin_bit_cast(bf16x4v, lo), __builtin_bit_cast(bf16x4v, hi), 0, 1, 2,
                                           3, 4, 5, 6, 7);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// forward of one 16-contact tile from register operands: the products of forward() (fp32 accumulation of the
// same bf16 operands, in another order) with the K = 32 instruction for layers 2-4
__device__ inline f32x4 forwardRegs(const InferOps& o, bf16x4 h0) {
  bf16x4 h1[4], h2[4], h3[4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++) h1[mt] = toBRelu(mfma(o.w1[mt], h0, o.b[mt]));
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    f32x4 acc = mfma32(o.w[2 * mt], h1[0], h1[1], o.b[4 + mt]);
    h2[mt] = toBRelu(mfma32(o.w[2 * mt + 1], h1[2], h1[3], acc));
  }
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    f32x4 acc = mfma32(o.w[8 + 2 * mt], h2[0], h2[1], o.b[8 + mt]);
    h3[mt] = toBRelu(mfma32(o.w[8 + 2 * mt + 1], h2[2], h2[3], acc));
  }
  return mfma32(o.w[17], h3[2], h3[3], mfma32(o.w[16], h3[0], h3[1], o.b4));
}

// input fragment of contact c (lane holds features 4(l/16)..4(l/16)+3), normalised
__device__ inline bf16x4 inputFrag(const float* __restrict__ x, bool ok, const float* __restrict__ norm, int lane) {
  const int k = 4 * (lane >> 4);
  float4 v = ok ? *reinterpret_cast<const float4*>(x + k) : float4{0.0f, 0.0f, 0.0f, 0.0f};
  if (!ok) return pack4(0.0f, 0.0f, 0.0f, 0.0f);
  return pack4((v.x - norm[k]) * norm[16 + k], (v.y - norm[k + 1]) * norm[17 + k], (v.z - norm[k + 2]) * norm[18 + k],
               (v.w - norm[k + 3]) * norm[19 + k]);
}

constexpr int kInferAhead = 8;  // tiles whose features a wave has in flight (HBM latency is ~10 tiles of compute)

// features and slot header (validity bits 0-3, first dense row from bit 4) of one contact, loaded ahead of use
struct InferIn {
  float4 v;
  int hx;
};

// branchless (a load under a lane mask makes the compiler wait for every load in flight): contacts past n read
// contact n - 1 and are masked by the caller; features of every contact c < n are readable, slots without a
// point included
template <bool Fused>
__device__ inline InferIn inferLoad(int64_t c, int64_t n, const float* __restrict__ X, const int4* __restrict__ hdr,
                                    int k4) {
  const int64_t cc = c < n ? c : n - 1;
  InferIn r;
  r.v = *reinterpret_cast<const float4*>(X + 16 * cc + k4);
  r.hx = Fused ? hdr[cc >> 2].x : 15;
  return r;
}

// one tile of a wave: contacts 16t..16t+15 from a ring slot, which is refilled (Refill) with the features of the
// wave's tile kInferAhead tiles later before the products are computed
// NewOnly (fused): only contacts without cached impulses (features 12-14 all zero, e.g. points new this step)
// take the prediction; the others keep the scaled cached impulses already in their rows
template <bool Fused, bool Refill, bool NewOnly>
__device__ inline void inferTile(InferIn& slot, int64_t t, int64_t n, int64_t ahead, const InferOps& o,
                                 const float4& mean, const float4& scale, const float* __restrict__ X,
                                 const float* __restrict__ norm, float* __restrict__ Y, const int4* __restrict__ hdr,
                                 float* nrow, float* frow, int lane, int k4) {
  const int64_t c = 16 * t + (lane & 15);
  const int hx = slot.hx;
  const bool ok = c < n && ((hx >> (int)(c & 3)) & 1) != 0;
  // the slot is consumed before it is refilled, so the refill lands in the same registers; the invalid
  // contacts' inputs are cleared by a mask (a select here becomes a branch)
  const bf16x4 hv = pack4((slot.v.x - mean.x) * scale.x, (slot.v.y - mean.y) * scale.y,
                          (slot.v.z - mean.z) * scale.z, (slot.v.w - mean.w) * scale.w);
  const u32x2 keep = ok ? u32x2{~0u, ~0u} : u32x2{0u, 0u};
  const bf16x4 h0 = __builtin_bit_cast(bf16x4, __builtin_bit_cast(u32x2, hv) & keep);
  int fresh = 1;
  if (NewOnly) {  // lanes 48-63 hold features 12-15 of the tile's contacts: cached impulses, friction
    const int none = (slot.v.x == 0.0f && slot.v.y == 0.0f && slot.v.z == 0.0f) ? 1 : 0;
    fresh = __shfl(none, (lane & 15) + 48);
  }
  if (Refill) slot = inferLoad<Fused>(c + ahead, n, X, hdr, k4);
  const f32x4 y = forwardRegs(o, h0);
  if (lane < 16 && ok && fresh) {
    float y0 = y.x * norm[32] + norm[35], y1 = y.y * norm[33] + norm[36], y2 = y.z * norm[34] + norm[37];
    if (Fused) {  // dense rows: the slot's first row, then one row per valid point in point order
      const int j = (int)(c & 3);
      const size_t row = (size_t)((unsigned)hx >> 4) + __popc(hx & 15 & ((1 << j) - 1));
      nrow[kRowWords * row + 19] = y0;
      frow[kRowWords * (2 * row) + 19] = y1;
      frow[kRowWords * (2 * row + 1) + 19] = y2;
    } else {
      Y[3 * c] = y0;
      Y[3 * c + 1] = y1;
      Y[3 * c + 2] = y2;
    }
  }
}

// inference: 256-thread blocks, one 16-contact tile per wave and iteration, operands in registers, the
// features of the wave's next kInferAhead tiles in flight.  Fused: the rows of valid slots (hdr) receive the
// predictions (warm start); otherwise y[n][3] is written.
template <bool Fused, bool NewOnly = false>
__global__ __launch_bounds__(256) void kMlpInfer(int64_t n, const float* __restrict__ X, const uint16_t* __restrict__ P,
                                                 const float* __restrict__ W, const float* __restrict__ norm,
                                                 float* __restrict__ Y, const int4* __restrict__ hdr, float* nrow,
                                                 float* frow) {
  const int lane = threadIdx.x & 63;
  // tile indices are wave-uniform (scalar registers): the loop exits are scalar branches, so the compiler
  // keeps the prefetch ring in place instead of copying it under a lane mask
  const int wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
  const int64_t ntiles = (n + 15) >> 4;
  if (wave >= ntiles) return;
  const int k4 = 4 * (lane >> 4);
  const int lc = lane & 15;
  InferOps o;
  loadInferOps(P, W, lane, o);
  const float4 mean = *reinterpret_cast<const float4*>(norm + k4);
  const float4 scale = *reinterpret_cast<const float4*>(norm + 16 + k4);
  // the once-per-wave operand loads complete before the ring is issued, so the loop's waits count ring loads
  // only (the empty asm pins the normalisation loads here; invariant loads otherwise sink to their first use)
  asm volatile("" ::"v"(mean.x), "v"(mean.y), "v"(mean.z), "v"(mean.w), "v"(scale.x), "v"(scale.y), "v"(scale.z),
               "v"(scale.w));
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_sched_barrier(0);
  InferIn ring[kInferAhead];
#pragma unroll
  for (int i = 0; i < kInferAhead; i++) {
    ring[i] = inferLoad<Fused>(16 * ((int64_t)wave + i * nwaves) + lc, n, X, hdr, k4);
    // issue order = slot order: the loop's wait for slot i then counts the loads issued after it (the wait
    // instruction is static, so a slot 0 issued last would cost a full drain every iteration)
    __builtin_amdgcn_sched_barrier(0);
  }
  // the wave's tiles are wave, wave + nwaves, ...: whole groups of kInferAhead refill their slots; the last,
  // partial group is a separate copy without refills, so no loop path re-enters with a fresh load in slot 0
  const int64_t tilesOfWave = (ntiles - wave + nwaves - 1) / nwaves;
  const int64_t ahead = 16 * (int64_t)kInferAhead * nwaves;
  int64_t t = wave;
  for (int64_t g = tilesOfWave / kInferAhead; g > 0; g--) {
#pragma unroll
    for (int i = 0; i < kInferAhead; i++, t += nwaves)
      inferTile<Fused, true, NewOnly>(ring[i], t, n, ahead, o, mean, scale, X, norm, Y, hdr, nrow, frow, lane, k4);
  }
  const int rest = (int)(tilesOfWave % kInferAhead);
#pragma unroll
  for (int i = 0; i < kInferAhead - 1; i++, t += nwaves) {
    if (i >= rest) break;
    inferTile<Fused, false, NewOnly>(ring[i], t, n, ahead, o, mean, scale, X, norm, Y, hdr, nrow, frow, lane, k4);
  }
}

// D-layout tile(s) M[c = l%16][16*mt + 4(l/16) + i] -> fragment-ready F[nt][lane][i] = M[4(l/16)+i][16*nt + l%16]
template <int NT>
__device__ inline void storeTransposed(uint16_t* lds, const bf16x4* m, uint16_t* __restrict__ out, int lane) {
  constexpr int Wd = 16 * NT;
#pragma unroll
  for (int mt = 0; mt < NT; mt++) {
    uint16_t* row = lds + (lane & 15) * Wd + 16 * mt + 4 * (lane >> 4);
    row[0] = (uint16_t)m[mt].x;
    row[1] = (uint16_t)m[mt].y;
    row[2] = (uint16_t)m[mt].z;
    row[3] = (uint16_t)m[mt].w;
  }
  __syncthreads();
#pragma unroll
  for (int nt = 0; nt < NT; nt++) {
    bf16x4 v;
    const int c0 = 4 * (lane >> 4), col = 16 * nt + (lane & 15);
    v.x = (short)lds[(c0 + 0) * Wd + col];
    v.y = (short)lds[(c0 + 1) * Wd + col];
    v.z = (short)lds[(c0 + 2) * Wd + col];
    v.w = (short)lds[(c0 + 3) * Wd + col];
    *reinterpret_cast<bf16x4*>(out + ((size_t)nt * 64 + lane) * 4) = v;
  }
  __syncthreads();
}

__device__ inline f32x4 reluGrad(f32x4 g, f32x4 z) {
  return f32x4{z.x > 0.0f ? g.x : 0.0f, z.y > 0.0f ? g.y : 0.0f, z.z > 0.0f ? g.z : 0.0f, z.w > 0.0f ? g.w : 0.0f};
}

// forward + backward of one 16-sample tile per block (64 threads); per-tile activations and
// pre-activation gradients stored fragment-ready for the weight-gradient contraction
__global__ __launch_bounds__(64) void kMlpFwdBwd(int64_t nb, const int* __restrict__ idx, const float* __restrict__ X,
                                                 const float* __restrict__ Yt, const float* __restrict__ norm,
                                                 const uint16_t* __restrict__ P, const float* __restrict__ W,
                                                 uint16_t* __restrict__ F, float* lossAcc, float gscale) {
  __shared__ uint16_t lds[16 * 64];
  const int lane = threadIdx.x;
  const int64_t ntiles = (nb + 15) >> 4;
  float lossLane = 0.0f;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t r = t * 16 + (lane & 15);
    const bool ok = r < nb;
    const int64_t s = ok ? idx[r] : 0;
    Acts a;
    f32x4 y = forward<true>(P, W, inputFrag(X + 16 * s, ok, norm, lane), lane, &a);
    // dL/dy' for the mean squared error over nb x 3 normalised outputs
    f32x4 g = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    if (lane < 16 && ok) {
      const float* yt = Yt + 3 * s;
      float d0 = y.x - (yt[0] - norm[35]) / norm[32];
      float d1 = y.y - (yt[1] - norm[36]) / norm[33];
      float d2 = y.z - (yt[2] - norm[37]) / norm[34];
      lossLane += d0 * d0 + d1 * d1 + d2 * d2;
      g = f32x4{gscale * d0, gscale * d1, gscale * d2, 0.0f};
    }
    bf16x4 g4 = toB(g);
    bf16x4 g3[4], g2[4], g1[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) g3[mt] = toB(reluGrad(mfma(frag(P, 40 + mt, lane), g4, f32x4{0, 0, 0, 0}), a.z3[mt]));
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      f32x4 acc = f32x4{0, 0, 0, 0};
#pragma unroll
      for (int kt = 0; kt < 4; kt++) acc = mfma(frag(P, 44 + 4 * mt + kt, lane), g3[kt], acc);
      g2[mt] = toB(reluGrad(acc, a.z2[mt]));
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      f32x4 acc = f32x4{0, 0, 0, 0};
#pragma unroll
      for (int kt = 0; kt < 4; kt++) acc = mfma(frag(P, 60 + 4 * mt + kt, lane), g2[kt], acc);
      g1[mt] = toB(reluGrad(acc, a.z1[mt]));
    }
    uint16_t* out = F + (size_t)t * kTileFrags * 256;
    storeTransposed<1>(lds, &a.h0, out, lane);
    storeTransposed<4>(lds, a.h1, out + 1 * 256, lane);
    storeTransposed<4>(lds, a.h2, out + 5 * 256, lane);
    storeTransposed<4>(lds, a.h3, out + 9 * 256, lane);
    storeTransposed<4>(lds, g1, out + 13 * 256, lane);
    storeTransposed<4>(lds, g2, out + 17 * 256, lane);
    storeTransposed<4>(lds, g3, out + 21 * 256, lane);
    storeTransposed<1>(lds, &g4, out + 25 * 256, lane);
  }
  for (int o = 32; o > 0; o >>= 1) lossLane += __shfl_xor(lossLane, o);
  if (lane == 0) atomicAdd(lossAcc, lossLane);
}

// weight gradients dW = sum_c dZ[c]^T H[c] (contraction over the samples of the batch).
// One wave per (output tile, tile chunk): 40 output tiles over the four layers.
__global__ __launch_bounds__(64) void kMlpWgrad(int64_t ntiles, int chunk, const uint16_t* __restrict__ F,
                                                float* grad) {
  const int lane = threadIdx.x;
  const int item = blockIdx.x % 40;
  const int64_t t0 = (int64_t)(blockIdx.x / 40) * chunk;
  int dz, h, mt, nt, wOff, bOff, inDim, outDim;
  if (item < 4) {
    dz = 13; h = 0; mt = item; nt = 0; wOff = kMlpW1; bOff = kMlpB1; inDim = kMlpIn; outDim = kMlpHid;
  } else if (item < 20) {
    dz = 17; h = 1; mt = (item - 4) >> 2; nt = (item - 4) & 3; wOff = kMlpW2; bOff = kMlpB2; inDim = kMlpHid; outDim = kMlpHid;
  } else if (item < 36) {
    dz = 21; h = 5; mt = (item - 20) >> 2; nt = (item - 20) & 3; wOff = kMlpW3; bOff = kMlpB3; inDim = kMlpHid; outDim = kMlpHid;
  } else {
    dz = 25; h = 9; mt = 0; nt = item - 36; wOff = kMlpW4; bOff = kMlpB4; inDim = kMlpHid; outDim = kMlpOut;
  }
  const int dzFrag = dz + (dz == 25 ? 0 : mt), hFrag = h + (h == 0 ? 0 : nt);
  const bf16x4 ones = bf16x4{(short)0x3f80, (short)0x3f80, (short)0x3f80, (short)0x3f80};
  f32x4 acc = f32x4{0, 0, 0, 0}, accB = f32x4{0, 0, 0, 0};
  const int64_t t1 = t0 + chunk < ntiles ? t0 + chunk : ntiles;
  for (int64_t t = t0; t < t1; t++) {
    const uint16_t* base = F + (size_t)t * kTileFrags * 256;
    bf16x4 A = frag(base, dzFrag, lane);
    acc = mfma(A, frag(base, hFrag, lane), acc);
    if (nt == 0) accB = mfma(A, ones, accB);
  }
  const int col = 16 * nt + (lane & 15);
  const float av[4] = {acc.x, acc.y, acc.z, acc.w};
  const float bv[4] = {accB.x, accB.y, accB.z, accB.w};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int o = 16 * mt + 4 * (lane >> 4) + i;
    if (o >= outDim) continue;
    atomicAdd(grad + wOff + o * inDim + col, av[i]);
    if (nt == 0 && (lane & 15) == 0) atomicAdd(grad + bOff + o, bv[i]);
  }
}

__global__ void kMlpAdam(const float* __restrict__ g, float* W, float* m, float* v, float lr, float bc1, float bc2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  W[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ void kFillF(float* p, int n, float v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void ensureModel(Ctx& c) {
  c.mlpW.ensure(kMlpParams);
  c.mlpWb.ensure((size_t)kMlpPackedWords);
  c.mlpNorm.ensure(kMlpNormWords);
}

void repack(Ctx& c) {
  hipLaunchKernelGGL(kMlpPack, dim3(gridFor(kMlpPackedWords)), dim3(256), 0, c.stream, c.mlpW.p, c.mlpWb.p);
  JB_CHECK(hipGetLastError());
}

void checkNorm(const float* norm) {
  for (int k = 0; k < kMlpNormWords; k++)
    if (!std::isfinite(norm[k])) throw std::invalid_argument("normalisation words must be finite");
  for (int k = 0; k < kMlpOut; k++)
    if (!(norm[32 + k] > 0.0f)) throw std::invalid_argument("output scales must be positive");
}

int inferGrid(int64_t n) {  // about as many waves as the device keeps resident (two per SIMD), four per block
  int64_t tiles = (n + 15) / 16;
  int64_t blocks = (tiles + 3) / 4;
  return (int)std::max<int64_t>(1, std::min<int64_t>(blocks, 512));
}

}  // namespace

void mlpSetParams(Ctx& c, const float* params, const float* norm) {
  if (!params || !norm) throw std::invalid_argument("null parameters");
  checkNorm(norm);
  ensureModel(c);
  hipStream_t s = c.stream;
  JB_CHECK(hipMemcpyAsync(c.mlpW.p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, s));
  JB_CHECK(hipMemcpyAsync(c.mlpNorm.p, norm, kMlpNormWords * sizeof(float), hipMemcpyHostToDevice, s));
  repack(c);
  JB_CHECK(hipStreamSynchronize(s));
  c.mlpReady = true;
}

void mlpGetParams(Ctx& c, float* params, float* norm) {
  if (!c.mlpReady) throw std::logic_error("no model loaded or trained");
  hipStream_t s = c.stream;
  if (params) JB_CHECK(hipMemcpyAsync(params, c.mlpW.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, s));
  if (norm) JB_CHECK(hipMemcpyAsync(norm, c.mlpNorm.p, kMlpNormWords * sizeof(float), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
}

void mlpInfer(Ctx& c, const float* x, int64_t n, float* y) {
  if (!c.mlpReady) throw std::logic_error("no model loaded or trained");
  if (n < 0 || (n > 0 && (!x || !y))) throw std::invalid_argument("bad feature arrays");
  if (n == 0) return;
  hipStream_t s = c.stream;
  c.mlpX.ensure((size_t)16 * n);
  c.mlpY.ensure((size_t)3 * n);
  JB_CHECK(hipMemcpyAsync(c.mlpX.p, x, 16 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
  tBegin(c, T_MLP);
  hipLaunchKernelGGL(kMlpInfer<false>, dim3(inferGrid(n)), dim3(256), 0, s, n, c.mlpX.p, c.mlpWb.p, c.mlpW.p, c.mlpNorm.p,
                     c.mlpY.p, nullptr, nullptr, nullptr);
  JB_CHECK(hipGetLastError());
  tEnd(c, T_MLP);
  JB_CHECK(hipMemcpyAsync(y, c.mlpY.p, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
  if (c.timing) tCollect(c);  // outside a step: the kernel's timing region is read here
}

// ---- inference without device allocation ---------------------------------------------------------
// Staging of jbgpu_mlp_infer_static: features (16 floats) then predictions (3 floats) of at most
// kStaticInferMax contacts, 4.98 MB, held by every device the library runs on until the process ends.  In host
// code `scratch` names a host-side stand-in, not device memory: kernels and copies use the device address that
// hipGetSymbolAddress returns for the current device (kept per device in g_scratchDev).
static __device__ __align__(256) float scratch[kStaticInferMax * (kMlpIn + kMlpOut)];
namespace {
std::mutex g_scratchMu;
std::vector<float*> g_scratchDev;  // device address of `scratch` per device id (nullptr: not looked up yet)
}  // namespace

void mlpStaticInit() {
  int dev = 0;
  JB_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if ((int)g_scratchDev.size() <= dev) g_scratchDev.resize((size_t)dev + 1, nullptr);
  if (g_scratchDev[dev]) return;
  void* p = nullptr;
  JB_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratchDev[dev] = (float*)p;
}

void mlpInferStatic(Ctx& c, const float* x, int64_t n, float* y) {
  if (!c.mlpReady) throw std::logic_error("no model loaded or trained");
  if (n < 0 || (n > 0 && (!x || !y))) throw std::invalid_argument("bad feature arrays");
  if (n > kStaticInferMax)
    throw std::length_error("jbgpu_mlp_infer_static takes at most " + std::to_string(kStaticInferMax) +
                            " contacts per call");
  if (n == 0) return;
  float* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (c.device < (int)g_scratchDev.size()) base = g_scratchDev[c.device];
  }
  if (!base) throw std::logic_error("static inference buffer not set up on this device");
  float* dX = base;                                  // n x 16 features
  float* dY = base + (size_t)kMlpIn * kStaticInferMax;  // n x 3 predictions
  hipStream_t s = c.stream;
  JB_CHECK(hipMemcpyAsync(dX, x, (size_t)kMlpIn * n * sizeof(float), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(kMlpInfer<false>, dim3(inferGrid(n)), dim3(256), 0, s, n, (const float*)dX, c.mlpWb.p,
                     c.mlpW.p, c.mlpNorm.p, dY, nullptr, nullptr, nullptr);
  JB_CHECK(hipGetLastError());
  JB_CHECK(hipMemcpyAsync(y, dY, (size_t)kMlpOut * n * sizeof(float), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
}

namespace {
// gradient of the loss for the nb samples listed at idx, accumulated into c.mlpG (zeroed here)
void gradBatch(Ctx& c, const int* idx, int64_t nb, float gscale) {
  hipStream_t s = c.stream;
  const int64_t ntiles = (nb + 15) / 16;
  c.mlpFrag.ensure((size_t)ntiles * kTileFrags * 256);
  hipLaunchKernelGGL(kFillF, dim3(gridFor(kMlpParams)), dim3(256), 0, s, c.mlpG.p, kMlpParams, 0.0f);
  const int fb = (int)std::min<int64_t>(ntiles, 8192);
  hipLaunchKernelGGL(kMlpFwdBwd, dim3(fb), dim3(64), 0, s, nb, idx, c.mlpX.p, c.mlpY.p, c.mlpNorm.p, c.mlpWb.p,
                     c.mlpW.p, c.mlpFrag.p, c.mlpLoss.p, gscale);
  const int chunk = 32;
  const int64_t nchunks = (ntiles + chunk - 1) / chunk;
  hipLaunchKernelGGL(kMlpWgrad, dim3((unsigned)(40 * nchunks)), dim3(64), 0, s, ntiles, chunk, c.mlpFrag.p, c.mlpG.p);
  JB_CHECK(hipGetLastError());
}

void uploadData(Ctx& c, const float* x, const float* y, int64_t n) {
  hipStream_t s = c.stream;
  c.mlpX.ensure((size_t)16 * n);
  c.mlpY.ensure((size_t)3 * n);
  c.mlpIdx.ensure((size_t)n);
  c.mlpG.ensure(kMlpParams);
  c.mlpLoss.ensure(1);
  JB_CHECK(hipMemcpyAsync(c.mlpX.p, x, 16 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
  JB_CHECK(hipMemcpyAsync(c.mlpY.p, y, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
}
}  // namespace

void mlpGrad(Ctx& c, const float* x, const float* y, int64_t n, float* grad, float* loss) {
  if (!c.mlpReady) throw std::logic_error("no model loaded or trained");
  if (n <= 0 || !x || !y || !grad) throw std::invalid_argument("bad training arrays");
  if (n > (int64_t)1 << 30) throw std::length_error("too many samples");
  hipStream_t s = c.stream;
  uploadData(c, x, y, n);
  std::vector<int> idx(n);
  for (int64_t i = 0; i < n; i++) idx[i] = (int)i;
  JB_CHECK(hipMemcpyAsync(c.mlpIdx.p, idx.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(kFillF, dim3(1), dim3(64), 0, s, c.mlpLoss.p, 1, 0.0f);
  gradBatch(c, c.mlpIdx.p, n, 2.0f / (3.0f * (float)n));
  float l = 0;
  JB_CHECK(hipMemcpyAsync(grad, c.mlpG.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipMemcpyAsync(&l, c.mlpLoss.p, sizeof(float), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
  if (loss) *loss = l / (3.0f * (float)n);
}

void mlpTrain(Ctx& c, const float* x, const float* y, int64_t n, int epochs, int batch, float lr, uint64_t seed,
              float* lossPerEpoch) {
  if (n <= 0 || !x || !y) throw std::invalid_argument("bad training arrays");
  if (epochs < 1 || batch < 16 || !(lr > 0.0f)) throw std::invalid_argument("bad training settings");
  if (n > (int64_t)1 << 30) throw std::length_error("too many samples");
  hipStream_t s = c.stream;
  // normalisation from the data (double accumulation); scales floored to keep constant features finite
  std::vector<float> norm(kMlpNormWords);
  for (int k = 0; k < kMlpIn; k++) {
    double m = 0, q = 0;
    for (int64_t i = 0; i < n; i++) m += x[16 * i + k];
    m /= (double)n;
    for (int64_t i = 0; i < n; i++) q += (x[16 * i + k] - m) * (x[16 * i + k] - m);
    double sd = std::sqrt(q / (double)n);
    norm[k] = (float)m;
    norm[16 + k] = (float)(1.0 / std::max(sd, 1e-6));
  }
  for (int k = 0; k < kMlpOut; k++) {
    double m = 0, q = 0;
    for (int64_t i = 0; i < n; i++) m += y[3 * i + k];
    m /= (double)n;
    for (int64_t i = 0; i < n; i++) q += (y[3 * i + k] - m) * (y[3 * i + k] - m);
    norm[32 + k] = (float)std::max(std::sqrt(q / (double)n), 1e-6);
    norm[35 + k] = (float)m;
  }
  // He-uniform initial weights, zero biases, from a seeded generator
  std::mt19937_64 rng(seed);
  std::vector<float> W(kMlpParams, 0.0f);
  auto fill = [&](int off, int rows, int cols) {
    std::uniform_real_distribution<float> u(-std::sqrt(6.0f / cols), std::sqrt(6.0f / cols));
    for (int i = 0; i < rows * cols; i++) W[off + i] = u(rng);
  };
  fill(kMlpW1, kMlpHid, kMlpIn);
  fill(kMlpW2, kMlpHid, kMlpHid);
  fill(kMlpW3, kMlpHid, kMlpHid);
  fill(kMlpW4, kMlpOut, kMlpHid);
  for (int i = 0; i < kMlpOut * kMlpHid; i++) W[kMlpW4 + i] *= 0.1f;
  mlpSetParams(c, W.data(), norm.data());
  uploadData(c, x, y, n);
  c.mlpM.ensure(kMlpParams);
  c.mlpV.ensure(kMlpParams);
  hipLaunchKernelGGL(kFillF, dim3(gridFor(kMlpParams)), dim3(256), 0, s, c.mlpM.p, kMlpParams, 0.0f);
  hipLaunchKernelGGL(kFillF, dim3(gridFor(kMlpParams)), dim3(256), 0, s, c.mlpV.p, kMlpParams, 0.0f);
  std::vector<int> perm(n);
  for (int64_t i = 0; i < n; i++) perm[i] = (int)i;
  int64_t step = 0;
  for (int e = 0; e < epochs; e++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    JB_CHECK(hipMemcpyAsync(c.mlpIdx.p, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(kFillF, dim3(1), dim3(64), 0, s, c.mlpLoss.p, 1, 0.0f);
    for (int64_t b0 = 0; b0 < n; b0 += batch) {
      const int64_t nb = std::min<int64_t>(batch, n - b0);
      // the batch is the window [b0, b0 + nb) of the permutation
      gradBatch(c, c.mlpIdx.p + b0, nb, 2.0f / (3.0f * (float)nb));
      step++;
      const float bc1 = 1.0f - std::pow(0.9f, (float)step), bc2 = 1.0f - std::pow(0.999f, (float)step);
      hipLaunchKernelGGL(kMlpAdam, dim3(gridFor(kMlpParams)), dim3(256), 0, s, c.mlpG.p, c.mlpW.p, c.mlpM.p, c.mlpV.p,
                         lr, bc1, bc2);
      repack(c);
    }
    float l = 0;
    JB_CHECK(hipMemcpyAsync(&l, c.mlpLoss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipStreamSynchronize(s));
    if (lossPerEpoch) lossPerEpoch[e] = l / (3.0f * (float)n);
  }
  JB_CHECK(hipStreamSynchronize(s));
  c.mlpReady = true;
}

void mlpWarmStart(Ctx& c, int na) {
  if (!c.mlpReady) throw std::logic_error("no model loaded or trained");
  const int64_t n = 4 * (int64_t)na;
  tBegin(c, T_MLP);
  if (c.mlpNewOnly)
    hipLaunchKernelGGL((kMlpInfer<true, true>), dim3(inferGrid(n)), dim3(256), 0, c.stream, n, c.features.p, c.mlpWb.p,
                       c.mlpW.p, c.mlpNorm.p, nullptr, c.hdr.p, c.nrow.p, c.frow.p);
  else
    hipLaunchKernelGGL((kMlpInfer<true, false>), dim3(inferGrid(n)), dim3(256), 0, c.stream, n, c.features.p, c.mlpWb.p,
                       c.mlpW.p, c.mlpNorm.p, nullptr, c.hdr.p, c.nrow.p, c.frow.p);
  JB_CHECK(hipGetLastError());
  tEnd(c, T_MLP);
}

}  // namespace jbg
