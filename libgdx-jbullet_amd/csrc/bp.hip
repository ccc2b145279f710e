// bp.hip — broadphase kernel set for gfx950.
//
// Per proxy (one thread each): the world AABB with the contact threshold
// (collision/dispatch/CollisionWorld.java:195-245), the broadphase's setAabb
// semantics — DbvtBroadphase's stored-AABB fattening state machine
// (bp/DbvtBroadphase.java:196-228, bp/Dbvt.java:157-169), AxisSweep3
// quantisation (bp/AxisSweep3Internal.java:201-216) or raw boxes
// (bp/SimpleBroadphase.java) — and the stage rotation of DbvtBroadphase.collide
// (:96-113).  Pair discovery is a linear BVH over the stored boxes (Morton
// codes, radix sort, Karras hierarchy, atomic bottom-up refit, per-leaf stack
// traversal reporting only higher sorted leaves); boxes larger than 1e6 m (the
// StaticPlaneShape's +-1e30 box) stay out of the tree and are tested against
// every proxy of their world by a flat kernel.  Every reported pair passes the
// reference predicate exactly (closed-interval float test DbvtAabbMm.java:209-212,
// or the quantised edge test), the group/mask filter
// (HashedOverlappingPairCache.java:179-188) and the world-id filter, so the
// sorted (uid_lo, uid_hi) list equals the reference's overlapping-pair set.

#include "internal.h"
#include "shapes.h"

namespace jbg {

namespace {

constexpr int STAGECOUNT = 2;
constexpr float DBVT_BP_MARGIN = 0.05f;
constexpr float PREDICTED_FRAMES = 2.0f;
constexpr float HUGE_EXTENT = 1e6f;
constexpr float kLargeFactor = 32.0f;  // relative size that also sends a box to the flat huge-proxy test
constexpr int kLargeCap = 64;          // ... when no more than this many boxes qualify

struct BpParams {
  int mode;
  int stageCurrent;  // before rotation
  int newStage;
  float wmin[3], wmax[3], quant[3];
  uint32_t handleMask, qmask, sentinel;
  double pad[3];
};

__device__ __forceinline__ void quantizeDev(const BpParams& P, V3 p, int isMax, int* out) {
  V3 c = p;
  c.x = jmax(c.x, P.wmin[0]); c.y = jmax(c.y, P.wmin[1]); c.z = jmax(c.z, P.wmin[2]);
  c.x = jmin(c.x, P.wmax[0]); c.y = jmin(c.y, P.wmax[1]); c.z = jmin(c.z, P.wmax[2]);
  V3 v = v3((c.x - P.wmin[0]) * P.quant[0], (c.y - P.wmin[1]) * P.quant[1], (c.z - P.wmin[2]) * P.quant[2]);
  out[0] = (int)((((uint32_t)jint(v.x) & P.handleMask) | (uint32_t)isMax) & P.qmask);
  out[1] = (int)((((uint32_t)jint(v.y) & P.handleMask) | (uint32_t)isMax) & P.qmask);
  out[2] = (int)((((uint32_t)jint(v.z) & P.handleMask) | (uint32_t)isMax) & P.qmask);
}

// setAabb of the selected broadphase for proxy i, then the collide() stage rotation
__device__ void setAabbDev(const BpParams& P, int i, V3 mn, V3 mx, bool update, int* stage, float* leaf, float* aabb,
                           int* q) {
  if (update) {
    if (P.mode == JBGPU_BP_SAP16 || P.mode == JBGPU_BP_SAP32) {
      quantizeDev(P, mn, 0, q + 6 * i);
      quantizeDev(P, mx, 1, q + 6 * i + 3);
    } else if (P.mode == JBGPU_BP_DBVT) {
      float* L = leaf + 6 * i;
      float* S = aabb + 6 * i;
      int st = stage[i];
      if (st == STAGECOUNT) {
        st3(L, mn);
        st3(L + 3, mx);
      } else {
        V3 lmn = ld3(L), lmx = ld3(L + 3);
        bool inter = (lmn.x <= mx.x) && (lmx.x >= mn.x) && (lmn.y <= mx.y) && (lmx.y >= mn.y) && (lmn.z <= mx.z) &&
                     (lmx.z >= mn.z);
        if (inter) {
          V3 delta = scl(add(mn, mx), 0.5f);
          V3 c = scl(add(ld3(S), ld3(S + 3)), 0.5f);
          delta = sub(delta, c);
          delta = scl(delta, PREDICTED_FRAMES);
          bool contained = (lmn.x <= mn.x) && (lmn.y <= mn.y) && (lmn.z <= mn.z) && (lmx.x >= mx.x) &&
                           (lmx.y >= mx.y) && (lmx.z >= mx.z);
          if (!contained) {
            mn = sub(mn, v3(DBVT_BP_MARGIN, DBVT_BP_MARGIN, DBVT_BP_MARGIN));
            mx = add(mx, v3(DBVT_BP_MARGIN, DBVT_BP_MARGIN, DBVT_BP_MARGIN));
            if (delta.x > 0) mx.x += delta.x; else mn.x += delta.x;
            if (delta.y > 0) mx.y += delta.y; else mn.y += delta.y;
            if (delta.z > 0) mx.z += delta.z; else mn.z += delta.z;
            st3(L, mn);
            st3(L + 3, mx);
          }
        } else {
          st3(L, mn);
          st3(L + 3, mx);
        }
      }
      stage[i] = P.stageCurrent;
    }
    st3(aabb + 6 * i, mn);
    st3(aabb + 6 * i + 3, mx);
  }
  if (P.mode == JBGPU_BP_DBVT && stage[i] == P.newStage) {
    for (int k = 0; k < 6; k++) leaf[6 * i + k] = aabb[6 * i + k];
    stage[i] = STAGECOUNT;
  }
}

// candidate box of proxy i: its stored AABB (Dbvt, Simple) or dequantised box (AxisSweep3); the spare lanes carry
// its world id and filter word, so the leaf gather reads one 32-byte record per leaf
__device__ void writeCandidate(const BpParams& P, int i, const float* aabb, const int* q, float* cand, int* isHuge,
                               const int* __restrict__ world, const int* __restrict__ filt) {
  V3 mn = ld3(aabb + 6 * i), mx = ld3(aabb + 6 * i + 3);
  if (P.mode == JBGPU_BP_SAP16 || P.mode == JBGPU_BP_SAP32) {
    const int* Q = q + 6 * i;
    float r[6];
    for (int ax = 0; ax < 3; ax++) {
      double s = (double)P.quant[ax], w0 = (double)P.wmin[ax];
      r[ax] = (float)(w0 + (double)(uint32_t)Q[ax] / s - P.pad[ax]);
      r[3 + ax] = (float)(w0 + ((double)(uint32_t)Q[3 + ax] + 1.0) / s + P.pad[ax]);
    }
    mn = v3(r[0], r[1], r[2]);
    mx = v3(r[3], r[4], r[5]);
  }
  float* C = cand + 8 * i;
  C[0] = mn.x; C[1] = mn.y; C[2] = mn.z; C[3] = __int_as_float(world[i]);
  C[4] = mx.x; C[5] = mx.y; C[6] = mx.z; C[7] = __int_as_float(filt[i]);
  float e = fmaxf(fmaxf(mx.x - mn.x, mx.y - mn.y), mx.z - mn.z);
  isHuge[i] = !(e < HUGE_EXTENT) ? 1 : 0;
}

__global__ void kSetAabbBodies(BpParams P, int n, const float* __restrict__ shapes, const float* __restrict__ tr,
                               const float* __restrict__ props, int* act, int* stage, float* leaf, float* aabb, int* q,
                               float* cand, int* isHuge, const int* __restrict__ world,
                               const int* __restrict__ filt, const int* __restrict__ abortFlag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (abortFlag && *abortFlag)) return;
  int a = act[i];
  bool active = (a != JBGPU_ISLAND_SLEEPING) && (a != JBGPU_DISABLE_SIMULATION);
  bool update = false;
  V3 mn = v3(0, 0, 0), mx = v3(0, 0, 0);
  if (active) {
    const float* pr = props + 8 * i;
    Shape s = loadShape(shapes + JBGPU_SHAPE_WORDS * ldi(pr + 6));
    Tr t;
    for (int k = 0; k < 9; k++) t.b.m[k] = tr[12 * i + k];
    t.o = ld3(tr + 12 * i + 9);
    shapeAabb(s, t, mn, mx);
    V3 ct = v3(kContactThreshold, kContactThreshold, kContactThreshold);
    mn = sub(mn, ct);
    mx = add(mx, ct);
    bool isStatic = (ldi(pr + 7) & JBGPU_BODY_STATIC) != 0;
    if (isStatic || len2(sub(mx, mn)) < 1e12f) {
      update = true;
    } else if (a != JBGPU_DISABLE_DEACTIVATION && a != JBGPU_DISABLE_SIMULATION) {
      act[i] = JBGPU_DISABLE_SIMULATION;
    }
  }
  setAabbDev(P, i, mn, mx, update, stage, leaf, aabb, q);
  writeCandidate(P, i, aabb, q, cand, isHuge, world, filt);
}

__global__ void kSetAabbArray(BpParams P, int n, const float* __restrict__ in6, int* stage, float* leaf, float* aabb,
                              int* q, float* cand, int* isHuge, const int* __restrict__ world,
                              const int* __restrict__ filt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  setAabbDev(P, i, ld3(in6 + 6 * i), ld3(in6 + 6 * i + 3), true, stage, leaf, aabb, q);
  writeCandidate(P, i, aabb, q, cand, isHuge, world, filt);
}

// a pair query with no setAabb since the last one (e.g. right after createProxy): the stored boxes stay, the
// collide's stage rotation still runs (DbvtBroadphase.java:96-113), the candidate boxes come from the state
__global__ void kKeepAabbs(BpParams P, int n, int* stage, float* leaf, float* aabb, int* q, float* cand, int* isHuge,
                           const int* __restrict__ world, const int* __restrict__ filt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  setAabbDev(P, i, v3(0, 0, 0), v3(0, 0, 0), false, stage, leaf, aabb, q);
  writeCandidate(P, i, aabb, q, cand, isHuge, world, filt);
}

__device__ __forceinline__ unsigned int fOrd(float f) {  // order-preserving float -> uint
  unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ordF(unsigned int u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Per-step proxy statistics in two stages (per-block partials, then one block), with no atomics: a few
// addresses updated by every wave serialise.  Over the proxies that are not huge: the bounds of the candidate-box
// centres (order-preserving uint form) and the sum and count of the largest box extents.
constexpr int kStatBlocks = 256;
constexpr int kStatWords = 8;  // lo[3], hi[3] (uint), extent sum, count (float)

struct ProxyStat {
  unsigned int lo[3], hi[3];
  float esum, ecnt;
  __device__ void init() {
    for (int k = 0; k < 3; k++) {
      lo[k] = 0xffffffffu;
      hi[k] = 0u;
    }
    esum = 0.0f;
    ecnt = 0.0f;
  }
  __device__ void merge(const ProxyStat& o) {
    for (int k = 0; k < 3; k++) {
      lo[k] = min(lo[k], o.lo[k]);
      hi[k] = max(hi[k], o.hi[k]);
    }
    esum += o.esum;
    ecnt += o.ecnt;
  }
  __device__ void shfl(int off) {
    ProxyStat o;
    for (int k = 0; k < 3; k++) {
      o.lo[k] = (unsigned int)__shfl_xor((int)lo[k], off);
      o.hi[k] = (unsigned int)__shfl_xor((int)hi[k], off);
    }
    o.esum = __shfl_xor(esum, off);
    o.ecnt = __shfl_xor(ecnt, off);
    merge(o);
  }
  __device__ void store(unsigned int* w) const {
    for (int k = 0; k < 3; k++) {
      w[k] = lo[k];
      w[3 + k] = hi[k];
    }
    w[6] = __float_as_uint(esum);
    w[7] = __float_as_uint(ecnt);
  }
  __device__ void load(const unsigned int* w) {
    for (int k = 0; k < 3; k++) {
      lo[k] = w[k];
      hi[k] = w[3 + k];
    }
    esum = __uint_as_float(w[6]);
    ecnt = __uint_as_float(w[7]);
  }
};

// block-wide merge in a fixed order (waves by shuffles, then the block's waves in index order): deterministic
__device__ void blockMerge(ProxyStat& st, unsigned int* lds) {
  for (int off = 32; off > 0; off >>= 1) st.shfl(off);
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) st.store(lds + kStatWords * wave);
  __syncthreads();
  if (threadIdx.x == 0) {
    st.load(lds);
    for (int w = 1; w < nw; w++) {
      ProxyStat o;
      o.load(lds + kStatWords * w);
      st.merge(o);
    }
  }
}

__global__ __launch_bounds__(256) void kProxyStats(int n, const float* __restrict__ cand,
                                                   const int* __restrict__ isHuge, unsigned int* part) {
  __shared__ unsigned int lds[kStatWords * 4];
  ProxyStat st;
  st.init();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (isHuge[i]) continue;
    const float4 mn = ((const float4*)cand)[2 * i], mx = ((const float4*)cand)[2 * i + 1];
    const float c3[3] = {(mn.x + mx.x) * 0.5f, (mn.y + mx.y) * 0.5f, (mn.z + mx.z) * 0.5f};
    for (int k = 0; k < 3; k++)
      if (c3[k] == c3[k]) {
        st.lo[k] = min(st.lo[k], fOrd(c3[k]));
        st.hi[k] = max(st.hi[k], fOrd(c3[k]));
      }
    float e = fmaxf(fmaxf(mx.x - mn.x, mx.y - mn.y), mx.z - mn.z);
    if (!(e == e)) e = 0.0f;
    st.esum += e;
    st.ecnt += 1.0f;
  }
  blockMerge(st, lds);
  if (threadIdx.x == 0) st.store(part + kStatWords * blockIdx.x);
}

// bounds[0, 6): centre bounds; stats[0, 3): extent sum, count, and the large-box count (cleared)
__global__ __launch_bounds__(256) void kProxyStatsFinal(int nparts, const unsigned int* __restrict__ part,
                                                        unsigned int* bounds, float* stats) {
  __shared__ unsigned int lds[kStatWords * 4];
  ProxyStat st;
  st.init();
  if ((int)threadIdx.x < nparts) st.load(part + kStatWords * threadIdx.x);
  blockMerge(st, lds);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 3; k++) {
      bounds[k] = st.lo[k];
      bounds[3 + k] = st.hi[k];
    }
    stats[0] = st.esum;
    stats[1] = st.ecnt;
    stats[2] = 0.0f;
  }
}

// spreads the low 10 bits of v to every third bit
__device__ __forceinline__ uint32_t expand10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// sort key of proxy i: world id above a 30-bit Morton code of its candidate-box centre, so that every
// world's leaves are contiguous and its subtrees separate; huge proxies carry all ones and sort last
// Morton code of the candidate box centre with mb / 3 bits per axis, the world id above it (a world's leaves are
// consecutive in the order); huge proxies carry all ones and sort last
template <class K>
__global__ void kMorton(int n, const float* __restrict__ cand, const int* __restrict__ isHuge,
                        const int* __restrict__ world, const unsigned int* __restrict__ b, int mb, K* code,
                        int* idx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  idx[i] = i;
  if (isHuge[i]) {
    code[i] = (K)~0ull;
    return;
  }
  const float* C = cand + 8 * i;
  uint32_t q3[3];
  for (int k = 0; k < 3; k++) {
    float lo = ordF(b[k]), hi = ordF(b[3 + k]);
    float c = (C[k] + C[4 + k]) * 0.5f;
    float ext = hi - lo;
    float u = ext > 0.0f ? (c - lo) / ext : 0.0f;
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    if (!(u == u)) u = 0.0f;
    const uint32_t cells = 1u << (mb / 3);
    q3[k] = min((uint32_t)(u * (float)cells), cells - 1u);
  }
  const uint32_t m = (expand10(q3[0]) << 2) | (expand10(q3[1]) << 1) | expand10(q3[2]);
  code[i] = (K)(((uint64_t)(uint32_t)world[i] << mb) | m);
}

// sorted leaf boxes; the spare .w lanes carry the leaf's world id (int bits)
// leafExact[s]: the exact-test data of sorted leaf s in one 32-byte record: the stored AABB (Dbvt,
// Simple) or the quantised box (AxisSweep3) as 6 words, the group/mask word and the proxy index
// Packed: the candidate records carry the world id and filter word in their spare lanes (the broadphase's), and
// for the Dbvt and Simple modes their boxes are the stored AABBs bit for bit, so a leaf reads one 32-byte record;
// otherwise (the query tree) the world, filter and exact boxes come from their own arrays
template <bool Packed>
__global__ void kGatherLeaves(int n, const int* __restrict__ dn, const int* __restrict__ order, const float* __restrict__ cand,
                              const int* __restrict__ world, float* leafBox, int mode, const float* __restrict__ aabb,
                              const int* __restrict__ q, const int* __restrict__ filt, int4* __restrict__ leafExact) {
  if (dn) n = *dn;  // graph-captured steps: the count on the device
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int o = order[s];
  const bool sap = mode == JBGPU_BP_SAP16 || mode == JBGPU_BP_SAP32;
  const float4* src = (const float4*)(cand + 8 * o);
  float4 a = src[0], b = src[1];
  int fo;
  if (Packed) {
    fo = __float_as_int(b.w);
  } else {
    a.w = __int_as_float(world[o]);
    fo = filt[o];
  }
  if (Packed && !sap) {
    leafExact[2 * s] = make_int4(__float_as_int(a.x), __float_as_int(a.y), __float_as_int(a.z), __float_as_int(b.x));
    leafExact[2 * s + 1] = make_int4(__float_as_int(b.y), __float_as_int(b.z), fo, o);
  } else {
    const int* X = sap ? q + 6 * o : (const int*)aabb + 6 * o;
    leafExact[2 * s] = make_int4(X[0], X[1], X[2], X[3]);
    leafExact[2 * s + 1] = make_int4(X[4], X[5], fo, o);
  }
  b.w = a.w;
  float4* dst = (float4*)(leafBox + 8 * s);
  dst[0] = a;
  dst[1] = b;
}

// Morton bits of the sort key (a multiple of 3, at most 30, at least 12): as many as leave the key, world id
// included, within 31 bits (four 8-bit digit passes), whose top bit stays clear below the huge proxies' all-ones
static int mortonBits(int worldBits) { return std::min(30, std::max(12, 3 * ((31 - worldBits) / 3))); }
static int mortonKeyBits(int worldBits) { return worldBits + mortonBits(worldBits) + 1; }

// Morton codes of n proxies sorted with their indices.  The codes are 32 bits wide whenever the packed key (world
// id, Morton bits, huge marker) fits, which halves the key bytes of every sort pass and of the build's neighbour
// reads; otherwise they are 64 bits wide.  code/codeAlt hold n 64-bit words either way.
struct MortonOrder {
  const void* code;
  bool narrow;
  const int* order;
};

template <class K>
__device__ __forceinline__ int deltaK(const K* code, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const K a = code[i], b = code[j];
  if (a == b) return 8 * (int)sizeof(K) + __clz((unsigned)(i ^ j));
  return sizeof(K) == 8 ? __clzll((long long)(a ^ b)) : __clz((unsigned)(a ^ b));
}

// Karras (2012) hierarchy: internal node i in [0, n-2]; child < 0 encodes leaf -(child+1)
template <class K>
__global__ void kBuild(int n, const int* __restrict__ dn, const K* __restrict__ code, int* child, int* first, int* last) {
  if (dn) n = *dn;  // graph-captured steps: the count on the device
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = (deltaK(code, n, i, i + 1) - deltaK(code, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = deltaK(code, n, i, i - d);
  int lmax = 2;
  while (deltaK(code, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (deltaK(code, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = deltaK(code, n, i, j);
  int s = 0;
  int div = 2;
  int t = (l + div - 1) / div;
  while (true) {
    if (deltaK(code, n, i, i + (s + t) * d) > dnode) s += t;
    if (t == 1) break;
    div <<= 1;
    t = (l + div - 1) / div;
  }
  int gamma = i + s * d + min(d, 0);
  int lo = min(i, j), hi = max(i, j);
  int left = (lo == gamma) ? -(gamma + 1) : gamma;
  int right = (hi == gamma + 1) ? -(gamma + 2) : gamma + 1;
  child[2 * i] = left;
  child[2 * i + 1] = right;
  first[i] = lo;
  last[i] = hi;
}

// union of 32 consecutive boxes (8 floats each) per output box; the last group may be partial
__global__ __launch_bounds__(256) void kChunk(int n, const int* __restrict__ dn, const float* __restrict__ src, float* __restrict__ dst) {
  if (dn) n = *dn;  // graph-captured steps: the count on the device
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const float inf = __builtin_huge_valf();
  float4 mn = make_float4(inf, inf, inf, 0.0f), mx = make_float4(-inf, -inf, -inf, 0.0f);
  if (i < n) {
    mn = ((const float4*)src)[2 * i];
    mx = ((const float4*)src)[2 * i + 1];
  }
  for (int o = 16; o > 0; o >>= 1) {  // stays inside each 32-lane half of the wave
    mn.x = fminf(mn.x, __shfl_xor(mn.x, o));
    mn.y = fminf(mn.y, __shfl_xor(mn.y, o));
    mn.z = fminf(mn.z, __shfl_xor(mn.z, o));
    mx.x = fmaxf(mx.x, __shfl_xor(mx.x, o));
    mx.y = fmaxf(mx.y, __shfl_xor(mx.y, o));
    mx.z = fmaxf(mx.z, __shfl_xor(mx.z, o));
  }
  if ((i & 31) == 0 && i < n) {
    ((float4*)dst)[2 * (i >> 5)] = make_float4(mn.x, mn.y, mn.z, 0.0f);
    ((float4*)dst)[2 * (i >> 5) + 1] = make_float4(mx.x, mx.y, mx.z, 0.0f);
  }
}

struct BoxAcc {
  float4 mn, mx;
  __device__ void add(const float* B) {
    float4 a = ((const float4*)B)[0], b = ((const float4*)B)[1];
    mn.x = fminf(mn.x, a.x); mn.y = fminf(mn.y, a.y); mn.z = fminf(mn.z, a.z);
    mx.x = fmaxf(mx.x, b.x); mx.y = fmaxf(mx.y, b.y); mx.z = fmaxf(mx.z, b.z);
  }
};

// internal node boxes: the union of the node's sorted-leaf range [first, last], assembled from
// leaves at the ragged ends and from full 32-, 1024- and 32768-leaf chunk boxes in between
__global__ __launch_bounds__(256) void kNodeBoxes(int n, const int* __restrict__ dn, const int* __restrict__ first, const int* __restrict__ last,
                                                  const float* __restrict__ leafBox, const float* __restrict__ c1,
                                                  const float* __restrict__ c2, const float* __restrict__ c3,
                                                  float* __restrict__ nodeBox) {
  if (dn) n = *dn;  // graph-captured steps: the count on the device
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const float inf = __builtin_huge_valf();
  BoxAcc B{make_float4(inf, inf, inf, 0.0f), make_float4(-inf, -inf, -inf, 0.0f)};
  int a = first[i], e = last[i] + 1;
  const float* lv[4] = {leafBox, c1, c2, c3};
  for (int L = 0; L < 3; L++) {
    while (a < e && (a & 31)) B.add(lv[L] + 8 * (a++));
    while (e > a && (e & 31)) B.add(lv[L] + 8 * (--e));
    a >>= 5;
    e >>= 5;
  }
  while (a < e) B.add(c3 + 8 * (a++));
  // leaves are sorted by world first: the node's worlds are those of its first and last leaf
  const float wlo = leafBox[8 * first[i] + 3], whi = leafBox[8 * last[i] + 3];
  ((float4*)nodeBox)[2 * i] = make_float4(B.mn.x, B.mn.y, B.mn.z, wlo);
  ((float4*)nodeBox)[2 * i + 1] = make_float4(B.mx.x, B.mx.y, B.mx.z, whi);
}

// traversal record of internal node i: both children's boxes (.w lanes: world range; a leaf's world in
// both), then (child0, child1, last0, last1) where child < 0 is leaf -child-1 and last is the highest
// sorted leaf below the child; one visit reads the five float4 at once
__global__ __launch_bounds__(256) void kNodeRecords(int n, const int* __restrict__ dn, const int* __restrict__ child, const int* __restrict__ last,
                                                    const float* __restrict__ leafBox,
                                                    const float* __restrict__ nodeBox, float4* __restrict__ rec) {
  if (dn) n = *dn;  // graph-captured steps: the count on the device
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int ch[2], lst[2];
  for (int c = 0; c < 2; c++) {
    ch[c] = child[2 * i + c];
    const bool leaf = ch[c] < 0;
    const int k = leaf ? -ch[c] - 1 : ch[c];
    const float4* src = (const float4*)(leaf ? leafBox : nodeBox) + 2 * k;
    rec[kRecF4 * i + 2 * c] = src[0];
    rec[kRecF4 * i + 2 * c + 1] = src[1];
    lst[c] = leaf ? k : last[k];
  }
  rec[kRecF4 * i + 4] = make_float4(__int_as_float(ch[0]), __int_as_float(ch[1]), __int_as_float(lst[0]),
                                    __int_as_float(lst[1]));
}

// 4-wide record of internal node i from the binary records: the two children of each internal child and every
// leaf child itself (2 to 4 entries), so a traversal step tests up to four boxes behind one dependent load and
// the descent takes half as many steps
__global__ __launch_bounds__(256) void kNodeRecords4(int n, const int* __restrict__ dn, const float4* __restrict__ rec, float4* __restrict__ rec4) {
  if (dn) n = *dn;  // graph-captured steps: the count on the device
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const float4* R = rec + kRecF4 * i;
  const float4 m = R[4];
  int ch[4] = {0, 0, 0, 0}, lst[4] = {-1, -1, -1, -1};
  float4* out = rec4 + kRec4F4 * i;
  int e = 0;
  for (int c = 0; c < 2; c++) {
    const int id = __float_as_int(c == 0 ? m.x : m.y);
    if (id < 0) {  // leaf child: itself
      out[2 * e] = R[2 * c];
      out[2 * e + 1] = R[2 * c + 1];
      ch[e] = id;
      lst[e] = __float_as_int(c == 0 ? m.z : m.w);
      e++;
    } else {  // internal child: its two children
      const float4* Q = rec + kRecF4 * id;
      const float4 q = Q[4];
      for (int g = 0; g < 2; g++, e++) {
        out[2 * e] = Q[2 * g];
        out[2 * e + 1] = Q[2 * g + 1];
        ch[e] = __float_as_int(g == 0 ? q.x : q.y);
        lst[e] = __float_as_int(g == 0 ? q.z : q.w);
      }
    }
  }
  for (; e < 4; e++) {
    out[2 * e] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    out[2 * e + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  ((int4*)out)[8] = make_int4(ch[0], ch[1], ch[2], ch[3]);
  ((int4*)out)[9] = make_int4(lst[0], lst[1], lst[2], lst[3]);
}

struct PairTest {
  int mode;
  const float* aabb;
  const int* q;
  const int* filt;
  const int* world;
};

__device__ __forceinline__ bool exactOverlap(const PairTest& T, int a, int b) {
  if ((T.world[a] != T.world[b])) return false;
  int fa = T.filt[a], fb = T.filt[b];
  short ga = (short)(fa & 0xffff), ma = (short)(fa >> 16), gb = (short)(fb & 0xffff), mb = (short)(fb >> 16);
  if (!(((ga & mb) != 0) && ((gb & ma) != 0))) return false;
  if (T.mode == JBGPU_BP_SAP16 || T.mode == JBGPU_BP_SAP32) {
    const int* A = T.q + 6 * a;
    const int* B = T.q + 6 * b;
    for (int ax = 0; ax < 3; ax++)
      if (A[3 + ax] < B[ax] || B[3 + ax] < A[ax]) return false;
    return true;
  }
  const float* A = T.aabb + 6 * a;
  const float* B = T.aabb + 6 * b;
  return (A[0] <= B[3]) && (A[3] >= B[0]) && (A[1] <= B[4]) && (A[4] >= B[1]) && (A[2] <= B[5]) && (A[5] >= B[2]);
}

// exactOverlap for two leaves of the same world from their leaf records (world equality is implied by the
// world lanes of the traversal record)
__device__ __forceinline__ bool leafExactOverlap(int mode, const int4& a0, const int4& a1, const int4& b0,
                                                 const int4& b1) {
  const int fa = a1.z, fb = b1.z;
  short ga = (short)(fa & 0xffff), ma = (short)(fa >> 16), gb = (short)(fb & 0xffff), mb = (short)(fb >> 16);
  if (!(((ga & mb) != 0) && ((gb & ma) != 0))) return false;
  if (mode == JBGPU_BP_SAP16 || mode == JBGPU_BP_SAP32) {
    const int A[6] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y}, B[6] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y};
    for (int ax = 0; ax < 3; ax++)
      if (A[3 + ax] < B[ax] || B[3 + ax] < A[ax]) return false;
    return true;
  }
  const float A[6] = {__int_as_float(a0.x), __int_as_float(a0.y), __int_as_float(a0.z),
                      __int_as_float(a0.w), __int_as_float(a1.x), __int_as_float(a1.y)};
  const float B[6] = {__int_as_float(b0.x), __int_as_float(b0.y), __int_as_float(b0.z),
                      __int_as_float(b0.w), __int_as_float(b1.x), __int_as_float(b1.y)};
  return (A[0] <= B[3]) && (A[3] >= B[0]) && (A[1] <= B[4]) && (A[4] >= B[1]) && (A[2] <= B[5]) && (A[5] >= B[2]);
}

__device__ __forceinline__ void emit(int a, int b, int bits, unsigned long long* counter, uint64_t* out, int64_t cap) {
  uint64_t lo = (uint64_t)min(a, b) + 1, hi = (uint64_t)max(a, b) + 1;
  unsigned long long k = atomicAdd(counter, 1ull);
  if ((int64_t)k < cap) out[k] = (lo << bits) | hi;
}

__device__ __forceinline__ bool boxOverlap(const float4& amn, const float4& amx, const float4& bmn, const float4& bmx) {
  return (amn.x <= bmx.x) && (amx.x >= bmn.x) && (amn.y <= bmx.y) && (amx.y >= bmn.y) && (amn.z <= bmx.z) &&
         (amx.z >= bmn.z);
}

// Pairs found by a block are first collected in its staging area (slots from an LDS counter), then copied
// to the output behind one global reservation per block; only staging overflow takes the global counter.
constexpr int kTraverseBlock = 256;
constexpr int kStagePerBlock = 16 * kTraverseBlock;
// traversal stack entries per thread: a 4-wide step pushes up to three entries and the 4-wide descent is half
// as deep as the binary tree (whose depth the 30-bit keys and index tie-breaks keep far below 64)
constexpr int kStackDepth = 96;
constexpr int kLdsStack = 16;  // of which in LDS
constexpr int kWidePairsPerProxy = 4;  // pair density from which the 4-wide traversal is used

// record layouts of the two traversals: W entries per step, entry c's box at R[2c], R[2c + 1]
template <int W>
struct TravRec;
template <>
struct TravRec<2> {  // kNodeRecords: (child0, child1, last0, last1) in R[4]
  static constexpr int F4 = kRecF4;
  __device__ static void meta(const float4* R, int* ch, int* lst) {
    const float4 m = R[4];
    ch[0] = __float_as_int(m.x);
    ch[1] = __float_as_int(m.y);
    lst[0] = __float_as_int(m.z);
    lst[1] = __float_as_int(m.w);
  }
};
template <>
struct TravRec<4> {  // kNodeRecords4: ids in R[8], last leaves in R[9]
  static constexpr int F4 = kRec4F4;
  __device__ static void meta(const float4* R, int* ch, int* lst) {
    const int4 a = ((const int4*)R)[8], b = ((const int4*)R)[9];
    ch[0] = a.x, ch[1] = a.y, ch[2] = a.z, ch[3] = a.w;
    lst[0] = b.x, lst[1] = b.y, lst[2] = b.z, lst[3] = b.w;
  }
};

template <int W>
__global__ __launch_bounds__(kTraverseBlock) void kTraverse(int n, const int* __restrict__ dn, const int* __restrict__ order,
                                                            const float* __restrict__ leafBox,
                                                            const float4* __restrict__ recs,
                                                            const int4* __restrict__ leafExact, int mode, int bits,
                                                            unsigned long long* counter, uint64_t* out, int64_t cap,
                                                            int* err, uint64_t* __restrict__ stage) {
  if (dn) n = *dn;  // graph-captured steps: the count on the device
  __shared__ unsigned int sCount;
  __shared__ unsigned long long sBase;
  // the top kLdsStack entries of every thread's stack live in LDS ([depth][thread]: no bank conflicts),
  // deeper ones in private memory
  __shared__ int sStack[kLdsStack][kTraverseBlock];
  if (threadIdx.x == 0) sCount = 0;
  __syncthreads();
  uint64_t* myStage = stage + (size_t)blockIdx.x * kStagePerBlock;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n && n >= 2) {
    const float4 qmn = ((const float4*)leafBox)[2 * s];
    const float4 qmx = ((const float4*)leafBox)[2 * s + 1];
    const int qw = __float_as_int(qmn.w);
    const int4 qe0 = leafExact[2 * s], qe1 = leafExact[2 * s + 1];
    const int pa = qe1.w;
    int deep[kStackDepth - kLdsStack];
    int sp = 0;
    int node = 0;
    // descend into the first hit internal entry directly, push the other hit ones
    while (true) {
      const float4* R = recs + TravRec<W>::F4 * node;
      int chs[W], lsts[W];
      TravRec<W>::meta(R, chs, lsts);
      // box tests of all entries first; then the exact-test records of the hit leaves, loaded together (one
      // round trip for up to W leaves instead of one per leaf), then the exact tests; then the descent
      bool hit[W];
      bool anyLeaf = false;
#pragma unroll
      for (int c = 0; c < W; c++) {
        const float4 bmn = R[2 * c], bmx = R[2 * c + 1];
        const bool worldIn = qw >= __float_as_int(bmn.w) && qw <= __float_as_int(bmx.w);
        // only leaves after s pair with s (unused entries have last leaf -1)
        hit[c] = lsts[c] > s && worldIn && boxOverlap(qmn, qmx, bmn, bmx);
        anyLeaf |= hit[c] && chs[c] < 0;
      }
      if (anyLeaf) {
        int4 e0[W], e1[W];
#pragma unroll
        for (int c = 0; c < W; c++) {  // entries that are not hit leaves read leaf 0's record, unused
          const int li = hit[c] && chs[c] < 0 ? -chs[c] - 1 : 0;
          e0[c] = leafExact[2 * li];
          e1[c] = leafExact[2 * li + 1];
        }
#pragma unroll
        for (int c = 0; c < W; c++) {
          if (!(hit[c] && chs[c] < 0)) continue;
          const int pb = e1[c].w;
          if (leafExactOverlap(mode, qe0, qe1, e0[c], e1[c])) {
            const uint64_t lo = (uint64_t)min(pa, pb) + 1, hi = (uint64_t)max(pa, pb) + 1;
            const unsigned slot = atomicAdd(&sCount, 1u);
            if (slot < (unsigned)kStagePerBlock)
              myStage[slot] = (lo << bits) | hi;
            else
              emit(pa, pb, bits, counter, out, cap);
          }
        }
      }
      int next = -1;
#pragma unroll
      for (int c = 0; c < W; c++) {
        const int ch = chs[c];
        if (!hit[c] || ch < 0) continue;
        if (next < 0) {
          next = ch;
        } else if (sp < kLdsStack) {
          sStack[sp++][threadIdx.x] = ch;
        } else if (sp < kStackDepth) {
          deep[sp++ - kLdsStack] = ch;
        } else {
          atomicOr(err, 1);
        }
      }
      if (next >= 0) {
        node = next;
      } else if (sp > 0) {
        --sp;
        node = sp < kLdsStack ? sStack[sp][threadIdx.x] : deep[sp - kLdsStack];
      } else {
        break;
      }
    }
  }
  __syncthreads();
  const unsigned staged = min(sCount, (unsigned)kStagePerBlock);
  if (threadIdx.x == 0) sBase = staged ? atomicAdd(counter, (unsigned long long)staged) : 0ull;
  __syncthreads();
  for (unsigned i = threadIdx.x; i < staged; i += kTraverseBlock) {
    const unsigned long long k = sBase + i;
    if ((int64_t)k < cap) out[k] = myStage[i];
  }
}

// per world w: range[2w], range[2w + 1] = its part of the huge list (sorted by world, then index); the
// table is cleared to empty ranges beforehand
__global__ void kHugeRanges(int nh, const int* __restrict__ dn, const int* __restrict__ hugeWorld, int* range) {
  if (dn) nh = *dn;  // graph-captured steps: the count on the device
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nh) return;
  const int w = hugeWorld[k];
  if (k == 0 || hugeWorld[k - 1] != w) range[2 * w] = k;
  if (k == nh - 1 || hugeWorld[k + 1] != w) range[2 * w + 1] = k + 1;
}

// every proxy against the huge proxies of its world; each thread counts its pairs, the block reserves room for
// all of them with one global atomic (one per pair on a single counter would serialise), then writes them
constexpr int kHugeBlock = 256;
__global__ __launch_bounds__(kHugeBlock) void kHuge(int n, const int* __restrict__ hugeRange,
                                                    const int* __restrict__ hugeList, const int* __restrict__ isHuge,
                                                    const float* __restrict__ cand, PairTest T, int bits,
                                                    unsigned long long* counter, uint64_t* out, int64_t cap) {
  __shared__ unsigned int sCount;
  __shared__ unsigned long long sBase;
  if (threadIdx.x == 0) sCount = 0;
  __syncthreads();
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  int lo = 0, hi = 0;
  float4 bmn, bmx;
  bool jh = false;
  if (j < n) {
    const int w = T.world[j];
    lo = hugeRange[2 * w];
    hi = hugeRange[2 * w + 1];
    bmn = ((const float4*)cand)[2 * j];
    bmx = ((const float4*)cand)[2 * j + 1];
    jh = isHuge[j] != 0;
  }
  auto hit = [&](int h) {
    if (h == j || (jh && h > j)) return false;
    const float4 amn = ((const float4*)cand)[2 * h];
    const float4 amx = ((const float4*)cand)[2 * h + 1];
    return boxOverlap(amn, amx, bmn, bmx) && exactOverlap(T, h, j);
  };
  unsigned int mine = 0;
  for (int k = lo; k < hi; k++) mine += hit(hugeList[k]) ? 1u : 0u;
  const unsigned int off = mine ? atomicAdd(&sCount, mine) : 0u;
  __syncthreads();
  if (threadIdx.x == 0) sBase = sCount ? atomicAdd(counter, (unsigned long long)sCount) : 0ull;
  __syncthreads();
  unsigned long long at = sBase + off;
  if (mine)
    for (int k = lo; k < hi; k++) {
      const int h = hugeList[k];
      if (!hit(h)) continue;
      const uint64_t a = (uint64_t)min(h, j) + 1, b = (uint64_t)max(h, j) + 1;
      if ((int64_t)at < cap) out[at] = (a << bits) | b;
      at++;
    }
}

// boxes far larger than the typical one (static walls, terrain slabs) leave the tree for the flat per-world
// test of kHuge: one such box overlaps thousands of others, which a single traversal thread would otherwise
// have to report one after another.  Two passes: count the candidates, then mark them only when there are at
// most kLargeCap (kHuge tests every proxy against each of them)
__device__ __forceinline__ bool isLarge(int i, float factor, const float* cand, const float* stats) {
  const float thr = factor * (stats[0] / stats[1]);
  const float* C = cand + 8 * i;
  const float e = fmaxf(fmaxf(C[4] - C[0], C[5] - C[1]), C[6] - C[2]);
  return e > thr;
}
__global__ void kCountLarge(int n, float factor, const float* __restrict__ cand, float* stats,
                            const int* __restrict__ isHuge) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool large = i < n && !isHuge[i] && stats[1] > 0.0f && isLarge(i, factor, cand, stats);
  const unsigned long long b = __ballot(large);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(stats + 2, (float)__popcll(b));
}
__global__ void kMarkLarge(int n, float factor, int cap, const float* __restrict__ cand,
                           const float* __restrict__ stats, int* isHuge) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || isHuge[i] || !(stats[1] > 0.0f) || stats[2] > (float)cap) return;
  if (isLarge(i, factor, cand, stats)) isHuge[i] = 2;
}

__global__ void kHugeWorlds(int nh, const int* __restrict__ dn, int* leafCounts, const int* __restrict__ list, const int* __restrict__ world, int* wk) {
  const int np = nh;
  if (dn) nh = *dn;  // graph-captured steps: the count on the device
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (leafCounts && k == 0) {  // graph-captured steps: tree leaves (proxies - huge) and the chunk counts
    const int nn = np - nh;
    leafCounts[0] = nn;
    leafCounts[1] = (nn + 31) / 32;
    leafCounts[2] = (leafCounts[1] + 31) / 32;
  }
  if (k < nh) wk[k] = world[list[k]];
}

// remap: unpack sorted keys, carry manifolds of surviving pairs, initialise new ones
// collision algorithm of a pair (DefaultCollisionConfiguration.getCollisionAlgorithmCreateFunc,
// collision/dispatch/DefaultCollisionConfiguration.java:149-213); planeFirst: a convex-plane pair whose lower
// uid is the plane (its manifold lists the convex body first)
__device__ __forceinline__ int pairTypeOf(int a, int b, const float* __restrict__ shapes,
                                          const float* __restrict__ props, bool& planeFirst) {
  const int ta = ldi(shapes + JBGPU_SHAPE_WORDS * ldi(props + 8 * a + 6));
  const int tb = ldi(shapes + JBGPU_SHAPE_WORDS * ldi(props + 8 * b + 6));
  int type;
  if (ta == JBGPU_SHAPE_SPHERE && tb == JBGPU_SHAPE_SPHERE) type = JBGPU_PAIR_SPHERE_SPHERE;
  else if (ta == JBGPU_SHAPE_PLANE && tb == JBGPU_SHAPE_PLANE) type = JBGPU_PAIR_NONE;
  else if (ta == JBGPU_SHAPE_PLANE || tb == JBGPU_SHAPE_PLANE) type = JBGPU_PAIR_CONVEX_PLANE;
  else type = JBGPU_PAIR_CONVEX_CONVEX;
  planeFirst = type == JBGPU_PAIR_CONVEX_PLANE && ta == JBGPU_SHAPE_PLANE;
  return type;
}

// pair types of a caller-supplied pair set (the remap computes them for the pairs the broadphase finds)
__global__ void kPairTypes(int64_t np, const uint64_t* __restrict__ keys, const float* __restrict__ shapes,
                           const float* __restrict__ props, int* ptype) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const uint64_t key = keys[i];
  bool planeFirst;
  ptype[i] = pairTypeOf((int)(key >> 32) - 1, (int)(key & 0xffffffffu) - 1, shapes, props, planeFirst);
}

__global__ void kRemap(int64_t np, const int* __restrict__ dnp, const uint64_t* __restrict__ packed, int bits,
                       uint64_t* keysNew, float* manNew, int64_t nOld, const int* __restrict__ dnOld,
                       const uint64_t* __restrict__ keysOld, const float* __restrict__ manOld,
                       const float* __restrict__ shapes, const float* __restrict__ props, int* ptype) {
  if (dnp) np = *dnp;  // graph-captured steps: both counts on the device
  if (dnOld) nOld = *dnOld;
  if ((int64_t)blockIdx.x * blockDim.x >= np) return;  // whole blocks past the pairs (np = 0 included)
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < np;  // every lane stays to the end: the point copy below is done by the whole wave
  if (!live) i = np - 1;
  uint64_t p = packed[i];
  uint64_t key = ((p >> bits) << 32) | (p & ((1ull << bits) - 1ull));
  if (live) keysNew[i] = key;
  int a = (int)(key >> 32) - 1, b = (int)(key & 0xffffffffu) - 1;
  bool planeFirst;
  const int type = pairTypeOf(a, b, shapes, props, planeFirst);
  int64_t lo = 0, hi = nOld;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (keysOld[mid] < key) lo = mid + 1; else hi = mid;
  }
  const bool found = live && lo < nOld && keysOld[lo] == key;
  // only the header and the live points travel (slots past the point count are never read)
  float4* dst = (float4*)(manNew + (size_t)JBGPU_MANIFOLD_WORDS * i);
  int nf4 = 0;  // 16-byte words of live points behind the header
  if (found) {
    const float4 h = ((const float4*)(manOld + (size_t)JBGPU_MANIFOLD_WORDS * lo))[0];
    dst[0] = h;
    nf4 = JBGPU_POINT_WORDS / 4 * min(max(__builtin_bit_cast(int, h.x), 0), 4);
  } else if (live) {
    int b0 = a, b1 = b;
    if (planeFirst) { b0 = b; b1 = a; }
    dst[0] = make_float4(__builtin_bit_cast(float, 0), __builtin_bit_cast(float, b0), __builtin_bit_cast(float, b1),
                         __builtin_bit_cast(float, type));
  }
  if (live) ptype[i] = type;
  // the points of the wave's carried records, two records per pass (lanes 0-31 and 32-63, 512 B each), so every
  // load and store instruction covers contiguous memory instead of 64 records 528 B apart
  uint64_t todo = __ballot(nf4 > 0);
  const int lane = threadIdx.x & 63, k = lane & 31;
  while (todo) {
    const int r0 = __ffsll((unsigned long long)todo) - 1;
    todo &= todo - 1;
    const int r1 = todo ? __ffsll((unsigned long long)todo) - 1 : -1;
    if (r1 >= 0) todo &= todo - 1;
    const int r = lane < 32 ? r0 : r1;
    const int n = __shfl(nf4, r < 0 ? r0 : r);
    const long long so = __shfl((long long)lo, r < 0 ? r0 : r), dn = __shfl((long long)i, r < 0 ? r0 : r);
    if (r >= 0 && k < n)
      ((float4*)(manNew + (size_t)JBGPU_MANIFOLD_WORDS * dn))[1 + k] =
          ((const float4*)(manOld + (size_t)JBGPU_MANIFOLD_WORDS * so))[1 + k];
  }
}

// broadphase-only contexts (no bodies, no shapes): unpack the sorted keys, no manifolds
__global__ void kUnpackKeys(int64_t np, const uint64_t* __restrict__ packed, int bits, uint64_t* keysNew) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  uint64_t p = packed[i];
  uint64_t mask = (1ull << bits) - 1ull;
  keysNew[i] = ((p >> bits) << 32) | (p & mask);
}

BpParams makeParams(const Ctx& c) {
  BpParams P{};
  P.mode = c.bpMode;
  P.stageCurrent = c.stageCurrent;
  P.newStage = (c.stageCurrent + 1) % STAGECOUNT;
  for (int k = 0; k < 3; k++) {
    P.wmin[k] = c.wmin[k];
    P.wmax[k] = c.wmax[k];
    P.quant[k] = c.quant[k];
    P.pad[k] = 2.5 * ((double)c.wmax[k] - (double)c.wmin[k]) / (double)c.sentinel;
  }
  P.handleMask = c.handleMask;
  P.qmask = c.qmask;
  P.sentinel = c.sentinel;
  return P;
}

void ensureProxyBuffers(Ctx& c) {
  size_t n = (size_t)c.np;
  c.cand.ensure(8 * n);
  c.isHuge.ensure(n);
  c.morton.ensure(n);
  c.mortonAlt.ensure(n);
  c.order.ensure(n);
  c.orderAlt.ensure(n);
  c.leafBox.ensure(8 * n);
  c.nodeBox.ensure(8 * n);
  c.nodeRec.ensure(4 * kRecF4 * (size_t)n);
  c.nodeRec4.ensure(4 * kRec4F4 * (size_t)n);
  c.leafExact.ensure(8 * (size_t)n);
  c.nodeChild.ensure(2 * n);
  c.nodeLast.ensure(n);
  c.nodeFirst.ensure(n);
  c.pairStage.ensure((size_t)gridFor(n, kTraverseBlock) * kStagePerBlock);
  c.chunkBox.ensure(8 * ((size_t)n / 31 + 64));
  c.flags.ensure(n);
  c.hugeList.ensure(n);
  c.hugeListAlt.ensure(n);
  c.hugeWorld.ensure(n);
  c.hugeWorldAlt.ensure(n);
  c.bounds.ensure(16 + kStatWords * kStatBlocks);  // [0, 6) centre bounds, [8, 11) extent statistics, partials
  c.counter.ensure(8);
}

}  // namespace

void bpInitQuant(Ctx& c) {
  if (c.bpMode == JBGPU_BP_SAP16) {
    c.handleMask = 0xfffe;
    c.sentinel = 0xffff;
    c.qmask = 0xffff;
  } else {
    c.handleMask = 0xfffffffe;
    c.sentinel = 0x7fffffff;
    c.qmask = 0xffffffff;
  }
  float maxInt = (float)(int32_t)c.sentinel;
  for (int k = 0; k < 3; k++) c.quant[k] = maxInt / (c.wmax[k] - c.wmin[k]);
}

void bpUpdateFromBodies(Ctx& c) {
  ensureProxyBuffers(c);
  BpParams P = makeParams(c);
  tBegin(c, T_SETAABB);
  hipLaunchKernelGGL(kSetAabbBodies, dim3(gridFor(c.np)), dim3(256), 0, c.stream, P, c.np, c.shapes.p, c.tr.p,
                     c.props.p, c.act.p, c.stage.p, c.leaf.p, c.aabb.p, c.q.p, c.cand.p, c.isHuge.p,
                     (const int*)c.world.p, (const int*)c.filt.p, (const int*)nullptr);
  tEnd(c, T_SETAABB);
  c.stageCurrent = P.newStage;
  c.rotationPending = true;
}

void bpUpdateFromAabbs(Ctx& c, const float* dAabb6) {
  ensureProxyBuffers(c);
  BpParams P = makeParams(c);
  tBegin(c, T_SETAABB);
  hipLaunchKernelGGL(kSetAabbArray, dim3(gridFor(c.np)), dim3(256), 0, c.stream, P, c.np, dAabb6, c.stage.p, c.leaf.p,
                     c.aabb.p, c.q.p, c.cand.p, c.isHuge.p, (const int*)c.world.p, (const int*)c.filt.p);
  tEnd(c, T_SETAABB);
  c.stageCurrent = P.newStage;
  c.rotationPending = true;
}

static int bitsFor(int64_t n) {
  int b = 1;
  while ((1ll << b) <= n) b++;
  return b;
}

// ---- query tree over the current shape AABBs (world queries) -----------------------------------------
// CollisionObject AABBs as rayTest computes them: getCollisionShape().getAabb(getWorldTransform())
__global__ void kRayAabbs(int n, const float* __restrict__ shapes, const float* __restrict__ props,
                          const float* __restrict__ tr, float* cand, float* aabb, int* isHuge) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Shape sh = loadShape(shapes + JBGPU_SHAPE_WORDS * ldi(props + 8 * i + 6));
  Tr t;
  for (int k = 0; k < 9; k++) t.b.m[k] = tr[12 * i + k];
  t.o = v3(tr[12 * i + 9], tr[12 * i + 10], tr[12 * i + 11]);
  V3 mn, mx;
  shapeAabb(sh, t, mn, mx);
  ((float4*)cand)[2 * i] = make_float4(mn.x, mn.y, mn.z, 0.0f);
  ((float4*)cand)[2 * i + 1] = make_float4(mx.x, mx.y, mx.z, 0.0f);
  float* A = aabb + 6 * i;
  A[0] = mn.x; A[1] = mn.y; A[2] = mn.z; A[3] = mx.x; A[4] = mx.y; A[5] = mx.z;
  const float e = fmaxf(fmaxf(mx.x - mn.x, mx.y - mn.y), mx.z - mn.z);
  isHuge[i] = !(e < HUGE_EXTENT) ? 1 : 0;  // planes (+-1e30 boxes) are tested per world, outside the tree
}

static MortonOrder mortonOrder(Ctx& c, hipStream_t s, int n, const float* cand, const int* isHuge,
                               const unsigned int* bounds, uint64_t* code, uint64_t* codeAlt, int* order, int* orderAlt) {
  const int mb = mortonBits(c.worldBits), kb = mortonKeyBits(c.worldBits);
  Db<int> dv(order, orderAlt);
  if (kb <= 32) {
    hipLaunchKernelGGL(kMorton<uint32_t>, dim3(gridFor(n)), dim3(256), 0, s, n, cand, isHuge, (const int*)c.world.p,
                       bounds, mb, (uint32_t*)code, order);
    Db<int> dk((int*)code, (int*)codeAlt);
    primSortPairs(c, dk, dv, n, 0, kb);
    return {dk.Current(), true, dv.Current()};
  }
  hipLaunchKernelGGL(kMorton<uint64_t>, dim3(gridFor(n)), dim3(256), 0, s, n, cand, isHuge, (const int*)c.world.p,
                     bounds, mb, code, order);
  Db<uint64_t> dk(code, codeAlt);
  primSortPairs(c, dk, dv, n, 0, kb);
  return {dk.Current(), false, dv.Current()};
}

static void launchBuild(hipStream_t s, int grid, int n, const int* dn, const MortonOrder& m, int* child, int* first,
                        int* last) {
  if (m.narrow)
    hipLaunchKernelGGL(kBuild<uint32_t>, dim3(grid), dim3(256), 0, s, n, dn, (const uint32_t*)m.code, child, first,
                       last);
  else
    hipLaunchKernelGGL(kBuild<uint64_t>, dim3(grid), dim3(256), 0, s, n, dn, (const uint64_t*)m.code, child, first,
                       last);
}

void rayBuildTree(Ctx& c) {
  RayTree& t = c.rt;
  const int n = c.nb;
  hipStream_t s = c.stream;
  t.n = n;
  t.cand.ensure(8 * (size_t)n);
  t.aabb.ensure(6 * (size_t)n);
  t.isHuge.ensure(n);
  t.order.ensure(n);
  t.orderAlt.ensure(n);
  t.morton.ensure(n);
  t.mortonAlt.ensure(n);
  t.hugeList.ensure(n);
  t.hugeListAlt.ensure(n);
  t.hugeWorld.ensure(n);
  t.hugeWorldAlt.ensure(n);
  t.leafBox.ensure(8 * (size_t)n);
  t.leafExact.ensure(8 * (size_t)n);
  t.nodeBox.ensure(8 * (size_t)n);
  t.nodeRec.ensure(4 * kRecF4 * (size_t)n);
  t.nodeChild.ensure(2 * (size_t)n);
  t.nodeFirst.ensure(n);
  t.nodeLast.ensure(n);
  t.chunkBox.ensure(8 * ((size_t)n / 31 + 64));
  t.bounds.ensure(16 + kStatWords * kStatBlocks);
  c.counter.ensure(8);
  hipLaunchKernelGGL(kRayAabbs, dim3(gridFor(n)), dim3(256), 0, s, n, (const float*)c.shapes.p, (const float*)c.props.p,
                     (const float*)c.tr.p, t.cand.p, t.aabb.p, t.isHuge.p);
  unsigned int* bounds = t.bounds.p;
  float* stats = (float*)(t.bounds.p + 8);
  unsigned int* part = t.bounds.p + 16;
  const int nparts = std::min(gridFor(n), kStatBlocks);
  hipLaunchKernelGGL(kProxyStats, dim3(nparts), dim3(256), 0, s, n, (const float*)t.cand.p, (const int*)t.isHuge.p,
                     part);
  hipLaunchKernelGGL(kProxyStatsFinal, dim3(1), dim3(256), 0, s, nparts, (const unsigned int*)part, bounds, stats);
  int* hugeCount = (int*)(c.counter.p + 7);
  primSelectIndex(c, (const int*)t.isHuge.p, n, t.hugeList.p, hugeCount);
  const MortonOrder mo = mortonOrder(c, s, n, (const float*)t.cand.p, (const int*)t.isHuge.p,
                                     (const unsigned int*)bounds, t.morton.p, t.mortonAlt.p, t.order.p, t.orderAlt.p);
  int hcount = 0;
  JB_CHECK(hipMemcpyAsync(&hcount, hugeCount, sizeof(hcount), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
  t.nHuge = hcount;
  t.nn = n - hcount;
  const int* order = mo.order;
  t.sortedOrder = order;
  t.hugeRange.ensure(2 * ((size_t)c.maxWorld + 1));
  JB_CHECK(hipMemsetAsync(t.hugeRange.p, 0, 2 * ((size_t)c.maxWorld + 1) * sizeof(int), s));
  if (t.nHuge > 0) {
    hipLaunchKernelGGL(kHugeWorlds, dim3(gridFor(t.nHuge)), dim3(256), 0, s, t.nHuge, (const int*)nullptr, (int*)nullptr, (const int*)t.hugeList.p,
                       (const int*)c.world.p, t.hugeWorld.p);
    if (c.worldBits > 0) {
      Db<int> hk(t.hugeWorld.p, t.hugeWorldAlt.p);
      Db<int> hv(t.hugeList.p, t.hugeListAlt.p);
      primSortPairs(c, hk, hv, t.nHuge, 0, c.worldBits);
      if (hk.Current() != t.hugeWorld.p)
        JB_CHECK(hipMemcpyAsync(t.hugeWorld.p, hk.Current(), t.nHuge * sizeof(int), hipMemcpyDeviceToDevice, s));
      if (hv.Current() != t.hugeList.p)
        JB_CHECK(hipMemcpyAsync(t.hugeList.p, hv.Current(), t.nHuge * sizeof(int), hipMemcpyDeviceToDevice, s));
    }
    hipLaunchKernelGGL(kHugeRanges, dim3(gridFor(t.nHuge)), dim3(256), 0, s, t.nHuge, (const int*)nullptr, (const int*)t.hugeWorld.p,
                       t.hugeRange.p);
  }
  const int nn = t.nn;
  if (nn >= 1)
    hipLaunchKernelGGL(kGatherLeaves<false>, dim3(gridFor(nn)), dim3(256), 0, s, nn, (const int*)nullptr, order, (const float*)t.cand.p,
                       (const int*)c.world.p, t.leafBox.p, (int)JBGPU_BP_DBVT, (const float*)t.aabb.p,
                       (const int*)nullptr, (const int*)c.filt.p, (int4*)t.leafExact.p);
  if (nn >= 2) {
    launchBuild(s, gridFor(nn - 1), nn, nullptr, mo, t.nodeChild.p, t.nodeFirst.p, t.nodeLast.p);
    const int n1 = (nn + 31) / 32, n2 = (n1 + 31) / 32;
    float* c1 = t.chunkBox.p;
    float* c2 = c1 + 8 * (size_t)n1;
    float* c3 = c2 + 8 * (size_t)n2;
    hipLaunchKernelGGL(kChunk, dim3(gridFor(nn)), dim3(256), 0, s, nn, (const int*)nullptr, t.leafBox.p, c1);
    hipLaunchKernelGGL(kChunk, dim3(gridFor(n1)), dim3(256), 0, s, n1, (const int*)nullptr, c1, c2);
    hipLaunchKernelGGL(kChunk, dim3(gridFor(n2)), dim3(256), 0, s, n2, (const int*)nullptr, c2, c3);
    hipLaunchKernelGGL(kNodeBoxes, dim3(gridFor(nn - 1)), dim3(256), 0, s, nn, (const int*)nullptr, t.nodeFirst.p, t.nodeLast.p,
                       t.leafBox.p, c1, c2, c3, t.nodeBox.p);
    hipLaunchKernelGGL(kNodeRecords, dim3(gridFor(nn - 1)), dim3(256), 0, s, nn, (const int*)nullptr, (const int*)t.nodeChild.p,
                       (const int*)t.nodeLast.p, (const float*)t.leafBox.p, (const float*)t.nodeBox.p,
                       (float4*)t.nodeRec.p);
  }
}

void bpFindPairs(Ctx& c) {
  const int n = c.np;
  const int bits = bitsFor((int64_t)n + 1);
  hipStream_t s = c.stream;
  // 4-wide records halve the dependent steps of a traversal but cost ~400 B per node to build: worth it when the
  // previous frame found several pairs per proxy (a settled pile: 10), not in sparse scenes (1M scattered boxes:
  // 0.07), where the traversals are short.  Both traversals report the same pairs.
  const bool wide = c.npairs >= kWidePairsPerProxy * (int64_t)n;
  if (!c.rotationPending) {  // no setAabb since the last pair query: the boxes stay, the stages rotate
    ensureProxyBuffers(c);
    BpParams P = makeParams(c);
    hipLaunchKernelGGL(kKeepAabbs, dim3(gridFor(n)), dim3(256), 0, s, P, n, c.stage.p, c.leaf.p, c.aabb.p, c.q.p,
                       c.cand.p, c.isHuge.p, (const int*)c.world.p, (const int*)c.filt.p);
    c.stageCurrent = P.newStage;
  }
  c.rotationPending = false;
  JB_CHECK(hipMemsetAsync(c.counter.p, 0, 8 * sizeof(unsigned long long), s));
  // centre bounds and extent statistics (two-stage reduction)
  tBegin(c, T_BVH);
  unsigned int* bounds = (unsigned int*)c.bounds.p;
  float* stats = c.bounds.p + 8;
  unsigned int* part = (unsigned int*)c.bounds.p + 16;
  const int nparts = std::min(gridFor(n), kStatBlocks);
  hipLaunchKernelGGL(kProxyStats, dim3(nparts), dim3(256), 0, s, n, (const float*)c.cand.p, (const int*)c.isHuge.p,
                     part);
  hipLaunchKernelGGL(kProxyStatsFinal, dim3(1), dim3(256), 0, s, nparts, (const unsigned int*)part, bounds, stats);
  // huge proxies (absolutely or relatively large), listed in index order
  hipLaunchKernelGGL(kCountLarge, dim3(gridFor(n)), dim3(256), 0, s, n, kLargeFactor, (const float*)c.cand.p,
                     stats, (const int*)c.isHuge.p);
  hipLaunchKernelGGL(kMarkLarge, dim3(gridFor(n)), dim3(256), 0, s, n, kLargeFactor, kLargeCap,
                     (const float*)c.cand.p, (const float*)stats, c.isHuge.p);
  int* hugeCount = (int*)(c.counter.p + 1);
  primSelectIndex(c, (const int*)c.isHuge.p, n, c.hugeList.p, hugeCount);
  // Morton codes
  const MortonOrder mo = mortonOrder(c, s, n, c.cand.p, c.isHuge.p, (const unsigned int*)bounds, c.morton.p,
                                     c.mortonAlt.p, c.order.p, c.orderAlt.p);
  int hcount = 0;
  JB_CHECK(hipMemcpyAsync(&hcount, hugeCount, sizeof(hcount), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
  c.nHuge = hcount;
  c.st.huge = c.nHuge;
  const int* order = mo.order;
  c.sortedOrder = order;
  const int nn = n - c.nHuge;  // leaves in the tree
  if (c.nHuge > 0) {  // per-world lookup: the index-ordered list sorted stably by world, i.e. by (world, index)
    hipLaunchKernelGGL(kHugeWorlds, dim3(gridFor(c.nHuge)), dim3(256), 0, s, c.nHuge, (const int*)nullptr, (int*)nullptr, (const int*)c.hugeList.p,
                       (const int*)c.world.p, c.hugeWorld.p);
    if (c.worldBits > 0) {
      Db<int> hk(c.hugeWorld.p, c.hugeWorldAlt.p);
      Db<int> hv(c.hugeList.p, c.hugeListAlt.p);
      primSortPairs(c, hk, hv, c.nHuge, 0, c.worldBits);
      if (hk.Current() != c.hugeWorld.p)
        JB_CHECK(hipMemcpyAsync(c.hugeWorld.p, hk.Current(), c.nHuge * sizeof(int), hipMemcpyDeviceToDevice, s));
      if (hv.Current() != c.hugeList.p)
        JB_CHECK(hipMemcpyAsync(c.hugeList.p, hv.Current(), c.nHuge * sizeof(int), hipMemcpyDeviceToDevice, s));
    }
  }
  if (c.nHuge > 0) {
    c.hugeRange.ensure(2 * ((size_t)c.maxWorld + 1));
    JB_CHECK(hipMemsetAsync(c.hugeRange.p, 0, 2 * ((size_t)c.maxWorld + 1) * sizeof(int), s));
    hipLaunchKernelGGL(kHugeRanges, dim3(gridFor(c.nHuge)), dim3(256), 0, s, c.nHuge, (const int*)nullptr, (const int*)c.hugeWorld.p,
                       c.hugeRange.p);
  }
  if (nn >= 2) {
    hipLaunchKernelGGL(kGatherLeaves<true>, dim3(gridFor(nn)), dim3(256), 0, s, nn, (const int*)nullptr, order, c.cand.p, c.world.p,
                       c.leafBox.p, c.bpMode, (const float*)c.aabb.p, (const int*)c.q.p, (const int*)c.filt.p,
                       (int4*)c.leafExact.p);
    launchBuild(s, gridFor(nn - 1), nn, nullptr, mo, c.nodeChild.p, c.nodeFirst.p, c.nodeLast.p);
    const int n1 = (nn + 31) / 32, n2 = (n1 + 31) / 32, n3 = (n2 + 31) / 32;
    float* c1 = c.chunkBox.p;
    float* c2 = c1 + 8 * (size_t)n1;
    float* c3 = c2 + 8 * (size_t)n2;
    hipLaunchKernelGGL(kChunk, dim3(gridFor(nn)), dim3(256), 0, s, nn, (const int*)nullptr, c.leafBox.p, c1);
    hipLaunchKernelGGL(kChunk, dim3(gridFor(n1)), dim3(256), 0, s, n1, (const int*)nullptr, c1, c2);
    hipLaunchKernelGGL(kChunk, dim3(gridFor(n2)), dim3(256), 0, s, n2, (const int*)nullptr, c2, c3);
    (void)n3;
    hipLaunchKernelGGL(kNodeBoxes, dim3(gridFor(nn - 1)), dim3(256), 0, s, nn, (const int*)nullptr, c.nodeFirst.p, c.nodeLast.p,
                       c.leafBox.p, c1, c2, c3, c.nodeBox.p);
    hipLaunchKernelGGL(kNodeRecords, dim3(gridFor(nn - 1)), dim3(256), 0, s, nn, (const int*)nullptr, (const int*)c.nodeChild.p,
                       (const int*)c.nodeLast.p, (const float*)c.leafBox.p, (const float*)c.nodeBox.p,
                       (float4*)c.nodeRec.p);
    if (wide)
      hipLaunchKernelGGL(kNodeRecords4, dim3(gridFor(nn - 1)), dim3(256), 0, s, nn, (const int*)nullptr, (const float4*)c.nodeRec.p,
                         (float4*)c.nodeRec4.p);
  }
  tEnd(c, T_BVH);
  PairTest T{c.bpMode, c.aabb.p, c.q.p, c.filt.p, c.world.p};
  for (int attempt = 0; attempt < 3; attempt++) {
    if (c.pairCap < 1024) c.pairCap = std::max<int64_t>(1024, 8 * (int64_t)n);
    c.rawKeys.ensure(c.pairCap);
    c.rawKeysAlt.ensure(c.pairCap);
    JB_CHECK(hipMemsetAsync(c.counter.p, 0, 2 * sizeof(unsigned long long), s));
    if (attempt == 0) tBegin(c, T_TRAVERSE);
    if (nn >= 2)
      hipLaunchKernelGGL(wide ? kTraverse<4> : kTraverse<2>, dim3(gridFor(nn, kTraverseBlock)), dim3(kTraverseBlock),
                         0, s, nn, (const int*)nullptr, order, c.leafBox.p, (const float4*)(wide ? c.nodeRec4.p : c.nodeRec.p),
                         (const int4*)c.leafExact.p, c.bpMode, bits, c.counter.p, c.rawKeys.p, c.pairCap,
                         (int*)(c.counter.p + 2), c.pairStage.p);
    if (attempt == 0) tEnd(c, T_TRAVERSE);
    if (c.nHuge > 0)
      hipLaunchKernelGGL(kHuge, dim3(gridFor(n, kHugeBlock)), dim3(kHugeBlock), 0, s, n, (const int*)c.hugeRange.p,
                         (const int*)c.hugeList.p, c.isHuge.p, c.cand.p, T, bits, c.counter.p, c.rawKeys.p,
                         c.pairCap);
    unsigned long long cnt[3] = {0, 0, 0};
    JB_CHECK(hipMemcpyAsync(cnt, c.counter.p, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    JB_CHECK(hipStreamSynchronize(s));
    if (cnt[2]) throw std::runtime_error("broadphase traversal stack overflow");
    if ((int64_t)cnt[0] <= c.pairCap) {
      c.npairsOld = c.npairs;
      c.npairs = (int64_t)cnt[0];
      break;
    }
    c.pairCap = (int64_t)cnt[0] + (int64_t)cnt[0] / 2;
    if (attempt == 2) throw std::runtime_error("broadphase pair capacity");
  }
  // sort the packed keys (2*bits significant bits)
  tBegin(c, T_KEYSORT);
  if (c.npairs > 1) {
    Db<uint64_t> kk(c.rawKeys.p, c.rawKeysAlt.p);
    primSortKeys(c, kk, (int)c.npairs, 0, 2 * bits);
    if (kk.Current() != c.rawKeys.p)
      JB_CHECK(hipMemcpyAsync(c.rawKeys.p, kk.Current(), c.npairs * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
  }
  tEnd(c, T_KEYSORT);
  c.st.pairs = c.npairs;
}

// ---- proxy removal and pair deltas ------------------------------------------------------------------
__global__ void kKillProxies(int n, const int* __restrict__ uids, int* filt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) filt[uids[i] - 1] = 0;
}
// 1 for the keys whose proxies both still pair (a destroyed proxy has group and mask 0)
__global__ void kPairLive(int64_t n, const uint64_t* __restrict__ keys, const int* __restrict__ filt, int* flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  flag[i] = filt[(int)(k >> 32) - 1] != 0 && filt[(int)(k & 0xffffffffu) - 1] != 0;
}
// 1 for the keys of a that are not in the sorted keys b
__global__ void kKeyAbsent(int64_t n, const uint64_t* __restrict__ a, int64_t m, const uint64_t* __restrict__ b,
                           int* flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = a[i];
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (b[mid] < k) lo = mid + 1; else hi = mid;
  }
  flag[i] = !(lo < m && b[lo] == k);
}

// keys[0, n) := the flagged keys, in order; returns their count
static int64_t compactKeys(Ctx& c, uint64_t* keys, int64_t n, const int* flag) {
  if (n == 0) return 0;
  hipStream_t s = c.stream;
  c.deltaKeys.ensure(n);
  int* cnt = (int*)(c.counter.p + 7);
  primSelectValues<uint64_t>(c, keys, flag, (int)n, c.deltaKeys.p, cnt);
  int k = 0;
  JB_CHECK(hipMemcpyAsync(&k, cnt, sizeof(int), hipMemcpyDeviceToHost, s));
  JB_CHECK(hipStreamSynchronize(s));
  if (k) JB_CHECK(hipMemcpyAsync(keys, c.deltaKeys.p, (size_t)k * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
  return k;
}

void bpDestroyProxies(Ctx& c, int n, const int32_t* uids) {
  std::vector<int> u(uids, uids + n);
  // every uid is checked (range, already destroyed, repeated in this call) before any state changes, so a
  // rejected call leaves the host and device state as they were
  std::vector<char> seen(c.np, 0);
  for (int i = 0; i < n; i++) {
    if (u[i] < 1 || u[i] > c.np) throw std::invalid_argument("proxy uid out of range");
    if (c.proxyDead[u[i] - 1]) throw std::invalid_argument("proxy already destroyed");
    if (seen[u[i] - 1]) throw std::invalid_argument("proxy uid repeated in one call");
    seen[u[i] - 1] = 1;
  }
  if (n == 0) return;
  for (int i = 0; i < n; i++) c.proxyDead[u[i] - 1] = 1;
  hipStream_t s = c.stream;
  c.counter.ensure(8);
  c.deltaFlag.ensure(n);
  JB_CHECK(hipMemcpyAsync(c.deltaFlag.p, u.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(kKillProxies, dim3(gridFor(n)), dim3(256), 0, s, n, (const int*)c.deltaFlag.p, c.filt.p);
  // removeOverlappingPairsContainingProxy: the pairs leave the current set and the previous one (so
  // the next delta lists them in neither direction)
  const int64_t sizes[2] = {c.npairs, c.npairsOld};
  int64_t kept[2] = {0, 0};
  for (int w = 0; w < 2; w++) {
    const int64_t m = sizes[w];
    uint64_t* keys = c.keys[w == 0 ? c.cur : 1 - c.cur].p;
    if (m == 0 || !keys) continue;
    c.deltaFlag.ensure(m);
    hipLaunchKernelGGL(kPairLive, dim3(gridFor(m)), dim3(256), 0, s, m, (const uint64_t*)keys, (const int*)c.filt.p,
                       c.deltaFlag.p);
    kept[w] = compactKeys(c, keys, m, c.deltaFlag.p);
  }
  JB_CHECK(hipStreamSynchronize(s));
  c.npairs = kept[0];
  c.npairsOld = kept[1];
  c.st.pairs = c.npairs;
}

// pairs the last pair computation added (in the current set, not the previous) and removed
void bpPairDeltas(Ctx& c, uint64_t* added, int64_t capAdded, int64_t* nAdded, uint64_t* removed, int64_t capRemoved,
                  int64_t* nRemoved) {
  hipStream_t s = c.stream;
  c.counter.ensure(8);
  for (int w = 0; w < 2; w++) {
    const int64_t n = w == 0 ? c.npairs : c.npairsOld, m = w == 0 ? c.npairsOld : c.npairs;
    const uint64_t* a = c.keys[w == 0 ? c.cur : 1 - c.cur].p;
    const uint64_t* b = c.keys[w == 0 ? 1 - c.cur : c.cur].p;
    uint64_t* out = w == 0 ? added : removed;
    const int64_t cap = w == 0 ? capAdded : capRemoved;
    int64_t* cnt = w == 0 ? nAdded : nRemoved;
    int64_t k = 0;
    if (n > 0) {
      c.deltaFlag.ensure(n);
      c.deltaKeys.ensure(n);
      if (m > 0) {
        hipLaunchKernelGGL(kKeyAbsent, dim3(gridFor(n)), dim3(256), 0, s, n, a, m, b, c.deltaFlag.p);
        int* dc = (int*)(c.counter.p + 7);
        primSelectValues<uint64_t>(c, a, (const int*)c.deltaFlag.p, (int)n, c.deltaKeys.p, dc);
        int kk = 0;
        JB_CHECK(hipMemcpyAsync(&kk, dc, sizeof(int), hipMemcpyDeviceToHost, s));
        JB_CHECK(hipStreamSynchronize(s));
        k = kk;
        if (out && k > cap) throw std::length_error("pair delta buffer too small");
        if (out && k) JB_CHECK(hipMemcpy(out, c.deltaKeys.p, (size_t)k * sizeof(uint64_t), hipMemcpyDeviceToHost));
      } else {  // nothing before: every key is new
        k = n;
        if (out && k > cap) throw std::length_error("pair delta buffer too small");
        if (out) JB_CHECK(hipMemcpy(out, a, (size_t)k * sizeof(uint64_t), hipMemcpyDeviceToHost));
      }
    }
    if (cnt) *cnt = k;
  }
}

void bpPairTypes(Ctx& c) {
  c.ptype.ensure(std::max<int64_t>(c.npairs, 1));
  if (c.npairs > 0)
    hipLaunchKernelGGL(kPairTypes, dim3(gridFor(c.npairs)), dim3(256), 0, c.stream, c.npairs,
                       (const uint64_t*)c.keys[c.cur].p, (const float*)c.shapes.p, (const float*)c.props.p, c.ptype.p);
  JB_CHECK(hipGetLastError());
}

void bpRemapManifolds(Ctx& c) {
  const int bits = bitsFor((int64_t)c.np + 1);
  int nxt = 1 - c.cur;
  c.keys[nxt].ensure(std::max<int64_t>(c.npairs, 1));
  if (!c.haveWorld) {  // no bodies: the pair set alone
    tBegin(c, T_REMAP);
    if (c.npairs > 0)
      hipLaunchKernelGGL(kUnpackKeys, dim3(gridFor(c.npairs)), dim3(256), 0, c.stream, c.npairs, c.rawKeys.p, bits,
                         c.keys[nxt].p);
    tEnd(c, T_REMAP);
    c.cur = nxt;
    return;
  }
  c.man[nxt].ensure((size_t)JBGPU_MANIFOLD_WORDS * std::max<int64_t>(c.npairs, 1));
  c.ptype.ensure(std::max<int64_t>(c.npairs, 1));
  tBegin(c, T_REMAP);
  if (c.npairs > 0)
    hipLaunchKernelGGL(kRemap, dim3(gridFor(c.npairs)), dim3(256), 0, c.stream, c.npairs, (const int*)nullptr, c.rawKeys.p, bits,
                       c.keys[nxt].p, c.man[nxt].p, c.npairsOld, (const int*)nullptr, c.keys[c.cur].p, c.man[c.cur].p, c.shapes.p,
                       c.props.p, c.ptype.p);
  tEnd(c, T_REMAP);
  c.cur = nxt;
}

// ---- the broadphase of a graph-captured step --------------------------------------------------------
// Every launch is shaped by the proxy count and the pair capacity; the huge-proxy, leaf and pair counts stay on
// the device (Ctx::gc); the counters, the huge ranges and the narrowphase and solver flags are cleared by one
// fill at the start of the step.  A pair set larger than the capacity, or a traversal stack overflow, raises the sticky
// abort flag and leaves the step's pair count 0, so nothing after the broadphase changes the world; the host
// then redoes that step on the ordinary path.
namespace {
__global__ void kGraphPairsFound(const unsigned long long* __restrict__ counter, int64_t cap, int* gc) {
  const unsigned long long found = counter[0];
  const int err = ((const int*)(counter + 2))[0];
  if (gc[kGcAbort] || err || (int64_t)found > cap) {
    gc[kGcAbort] = 1;
    gc[kGcPairs] = 0;
  } else {
    gc[kGcPairs] = (int)found;
  }
}
}  // namespace

void bpGraphReserve(Ctx& c) {
  ensureProxyBuffers(c);
  const size_t cap = (size_t)std::max<int64_t>(c.pairCap, 1);
  c.rawKeys.ensure(cap);
  c.rawKeysAlt.ensure(cap);
  // the pair sets of the last two steps (keys and manifolds) carry over: grown, not replaced
  c.keys[0].ensureKeep(cap);
  c.keys[1].ensureKeep(cap);
  c.man[0].ensureKeep((size_t)JBGPU_MANIFOLD_WORDS * cap);
  c.man[1].ensureKeep((size_t)JBGPU_MANIFOLD_WORDS * cap);
  c.ptype.ensure(cap);
  c.hugeRange.ensure(2 * ((size_t)c.maxWorld + 1));
  c.gc.ensure(kGcWords);
  primReserve(c, (int)std::max<int64_t>(c.np, c.pairCap));
}

const uint64_t* bpStepGraph(Ctx& c) {
  const int n = c.np;
  const int bits = bitsFor((int64_t)n + 1);
  hipStream_t s = c.stream;
  int* gc = c.gc.p;
  const bool wide = c.npairs >= kWidePairsPerProxy * (int64_t)n;
  BpParams P = makeParams(c);
  hipLaunchKernelGGL(kSetAabbBodies, dim3(gridFor(n)), dim3(256), 0, s, P, n, c.shapes.p, c.tr.p, c.props.p, c.act.p,
                     c.stage.p, c.leaf.p, c.aabb.p, c.q.p, c.cand.p, c.isHuge.p, (const int*)c.world.p,
                     (const int*)c.filt.p, (const int*)(gc + kGcAbort));
  c.stageCurrent = P.newStage;
  c.rotationPending = false;
  unsigned int* bounds = (unsigned int*)c.bounds.p;
  float* stats = c.bounds.p + 8;
  unsigned int* part = (unsigned int*)c.bounds.p + 16;
  const int nparts = std::min(gridFor(n), kStatBlocks);
  hipLaunchKernelGGL(kProxyStats, dim3(nparts), dim3(256), 0, s, n, (const float*)c.cand.p, (const int*)c.isHuge.p,
                     part);
  hipLaunchKernelGGL(kProxyStatsFinal, dim3(1), dim3(256), 0, s, nparts, (const unsigned int*)part, bounds, stats);
  hipLaunchKernelGGL(kCountLarge, dim3(gridFor(n)), dim3(256), 0, s, n, kLargeFactor, (const float*)c.cand.p,
                     stats, (const int*)c.isHuge.p);
  hipLaunchKernelGGL(kMarkLarge, dim3(gridFor(n)), dim3(256), 0, s, n, kLargeFactor, kLargeCap,
                     (const float*)c.cand.p, (const float*)stats, c.isHuge.p);
  primSelectIndex(c, (const int*)c.isHuge.p, n, c.hugeList.p, gc + kGcHuge);
  const MortonOrder mo = mortonOrder(c, s, n, c.cand.p, c.isHuge.p, (const unsigned int*)bounds, c.morton.p,
                                     c.mortonAlt.p, c.order.p, c.orderAlt.p);
  const int* order = mo.order;
  c.sortedOrder = order;
  // huge proxies by world (and the tree's leaf counts)
  hipLaunchKernelGGL(kHugeWorlds, dim3(gridFor(n)), dim3(256), 0, s, n, (const int*)(gc + kGcHuge), gc + kGcLeaves,
                     (const int*)c.hugeList.p, (const int*)c.world.p, c.hugeWorld.p);
  Db<int> hk(c.hugeWorld.p, c.hugeWorldAlt.p);
  Db<int> hv(c.hugeList.p, c.hugeListAlt.p);
  if (c.worldBits > 0) primSortPairsDev(c, hk, hv, (const int*)(gc + kGcHuge), n, 0, c.worldBits);
  hipLaunchKernelGGL(kHugeRanges, dim3(gridFor(n)), dim3(256), 0, s, n, (const int*)(gc + kGcHuge),
                     (const int*)hk.Current(), c.hugeRange.p);
  // the tree over the remaining leaves (chunk levels at the offsets of a full proxy set)
  const int* nl = gc + kGcLeaves;
  const int n1 = (n + 31) / 32, n2 = (n1 + 31) / 32;
  float* c1 = c.chunkBox.p;
  float* c2 = c1 + 8 * (size_t)n1;
  float* c3 = c2 + 8 * (size_t)n2;
  hipLaunchKernelGGL(kGatherLeaves<true>, dim3(gridFor(n)), dim3(256), 0, s, n, nl, order, c.cand.p, c.world.p, c.leafBox.p,
                     c.bpMode, (const float*)c.aabb.p, (const int*)c.q.p, (const int*)c.filt.p, (int4*)c.leafExact.p);
  launchBuild(s, gridFor(n), n, nl, mo, c.nodeChild.p, c.nodeFirst.p, c.nodeLast.p);
  hipLaunchKernelGGL(kChunk, dim3(gridFor(n)), dim3(256), 0, s, n, nl, c.leafBox.p, c1);
  hipLaunchKernelGGL(kChunk, dim3(gridFor(n1)), dim3(256), 0, s, n1, (const int*)(gc + kGcChunk1), c1, c2);
  hipLaunchKernelGGL(kChunk, dim3(gridFor(n2)), dim3(256), 0, s, n2, (const int*)(gc + kGcChunk2), c2, c3);
  hipLaunchKernelGGL(kNodeBoxes, dim3(gridFor(n)), dim3(256), 0, s, n, nl, c.nodeFirst.p, c.nodeLast.p, c.leafBox.p,
                     c1, c2, c3, c.nodeBox.p);
  hipLaunchKernelGGL(kNodeRecords, dim3(gridFor(n)), dim3(256), 0, s, n, nl, (const int*)c.nodeChild.p,
                     (const int*)c.nodeLast.p, (const float*)c.leafBox.p, (const float*)c.nodeBox.p,
                     (float4*)c.nodeRec.p);
  if (wide)
    hipLaunchKernelGGL(kNodeRecords4, dim3(gridFor(n)), dim3(256), 0, s, n, nl, (const float4*)c.nodeRec.p,
                       (float4*)c.nodeRec4.p);
  // pairs: traversal, then every proxy against its world's huge proxies
  PairTest T{c.bpMode, c.aabb.p, c.q.p, c.filt.p, c.world.p};
  hipLaunchKernelGGL(wide ? kTraverse<4> : kTraverse<2>, dim3(gridFor(n, kTraverseBlock)), dim3(kTraverseBlock), 0, s,
                     n, nl, order, c.leafBox.p, (const float4*)(wide ? c.nodeRec4.p : c.nodeRec.p),
                     (const int4*)c.leafExact.p, c.bpMode, bits, c.counter.p, c.rawKeys.p, c.pairCap,
                     (int*)(c.counter.p + 2), c.pairStage.p);
  hipLaunchKernelGGL(kHuge, dim3(gridFor(n, kHugeBlock)), dim3(kHugeBlock), 0, s, n, (const int*)c.hugeRange.p,
                     (const int*)hv.Current(), c.isHuge.p, c.cand.p, T, bits, c.counter.p, c.rawKeys.p, c.pairCap);
  hipLaunchKernelGGL(kGraphPairsFound, dim3(1), dim3(1), 0, s, (const unsigned long long*)c.counter.p, c.pairCap, gc);
  Db<uint64_t> kk(c.rawKeys.p, c.rawKeysAlt.p);
  primSortKeysDev(c, kk, (const int*)(gc + kGcPairs), (int)c.pairCap, 0, 2 * bits);
  return kk.Current();
}

void bpRemapGraph(Ctx& c, const uint64_t* sorted) {
  const int bits = bitsFor((int64_t)c.np + 1);
  const int nxt = 1 - c.cur;
  int* gc = c.gc.p;
  hipLaunchKernelGGL(kRemap, dim3(gridFor(c.pairCap)), dim3(256), 0, c.stream, c.pairCap, (const int*)(gc + kGcPairs),
                     sorted, bits, c.keys[nxt].p, c.man[nxt].p, (int64_t)0, (const int*)(gc + kGcPairCount + c.cur),
                     (const uint64_t*)c.keys[c.cur].p, (const float*)c.man[c.cur].p, (const float*)c.shapes.p,
                     (const float*)c.props.p, c.ptype.p);
  c.cur = nxt;  // the end-of-step kernel records the new set's count
}

}  // namespace jbg
