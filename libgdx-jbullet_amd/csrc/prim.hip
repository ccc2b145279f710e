// prim.hip — device-wide exclusive scan, flag compaction and stable LSD radix sort for gfx950 (wave64).
//
// The element count is read from device memory (clamped to a capacity fixed by the caller), so every launch
// has a shape that depends only on the capacity: no host wait for the count, no memset, and a sequence of
// these launches can be captured once in a hipGraph and replayed for any count up to the capacity.
//
// Radix sort: 8-bit digits, least significant first, three launches per digit: per-tile digit histograms
// (LDS atomics), one exclusive scan over the digit-major histogram table, and a stable scatter in which every
// tile ranks its keys chunk by chunk (256 keys per chunk, in thread order): lanes of a wave holding the same
// digit are found with eight ballots, the waves' counts are combined through LDS, and a per-digit running
// offset carries the order from chunk to chunk.  Equal digits keep their input order, so the sort is stable.
#include <type_traits>

#include "internal.h"

namespace jbg {

namespace {

constexpr int kPT = 256;                // threads per block
constexpr int kPerThread = 4;           // items per thread and tile
constexpr int kTile = kPT * kPerThread; // items per tile (1024: a 100k-key sort still spreads over ~100 blocks)
constexpr int kRadix = 256;
constexpr int kScanTop = 1024;          // threads of the single-block scan of the tile sums

__device__ __forceinline__ int countOf(const int* count, int hostCount, int cap) {
  const int n = count ? *count : hostCount;
  return n < 0 ? 0 : (n > cap ? cap : n);
}

// inclusive scan over the 64 lanes of a wave
__device__ __forceinline__ int waveInclusive(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off);
    if (lane >= off) v += u;
  }
  return v;
}

// exclusive scan of one value per thread over a block of T threads; *total gets the block sum
template <int T>
__device__ __forceinline__ int blockExclusive(int v, int* waveSums, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int inc = waveInclusive(v);
  if (lane == 63) waveSums[w] = inc;
  __syncthreads();
  if (w == 0) {
    const int s = lane < T / 64 ? waveSums[lane] : 0;
    const int si = waveInclusive(s);
    if (lane < T / 64) waveSums[lane] = si - s;
    if (lane == T / 64 - 1) *total = si;
  }
  __syncthreads();
  return inc - v + waveSums[w];
}

// ---- scan --------------------------------------------------------------------------------------------------
// tile sums of in[0, n)
__global__ __launch_bounds__(kPT) void kTileSums(const int* __restrict__ in, const int* __restrict__ count,
                                                 int hostCount, int cap, int* __restrict__ tileSum) {
  __shared__ int waveSums[kPT / 64];
  __shared__ int total;
  const int n = countOf(count, hostCount, cap);
  const int base = blockIdx.x * kTile;
  int s = 0;
  for (int i = base + (int)threadIdx.x; i < base + kTile && i < n; i += kPT) s += in[i];
  (void)blockExclusive<kPT>(s, waveSums, &total);
  if (threadIdx.x == 0) tileSum[blockIdx.x] = total;
}

// exclusive scan of the tile sums in place (one block); *total = sum of all
__global__ __launch_bounds__(kScanTop) void kScanTileSums(int nTiles, int* tileSum, int* total) {
  __shared__ int waveSums[kScanTop / 64];
  __shared__ int chunkTotal;
  int carry = 0;
  for (int c0 = 0; c0 < nTiles; c0 += kScanTop) {
    const int i = c0 + (int)threadIdx.x;
    const int v = i < nTiles ? tileSum[i] : 0;
    const int ex = blockExclusive<kScanTop>(v, waveSums, &chunkTotal);
    if (i < nTiles) tileSum[i] = carry + ex;
    carry += chunkTotal;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

// out[i] = tileOffset + exclusive prefix of in inside the tile (tile staged in LDS for coalesced access)
__global__ __launch_bounds__(kPT) void kScanTiles(const int* __restrict__ in, const int* __restrict__ count,
                                                  int hostCount, int cap, const int* __restrict__ tileOffset,
                                                  int* __restrict__ out) {
  __shared__ int tile[kTile];
  __shared__ int waveSums[kPT / 64];
  __shared__ int total;
  const int n = countOf(count, hostCount, cap);
  const int base = blockIdx.x * kTile;
  if (base >= n) return;
  for (int k = threadIdx.x; k < kTile; k += kPT) tile[k] = base + k < n ? in[base + k] : 0;
  __syncthreads();
  int s = 0;
  const int t0 = threadIdx.x * kPerThread;
#pragma unroll
  for (int k = 0; k < kPerThread; k++) s += tile[t0 + k];
  int run = blockExclusive<kPT>(s, waveSums, &total) + tileOffset[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kPerThread; k++) {
    const int v = tile[t0 + k];
    tile[t0 + k] = run;
    run += v;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kTile && base + k < n; k += kPT) out[base + k] = tile[k];
}

// ---- compaction ----------------------------------------------------------------------------------------------
// selected[ tileOffset + rank ] = i (or vals[i]) for every i < n with flag[i] != 0, in index order
template <class T, bool Vals>
__global__ __launch_bounds__(kPT) void kSelectTiles(const int* __restrict__ flag, const T* __restrict__ vals,
                                                    const int* __restrict__ count, int hostCount, int cap,
                                                    const int* __restrict__ tileOffset, T* __restrict__ out) {
  __shared__ int tile[kTile];
  __shared__ int waveSums[kPT / 64];
  __shared__ int total;
  const int n = countOf(count, hostCount, cap);
  const int base = blockIdx.x * kTile;
  if (base >= n) return;
  for (int k = threadIdx.x; k < kTile; k += kPT) tile[k] = base + k < n ? (flag[base + k] != 0) : 0;
  __syncthreads();
  int s = 0;
  const int t0 = threadIdx.x * kPerThread;
#pragma unroll
  for (int k = 0; k < kPerThread; k++) s += tile[t0 + k];
  int run = blockExclusive<kPT>(s, waveSums, &total) + tileOffset[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kPerThread; k++) {
    const int v = tile[t0 + k];
    tile[t0 + k] = v ? run : -1;
    run += v;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kTile && base + k < n; k += kPT) {
    const int pos = tile[k];
    if (pos >= 0) out[pos] = Vals ? vals[base + k] : (T)(base + k);
  }
}

// flags to 0/1 counts for the tile sums of a compaction
__global__ __launch_bounds__(kPT) void kFlagTileSums(const int* __restrict__ flag, const int* __restrict__ count,
                                                     int hostCount, int cap, int* __restrict__ tileSum) {
  __shared__ int waveSums[kPT / 64];
  __shared__ int total;
  const int n = countOf(count, hostCount, cap);
  const int base = blockIdx.x * kTile;
  int s = 0;
  for (int i = base + (int)threadIdx.x; i < base + kTile && i < n; i += kPT) s += flag[i] != 0;
  (void)blockExclusive<kPT>(s, waveSums, &total);
  if (threadIdx.x == 0) tileSum[blockIdx.x] = total;
}

// ---- radix sort ----------------------------------------------------------------------------------------------
// the digit of a key: bits [shift, shift + 8) under mask (the last digit of a bit range may be narrower)
template <class K>
__device__ __forceinline__ int digitOf(K k, int shift, int mask) {
  return (int)((k >> shift) & (K)mask);
}

// hist[d * nTiles + tile] = keys of the tile with digit d
template <class K>
__global__ __launch_bounds__(kPT) void kRadixHist(const K* __restrict__ keys, const int* __restrict__ count,
                                                  int hostCount, int cap, int shift, int mask, int* __restrict__ hist) {
  __shared__ int h[kRadix];
  const int n = countOf(count, hostCount, cap);
  const int base = blockIdx.x * kTile;
  h[threadIdx.x] = 0;
  __syncthreads();
  for (int i = base + (int)threadIdx.x; i < base + kTile && i < n; i += kPT)
    atomicAdd(&h[digitOf(keys[i], shift, mask)], 1);
  __syncthreads();
  hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// stable scatter of one tile: chunks of kPT keys in thread order
template <class K, bool Vals>
__global__ __launch_bounds__(kPT) void kRadixScatter(const K* __restrict__ kin, const int* __restrict__ vin,
                                                     K* __restrict__ kout, int* __restrict__ vout,
                                                     const int* __restrict__ count, int hostCount, int cap, int shift,
                                                     int mask, const int* __restrict__ offs) {
  __shared__ int running[kRadix];
  __shared__ int cnt[kPT / 64][kRadix];
  const int n = countOf(count, hostCount, cap);
  const int base = blockIdx.x * kTile;
  if (base >= n) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  running[t] = offs[t * gridDim.x + blockIdx.x];
#pragma unroll
  for (int q = 0; q < kPT / 64; q++) cnt[q][t] = 0;
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int c0 = base; c0 < base + kTile && c0 < n; c0 += kPT) {
    const int i = c0 + t;
    const bool valid = i < n;
    K k = valid ? kin[i] : (K)0;
    const int v = (Vals && valid) ? vin[i] : 0;
    const int d = digitOf(k, shift, mask);
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const unsigned long long bb = __ballot((d >> b) & 1);
      m &= ((d >> b) & 1) ? bb : ~bb;
    }
    const int rank = __popcll(m & below);
    if (valid && rank == 0) cnt[w][d] = __popcll(m);
    __syncthreads();
    if (valid) {
      int pos = running[d] + rank;
      for (int q = 0; q < w; q++) pos += cnt[q][d];
      kout[pos] = k;
      if (Vals) vout[pos] = v;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int q = 0; q < kPT / 64; q++) {
      add += cnt[q][t];
      cnt[q][t] = 0;
    }
    running[t] += add;
    __syncthreads();
  }
}

}  // namespace

int primTiles(int cap) { return (std::max(cap, 1) + kTile - 1) / kTile; }

size_t primScratchInts(int cap, int keyBytes) {
  (void)keyBytes;
  const int tiles = primTiles(cap);
  const size_t table = (size_t)kRadix * tiles;  // the radix sort's digit table, then its scan's tile sums
  return table + (size_t)primTiles((int)std::min<size_t>(table, 1u << 30)) + (size_t)tiles + 64;
}

void primExclusiveScan(hipStream_t s, const int* in, int* out, const int* count, int hostCount, int cap, int* scratch,
                       int* total) {
  const int tiles = primTiles(cap);
  hipLaunchKernelGGL(kTileSums, dim3(tiles), dim3(kPT), 0, s, in, count, hostCount, cap, scratch);
  hipLaunchKernelGGL(kScanTileSums, dim3(1), dim3(kScanTop), 0, s, tiles, scratch, total);
  hipLaunchKernelGGL(kScanTiles, dim3(tiles), dim3(kPT), 0, s, in, count, hostCount, cap, (const int*)scratch, out);
}

void primSelectFlagged(hipStream_t s, const int* flag, const int* vals, int* out, const int* count, int hostCount,
                       int cap, int* scratch, int* selectedCount) {
  const int tiles = primTiles(cap);
  hipLaunchKernelGGL(kFlagTileSums, dim3(tiles), dim3(kPT), 0, s, flag, count, hostCount, cap, scratch);
  hipLaunchKernelGGL(kScanTileSums, dim3(1), dim3(kScanTop), 0, s, tiles, scratch, selectedCount);
  if (vals)
    hipLaunchKernelGGL((kSelectTiles<int, true>), dim3(tiles), dim3(kPT), 0, s, flag, vals, count, hostCount, cap,
                       (const int*)scratch, out);
  else
    hipLaunchKernelGGL((kSelectTiles<int, false>), dim3(tiles), dim3(kPT), 0, s, flag, (const int*)nullptr, count,
                       hostCount, cap, (const int*)scratch, out);
}

template <class K>
bool primRadixSort(hipStream_t s, K* k0, int* v0, K* k1, int* v1, const int* count, int hostCount, int cap, int bitLo,
                   int bitHi, int* scratch) {
  const int tiles = primTiles(cap);
  int* hist = scratch;
  int* top = scratch + (size_t)kRadix * tiles;  // tile sums of the table's scan
  bool alt = false;
  for (int shift = bitLo; shift < bitHi; shift += 8) {
    K* ki = alt ? k1 : k0;
    K* ko = alt ? k0 : k1;
    int* vi = alt ? v1 : v0;
    int* vo = alt ? v0 : v1;
    const int mask = (1 << std::min(8, bitHi - shift)) - 1;  // bits at or above bitHi take no part
    hipLaunchKernelGGL(kRadixHist<K>, dim3(tiles), dim3(kPT), 0, s, (const K*)ki, count, hostCount, cap, shift, mask,
                       hist);
    // digit-major table (digit, tile): one exclusive scan gives every tile's first position per digit (a small
    // table in one single-block launch, a large one as tile sums, their scan and the tiles)
    const int entries = kRadix * tiles;
    if (entries <= 4 * kScanTop) {
      hipLaunchKernelGGL(kScanTileSums, dim3(1), dim3(kScanTop), 0, s, entries, hist, (int*)nullptr);
    } else {
      const int scanTiles = primTiles(entries);
      hipLaunchKernelGGL(kTileSums, dim3(scanTiles), dim3(kPT), 0, s, (const int*)hist, (const int*)nullptr, entries,
                         entries, top);
      hipLaunchKernelGGL(kScanTileSums, dim3(1), dim3(kScanTop), 0, s, scanTiles, top, (int*)nullptr);
      hipLaunchKernelGGL(kScanTiles, dim3(scanTiles), dim3(kPT), 0, s, (const int*)hist, (const int*)nullptr, entries,
                         entries, (const int*)top, hist);
    }
    if (v0)
      hipLaunchKernelGGL((kRadixScatter<K, true>), dim3(tiles), dim3(kPT), 0, s, (const K*)ki, (const int*)vi, ko, vo,
                         count, hostCount, cap, shift, mask, (const int*)hist);
    else
      hipLaunchKernelGGL((kRadixScatter<K, false>), dim3(tiles), dim3(kPT), 0, s, (const K*)ki, (const int*)nullptr,
                         ko, (int*)nullptr, count, hostCount, cap, shift, mask, (const int*)hist);
    alt = !alt;
  }
  return alt;
}

template bool primRadixSort<uint32_t>(hipStream_t, uint32_t*, int*, uint32_t*, int*, const int*, int, int, int, int,
                                      int*);
template bool primRadixSort<uint64_t>(hipStream_t, uint64_t*, int*, uint64_t*, int*, const int*, int, int, int, int,
                                      int*);

// ---- the context-level forms used by the step: host-known counts, scratch in the context ----------------------
static int* primScratch(Ctx& c, int cap) {
  c.primTmp.ensure(primScratchInts(cap) + 64);
  return c.primTmp.p;
}

template <class K>
void primSortPairs(Ctx& c, Db<K>& k, Db<int>& v, int n, int bitLo, int bitHi) {
  if (n <= 1 || bitHi <= bitLo) return;
  using U = typename std::conditional<sizeof(K) == 8, uint64_t, uint32_t>::type;  // keys sort as unsigned words
  const bool alt = primRadixSort<U>(c.stream, (U*)k.cur, v.cur, (U*)k.alt, v.alt, nullptr, n, n, bitLo, bitHi,
                                    primScratch(c, n));
  if (alt) {
    std::swap(k.cur, k.alt);
    std::swap(v.cur, v.alt);
  }
}
template <class K>
void primSortKeys(Ctx& c, Db<K>& k, int n, int bitLo, int bitHi) {
  if (n <= 1 || bitHi <= bitLo) return;
  using U = typename std::conditional<sizeof(K) == 8, uint64_t, uint32_t>::type;
  const bool alt = primRadixSort<U>(c.stream, (U*)k.cur, nullptr, (U*)k.alt, nullptr, nullptr, n, n, bitLo, bitHi,
                                    primScratch(c, n));
  if (alt) std::swap(k.cur, k.alt);
}
template void primSortPairs<uint64_t>(Ctx&, Db<uint64_t>&, Db<int>&, int, int, int);
template void primSortPairs<unsigned long long>(Ctx&, Db<unsigned long long>&, Db<int>&, int, int, int);
template void primSortPairs<int>(Ctx&, Db<int>&, Db<int>&, int, int, int);
template void primSortKeys<uint64_t>(Ctx&, Db<uint64_t>&, int, int, int);

void primSelectIndex(Ctx& c, const int* flag, int n, int* out, int* count) {
  if (n <= 0) {
    JB_CHECK(hipMemsetAsync(count, 0, sizeof(int), c.stream));
    return;
  }
  primSelectFlagged(c.stream, flag, nullptr, out, nullptr, n, n, primScratch(c, n), count);
}
template <class T>
void primSelectValues(Ctx& c, const T* vals, const int* flag, int n, T* out, int* count) {
  if (n <= 0) {
    JB_CHECK(hipMemsetAsync(count, 0, sizeof(int), c.stream));
    return;
  }
  const int tiles = primTiles(n);
  int* scratch = primScratch(c, n);
  hipLaunchKernelGGL(kFlagTileSums, dim3(tiles), dim3(kPT), 0, c.stream, flag, (const int*)nullptr, n, n, scratch);
  hipLaunchKernelGGL(kScanTileSums, dim3(1), dim3(kScanTop), 0, c.stream, tiles, scratch, count);
  hipLaunchKernelGGL((kSelectTiles<T, true>), dim3(tiles), dim3(kPT), 0, c.stream, flag, vals, (const int*)nullptr, n,
                     n, (const int*)scratch, out);
}
template void primSelectValues<uint64_t>(Ctx&, const uint64_t*, const int*, int, uint64_t*, int*);
template void primSelectValues<int>(Ctx&, const int*, const int*, int, int*, int*);

void primScan(Ctx& c, const int* in, int* out, int n) {
  if (n <= 0) return;
  primExclusiveScan(c.stream, in, out, nullptr, n, n, primScratch(c, n), nullptr);
}

// ---- device-count forms (graph-captured steps) -----------------------------------------------------------------
void primReserve(Ctx& c, int cap) { (void)primScratch(c, std::max(cap, 1)); }

template <class K>
void primSortPairsDev(Ctx& c, Db<K>& k, Db<int>& v, const int* count, int cap, int bitLo, int bitHi) {
  if (bitHi <= bitLo) return;
  cap = std::max(cap, 1);
  using U = typename std::conditional<sizeof(K) == 8, uint64_t, uint32_t>::type;
  const bool alt = primRadixSort<U>(c.stream, (U*)k.cur, v.cur, (U*)k.alt, v.alt, count, 0, cap, bitLo, bitHi,
                                    primScratch(c, cap));
  if (alt) {
    std::swap(k.cur, k.alt);
    std::swap(v.cur, v.alt);
  }
}
template <class K>
void primSortKeysDev(Ctx& c, Db<K>& k, const int* count, int cap, int bitLo, int bitHi) {
  if (bitHi <= bitLo) return;
  cap = std::max(cap, 1);
  using U = typename std::conditional<sizeof(K) == 8, uint64_t, uint32_t>::type;
  const bool alt = primRadixSort<U>(c.stream, (U*)k.cur, nullptr, (U*)k.alt, nullptr, count, 0, cap, bitLo, bitHi,
                                    primScratch(c, cap));
  if (alt) std::swap(k.cur, k.alt);
}
template void primSortPairsDev<uint64_t>(Ctx&, Db<uint64_t>&, Db<int>&, const int*, int, int, int);
template void primSortPairsDev<int>(Ctx&, Db<int>&, Db<int>&, const int*, int, int, int);
template void primSortKeysDev<uint64_t>(Ctx&, Db<uint64_t>&, const int*, int, int, int);

void primSelectIndexDev(Ctx& c, const int* flag, const int* count, int cap, int* out, int* selected) {
  cap = std::max(cap, 1);
  primSelectFlagged(c.stream, flag, nullptr, out, count, 0, cap, primScratch(c, cap), selected);
}
void primScanDev(Ctx& c, const int* in, int* out, const int* count, int cap) {
  cap = std::max(cap, 1);
  primExclusiveScan(c.stream, in, out, count, 0, cap, primScratch(c, cap), nullptr);
}

__global__ void kFill32(size_t n, uint32_t* __restrict__ p, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void kFillList(FillList f) {
  for (int e = 0; e < f.count; e++)
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < f.n[e]; i += gridDim.x * blockDim.x) f.p[e][i] = f.v[e];
}
void graphFillList(hipStream_t s, const FillList& f) {
  if (f.count == 0) return;
  uint32_t most = 0;
  for (int e = 0; e < f.count; e++) most = std::max(most, f.n[e]);
  const int grid = (int)std::min<uint32_t>((most + 255) / 256, 256);
  hipLaunchKernelGGL(kFillList, dim3(std::max(grid, 1)), dim3(256), 0, s, f);
}
void graphFill(hipStream_t s, void* p, size_t words, uint32_t v) {
  if (words == 0) return;
  const int grid = (int)std::min<size_t>((words + 255) / 256, 1024);
  hipLaunchKernelGGL(kFill32, dim3(grid), dim3(256), 0, s, words, (uint32_t*)p, v);
}

__global__ void kWiden(int n, const int* __restrict__ in, int64_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}
void primScanWide(Ctx& c, const int* in, int64_t* out, int n) {
  if (n <= 0) return;
  c.primWide.ensure(n);
  primExclusiveScan(c.stream, in, c.primWide.p, nullptr, n, n, primScratch(c, n), nullptr);
  hipLaunchKernelGGL(kWiden, dim3(gridFor(n)), dim3(256), 0, c.stream, n, (const int*)c.primWide.p, out);
}

}  // namespace jbg
