// prim.hip — device-wide exclusive scan, flag compaction and stable LSD radix sort for gfx950 (wave64).
//
// The element count is read from device memory (clamped to a capacity fixed by the caller), so every launch
// has a shape that depends only on the capacity: no host wait for the count, no memset, and a sequence of
// these launches can be captured once in a hipGraph and replayed for any count up to the capacity.
//
// Radix sort: 8-bit digits, least significant first.  One launch counts the digits of every pass at once (and
// clears the look-back state), one scans those counts; then one launch per digit: each workgroup takes the next
// tile by an atomic ticket, publishes its per-digit counts, adds up the earlier tiles' counts by decoupled
// look-back (a tile's inclusive prefix ends the walk), and scatters its keys stably, chunk by chunk (256 keys per
// chunk, in thread order): lanes of a wave holding the same digit are found with eight ballots, the waves' counts
// are combined through LDS, and a per-digit running offset carries the order from chunk to chunk.  Equal digits
// keep their input order, so the sort is stable.  A tile only ever waits for tiles with smaller tickets, which
// were taken by workgroups already running; the wait is bounded all the same and sets an error word.
#include <type_traits>

#include "internal.h"

namespace jbg {

namespace {

constexpr int kPT = 256;                // threads per block
constexpr int kPerThread = 4;           // items per thread and tile
constexpr int kTile = kPT * kPerThread; // items per tile (1024: a 100k-key sort still spreads over ~100 blocks)
constexpr int kRadix = 256;
constexpr int kScanTop = 1024;          // threads of the single-block scan of the tile sums
// sort header at the start of the scratch (fixed offsets; kept zero between sorts where noted)
constexpr int kMaxPasses = 8;
constexpr int kHdrHist = 0;                            // digit counts per pass (zero between sorts)
constexpr int kHdrOfs = kMaxPasses * kRadix;           // their exclusive scan
constexpr int kHdrTicket = 2 * kMaxPasses * kRadix;    // tile tickets per pass
constexpr int kHdrErr = kHdrTicket + 16;               // sticky: a look-back wait gave up
constexpr int kSortHdr = kHdrErr + 16;
constexpr uint32_t kStAgg = 1u << 30, kStIncl = 2u << 30, kStMask = (1u << 30) - 1u;  // look-back status word
constexpr int kLookback = 8;  // earlier tiles' status words read per look-back round trip
// keys per thread of a digit pass: 16 from 512k keys, 8 from 256k (fewer tiles: shorter look-back walks), 4 for
// small sorts (more workgroups)
inline int sweepPer(int cap) { return cap >= (1 << 19) ? 16 : cap >= (1 << 18) ? 8 : 4; }
inline int sweepTiles(int cap) { return (std::max(cap, 1) + kPT * sweepPer(cap) - 1) / (kPT * sweepPer(cap)); }

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

// digit counts of every pass (shift bitLo + 8p) over all keys into hdr[kHdrHist]; clears the look-back status
template <class K>
__global__ __launch_bounds__(kPT) void kDigitHist(const K* __restrict__ keys, const int* __restrict__ count,
                                                  int hostCount, int cap, int bitLo, int bitHi, int* hdr,
                                                  uint32_t* __restrict__ status, size_t statusWords) {
  __shared__ int h[kMaxPasses][kRadix];
  const int n = countOf(count, hostCount, cap);
  const int t = threadIdx.x;
  const int passes = (bitHi - bitLo + 7) / 8;
  for (int p = 0; p < passes; p++) h[p][t] = 0;
  __syncthreads();
  for (int i = blockIdx.x * kPT + t; i < n; i += gridDim.x * kPT) {
    const K k = keys[i];
    for (int p = 0; p < passes; p++) {
      const int shift = bitLo + 8 * p;
      atomicAdd(&h[p][digitOf(k, shift, (1 << min(8, bitHi - shift)) - 1)], 1);
    }
  }
  __syncthreads();
  for (int p = 0; p < passes; p++)
    if (h[p][t]) atomicAdd(&hdr[kHdrHist + p * kRadix + t], h[p][t]);
  for (size_t w = (size_t)blockIdx.x * kPT + t; w < statusWords; w += (size_t)gridDim.x * kPT) status[w] = 0u;
}

// exclusive scan of every pass's digit counts into hdr[kHdrOfs]; counts and tickets left zero
__global__ __launch_bounds__(kPT) void kDigitScan(int passes, int* hdr) {
  __shared__ int ws[kPT / 64];
  __shared__ int tot;
  const int t = threadIdx.x;
  for (int p = 0; p < passes; p++) {
    const int v = hdr[kHdrHist + p * kRadix + t];
    hdr[kHdrOfs + p * kRadix + t] = blockExclusive<kPT>(v, ws, &tot);
    hdr[kHdrHist + p * kRadix + t] = 0;
    __syncthreads();
  }
  if (t < kMaxPasses) hdr[kHdrTicket + t] = 0;
}

// one digit pass: the tile of this workgroup's ticket, its keys in registers, look-back, stable scatter.
// Every 256-key chunk of the tile is ranked up front (eight ballots per wave give a key's rank among the wave's
// keys of its digit, and the wave's count of each digit), with no barrier between chunks; one scan per digit
// over the (chunk, wave) counts then gives the tile's digit counts (for the look-back) and every key's offset
// inside the tile, in chunk, wave, lane order: the order of a stable sort.
template <class K, bool Vals, int Per>
__global__ __launch_bounds__(kPT) void kOneSweep(const K* __restrict__ kin, const int* __restrict__ vin,
                                                 K* __restrict__ kout, int* __restrict__ vout,
                                                 const int* __restrict__ count, int hostCount, int cap, int shift,
                                                 int mask, int pass, int* hdr, uint32_t* status) {
  constexpr int kWaves = kPT / 64;
  __shared__ int sTile;
  __shared__ int running[kRadix];
  // per (chunk, wave, digit): the wave's count of the digit in the chunk, then its offset inside the tile
  __shared__ __align__(16) unsigned short wc[Per][kWaves][kRadix];
  const int n = countOf(count, hostCount, cap);
  constexpr int kTileKeys = kPT * Per;
  const int used = (n + kTileKeys - 1) / kTileKeys;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) sTile = atomicAdd(&hdr[kHdrTicket + pass], 1);
  for (int i = t; i < Per * kWaves * kRadix / 8; i += kPT) reinterpret_cast<uint4*>(&wc[0][0][0])[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  const int tile = sTile;
  if (tile >= used) return;  // the whole workgroup: no tile left
  const int base = tile * kTileKeys;
  K key[Per];
  int val[Per], dig[Per], rank[Per];
#pragma unroll
  for (int q = 0; q < Per; q++) {
    const int i = base + q * kPT + t;
    const bool ok = i < n;
    key[q] = ok ? kin[i] : (K)0;
    val[q] = (Vals && ok) ? vin[i] : 0;
    dig[q] = digitOf(key[q], shift, mask);
  }
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int q = 0; q < Per; q++) {
    const bool valid = base + q * kPT + t < n;
    const int d = dig[q];
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const unsigned long long bb = __ballot((d >> b) & 1);
      m &= ((d >> b) & 1) ? bb : ~bb;
    }
    rank[q] = __popcll(m & below);
    if (valid && rank[q] == 0) wc[q][w][d] = (unsigned short)__popcll(m);
  }
  __syncthreads();
  int mine = 0;
#pragma unroll
  for (int q = 0; q < Per; q++)
#pragma unroll
    for (int v = 0; v < kWaves; v++) {
      const int c = wc[q][v][t];
      wc[q][v][t] = (unsigned short)mine;
      mine += c;
    }
  // status[pass][tile][digit]: the tile's count at once (its inclusive prefix for tile 0), then the inclusive
  // prefix once the earlier tiles' counts are summed
  uint32_t* st = status + (size_t)pass * gridDim.x * kRadix;
  __hip_atomic_store(st + (size_t)tile * kRadix + t, (tile == 0 ? kStIncl : kStAgg) | (uint32_t)mine,
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int prefix = 0;
  if (tile > 0) {
    // the walk reads kLookback earlier tiles' words per round trip (independent loads), then sums them in order
    // until an inclusive prefix, or stops at a word not yet published and reads again from there
    int j = tile - 1;
    unsigned spins = 0;
    bool found = false;
    while (!found) {
      uint32_t sv[kLookback];
#pragma unroll
      for (int k = 0; k < kLookback; k++)
        sv[k] = j - k >= 0 ? __hip_atomic_load(st + (size_t)(j - k) * kRadix + t, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)
                           : 0u;  // never summed: tile 0's word is inclusive
      int taken = 0;
      bool stalled = false;
#pragma unroll
      for (int k = 0; k < kLookback; k++) {
        if (found || stalled) continue;
        const uint32_t f = sv[k] & ~kStMask;
        if (f == 0u) {
          stalled = true;
          continue;
        }
        prefix += (int)(sv[k] & kStMask);
        taken++;
        if (f == kStIncl) found = true;
      }
      j -= taken;
      if (stalled && !found) {
        if (++spins > (1u << 24)) {
          atomicOr(&hdr[kHdrErr], 1);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __hip_atomic_store(st + (size_t)tile * kRadix + t, kStIncl | (uint32_t)(prefix + mine), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  running[t] = hdr[kHdrOfs + pass * kRadix + t] + prefix;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Per; q++) {
    if (base + q * kPT + t < n) {
      const int d = dig[q];
      const int pos = running[d] + wc[q][w][d] + rank[q];
      kout[pos] = key[q];
      if (Vals) vout[pos] = val[q];
    }
  }
}

}  // namespace

int primTiles(int cap) { return (std::max(cap, 1) + kTile - 1) / kTile; }

// scratch: the sort header, the scans' tile sums (tiles + 64), the sort's look-back status (passes x tiles x 256)
size_t primScratchInts(int cap, int keyBytes) {
  (void)keyBytes;
  const int tiles = primTiles(cap);
  return (size_t)kSortHdr + (size_t)tiles + 64 + (size_t)kMaxPasses * sweepTiles(cap) * kRadix;
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
  if (bitHi <= bitLo) return false;
  if (cap >= (1 << 30)) throw std::length_error("radix sort of 2^30 keys or more");
  const int tiles = sweepTiles(cap);
  const int passes = (bitHi - bitLo + 7) / 8;
  if (passes > kMaxPasses) throw std::logic_error("radix sort of more than 64 bits");
  int* hdr = scratch;
  uint32_t* status = (uint32_t*)(scratch + kSortHdr + primTiles(cap) + 64);
  const size_t statusWords = (size_t)passes * tiles * kRadix;
  hipLaunchKernelGGL(kDigitHist<K>, dim3(std::min(primTiles(cap), 512)), dim3(kPT), 0, s, (const K*)k0, count, hostCount,
                     cap, bitLo, bitHi, hdr, status, statusWords);
  hipLaunchKernelGGL(kDigitScan, dim3(1), dim3(kPT), 0, s, passes, hdr);
  bool alt = false;
  for (int p = 0; p < passes; p++) {
    const int shift = bitLo + 8 * p;
    const int mask = (1 << std::min(8, bitHi - shift)) - 1;  // bits at or above bitHi take no part
    K* ki = alt ? k1 : k0;
    K* ko = alt ? k0 : k1;
    int* vi = alt ? v1 : v0;
    int* vo = alt ? v0 : v1;
    const int per = sweepPer(cap);
    const bool wide = per == 8;
    if (v0 && per == 16)
      hipLaunchKernelGGL((kOneSweep<K, true, 16>), dim3(tiles), dim3(kPT), 0, s, (const K*)ki, (const int*)vi, ko, vo,
                         count, hostCount, cap, shift, mask, p, hdr, status);
    else if (per == 16)
      hipLaunchKernelGGL((kOneSweep<K, false, 16>), dim3(tiles), dim3(kPT), 0, s, (const K*)ki, (const int*)nullptr, ko,
                         (int*)nullptr, count, hostCount, cap, shift, mask, p, hdr, status);
    else if (v0 && wide)
      hipLaunchKernelGGL((kOneSweep<K, true, 8>), dim3(tiles), dim3(kPT), 0, s, (const K*)ki, (const int*)vi, ko, vo,
                         count, hostCount, cap, shift, mask, p, hdr, status);
    else if (v0)
      hipLaunchKernelGGL((kOneSweep<K, true, 4>), dim3(tiles), dim3(kPT), 0, s, (const K*)ki, (const int*)vi, ko, vo,
                         count, hostCount, cap, shift, mask, p, hdr, status);
    else if (wide)
      hipLaunchKernelGGL((kOneSweep<K, false, 8>), dim3(tiles), dim3(kPT), 0, s, (const K*)ki, (const int*)nullptr, ko,
                         (int*)nullptr, count, hostCount, cap, shift, mask, p, hdr, status);
    else
      hipLaunchKernelGGL((kOneSweep<K, false, 4>), dim3(tiles), dim3(kPT), 0, s, (const K*)ki, (const int*)nullptr, ko,
                         (int*)nullptr, count, hostCount, cap, shift, mask, p, hdr, status);
    alt = !alt;
  }
  return alt;
}

template bool primRadixSort<uint32_t>(hipStream_t, uint32_t*, int*, uint32_t*, int*, const int*, int, int, int, int,
                                      int*);
template bool primRadixSort<uint64_t>(hipStream_t, uint64_t*, int*, uint64_t*, int*, const int*, int, int, int, int,
                                      int*);

// ---- the context-level forms used by the step: host-known counts, scratch in the context ----------------------
// the sort header starts the scratch (its counts must start zero: cleared whenever the buffer is new); the scans
// and compactions use the part behind it
static int* primScratch(Ctx& c, int cap) {
  const size_t before = c.primTmp.n;  // a new buffer always has a new capacity (its address may repeat)
  c.primTmp.ensure(primScratchInts(cap) + 64);
  if (c.primTmp.n != before) JB_CHECK(hipMemsetAsync(c.primTmp.p, 0, kSortHdr * sizeof(int), c.stream));
  return c.primTmp.p;
}
static int* primScanScratch(Ctx& c, int cap) { return primScratch(c, cap) + kSortHdr; }

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
template void primSortKeys<int>(Ctx&, Db<int>&, int, int, int);

void primSelectIndex(Ctx& c, const int* flag, int n, int* out, int* count) {
  if (n <= 0) {
    JB_CHECK(hipMemsetAsync(count, 0, sizeof(int), c.stream));
    return;
  }
  primSelectFlagged(c.stream, flag, nullptr, out, nullptr, n, n, primScanScratch(c, n), count);
}
template <class T>
void primSelectValues(Ctx& c, const T* vals, const int* flag, int n, T* out, int* count) {
  if (n <= 0) {
    JB_CHECK(hipMemsetAsync(count, 0, sizeof(int), c.stream));
    return;
  }
  const int tiles = primTiles(n);
  int* scratch = primScanScratch(c, n);
  hipLaunchKernelGGL(kFlagTileSums, dim3(tiles), dim3(kPT), 0, c.stream, flag, (const int*)nullptr, n, n, scratch);
  hipLaunchKernelGGL(kScanTileSums, dim3(1), dim3(kScanTop), 0, c.stream, tiles, scratch, count);
  hipLaunchKernelGGL((kSelectTiles<T, true>), dim3(tiles), dim3(kPT), 0, c.stream, flag, vals, (const int*)nullptr, n,
                     n, (const int*)scratch, out);
}
template void primSelectValues<uint64_t>(Ctx&, const uint64_t*, const int*, int, uint64_t*, int*);
template void primSelectValues<int>(Ctx&, const int*, const int*, int, int*, int*);

void primScan(Ctx& c, const int* in, int* out, int n) {
  if (n <= 0) return;
  primExclusiveScan(c.stream, in, out, nullptr, n, n, primScanScratch(c, n), nullptr);
}

// ---- device-count forms (graph-captured steps) -----------------------------------------------------------------
void primReserve(Ctx& c, int cap) { (void)primScratch(c, std::max(cap, 1)); }

void primErrCopy(Ctx& c) {
  c.primErrHost.ensure(1);
  c.primErrHost.p[0] = 0;
  if (c.primTmp.p)
    JB_CHECK(hipMemcpyAsync(c.primErrHost.p, c.primTmp.p + kHdrErr, sizeof(int), hipMemcpyDeviceToHost, c.stream));
}
void primErrCheck(Ctx& c) {
  if (c.primErrHost.p && c.primErrHost.p[0])
    throw std::runtime_error("a radix sort's look-back wait gave up (the step's results are not valid)");
}

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
  primSelectFlagged(c.stream, flag, nullptr, out, count, 0, cap, primScanScratch(c, cap), selected);
}
void primScanDev(Ctx& c, const int* in, int* out, const int* count, int cap) {
  cap = std::max(cap, 1);
  primExclusiveScan(c.stream, in, out, count, 0, cap, primScanScratch(c, cap), nullptr);
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
  primExclusiveScan(c.stream, in, c.primWide.p, nullptr, n, n, primScanScratch(c, n), nullptr);
  hipLaunchKernelGGL(kWiden, dim3(gridFor(n)), dim3(256), 0, c.stream, n, (const int*)c.primWide.p, out);
}

}  // namespace jbg
