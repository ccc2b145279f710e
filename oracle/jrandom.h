// Oracle (test infrastructure only): java.util.Random, restated from the
// algorithm documented in the Java SE API (48-bit LCG, multiplier 0x5DEECE66D,
// addend 0xB, seed scrambled with the multiplier).  The reference's solver draws
// its row shuffle from `new Random(0)` (SequentialImpulseConstraintSolver.java:108,
// randInt2 at :138-149).  Pinned by known answers in tests/test_oracle_pins.py.
#pragma once
#include <cstdint>

namespace jb {

struct JavaRandom {
  int64_t seed;
  explicit JavaRandom(int64_t s = 0) { setSeed(s); }
  void setSeed(int64_t s) { seed = (s ^ 0x5DEECE66DLL) & ((1LL << 48) - 1); }
  int32_t next(int bits) {
    seed = (int64_t)(((uint64_t)seed * 0x5DEECE66DULL + 0xBULL) & ((1ULL << 48) - 1));
    return (int32_t)((uint64_t)seed >> (48 - bits));
  }
  int32_t nextInt() { return next(32); }
  // nextInt(bound): power-of-two fast path, otherwise rejection loop
  int32_t nextInt(int32_t bound) {
    int32_t r = next(31);
    int32_t m = bound - 1;
    if ((bound & m) == 0) return (int32_t)(((int64_t)bound * (int64_t)r) >> 31);
    for (int32_t u = r; (int32_t)((uint32_t)u - (uint32_t)(r = u % bound) + (uint32_t)m) < 0; u = next(31)) {
    }
    return r;
  }
};

}  // namespace jb
