/* splitmix.h -- the synthetic-pixel generator shared by the oracle, the
 * reference-DSL driver, the tests and the GPU fill kernel.
 * SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014): the k-th output (k >= 1) of a generator seeded
 * with s is mix(s + k * 0x9e3779b97f4a7c15).  Pixel i (dense row-major) of an
 * image takes output k = i + 1.
 *   uint8  : x >> 56          float U[0,1) : (x >> 40) * 2^-24            */
#ifndef HIPACC_ORACLE_SPLITMIX_H
#define HIPACC_ORACLE_SPLITMIX_H
#include <stdint.h>

static inline uint64_t sm64_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + k * 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static inline uint8_t sm64_u8(uint64_t seed, uint64_t i) { return (uint8_t)(sm64_at(seed, i + 1) >> 56); }
static inline float sm64_f32(uint64_t seed, uint64_t i) {
  return (float)(sm64_at(seed, i + 1) >> 40) * (1.0f / 16777216.0f);
}
#endif
