// Counter-based SplitMix64 (Steele, Lea & Flood 2014; Vigna's splitmix64):
// output `counter` (0-based) of the stream seeded `seed`, so every draw is
// addressable without state. Used for the seeded mesh perturbation and the
// N(0,1) right-hand sides.
#pragma once
#include <cmath>
#include <cstdint>

namespace bpx {
namespace rng {

inline uint64_t u64(uint64_t seed, uint64_t counter) {
  uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// uniform in [0, 1) with 53 random bits
inline double u01(uint64_t seed, uint64_t counter) { return (double)(u64(seed, counter) >> 11) * 0x1.0p-53; }

// standard normal by Box-Muller on counters 2k and 2k+1
inline double normal(uint64_t seed, uint64_t counter) {
  const double u1 = 1.0 - u01(seed, 2 * counter);
  const double u2 = u01(seed, 2 * counter + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

}  // namespace rng
}  // namespace bpx
