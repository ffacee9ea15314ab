// Deviate sampling of exponential contraction, bit-compatible with the draw
// in RunContractionLP (lib/components/exponential_contraction.h:467-471):
//
//   std::exponential_distribution<double> distribution(beta);
//   std::default_random_engine generator((unsigned)(gid + seed + it*(rng_offset + seed)));
//   dev = (VertexID)(1.0 * distribution(generator));
//
// default_random_engine is minstd_rand0 (x <- 16807 x mod 2^31-1) in libstdc++.
// exponential_distribution draws u = generate_canonical<double,53>(g), which for
// this engine (range R = 2^31-2, 30 bits per call) takes two outputs:
//   u = ((x1-1) + (x2-1)*R) / (R*R), clamped below 1,
// and returns -log(1-u)/beta.  The sum is contracted to one fused multiply-add
// when the reference is compiled with -O3 -march=native on FMA hardware, so
// both forms are offered.
//
// The device never evaluates log: the host tabulates, for k = 1..K, the
// largest double W_k with trunc(-log(W_k)/beta) >= k (glibc's log, the one the
// reference calls), and the device counts the W_k >= 1-u.  The count equals
// the truncated deviate for every w because -log(w)/beta is non-increasing in w.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define KC_HD __host__ __device__
#else
#define KC_HD
#endif

namespace kc {

constexpr uint32_t kMinstdModulus = 2147483647u;  // 2^31 - 1

KC_HD inline uint32_t minstd_seed(uint32_t s) {
  uint32_t x = s % kMinstdModulus;
  return x == 0 ? 1u : x;
}

// x * 16807 mod (2^31 - 1) without a 64-bit division (Mersenne folding).
KC_HD inline uint32_t minstd_next(uint32_t x) {
  uint64_t t = static_cast<uint64_t>(x) * 16807ull;
  uint64_t r = (t & kMinstdModulus) + (t >> 31);
  if (r >= kMinstdModulus) r -= kMinstdModulus;
  return static_cast<uint32_t>(r);
}

// The engine's seed: (unsigned)(gid + seed + iteration*(rng_offset + seed)),
// all arithmetic modulo 2^64 in the reference, so modulo 2^32 here.
KC_HD inline uint32_t deviate_engine_seed(uint32_t gid, int32_t seed, uint32_t iteration,
                                          uint64_t rng_offset) {
  uint32_t s = static_cast<uint32_t>(seed);
  return gid + s + iteration * (static_cast<uint32_t>(rng_offset) + s);
}

KC_HD inline double canonical_draw(uint32_t engine_seed, bool fused) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const double r = 2147483646.0;  // max - min + 1 of minstd_rand0
  uint32_t x = minstd_seed(engine_seed);
  x = minstd_next(x);
  const double lo = static_cast<double>(x - 1u);
  x = minstd_next(x);
  const double hi = static_cast<double>(x - 1u);
  double sum;
  if (fused) {
    sum = __builtin_fma(hi, r, lo);
  } else {
    const double prod = hi * r;
    sum = lo + prod;
  }
  const double scale = r * r;
  double u = sum / scale;
  if (u >= 1.0) u = 0x1.fffffffffffffp-1;  // nextafter(1, 0)
  return u;
}

// Count of table entries >= w (the table is non-increasing).  The entries
// fall geometrically (entry k - 1 is about table[0]^k), so the count is about
// log(w) / log(table[0]); the search starts there and steps to the exact
// boundary, which takes a probe or two instead of a binary search's
// log2(k_max) dependent loads.  The comparisons alone fix the result, so the
// start is taken in single precision (a guess off by one costs a probe).
KC_HD inline uint32_t deviate_from_table(double w, const double *table, uint32_t k_max) {
  if (k_max == 0) return 0;
  const double t0 = table[0];
  uint32_t k = 0;
  if (w <= t0 && t0 < 1.0) {
    // (t0 within a float ulp of 1 gives an infinite or undefined quotient: start at 0)
    const float g = logf(static_cast<float>(w)) / logf(static_cast<float>(t0));
    k = g > 0.f ? (g < static_cast<float>(k_max) ? static_cast<uint32_t>(g) : k_max) : 0u;
  }
  while (k > 0 && table[k - 1] < w) --k;
  while (k < k_max && table[k] >= w) ++k;
  return k;
}

// Hashed stream: u from a SplitMix64 finaliser over (seed, iteration, gid),
// 53 bits, then the same exponential law through the same table.  Unlike
// the reference's per-vertex minstd_rand0 seeds (whose first outputs are
// linear in the seed, so the deviates of consecutive ids and iterations are
// correlated), the hashed draws are independent across vertices and levels.
// SplitMix64 output function (Steele, Lea, Flood 2014).
KC_HD inline uint64_t splitmix64_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
KC_HD inline double hashed_uniform(uint32_t gid, int32_t seed, uint32_t iteration) {
  const uint64_t key = (static_cast<uint64_t>(static_cast<uint32_t>(seed)) << 32 | iteration) * 0x9e3779b97f4a7c15ull;
  const uint64_t z = splitmix64_mix(splitmix64_mix(key) + static_cast<uint64_t>(gid) * 0xd1b54a32d192ed03ull);
  return static_cast<double>(z >> 11) * 0x1p-53;
}

// Uniform deviate in [0, 99] for local contraction: SplitMix64 over
// (seed, iteration, gid), reduced by Lemire's multiply-shift.
KC_HD inline uint32_t uniform_deviate_100(uint32_t gid, int32_t seed, uint32_t iteration) {
  uint64_t z = splitmix64_mix(0x9e3779b97f4a7c15ull * (static_cast<uint64_t>(static_cast<uint32_t>(seed)) + 1) ^
                              (static_cast<uint64_t>(iteration) << 32) ^ gid);
  return static_cast<uint32_t>(((z >> 32) * 100ull) >> 32);
}

// Proxy slot in [0, k) of the edge (v, w) of a hub v split into k proxies
// (kc_options.replicate): SplitMix64 over the ids and the seed, reduced by a
// multiply-shift of the upper 32 bits.  The reference draws it from
// default_random_engine(seed + v * n + w) (exponential_contraction.h:1066-1072);
// only the spread over the proxies matters, not the stream.
KC_HD inline uint32_t replica_pick(uint32_t gv, uint32_t gw, int32_t seed, uint32_t k) {
  const uint64_t z = splitmix64_mix((static_cast<uint64_t>(gv) << 32 | gw) ^
                                    (0x9e3779b97f4a7c15ull * (static_cast<uint64_t>(static_cast<uint32_t>(seed)) + 1)));
  return static_cast<uint32_t>(((z >> 32) * k) >> 32);
}

// Deviate of one vertex: stream 0 = reference draw, 1 = hashed draw.
KC_HD inline uint32_t exp_deviate(uint32_t gid, int32_t seed, uint32_t iteration, uint64_t rng_offset,
                                  bool fused, int stream, const double *table, uint32_t k_max) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const double u = stream == 0 ? canonical_draw(deviate_engine_seed(gid, seed, iteration, rng_offset), fused)
                               : hashed_uniform(gid, seed, iteration);
  const double w = 1.0 - u;
  return deviate_from_table(w, table, k_max);
}

}  // namespace kc
