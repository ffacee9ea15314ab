// Counter-based random streams of the synthetic generators, shared by the
// host generators (host/generators.cpp) and the device generators
// (csrc/gen_kernels.hip), so both produce the same edge multiset.  Every value
// is SplitMix64's output function (Steele, Lea, Flood 2014) applied to
// base + (i+1)*golden, with base derived from (seed, stream id).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define KC_GEN_HD __host__ __device__
#else
#define KC_GEN_HD
#endif

namespace kc {

KC_GEN_HD inline uint64_t gen_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct GenStream {
  uint64_t base;
};

// Stream ids: 1 gnm, 2 rmat, 3 grid bonds, 4 rgg points, 5 ba, 6 rhg points.
KC_GEN_HD inline GenStream gen_stream(int32_t seed, uint64_t stream) {
  return GenStream{gen_mix64(gen_mix64(static_cast<uint64_t>(static_cast<uint32_t>(seed)) ^ (stream << 40)) + stream)};
}

KC_GEN_HD inline uint64_t gen_value(GenStream s, uint64_t i) { return gen_mix64(s.base + (i + 1) * 0x9e3779b97f4a7c15ull); }

// floor(z * n / 2^64)
KC_GEN_HD inline uint64_t gen_bounded(uint64_t z, uint64_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(z, n);
#else
  return static_cast<uint64_t>((static_cast<unsigned __int128>(z) * n) >> 64);
#endif
}

KC_GEN_HD inline double gen_unit(uint64_t z) { return static_cast<double>(z >> 11) * 0x1p-53; }

// Edge i of G(n, m): two independent uniform endpoints.
KC_GEN_HD inline void gnm_edge(GenStream s, uint64_t i, uint64_t n, uint32_t *u, uint32_t *v) {
  *u = static_cast<uint32_t>(gen_bounded(gen_value(s, 2 * i), n));
  *v = static_cast<uint32_t>(gen_bounded(gen_value(s, 2 * i + 1), n));
}

// R-MAT quadrant thresholds in 1/2^32 units.
struct RmatThresholds {
  uint64_t a, ab, abc;
};
KC_GEN_HD inline RmatThresholds rmat_thresholds(double a, double b, double c) {
  return RmatThresholds{static_cast<uint64_t>(a * 4294967296.0), static_cast<uint64_t>((a + b) * 4294967296.0),
                        static_cast<uint64_t>((a + b + c) * 4294967296.0)};
}

// Edge i of R-MAT: scale recursive quadrant choices, most significant bit
// first, 32 random bits per choice (two choices per stream value).
KC_GEN_HD inline void rmat_edge(GenStream s, uint64_t i, uint32_t scale, RmatThresholds t, uint32_t *u, uint32_t *v) {
  const uint64_t words = (scale + 1) / 2;
  uint32_t uu = 0, vv = 0;
  uint64_t z = 0;
  for (uint32_t l = 0; l < scale; ++l) {
    if ((l & 1) == 0) z = gen_value(s, i * words + l / 2);
    const uint64_t r = (l & 1) ? (z >> 32) : (z & 0xffffffffull);
    uint32_t bu = 0, bv = 0;
    if (r < t.a) {
    } else if (r < t.ab) {
      bv = 1;
    } else if (r < t.abc) {
      bu = 1;
    } else {
      bu = 1;
      bv = 1;
    }
    uu = (uu << 1) | bu;
    vv = (vv << 1) | bv;
  }
  *u = uu;
  *v = vv;
}

// Bond keep test of the lattices: bond id = vertex id * dims + direction.
KC_GEN_HD inline uint64_t grid_keep_threshold(double p) {
  return p >= 1.0 ? ~0ull : static_cast<uint64_t>(p * 18446744073709551616.0);
}
KC_GEN_HD inline bool grid_bond_kept(GenStream s, uint64_t bond, double p, uint64_t thresh) {
  return p >= 1.0 || gen_value(s, bond) < thresh;
}

// Barabasi-Albert edge e (source vertex e / d): its target slot is uniform
// among the earlier slots; odd slots resolve recursively (Sanders, Schulz 2016).
KC_GEN_HD inline void ba_edge(GenStream s, uint64_t e, uint64_t d, uint32_t *u, uint32_t *v) {
  const uint64_t slot = 2 * e;
  uint64_t t = gen_bounded(gen_value(s, slot + 1), slot + 1);
  while (t & 1) t = gen_bounded(gen_value(s, t), t);
  *u = static_cast<uint32_t>(slot / 2 / d);
  *v = static_cast<uint32_t>(t / 2 / d);
}

}  // namespace kc
