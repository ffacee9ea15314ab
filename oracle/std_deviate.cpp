// The deviate draw of RunContractionLP written with the standard library
// objects the reference calls (lib/components/exponential_contraction.h:467-471).
// Compiled twice: once with -ffp-contract=off (ORACLE_STD_NAME =
// or_std_deviate_nofma) and once with -mfma -ffp-contract=fast
// (or_std_deviate_fma), the two ways a compiler may evaluate
// generate_canonical's sum.  TEST INFRASTRUCTURE ONLY.
#include <cstdint>
#include <random>

#ifndef ORACLE_STD_NAME
#error define ORACLE_STD_NAME
#endif

extern "C" uint64_t ORACLE_STD_NAME(uint32_t gid, int32_t seed, uint32_t iteration, uint64_t rng_offset, double beta) {
  using VertexID = unsigned long long;
  const VertexID g = gid;
  const VertexID offset = rng_offset;
  const unsigned int it = iteration;
  const int s = seed;
  std::exponential_distribution<double> distribution(beta);
  std::default_random_engine generator(static_cast<unsigned int>(g + s + it * (offset + s)));
  const double weight = 1.0;
  return static_cast<VertexID>(weight * distribution(generator));
}

#ifdef ORACLE_STD_REPORT_FMA
extern "C" int32_t or_std_fma_compiled(void) {
#if defined(__FMA__)
  return 1;
#else
  return 0;
#endif
}
#endif
