// Learned centre proposals: per-vertex features, the MLP 6 -> 64 -> 64 -> 1
// (bf16 weights, f32 accumulation on MFMA) and its training state.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "kacomp.h"

namespace kc {

constexpr int kMlpIn = 6;      // features per vertex
constexpr int kMlpInPad = 16;  // input width of the first MFMA (rows 6..15 of W1 are zero)
constexpr int kMlpHid = 64;

// Device parameters in f32 (the master copy the optimiser updates).
struct MlpMaster {
  float W1[kMlpInPad * kMlpHid];  // [input][hidden]
  float b1[kMlpHid];
  float W2[kMlpHid * kMlpHid];    // [hidden 1][hidden 2]
  float b2[kMlpHid];
  float w3[kMlpHid];
  float b3;
};
constexpr int kMlpMasterFloats = static_cast<int>(sizeof(MlpMaster) / sizeof(float));

// bf16 copies of the weight matrices read by the kernels (biases stay f32).
struct MlpBf16 {
  uint16_t W1[kMlpInPad * kMlpHid];
  uint16_t W2[kMlpHid * kMlpHid];
  uint16_t w3[kMlpHid];
};

__host__ __device__ inline uint16_t f32_to_bf16(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return static_cast<uint16_t>(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  return __builtin_bit_cast(float, static_cast<uint32_t>(h) << 16);
}

}  // namespace kc

// Model handle: parameters, optimiser state and the trace dataset, all in HBM.
struct kc_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  kc::MlpMaster *master = nullptr;  // f32 parameters
  kc::MlpBf16 *bf = nullptr;        // bf16 weights
  float *adam_m = nullptr, *adam_v = nullptr, *grad = nullptr, *loss = nullptr;
  float *partials = nullptr;  // per-wave gradient partials of one training pass
  uint64_t partial_waves = 0;
  uint64_t adam_t = 0;
  uint64_t version = 1;  // bumped by every parameter change
  // trace dataset: features [cap x 6] and targets [cap]
  float *X = nullptr, *y = nullptr;
  uint64_t count = 0, cap = 0;
  uint32_t *idx = nullptr;
  uint64_t idx_cap = 0;
  std::vector<void *> owned;
  template <class T>
  T *dalloc(uint64_t n);
  ~kc_mlp();
};

namespace kc {

// Per-vertex degree (u32) of a level CSR.
void level_degrees(uint64_t n, const uint64_t *off, uint32_t *deg, hipStream_t s);
// Features of every vertex of a level: log2(1+deg), log2(1+mean) and
// log2(1+max) of up to four sampled neighbour degrees, log2(1+deg*mean) (a
// 2-hop size estimate), the adjacency rate of two sampled neighbour pairs
// (scanning up to 8 entries of the first one's row), and the round index.
void level_features(uint64_t n, const uint64_t *off, const uint32_t *col, const uint32_t *deg, uint32_t round,
                    float *X, hipStream_t s);
// Targets of the plain trace: y[v] = log2(size of the cluster v ended in).
void cluster_targets(uint64_t n, const uint32_t *lab, uint32_t *size, float *y, hipStream_t s);
// Scores of n feature rows.
void mlp_infer(const kc_mlp *m, uint64_t n, const float *X, float *score, hipStream_t s);
// Scores, then delay[v] = min(255, round(alpha * max(score[v], 0))).
void mlp_delays(const kc_mlp *m, uint64_t n, const float *X, float *score, float alpha, uint8_t *delay, hipStream_t s);
// The delays alone, from scores computed earlier.
void score_delays(uint64_t n, const float *score, float alpha, uint8_t *delay, hipStream_t s);
// Appends rows (every stride-th of n) of X / y to the model's dataset.
void mlp_append(kc_mlp *m, uint64_t n, const float *X, const float *y, uint64_t stride, hipStream_t s);

}  // namespace kc
