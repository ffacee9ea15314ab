// CPU restatement of the learned centre proposals: per-vertex features and the
// MLP 6 -> 64 -> 64 -> 1 forward and backward in FP64.  TEST INFRASTRUCTURE
// ONLY (see oracle.h).  With bf16 set, values are rounded to bf16 exactly where
// the device rounds them (inputs, weights, first hidden layer, the two error
// signals fed to matrix cores), so the device result differs only by f32
// accumulation; without it the network is plain FP64, which the tests pin by
// finite differences.  The learned part has no reference counterpart.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

constexpr int kIn = 6, kHid = 64;

double bf16(double x) {
  float f = static_cast<float>(x);
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  std::memcpy(&f, &u, 4);
  return f;
}

struct Net {
  double W1[kIn][kHid], b1[kHid], W2[kHid][kHid], b2[kHid], w3[kHid], b3;
};

// Public parameter order: W1[6][64], b1, W2[64][64], b2, w3, b3.
Net unpack(const double *w, bool round) {
  Net n;
  const double *p = w;
  for (int k = 0; k < kIn; ++k)
    for (int c = 0; c < kHid; ++c) n.W1[k][c] = round ? bf16(*p++) : *p++;
  for (int c = 0; c < kHid; ++c) n.b1[c] = static_cast<float>(*p++);
  for (int k = 0; k < kHid; ++k)
    for (int c = 0; c < kHid; ++c) n.W2[k][c] = round ? bf16(*p++) : *p++;
  for (int c = 0; c < kHid; ++c) n.b2[c] = static_cast<float>(*p++);
  for (int c = 0; c < kHid; ++c) n.w3[c] = round ? bf16(*p++) : *p++;
  n.b3 = static_cast<float>(*p++);
  if (!round) {  // keep the FP64 values exactly (biases included)
    p = w + kIn * kHid;
    for (int c = 0; c < kHid; ++c) n.b1[c] = *p++;
    p += kHid * kHid;
    for (int c = 0; c < kHid; ++c) n.b2[c] = *p++;
    p += kHid;
    n.b3 = *p;
  }
  return n;
}

struct Act {
  double x[kIn], a1[kHid], h1[kHid], a2[kHid], h2[kHid], s;
};

void forward(const Net &n, const float *xin, bool round, Act &a) {
  for (int k = 0; k < kIn; ++k) a.x[k] = round ? bf16(xin[k]) : xin[k];
  for (int c = 0; c < kHid; ++c) {
    double z = 0;
    for (int k = 0; k < kIn; ++k) z += a.x[k] * n.W1[k][c];
    a.a1[c] = z + n.b1[c];
    const double h = a.a1[c] > 0 ? a.a1[c] : 0;
    a.h1[c] = round ? bf16(h) : h;
  }
  for (int c = 0; c < kHid; ++c) {
    double z = 0;
    for (int k = 0; k < kHid; ++k) z += a.h1[k] * n.W2[k][c];
    a.a2[c] = z + n.b2[c];
    a.h2[c] = a.a2[c] > 0 ? a.a2[c] : 0;
  }
  double s = 0;
  for (int c = 0; c < kHid; ++c) s += a.h2[c] * n.w3[c];
  a.s = s + n.b3;
}

}  // namespace

extern "C" {

void or_mlp_forward(const double *w, const float *X, uint64_t n, int32_t bf16_points, double *scores) {
  const bool round = bf16_points != 0;
  const Net net = unpack(w, round);
  Act a;
  for (uint64_t i = 0; i < n; ++i) {
    forward(net, X + i * kIn, round, a);
    scores[i] = a.s;
  }
}

// Gradient of 0.5 * mean((s - y)^2) in the public parameter order.
void or_mlp_grad(const double *w, const float *X, const float *y, uint64_t n, int32_t bf16_points, double *grads,
                 double *loss) {
  const bool round = bf16_points != 0;
  const Net net = unpack(w, round);
  std::vector<double> g(KC_ORACLE_MLP_PARAMS, 0.0);
  double *gW1 = g.data(), *gb1 = gW1 + kIn * kHid, *gW2 = gb1 + kHid, *gb2 = gW2 + kHid * kHid, *gw3 = gb2 + kHid,
         *gb3 = gw3 + kHid;
  double L = 0;
  Act a;
  const double inv = 1.0 / static_cast<double>(n);
  for (uint64_t i = 0; i < n; ++i) {
    forward(net, X + i * kIn, round, a);
    const double diff = a.s - y[i];
    L += 0.5 * diff * diff * inv;
    const double ds = diff * inv;
    *gb3 += ds;
    double dz2[kHid], dz2r[kHid];
    for (int c = 0; c < kHid; ++c) {
      gw3[c] += a.h2[c] * ds;
      dz2[c] = a.h2[c] > 0 ? ds * net.w3[c] : 0;
      gb2[c] += dz2[c];
      dz2r[c] = round ? bf16(dz2[c]) : dz2[c];
    }
    for (int k = 0; k < kHid; ++k) {
      double dh = 0;
      for (int c = 0; c < kHid; ++c) {
        gW2[k * kHid + c] += a.h1[k] * dz2r[c];
        dh += dz2r[c] * net.W2[k][c];
      }
      const double dz1 = a.a1[k] > 0 ? dh : 0;
      gb1[k] += dz1;
      const double dz1r = round ? bf16(dz1) : dz1;
      for (int f = 0; f < kIn; ++f) gW1[f * kHid + k] += a.x[f] * dz1r;
    }
  }
  std::memcpy(grads, g.data(), sizeof(double) * KC_ORACLE_MLP_PARAMS);
  if (loss) *loss = L;
}

// Features of every vertex (the device's k_features restated).
void or_level_features(uint64_t n, const uint64_t *off, const uint32_t *tgt, uint32_t round, float *X) {
  for (uint64_t v = 0; v < n; ++v) {
    const uint64_t e0 = off[v];
    const uint32_t d = static_cast<uint32_t>(off[v + 1] - off[v]);
    const uint32_t S = d < 4 ? d : 4;
    uint32_t u[4] = {0, 0, 0, 0};
    float sum = 0.f, mx = 0.f;
    for (uint32_t i = 0; i < S; ++i) {
      u[i] = tgt[e0 + (static_cast<uint64_t>(i) * d) / S];
      const float du = static_cast<float>(off[u[i] + 1] - off[u[i]]);
      sum += du;
      mx = du > mx ? du : mx;
    }
    const float mean = S ? sum / static_cast<float>(S) : 0.f;
    float clus = 0.f;
    const uint32_t pairs = S / 2;
    for (uint32_t p = 0; p < pairs; ++p) {
      const uint32_t a = u[2 * p], b = u[2 * p + 1];
      const uint64_t da = std::min<uint64_t>(off[a + 1] - off[a], 8);
      bool hit = false;
      for (uint64_t k = 0; k < da; ++k) hit |= tgt[off[a] + k] == b;
      clus += hit ? 1.f : 0.f;
    }
    float *x = X + v * kIn;
    x[0] = std::log2(1.f + static_cast<float>(d));
    x[1] = std::log2(1.f + mean);
    x[2] = std::log2(1.f + mx);
    x[3] = std::log2(1.f + static_cast<float>(d) * mean);
    x[4] = pairs ? clus / static_cast<float>(pairs) : 0.f;
    x[5] = static_cast<float>(round);
  }
}

}  // extern "C"
