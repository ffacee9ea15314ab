// Learned centre proposals (no counterpart in the reference, whose hook is the
// commented-out TIEBREAK_DEGREE weighting of RunContractionLP,
// lib/components/exponential_contraction.h:459-466).
//
// A per-vertex MLP 6 -> 64 -> 64 -> 1 with bf16 weights and f32 accumulation
// scores every vertex of a contraction level from local features.  The score
// is trained on traces of the plain method to predict log2 of the size of the
// cluster the vertex ends in; at run time round(alpha * max(score, 0)) is added
// to the vertex's deviate, so vertices expected to be absorbed by a large
// cluster start later and fewer small clusters survive to the next level.  Labels do
// not depend on it: any deviates yield the same components.
//
// Both layers of 64 run on v_mfma_f32_16x16x16_bf16 (wave64), 16 vertices per
// tile.  Fragment layout for lane L (r = L % 16, g = L / 16, j = 0..3):
//   A (16 x 16, row x k):  A[r][4g + j]
//   B (16 x 16, k x col):  B[4g + j][r]
//   D (16 x 16, row x col): D[4g + j][r]
// Hidden activations move between the D and A layouts through LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "gen_rng.h"
#include "kc_common.h"
#include "kernels.h"
#include "learned.h"

template <class T>
T *kc_mlp::dalloc(uint64_t n) {
  void *p = nullptr;
  const size_t bytes = std::max<uint64_t>(n, 1) * sizeof(T) + 64;
  if (hipMalloc(&p, bytes) != hipSuccess) kc::fail(KC_ERR_OOM, "mlp: hipMalloc of %zu bytes failed", bytes);
  owned.push_back(p);
  return static_cast<T *>(p);
}

kc_mlp::~kc_mlp() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (void *p : owned) (void)hipFree(p);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace kc {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kRowPad = 72;  // LDS row stride of a 16 x 64 bf16 tile (breaks bank aliasing)
constexpr uint64_t kPartialStride = kMlpMasterFloats + 3;  // floats per wave partial (parameters, loss)
constexpr int kXPad = 20;    // LDS row stride of a 16 x 16 bf16 input tile

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4s frag(uint16_t a, uint16_t b, uint16_t c, uint16_t d) {
  v4s r;
  r.x = static_cast<short>(a);
  r.y = static_cast<short>(b);
  r.z = static_cast<short>(c);
  r.w = static_cast<short>(d);
  return r;
}

__device__ __forceinline__ float sum16(float x) {  // over the 16 lanes of a row group
  for (int o = 1; o < 16; o <<= 1) x += __shfl_xor(x, o);
  return x;
}
__device__ __forceinline__ float sum_groups(float x) {  // over the 4 row groups (same r)
  x += __shfl_xor(x, 16);
  x += __shfl_xor(x, 32);
  return x;
}

// ---------------------------------------------------------------------------
// Features and targets.
__global__ void k_degrees(uint64_t n, const uint64_t *off, uint32_t *deg) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    deg[v] = static_cast<uint32_t>(off[v + 1] - off[v]);
}

// Loads are issued in four dependent phases (sampled neighbour ids, their
// degrees, the row starts of the pair heads, the first 8 entries of those
// rows) with no branch between them, so every phase's loads are in flight
// together; unused slots repeat sample 0 and are masked out afterwards.
__global__ void k_features(uint64_t n, const uint64_t *off, const uint32_t *col, const uint32_t *deg, float round,
                           float *X) {
  const uint64_t m2 = off[n];
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t e0 = off[v];
    const uint32_t d = deg[v];
    const uint32_t S = d < 4 ? d : 4;
    float sum = 0.f, mx = 0.f, clus = 0.f;
    const uint32_t pairs = S / 2;
    if (d > 0) {
      uint32_t u[4], du[4];
      // sample i sits at floor(i d / S): i d / 4 when d >= 4, else i (S = d)
#pragma unroll
      for (uint32_t i = 0; i < 4; ++i)
        u[i] = col[e0 + (d >= 4 ? (static_cast<uint64_t>(i) * d) >> 2 : (i < d ? i : 0u))];
#pragma unroll
      for (uint32_t i = 0; i < 4; ++i) du[i] = deg[u[i]];
#pragma unroll
      for (uint32_t i = 0; i < 4; ++i)
        if (i < S) {
          const float f = static_cast<float>(du[i]);
          sum += f;
          mx = f > mx ? f : mx;
        }
      // the first min(deg, 8) entries of a pair head's row lie inside the 12
      // entries from the row start rounded down to a multiple of 4: three
      // aligned 16-byte loads (entry-wise near the end of the array)
      uint64_t r0[2];
#pragma unroll
      for (uint32_t p = 0; p < 2; ++p) r0[p] = off[u[2 * p]];
      uint4 w[2][3];
      bool vec[2];
      const uint64_t last = m2 >= 12 ? (m2 - 12) & ~3ull : 0;  // a window start that stays inside col
#pragma unroll
      for (uint32_t p = 0; p < 2; ++p) {
        const uint64_t a0 = r0[p] & ~3ull;
        vec[p] = m2 >= 12 && a0 + 12 <= m2;
        const uint4 *src = reinterpret_cast<const uint4 *>(col + (vec[p] ? a0 : last));
        if (m2 >= 12) {
#pragma unroll
          for (int q = 0; q < 3; ++q) w[p][q] = src[q];
        }
      }
#pragma unroll
      for (uint32_t p = 0; p < 2; ++p) {
        if (p >= pairs) continue;
        const uint32_t b = u[2 * p + 1], da = du[2 * p] < 8 ? du[2 * p] : 8u;
        bool hit = false;
        if (vec[p]) {
          const uint32_t sh = static_cast<uint32_t>(r0[p] & 3ull), lim = sh + da;
          const uint32_t c[12] = {w[p][0].x, w[p][0].y, w[p][0].z, w[p][0].w, w[p][1].x, w[p][1].y,
                                  w[p][1].z, w[p][1].w, w[p][2].x, w[p][2].y, w[p][2].z, w[p][2].w};
#pragma unroll
          for (uint32_t j = 0; j < 12; ++j) hit |= j >= sh && j < lim && c[j] == b;
        } else {
          for (uint32_t k = 0; k < da; ++k) hit |= col[r0[p] + k] == b;
        }
        clus += hit ? 1.f : 0.f;
      }
    }
    const float mean = S ? sum / static_cast<float>(S) : 0.f;
    float *x = X + v * kMlpIn;
    x[0] = log2f(1.f + static_cast<float>(d));
    x[1] = log2f(1.f + mean);
    x[2] = log2f(1.f + mx);
    x[3] = log2f(1.f + static_cast<float>(d) * mean);
    x[4] = pairs ? clus / static_cast<float>(pairs) : 0.f;
    x[5] = round;
  }
}

__global__ void k_cluster_sizes(uint64_t n, const uint32_t *lab, uint32_t *size) {
  for (uint64_t base = blockIdx.x * static_cast<uint64_t>(blockDim.x); base < n;
       base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t v = base + threadIdx.x;
    const bool live = v < n;
    peeled_fetch_add(size, live ? lab[v] : 0u, 1u, live);
  }
}

__global__ void k_cluster_targets(uint64_t n, const uint32_t *lab, const uint32_t *size, float *y) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    y[v] = log2f(static_cast<float>(size[lab[v]]));
}

// ---------------------------------------------------------------------------
// Inference: 4 waves per block, one 16-vertex tile per wave per step, every
// layer computed transposed (weights as the A operand, vertices as columns).
// For v_mfma_f32_16x16x16_bf16 lane (r, g) holds A[r][4g..4g+3],
// B[4g..4g+3][r] and D[4g..4g+3][r]; so the transposed layer-1 output a lane
// holds, H[vertex r][hidden 4g..4g+3 of a block], is exactly its B fragment
// of that block for layer 2: the hidden layer never leaves the registers.
__global__ __launch_bounds__(256) void k_mlp_infer(uint64_t n, const float *X, const MlpBf16 *wb, const MlpMaster *wm,
                                                  float *score) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  v4s a1[4], a2[4][4];  // W1^T rows of hidden block nb; W2^T rows of block nb2 over hidden-1 block kb
  float b1v[4][4], b2v[4][4], w3v[4][4];
  for (int nb = 0; nb < 4; ++nb) {
    const int h = nb * 16 + r;
    a1[nb] = frag(wb->W1[(4 * g) * kMlpHid + h], wb->W1[(4 * g + 1) * kMlpHid + h], wb->W1[(4 * g + 2) * kMlpHid + h],
                  wb->W1[(4 * g + 3) * kMlpHid + h]);
    for (int kb = 0; kb < 4; ++kb) {
      const int k = kb * 16 + 4 * g;
      a2[kb][nb] = frag(wb->W2[k * kMlpHid + h], wb->W2[(k + 1) * kMlpHid + h], wb->W2[(k + 2) * kMlpHid + h],
                        wb->W2[(k + 3) * kMlpHid + h]);
    }
    for (int j = 0; j < 4; ++j) {
      const int c = nb * 16 + 4 * g + j;
      b1v[nb][j] = wm->b1[c];
      b2v[nb][j] = wm->b2[c];
      w3v[nb][j] = bf16_to_f32(wb->w3[c]);
    }
  }
  const float b3 = wm->b3;
  const v4f zero = {0.f, 0.f, 0.f, 0.f};
  const uint64_t tiles = (n + 15) / 16;
  const uint64_t step = static_cast<uint64_t>(gridDim.x) * 4;
  // the features of a wave's next tile are loaded while it computes this one
  auto load_x = [&](uint64_t tile, float (&xf)[4]) {
    const uint64_t row = tile * 16 + r;
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * g + j;
      xf[j] = row < n && k < kMlpIn ? X[row * kMlpIn + k] : 0.f;
    }
  };
  float xnext[4];
  load_x(static_cast<uint64_t>(blockIdx.x) * 4 + wave, xnext);
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * 4; base < tiles; base += step) {
    const uint64_t tile = base + wave;
    const uint64_t row = tile * 16 + r;
    const float xf[4] = {xnext[0], xnext[1], xnext[2], xnext[3]};
    load_x(tile + step, xnext);
    const v4s xb = frag(f32_to_bf16(xf[0]), f32_to_bf16(xf[1]), f32_to_bf16(xf[2]),
                        f32_to_bf16(xf[3]));  // X^T: inputs 4g.., vertex r
    v4s hb[4];                                         // H^T fragments, hidden block kb
    for (int nb = 0; nb < 4; ++nb) {
      const v4f z = mfma(a1[nb], xb, zero);
      uint16_t hv[4];
      for (int j = 0; j < 4; ++j) {
        const float h = z[j] + b1v[nb][j];
        hv[j] = f32_to_bf16(h > 0.f ? h : 0.f);
      }
      hb[nb] = frag(hv[0], hv[1], hv[2], hv[3]);
    }
    float acc = 0.f;
    for (int nb = 0; nb < 4; ++nb) {
      v4f z2 = zero;
      for (int kb = 0; kb < 4; ++kb) z2 = mfma(a2[kb][nb], hb[kb], z2);
      for (int j = 0; j < 4; ++j) {
        const float h = z2[j] + b2v[nb][j];
        acc += (h > 0.f ? h : 0.f) * w3v[nb][j];
      }
    }
    acc = sum_groups(acc);  // over the 4 lane groups holding vertex r
    if (g == 0 && row < n) score[row] = acc + b3;
  }
}

// delay[v] = min(255, round(alpha * max(score[v], 0))).
__global__ void k_delays(uint64_t n, const float *score, float alpha, uint8_t *delay) {
  for (uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; v < n;
       v += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const float d = score[v];
    const float dl = alpha * (d > 0.f ? d : 0.f) + 0.5f;
    delay[v] = static_cast<uint8_t>(dl >= 255.f ? 255.f : dl);
  }
}

// ---------------------------------------------------------------------------
// Training: forward, loss 0.5 * mean((s - y)^2) over the batch, backward.
// Every wave accumulates its gradient tiles in registers over its sample
// tiles and adds them to grad (MlpMaster layout) once at the end.
__global__ __launch_bounds__(256) void k_mlp_train(uint64_t bt, const uint32_t *idx, const float *X, const float *y,
                                                  const MlpBf16 *wb, const MlpMaster *wm, float *partials) {
  __shared__ uint16_t xs_all[4][16 * kXPad];
  __shared__ uint16_t h1_all[4][16 * kRowPad];
  __shared__ uint16_t dz2_all[4][16 * kRowPad];
  __shared__ uint16_t dz1_all[4][16 * kRowPad];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  uint16_t *xs = xs_all[wave], *h1 = h1_all[wave], *dz2s = dz2_all[wave], *dz1s = dz1_all[wave];
  v4s b1f[4], b2f[4][4], b2t[4][4];
  float b1v[4], b2v[4], w3v[4];
  for (int nb = 0; nb < 4; ++nb) {
    const int c = nb * 16 + r;
    b1f[nb] = frag(wb->W1[(4 * g) * kMlpHid + c], wb->W1[(4 * g + 1) * kMlpHid + c], wb->W1[(4 * g + 2) * kMlpHid + c],
                   wb->W1[(4 * g + 3) * kMlpHid + c]);
    for (int kb = 0; kb < 4; ++kb) {
      const int k = kb * 16 + 4 * g;
      b2f[kb][nb] = frag(wb->W2[k * kMlpHid + c], wb->W2[(k + 1) * kMlpHid + c], wb->W2[(k + 2) * kMlpHid + c],
                         wb->W2[(k + 3) * kMlpHid + c]);
      // W2 transposed: B[k = hidden 2][col = hidden 1] = W2[hidden 1][hidden 2]
      const uint16_t *row = wb->W2 + c * kMlpHid + k;
      b2t[kb][nb] = frag(row[0], row[1], row[2], row[3]);
    }
    b1v[nb] = wm->b1[c];
    b2v[nb] = wm->b2[c];
    w3v[nb] = bf16_to_f32(wb->w3[c]);
  }
  const float b3 = wm->b3;
  const float inv_bt = 1.f / static_cast<float>(bt);
  const v4f zero = {0.f, 0.f, 0.f, 0.f};
  v4f gW1[4] = {zero, zero, zero, zero};
  v4f gW2[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) gW2[a][b] = zero;
  float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gw3[4] = {0, 0, 0, 0}, gb3 = 0.f, lsum = 0.f;

  const uint64_t tiles = (bt + 15) / 16;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * 4; base < tiles; base += static_cast<uint64_t>(gridDim.x) * 4) {
    const uint64_t tile = base + wave;
    // input tile (row r of this lane's A fragment) and its copy in LDS
    uint16_t xa[4] = {0, 0, 0, 0};
    const uint64_t q = tile * 16 + r;
    if (q < bt) {
      const uint64_t sidx = idx[q];
      for (int j = 0; j < 4; ++j) {
        const int k = 4 * g + j;
        if (k < kMlpIn) xa[j] = f32_to_bf16(X[sidx * kMlpIn + k]);
      }
    }
    for (int j = 0; j < 4; ++j) xs[r * kXPad + 4 * g + j] = xa[j];
    const v4s a1 = frag(xa[0], xa[1], xa[2], xa[3]);
    bool m1[4][4];
    for (int nb = 0; nb < 4; ++nb) {
      const v4f z = mfma(a1, b1f[nb], zero);
      for (int j = 0; j < 4; ++j) {
        const float h = z[j] + b1v[nb];
        m1[nb][j] = h > 0.f;
        h1[(4 * g + j) * kRowPad + nb * 16 + r] = f32_to_bf16(h > 0.f ? h : 0.f);
      }
    }
    __syncthreads();
    v4f z2[4] = {zero, zero, zero, zero};
    for (int kb = 0; kb < 4; ++kb) {
      const uint16_t *src = h1 + r * kRowPad + kb * 16 + 4 * g;
      const v4s a2 = frag(src[0], src[1], src[2], src[3]);
      for (int nb = 0; nb < 4; ++nb) z2[nb] = mfma(a2, b2f[kb][nb], z2[nb]);
    }
    float h2[4][4];
    float ds[4];
    for (int j = 0; j < 4; ++j) {
      float acc = 0.f;
      for (int nb = 0; nb < 4; ++nb) {
        const float h = z2[nb][j] + b2v[nb];
        h2[nb][j] = h > 0.f ? h : 0.f;
        acc += h2[nb][j] * w3v[nb];
      }
      const float s = sum16(acc) + b3;
      const uint64_t qq = tile * 16 + 4 * g + j;
      const bool valid = qq < bt;
      const float diff = valid ? s - y[idx[qq]] : 0.f;
      ds[j] = diff * inv_bt;
      if (r == 0) {
        lsum += 0.5f * diff * diff * inv_bt;
        gb3 += ds[j];
      }
    }
    for (int nb = 0; nb < 4; ++nb)
      for (int j = 0; j < 4; ++j) {
        gw3[nb] += h2[nb][j] * ds[j];
        const float d2 = h2[nb][j] > 0.f ? ds[j] * w3v[nb] : 0.f;
        gb2[nb] += d2;
        dz2s[(4 * g + j) * kRowPad + nb * 16 + r] = f32_to_bf16(d2);
      }
    __syncthreads();
    // dW2 += H1^T dZ2 ; dH1 = dZ2 W2^T
    v4s bd[4];
    for (int nb = 0; nb < 4; ++nb) {
      const uint16_t *c0 = dz2s + (4 * g) * kRowPad + nb * 16 + r;
      bd[nb] = frag(c0[0], c0[kRowPad], c0[2 * kRowPad], c0[3 * kRowPad]);
    }
    for (int mb = 0; mb < 4; ++mb) {
      const uint16_t *c0 = h1 + (4 * g) * kRowPad + mb * 16 + r;
      const v4s at = frag(c0[0], c0[kRowPad], c0[2 * kRowPad], c0[3 * kRowPad]);
      for (int nb = 0; nb < 4; ++nb) gW2[mb][nb] = mfma(at, bd[nb], gW2[mb][nb]);
    }
    v4f dh1[4] = {zero, zero, zero, zero};
    for (int kb = 0; kb < 4; ++kb) {
      const uint16_t *src = dz2s + r * kRowPad + kb * 16 + 4 * g;
      const v4s ad = frag(src[0], src[1], src[2], src[3]);
      for (int nb = 0; nb < 4; ++nb) dh1[nb] = mfma(ad, b2t[kb][nb], dh1[nb]);
    }
    for (int nb = 0; nb < 4; ++nb)
      for (int j = 0; j < 4; ++j) {
        const float d1 = m1[nb][j] ? dh1[nb][j] : 0.f;
        gb1[nb] += d1;
        dz1s[(4 * g + j) * kRowPad + nb * 16 + r] = f32_to_bf16(d1);
      }
    __syncthreads();
    // dW1 += X^T dZ1
    const uint16_t *xc = xs + (4 * g) * kXPad + r;
    const v4s ax = frag(xc[0], xc[kXPad], xc[2 * kXPad], xc[3 * kXPad]);
    for (int nb = 0; nb < 4; ++nb) {
      const uint16_t *c0 = dz1s + (4 * g) * kRowPad + nb * 16 + r;
      const v4s bz = frag(c0[0], c0[kRowPad], c0[2 * kRowPad], c0[3 * kRowPad]);
      gW1[nb] = mfma(ax, bz, gW1[nb]);
    }
    __syncthreads();
  }
  // this wave's partial gradient (every entry written once; summed in a fixed
  // order by k_sum_partials, so training is reproducible run to run)
  MlpMaster *grad = reinterpret_cast<MlpMaster *>(partials + static_cast<uint64_t>(blockIdx.x * 4 + wave) *
                                                                  kPartialStride);
  for (int nb = 0; nb < 4; ++nb) {
    const int c = nb * 16 + r;
    for (int j = 0; j < 4; ++j) {
      const int k1 = 4 * g + j;
      grad->W1[k1 * kMlpHid + c] = gW1[nb][j];
      for (int mb = 0; mb < 4; ++mb) grad->W2[(mb * 16 + k1) * kMlpHid + c] = gW2[mb][nb][j];
    }
    const float v1 = sum_groups(gb1[nb]), v2 = sum_groups(gb2[nb]), v3 = sum_groups(gw3[nb]);
    if (g == 0) {
      grad->b1[c] = v1;
      grad->b2[c] = v2;
      grad->w3[c] = v3;
    }
  }
  // gb3 and the loss live on lanes r == 0 (one per row group)
  const float t3 = sum_groups(gb3), tl = sum_groups(lsum);
  if (lane == 0) {
    grad->b3 = t3;
    partials[static_cast<uint64_t>(blockIdx.x * 4 + wave) * kPartialStride + kMlpMasterFloats] = tl;
  }
}

// grad[i] = sum over waves of partials[w][i] (i < kMlpMasterFloats), loss
// likewise from the slot after the parameters.
__global__ void k_sum_partials(uint32_t waves, const float *partials, float *grad, float *loss) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= static_cast<uint32_t>(kMlpMasterFloats);
       i += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (uint32_t w = 0; w < waves; ++w) s += partials[static_cast<uint64_t>(w) * kPartialStride + i];
    if (i < static_cast<uint32_t>(kMlpMasterFloats))
      grad[i] = s;
    else
      *loss = s;
  }
}

// Adam on the f32 parameters, then the bf16 copies.
__global__ void k_adam(uint32_t count, float *p, float *m, float *v, const float *gr, float lr, float b1, float b2,
                       float eps, float bc1, float bc2) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const float gi = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  }
}

__global__ void k_quantize(const MlpMaster *w, MlpBf16 *b) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kMlpInPad * kMlpHid; i += gridDim.x * blockDim.x)
    b->W1[i] = f32_to_bf16(w->W1[i]);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kMlpHid * kMlpHid; i += gridDim.x * blockDim.x)
    b->W2[i] = f32_to_bf16(w->W2[i]);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kMlpHid; i += gridDim.x * blockDim.x)
    b->w3[i] = f32_to_bf16(w->w3[i]);
}

// Mini-batch indices: uniform over [0, count) from a counter-based stream.
__global__ void k_batch_idx(uint32_t bt, uint64_t count, GenStream s, uint64_t step, uint32_t *idx) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < bt; i += gridDim.x * blockDim.x)
    idx[i] = static_cast<uint32_t>(gen_bounded(gen_value(s, step * bt + i), count));
}

__global__ void k_append(uint64_t n, uint64_t stride, const float *X, const float *y, float *dX, float *dy) {
  const uint64_t rows = (n + stride - 1) / stride;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < rows;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t v = i * stride;
    for (int k = 0; k < kMlpIn; ++k) dX[i * kMlpIn + k] = X[v * kMlpIn + k];
    dy[i] = y[v];
  }
}

unsigned grid_of(uint64_t n, uint64_t per = kBlock, uint64_t cap = 65536) {
  uint64_t g = (n + per - 1) / per;
  if (g == 0) g = 1;
  return static_cast<unsigned>(g < cap ? g : cap);
}

// Public parameter order (KC_MLP_PARAMS floats): W1 [6][64], b1, W2 [64][64], b2, w3, b3.
void to_public(const MlpMaster &w, float *out) {
  float *o = out;
  for (int k = 0; k < kMlpIn; ++k)
    for (int c = 0; c < kMlpHid; ++c) *o++ = w.W1[k * kMlpHid + c];
  for (int c = 0; c < kMlpHid; ++c) *o++ = w.b1[c];
  for (int i = 0; i < kMlpHid * kMlpHid; ++i) *o++ = w.W2[i];
  for (int c = 0; c < kMlpHid; ++c) *o++ = w.b2[c];
  for (int c = 0; c < kMlpHid; ++c) *o++ = w.w3[c];
  *o++ = w.b3;
}

void from_public(const float *in, MlpMaster &w) {
  std::memset(&w, 0, sizeof(w));
  const float *p = in;
  for (int k = 0; k < kMlpIn; ++k)
    for (int c = 0; c < kMlpHid; ++c) w.W1[k * kMlpHid + c] = *p++;
  for (int c = 0; c < kMlpHid; ++c) w.b1[c] = *p++;
  for (int i = 0; i < kMlpHid * kMlpHid; ++i) w.W2[i] = *p++;
  for (int c = 0; c < kMlpHid; ++c) w.b2[c] = *p++;
  for (int c = 0; c < kMlpHid; ++c) w.w3[c] = *p++;
  w.b3 = *p++;
}

void upload(kc_mlp *m, const MlpMaster &w) {
  KC_HIP(hipMemcpyAsync(m->master, &w, sizeof(w), hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_quantize, dim3(16), dim3(kBlock), 0, m->stream, m->master, m->bf);
  KC_LAUNCH_CHECK();
  KC_HIP(hipStreamSynchronize(m->stream));
  ++m->version;
}

// One pass of k_mlp_train over bt rows of idx: per-wave partials, then their
// fixed-order sum into m->grad and m->loss.
void train_pass(kc_mlp *m, uint64_t bt, const uint32_t *idx, const float *X, const float *y, hipStream_t s) {
  const uint64_t tiles = (bt + 15) / 16;
  // about 8 tiles per wave
  const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>(1024, (tiles + 31) / 32));
  const uint64_t waves = blocks * 4;
  if (waves > m->partial_waves) {
    m->partials = m->dalloc<float>(waves * kPartialStride);
    m->partial_waves = waves;
  }
  hipLaunchKernelGGL(k_mlp_train, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, bt, idx, X, y, m->bf, m->master,
                     m->partials);
  KC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_partials, dim3(grid_of(kMlpMasterFloats + 1)), dim3(kBlock), 0, s,
                     static_cast<uint32_t>(waves), m->partials, m->grad, m->loss);
  KC_LAUNCH_CHECK();
}

}  // namespace

void level_degrees(uint64_t n, const uint64_t *off, uint32_t *deg, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_degrees, dim3(grid_of(n)), dim3(kBlock), 0, s, n, off, deg);
  KC_LAUNCH_CHECK();
}

void level_features(uint64_t n, const uint64_t *off, const uint32_t *col, const uint32_t *deg, uint32_t round,
                    float *X, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_features, dim3(grid_of(n)), dim3(kBlock), 0, s, n, off, col, deg, static_cast<float>(round), X);
  KC_LAUNCH_CHECK();
}

void cluster_targets(uint64_t n, const uint32_t *lab, uint32_t *size, float *y, hipStream_t s) {
  if (!n) return;
  fill_u32(size, n, 0u, s);
  hipLaunchKernelGGL(k_cluster_sizes, dim3(grid_of(n)), dim3(kBlock), 0, s, n, lab, size);
  KC_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cluster_targets, dim3(grid_of(n)), dim3(kBlock), 0, s, n, lab, size, y);
  KC_LAUNCH_CHECK();
}

void mlp_infer(const kc_mlp *m, uint64_t n, const float *X, float *score, hipStream_t s) {
  if (!n) return;
  const uint64_t tiles = (n + 15) / 16;
  hipLaunchKernelGGL(k_mlp_infer, dim3(grid_of(tiles, 4, 16384)), dim3(256), 0, s, n, X, m->bf, m->master, score);
  KC_LAUNCH_CHECK();
}

void mlp_delays(const kc_mlp *m, uint64_t n, const float *X, float *score, float alpha, uint8_t *delay,
                hipStream_t s) {
  if (!n) return;
  mlp_infer(m, n, X, score, s);
  hipLaunchKernelGGL(k_delays, dim3(grid_of(n)), dim3(kBlock), 0, s, n, score, alpha, delay);
  KC_LAUNCH_CHECK();
}

void score_delays(uint64_t n, const float *score, float alpha, uint8_t *delay, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_delays, dim3(grid_of(n)), dim3(kBlock), 0, s, n, score, alpha, delay);
  KC_LAUNCH_CHECK();
}

void mlp_append(kc_mlp *m, uint64_t n, const float *X, const float *y, uint64_t stride, hipStream_t s) {
  if (!n) return;
  if (stride < 1) stride = 1;
  const uint64_t rows = (n + stride - 1) / stride;
  if (m->count + rows > m->cap) {
    // grow: keep what is there
    const uint64_t cap = std::max<uint64_t>(2 * m->cap, m->count + rows + (1u << 20));
    float *nX = m->dalloc<float>(cap * kMlpIn), *ny = m->dalloc<float>(cap);
    if (m->count) {
      KC_HIP(hipMemcpyAsync(nX, m->X, m->count * kMlpIn * sizeof(float), hipMemcpyDeviceToDevice, s));
      KC_HIP(hipMemcpyAsync(ny, m->y, m->count * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    m->X = nX;
    m->y = ny;
    m->cap = cap;
  }
  hipLaunchKernelGGL(k_append, dim3(grid_of(rows)), dim3(kBlock), 0, s, n, stride, X, y, m->X + m->count * kMlpIn,
                     m->y + m->count);
  KC_LAUNCH_CHECK();
  m->count += rows;
}

}  // namespace kc

using namespace kc;

extern "C" {

int kc_mlp_create(int32_t seed, kc_mlp **out) {
  return guarded([&] {
    if (!out) fail(KC_ERR_INVALID, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) fail(KC_ERR_HIP, "no HIP device available");
    std::unique_ptr<kc_mlp> m(new kc_mlp());
    KC_HIP(hipGetDevice(&m->device));
    KC_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    m->master = m->dalloc<MlpMaster>(1);
    m->bf = m->dalloc<MlpBf16>(1);
    m->adam_m = m->dalloc<float>(kMlpMasterFloats);
    m->adam_v = m->dalloc<float>(kMlpMasterFloats);
    m->grad = m->dalloc<float>(kMlpMasterFloats);
    m->loss = m->dalloc<float>(1);
    fill_u32(reinterpret_cast<uint32_t *>(m->adam_m), kMlpMasterFloats, 0u, m->stream);
    fill_u32(reinterpret_cast<uint32_t *>(m->adam_v), kMlpMasterFloats, 0u, m->stream);
    // He-uniform initialisation from a counter-based stream
    MlpMaster w;
    std::memset(&w, 0, sizeof(w));
    const GenStream rs = gen_stream(seed, 11);
    uint64_t i = 0;
    auto uni = [&](float lim) { return static_cast<float>((2.0 * gen_unit(gen_value(rs, i++)) - 1.0) * lim); };
    const float l1 = std::sqrt(6.f / kMlpIn), l2 = std::sqrt(6.f / kMlpHid), l3 = std::sqrt(3.f / kMlpHid);
    for (int k = 0; k < kMlpIn; ++k)
      for (int c = 0; c < kMlpHid; ++c) w.W1[k * kMlpHid + c] = uni(l1);
    for (int k = 0; k < kMlpHid * kMlpHid; ++k) w.W2[k] = uni(l2);
    for (int c = 0; c < kMlpHid; ++c) w.w3[c] = uni(l3);
    upload(m.get(), w);
    *out = m.release();
  });
}

int kc_mlp_destroy(kc_mlp *m) {
  return guarded([&] { delete m; });
}

int kc_mlp_get_weights(const kc_mlp *m, float *w) {
  return guarded([&] {
    if (!m || !w) fail(KC_ERR_INVALID, "null argument");
    MlpMaster h;
    KC_HIP(hipStreamSynchronize(m->stream));
    KC_HIP(hipMemcpy(&h, m->master, sizeof(h), hipMemcpyDeviceToHost));
    to_public(h, w);
  });
}

int kc_mlp_set_weights(kc_mlp *m, const float *w) {
  return guarded([&] {
    if (!m || !w) fail(KC_ERR_INVALID, "null argument");
    MlpMaster h;
    from_public(w, h);
    upload(m, h);
  });
}

int kc_mlp_forward(kc_mlp *m, const float *features, uint64_t n, float *scores) {
  return guarded([&] {
    if (!m || (n && (!features || !scores))) fail(KC_ERR_INVALID, "null argument");
    if (!n) return;
    float *dX = m->dalloc<float>(n * kMlpIn);
    float *ds = m->dalloc<float>(n);
    KC_HIP(hipMemcpyAsync(dX, features, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, m->stream));
    mlp_infer(m, n, dX, ds, m->stream);
    KC_HIP(hipMemcpyAsync(scores, ds, n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    KC_HIP(hipStreamSynchronize(m->stream));
    // scratch of a test helper: released at once
    for (void *p : {static_cast<void *>(dX), static_cast<void *>(ds)}) {
      m->owned.erase(std::find(m->owned.begin(), m->owned.end(), p));
      KC_HIP(hipFree(p));
    }
  });
}

int kc_mlp_gradients(kc_mlp *m, const float *features, const float *targets, uint64_t n, float *grads, float *loss) {
  return guarded([&] {
    if (!m || !features || !targets || !grads || n == 0) fail(KC_ERR_INVALID, "null argument or empty batch");
    if (n >= (1ull << 32)) fail(KC_ERR_INVALID, "batch too large");
    float *dX = m->dalloc<float>(n * kMlpIn), *dy = m->dalloc<float>(n);
    uint32_t *di = m->dalloc<uint32_t>(n);
    std::vector<uint32_t> id(n);
    for (uint64_t i = 0; i < n; ++i) id[i] = static_cast<uint32_t>(i);
    KC_HIP(hipMemcpyAsync(dX, features, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, m->stream));
    KC_HIP(hipMemcpyAsync(dy, targets, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
    KC_HIP(hipMemcpyAsync(di, id.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, m->stream));
    train_pass(m, n, di, dX, dy, m->stream);
    MlpMaster g;
    KC_HIP(hipMemcpyAsync(&g, m->grad, sizeof(g), hipMemcpyDeviceToHost, m->stream));
    float l = 0.f;
    KC_HIP(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    KC_HIP(hipStreamSynchronize(m->stream));
    to_public(g, grads);
    if (loss) *loss = l;
    for (void *p : {static_cast<void *>(dX), static_cast<void *>(dy), static_cast<void *>(di)}) {
      m->owned.erase(std::find(m->owned.begin(), m->owned.end(), p));
      KC_HIP(hipFree(p));
    }
  });
}

int kc_mlp_dataset_size(const kc_mlp *m, uint64_t *n) {
  return guarded([&] {
    if (!m || !n) fail(KC_ERR_INVALID, "null argument");
    *n = m->count;
  });
}

int kc_mlp_get_dataset(const kc_mlp *m, uint64_t first, uint64_t count, float *features, float *targets) {
  return guarded([&] {
    if (!m || (count && (!features || !targets))) fail(KC_ERR_INVALID, "null argument");
    if (first > m->count || count > m->count - first) fail(KC_ERR_INVALID, "rows outside the dataset");
    if (!count) return;
    KC_HIP(hipStreamSynchronize(m->stream));
    KC_HIP(hipMemcpy(features, m->X + first * kMlpIn, count * kMlpIn * sizeof(float), hipMemcpyDeviceToHost));
    KC_HIP(hipMemcpy(targets, m->y + first, count * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int kc_mlp_add_dataset(kc_mlp *m, const float *features, const float *targets, uint64_t n) {
  return guarded([&] {
    if (!m || (n && (!features || !targets))) fail(KC_ERR_INVALID, "null argument");
    if (!n) return;
    float *dX = m->dalloc<float>(n * kMlpIn), *dy = m->dalloc<float>(n);
    KC_HIP(hipMemcpyAsync(dX, features, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, m->stream));
    KC_HIP(hipMemcpyAsync(dy, targets, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
    mlp_append(m, n, dX, dy, 1, m->stream);
    KC_HIP(hipStreamSynchronize(m->stream));
    for (void *p : {static_cast<void *>(dX), static_cast<void *>(dy)}) {
      m->owned.erase(std::find(m->owned.begin(), m->owned.end(), p));
      KC_HIP(hipFree(p));
    }
  });
}

int kc_mlp_clear_dataset(kc_mlp *m) {
  return guarded([&] {
    if (!m) fail(KC_ERR_INVALID, "null argument");
    m->count = 0;
  });
}

void kc_train_params_init(kc_train_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->steps = 400;
  p->batch = 4096;
  p->lr = 3e-3f;
  p->beta1 = 0.9f;
  p->beta2 = 0.999f;
  p->eps = 1e-8f;
  p->seed = 1;
}

int kc_mlp_train(kc_mlp *m, const kc_train_params *pp, kc_train_stats *st) {
  return guarded([&] {
    if (!m || !pp) fail(KC_ERR_INVALID, "null argument");
    const kc_train_params p = *pp;
    if (m->count == 0) fail(KC_ERR_INVALID, "mlp: empty dataset (collect traces first)");
    if (p.steps < 0 || p.batch < 16 || !(p.lr > 0.f)) fail(KC_ERR_INVALID, "mlp: bad training parameters");
    const uint32_t bt = static_cast<uint32_t>(p.batch);
    if (bt > m->idx_cap) {
      m->idx = m->dalloc<uint32_t>(bt);
      m->idx_cap = bt;
    }
    const GenStream rs = gen_stream(p.seed, 12);
    const auto t0 = std::chrono::steady_clock::now();
    float first = -1.f, last = -1.f;
    for (int32_t step = 0; step < p.steps; ++step) {
      hipLaunchKernelGGL(k_batch_idx, dim3(grid_of(bt)), dim3(kBlock), 0, m->stream, bt, m->count, rs,
                         static_cast<uint64_t>(step), m->idx);
      KC_LAUNCH_CHECK();
      train_pass(m, bt, m->idx, m->X, m->y, m->stream);
      ++m->adam_t;
      const float bc1 = 1.f - std::pow(p.beta1, static_cast<float>(m->adam_t));
      const float bc2 = 1.f - std::pow(p.beta2, static_cast<float>(m->adam_t));
      hipLaunchKernelGGL(k_adam, dim3(grid_of(kMlpMasterFloats)), dim3(kBlock), 0, m->stream,
                         static_cast<uint32_t>(kMlpMasterFloats), reinterpret_cast<float *>(m->master), m->adam_m,
                         m->adam_v, m->grad, p.lr, p.beta1, p.beta2, p.eps, bc1, bc2);
      KC_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_quantize, dim3(16), dim3(kBlock), 0, m->stream, m->master, m->bf);
      KC_LAUNCH_CHECK();
      if (step == 0 || step == p.steps - 1) {
        float l = 0.f;
        KC_HIP(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
        KC_HIP(hipStreamSynchronize(m->stream));
        if (step == 0) first = l;
        last = l;
      }
    }
    KC_HIP(hipStreamSynchronize(m->stream));
    ++m->version;
    if (st) {
      std::memset(st, 0, sizeof(*st));
      st->first_loss = first;
      st->last_loss = last;
      st->steps = p.steps;
      st->samples = m->count;
      st->time_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
  });
}

}  // extern "C"
