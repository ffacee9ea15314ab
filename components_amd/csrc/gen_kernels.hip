// Device generation of the synthetic inputs and device CSR construction.
// The edge streams are the host generators' (gen_rng.h), so a graph built
// here has the same rows as kc_generate's up to the order inside each row.
// CSR build: degree count, scan, scatter of both directions into row buckets,
// then the bucket dedup and compaction kernels of the contraction path.
#include <hip/hip_runtime.h>

#include "gen_rng.h"
#include "kc_common.h"
#include "kernels.h"

namespace kc {

namespace {

inline unsigned grid_cap(uint64_t n, uint64_t cap = 65536) {
  uint64_t g = ceil_div(n, kBlock);
  if (g == 0) g = 1;
  return static_cast<unsigned>(g < cap ? g : cap);
}

__global__ void k_gen_gnm(uint64_t count, uint64_t n, GenStream s, uint32_t *src, uint32_t *dst) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    gnm_edge(s, i, n, &src[i], &dst[i]);
}

__global__ void k_gen_rmat(uint64_t count, uint32_t scale, RmatThresholds t, GenStream s, uint32_t *src,
                           uint32_t *dst) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    rmat_edge(s, i, scale, t, &src[i], &dst[i]);
}

__global__ void k_gen_ba(uint64_t count, uint64_t d, GenStream s, uint32_t *src, uint32_t *dst) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    ba_edge(s, i, d, &src[i], &dst[i]);
}

// Lattice bonds: slot (vertex id, direction) of the output holds the bond when
// it is kept and the self-loop (id, id) otherwise; the CSR build drops
// self-loops.  Coordinates use 32-bit arithmetic (every extent is below 2^32).
__global__ void k_gen_grid(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t dims, uint32_t periodic, double p,
                           uint64_t thresh, GenStream s, uint32_t *src, uint32_t *dst) {
  const uint64_t n = static_cast<uint64_t>(nx) * ny * nz;
  for (uint64_t id = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; id < n;
       id += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t x = static_cast<uint32_t>(id % nx);
    const uint64_t yz = id / nx;
    const uint32_t y = static_cast<uint32_t>(yz % ny);
    const uint32_t z = static_cast<uint32_t>(yz / ny);
    for (uint32_t d = 0; d < dims; ++d) {
      const uint32_t c = d == 0 ? x : d == 1 ? y : z;
      const uint32_t lim = d == 0 ? nx : d == 1 ? ny : nz;
      const uint64_t bond = id * dims + d;
      uint32_t c2 = c + 1;
      bool ok = true;
      if (c2 >= lim) {
        ok = periodic && lim > 2;
        c2 = 0;
      }
      const bool keep = ok && grid_bond_kept(s, bond, p, thresh);
      const uint32_t tx = d == 0 ? c2 : x, ty = d == 1 ? c2 : y, tz = d == 2 ? c2 : z;
      const uint64_t t = (static_cast<uint64_t>(tz) * ny + ty) * nx + tx;
      src[bond] = static_cast<uint32_t>(id);
      dst[bond] = keep ? static_cast<uint32_t>(t) : static_cast<uint32_t>(id);
    }
  }
}

// Loops run in whole-wave steps so that the peeled atomics see every lane.
__global__ void k_deg_count(uint64_t count, const uint32_t *src, const uint32_t *dst, uint32_t *deg) {
  for (uint64_t base = blockIdx.x * static_cast<uint64_t>(blockDim.x); base < count;
       base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t i = base + threadIdx.x;
    const bool live = i < count;
    const uint32_t u = live ? src[i] : 0u, v = live ? dst[i] : 0u;
    const bool e = live && u != v;
    peeled_fetch_add(deg, u, 1u, e);
    peeled_fetch_add(deg, v, 1u, e);
  }
}

__global__ void k_edge_scatter(uint64_t count, const uint32_t *src, const uint32_t *dst, const uint64_t *off,
                               uint32_t *fill, uint32_t *bucket) {
  for (uint64_t base = blockIdx.x * static_cast<uint64_t>(blockDim.x); base < count;
       base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t i = base + threadIdx.x;
    const bool live = i < count;
    const uint32_t u = live ? src[i] : 0u, v = live ? dst[i] : 0u;
    const bool e = live && u != v;
    const uint32_t pu = peeled_fetch_add(fill, u, 1u, e);
    const uint32_t pv = peeled_fetch_add(fill, v, 1u, e);
    if (e) {
      bucket[off[u] + pu] = v;
      bucket[off[v] + pv] = u;
    }
  }
}

// Non-empty rows per block (one atomic per block, on a 64-bit counter).
__global__ __launch_bounds__(kBlock) void k_nonzero_rows(uint64_t n, const uint64_t *off, unsigned long long *counter) {
  __shared__ uint32_t wc[kBlock / kWave];
  uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  const bool nz = v < n && off[v + 1] > off[v];
  const uint32_t c = static_cast<uint32_t>(__popcll(__ballot(nz)));
  if (__lane_id() == 0) wc[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t t = wc[0] + wc[1] + wc[2] + wc[3];
    if (t) atomicAdd(counter, static_cast<unsigned long long>(t));
  }
}

// Edges of a sorted binary file (1-based (u, v) u64 pairs) as 0-based u32
// pairs; a pair outside [1, n] raises *bad and becomes the self-loop (0, 0),
// which the CSR build drops.
__global__ void k_binary_edges(const uint64_t *raw, uint64_t count, uint64_t n, uint32_t *src, uint32_t *dst,
                               int *bad) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t u = raw[2 * i], v = raw[2 * i + 1];
    const bool ok = u >= 1 && v >= 1 && u <= n && v <= n;
    if (!ok) *bad = 1;
    src[i] = ok ? static_cast<uint32_t>(u - 1) : 0u;
    dst[i] = ok ? static_cast<uint32_t>(v - 1) : 0u;
  }
}

}  // namespace

void binary_edges(const uint64_t *raw, uint64_t count, uint64_t n, uint32_t *src, uint32_t *dst, int *bad,
                  hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(k_binary_edges, dim3(grid_cap(count)), dim3(kBlock), 0, s, raw, count, n, src, dst, bad);
  KC_LAUNCH_CHECK();
}

void gen_gnm(uint64_t count, uint64_t n, int32_t seed, uint32_t *src, uint32_t *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_gnm, dim3(grid_cap(count)), dim3(kBlock), 0, s, count, n, gen_stream(seed, 1), src, dst);
  KC_LAUNCH_CHECK();
}

void gen_rmat(uint64_t count, uint32_t scale, double a, double b, double c, int32_t seed, uint32_t *src, uint32_t *dst,
              hipStream_t s) {
  hipLaunchKernelGGL(k_gen_rmat, dim3(grid_cap(count)), dim3(kBlock), 0, s, count, scale, rmat_thresholds(a, b, c),
                     gen_stream(seed, 2), src, dst);
  KC_LAUNCH_CHECK();
}

void gen_ba(uint64_t count, uint64_t d, int32_t seed, uint32_t *src, uint32_t *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_ba, dim3(grid_cap(count)), dim3(kBlock), 0, s, count, d, gen_stream(seed, 5), src, dst);
  KC_LAUNCH_CHECK();
}

void gen_grid(uint64_t nx, uint64_t ny, uint64_t nz, int dims, int periodic, double p, int32_t seed, uint32_t *src,
              uint32_t *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_grid, dim3(grid_cap(nx * ny * nz)), dim3(kBlock), 0, s, static_cast<uint32_t>(nx),
                     static_cast<uint32_t>(ny), static_cast<uint32_t>(nz), static_cast<uint32_t>(dims),
                     static_cast<uint32_t>(periodic != 0), p, grid_keep_threshold(p), gen_stream(seed, 3), src, dst);
  KC_LAUNCH_CHECK();
}

void deg_count(uint64_t count, const uint32_t *src, const uint32_t *dst, uint32_t *deg, hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(k_deg_count, dim3(grid_cap(count)), dim3(kBlock), 0, s, count, src, dst, deg);
  KC_LAUNCH_CHECK();
}

void edge_scatter(uint64_t count, const uint32_t *src, const uint32_t *dst, const uint64_t *off, uint32_t *fill,
                  uint32_t *bucket, hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(k_edge_scatter, dim3(grid_cap(count)), dim3(kBlock), 0, s, count, src, dst, off, fill, bucket);
  KC_LAUNCH_CHECK();
}

void nonzero_rows(uint64_t n, const uint64_t *off, unsigned long long *counter, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_nonzero_rows, dim3(static_cast<unsigned>(ceil_div(n, kBlock))), dim3(kBlock), 0, s, n, off,
                     counter);
  KC_LAUNCH_CHECK();
}

}  // namespace kc
