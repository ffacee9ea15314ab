// Synthetic graph generators with the KaGen-style parameters of
// IOUtility::GenerateSyntheticGraph (lib/io/io_utils.h:59-83) and
// ParseParameters (app/parse_parameters.h:83-99).  KaGen itself is an absent
// submodule, so these are the project's own generators: every random value is
// a pure function of (seed, stream, index) through SplitMix64 (Steele, Lea,
// Flood 2014), so the edge list does not depend on the thread count.
//
//   gnm_undirected  n vertices, m edges with iid uniform endpoints
//   rmat            2^scale vertices, m edges, recursive quadrant choice (a,b,c,d)
//   grid_2d         nx x ny lattice, each bond kept with probability p (periodic optional)
//   grid_3d         nx x ny x nz lattice, bond percolation with probability p
//   rgg_2d/rgg_3d   n uniform points in the unit square/cube, edge iff distance <= r;
//                   vertices numbered in cell order
//   rhg             threshold random hyperbolic graph (Krioukov et al. 2010), n points,
//                   power-law exponent gamma, target average degree d; band search of
//                   von Looz, Meyerhenke, Prutkin (2015); vertices numbered by angle
//   ba              Barabasi-Albert with d edges per vertex by recursive slot resolution
//                   (Sanders, Schulz 2016)
// Self-loops are dropped and duplicate edges merged by the CSR builder.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "csr_build.h"
#include "gen_rng.h"
#include "kacomp.h"

namespace kc {
void set_error(const char *fmt, ...);
}

namespace {

using Stream = kc::GenStream;
inline uint64_t bounded(uint64_t z, uint64_t n) { return kc::gen_bounded(z, n); }
inline double unit(uint64_t z) { return kc::gen_unit(z); }
struct StreamFn {
  Stream s;
  StreamFn(int32_t seed, uint64_t id) : s(kc::gen_stream(seed, id)) {}
  uint64_t operator()(uint64_t i) const { return kc::gen_value(s, i); }
};

struct EdgeList {
  std::vector<uint32_t> src, dst;
  void add(uint32_t u, uint32_t v) {
    src.push_back(u);
    dst.push_back(v);
  }
};

// Concatenate per-thread lists.
void merge(std::vector<EdgeList> &parts, EdgeList &out) {
  size_t total = 0;
  for (auto &p : parts) total += p.src.size();
  out.src.resize(total);
  out.dst.resize(total);
  size_t at = 0;
  for (auto &p : parts) {
    std::copy(p.src.begin(), p.src.end(), out.src.begin() + at);
    std::copy(p.dst.begin(), p.dst.end(), out.dst.begin() + at);
    at += p.src.size();
    std::vector<uint32_t>().swap(p.src);
    std::vector<uint32_t>().swap(p.dst);
  }
}

void gen_gnm(const kc_gen_params &p, EdgeList &el) {
  const uint64_t n = p.n, m = p.m;
  const Stream rs = kc::gen_stream(p.seed, 1);
  el.src.resize(m);
  el.dst.resize(m);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(m); ++i) kc::gnm_edge(rs, i, n, &el.src[i], &el.dst[i]);
}

void gen_rmat(const kc_gen_params &p, EdgeList &el, uint32_t scale) {
  const uint64_t m = p.m;
  const Stream rs = kc::gen_stream(p.seed, 2);
  const kc::RmatThresholds t = kc::rmat_thresholds(p.rmat_a, p.rmat_b, p.rmat_c);
  el.src.resize(m);
  el.dst.resize(m);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(m); ++i) kc::rmat_edge(rs, i, scale, t, &el.src[i], &el.dst[i]);
}

void gen_grid(const kc_gen_params &p, EdgeList &el, int dims) {
  const uint64_t nx = p.nx, ny = p.ny, nz = dims == 3 ? p.nz : 1;
  const uint64_t n = nx * ny * nz;
  const Stream rs = kc::gen_stream(p.seed, 3);
  const double prob = p.p;
  const uint64_t thresh = kc::grid_keep_threshold(prob);
  const int nt = omp_get_max_threads();
  std::vector<EdgeList> parts(nt);
#pragma omp parallel
  {
    EdgeList &mine = parts[omp_get_thread_num()];
#pragma omp for schedule(static)
    for (int64_t id = 0; id < static_cast<int64_t>(n); ++id) {
      const uint64_t x = id % nx, y = (id / nx) % ny, z = id / (nx * ny);
      for (int d = 0; d < dims; ++d) {
        uint64_t tx = x, ty = y, tz = z;
        const uint64_t lim = d == 0 ? nx : d == 1 ? ny : nz;
        uint64_t &t = d == 0 ? tx : d == 1 ? ty : tz;
        if (t + 1 < lim) {
          ++t;
        } else if (p.periodic && lim > 2) {
          t = 0;
        } else {
          continue;
        }
        const uint64_t bond = static_cast<uint64_t>(id) * dims + d;
        if (!kc::grid_bond_kept(rs, bond, prob, thresh)) continue;
        mine.add(static_cast<uint32_t>(id), static_cast<uint32_t>((tz * ny + ty) * nx + tx));
      }
    }
  }
  merge(parts, el);
}

void gen_rgg(const kc_gen_params &p, EdgeList &el, int dims) {
  const uint64_t n = p.n;
  const double r = p.r;
  StreamFn rs(p.seed, 4);
  uint64_t cells = r > 0 ? static_cast<uint64_t>(std::floor(1.0 / r)) : 1;
  if (cells < 1) cells = 1;
  const uint64_t max_cells = dims == 2 ? (1ull << 26) : (1ull << 24);
  while ((dims == 2 ? cells * cells : cells * cells * cells) > max_cells) cells /= 2;
  const uint64_t ncell = dims == 2 ? cells * cells : cells * cells * cells;
  std::vector<double> pt(n * dims);
  std::vector<uint64_t> cell_of(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
    uint64_t cid = 0;
    for (int d = dims - 1; d >= 0; --d) {
      const double x = unit(rs(static_cast<uint64_t>(i) * dims + d));
      pt[i * dims + d] = x;
      uint64_t c = static_cast<uint64_t>(x * cells);
      if (c >= cells) c = cells - 1;
      cid = cid * cells + c;
    }
    cell_of[i] = cid;
  }
  // counting sort by cell: vertex id = position in cell order
  std::vector<uint64_t> start(ncell + 1, 0);
  for (uint64_t i = 0; i < n; ++i) start[cell_of[i] + 1]++;
  for (uint64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
  std::vector<uint64_t> fillp(start.begin(), start.end() - 1);
  std::vector<double> sp(n * dims);
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t at = fillp[cell_of[i]]++;
    for (int d = 0; d < dims; ++d) sp[at * dims + d] = pt[i * dims + d];
  }
  std::vector<double>().swap(pt);
  std::vector<uint64_t>().swap(cell_of);
  const double r2 = r * r;
  // half of the neighbour cells (lexicographically larger offsets) plus the cell itself
  std::vector<std::vector<int>> offs;
  if (dims == 2) {
    offs = {{1, -1}, {1, 0}, {1, 1}, {0, 1}};
  } else {
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (dz > 0 || (dz == 0 && (dy > 0 || (dy == 0 && dx > 0)))) offs.push_back({dx, dy, dz});
        }
  }
  const int nt = omp_get_max_threads();
  std::vector<EdgeList> parts(nt);
#pragma omp parallel
  {
    EdgeList &mine = parts[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 256)
    for (int64_t c = 0; c < static_cast<int64_t>(ncell); ++c) {
      const uint64_t cx = c % cells, cy = (c / cells) % cells, cz = dims == 3 ? c / (cells * cells) : 0;
      for (uint64_t i = start[c]; i < start[c + 1]; ++i) {
        const double *pi = &sp[i * dims];
        for (uint64_t j = i + 1; j < start[c + 1]; ++j) {
          const double *pj = &sp[j * dims];
          double d2 = 0;
          for (int d = 0; d < dims; ++d) d2 += (pi[d] - pj[d]) * (pi[d] - pj[d]);
          if (d2 <= r2) mine.add(static_cast<uint32_t>(i), static_cast<uint32_t>(j));
        }
        for (auto &o : offs) {
          const int64_t nx = static_cast<int64_t>(cx) + o[0];
          const int64_t ny = static_cast<int64_t>(cy) + o[1];
          const int64_t nz = dims == 3 ? static_cast<int64_t>(cz) + o[2] : 0;
          if (nx < 0 || ny < 0 || nz < 0 || nx >= static_cast<int64_t>(cells) || ny >= static_cast<int64_t>(cells) ||
              nz >= static_cast<int64_t>(cells))
            continue;
          const uint64_t c2 = dims == 2 ? static_cast<uint64_t>(ny) * cells + nx
                                        : (static_cast<uint64_t>(nz) * cells + ny) * cells + nx;
          for (uint64_t j = start[c2]; j < start[c2 + 1]; ++j) {
            const double *pj = &sp[j * dims];
            double d2 = 0;
            for (int d = 0; d < dims; ++d) d2 += (pi[d] - pj[d]) * (pi[d] - pj[d]);
            if (d2 <= r2) mine.add(static_cast<uint32_t>(i), static_cast<uint32_t>(j));
          }
        }
      }
    }
  }
  merge(parts, el);
}

void gen_rhg(const kc_gen_params &p, EdgeList &el) {
  const uint64_t n = p.n;
  const double alpha = (p.gamma - 1.0) / 2.0;
  const double kbar = p.d;
  // target radius from the average-degree approximation of Krioukov et al. (2010)
  const double xi = 2.0 * alpha / (2.0 * alpha - 1.0);
  const double R = 2.0 * std::log(2.0 * xi * xi * static_cast<double>(n) / (M_PI * kbar));
  StreamFn rs(p.seed, 6);
  struct Pt {
    double theta, r;
  };
  std::vector<Pt> pts(n);
  const double cR = std::cosh(alpha * R) - 1.0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
    const double th = 2.0 * M_PI * unit(rs(2 * i));
    const double u = unit(rs(2 * i + 1));
    pts[i] = {th, std::acosh(1.0 + u * cR) / alpha};
  }
  std::sort(pts.begin(), pts.end(), [](const Pt &a, const Pt &b) { return a.theta < b.theta || (a.theta == b.theta && a.r < b.r); });
  // bands [b_i, b_{i+1}): b_0 = 0, b_i = R - R / 2^i
  int nb = 1;
  while ((1ull << nb) < n && nb < 40) ++nb;
  nb = std::max(2, nb / 2 + 1);
  std::vector<double> bnd(nb + 1);
  bnd[0] = 0.0;
  for (int i = 1; i < nb; ++i) bnd[i] = R - R / std::ldexp(1.0, i);
  bnd[nb] = R + 1e-9;
  struct BP {
    double theta, coshr, sinhr, cost, sint;
    uint32_t id;
  };
  std::vector<std::vector<BP>> bands(nb);
  for (uint64_t i = 0; i < n; ++i) {
    int b = static_cast<int>(std::upper_bound(bnd.begin(), bnd.end(), pts[i].r) - bnd.begin()) - 1;
    if (b < 0) b = 0;
    if (b >= nb) b = nb - 1;
    bands[b].push_back({pts[i].theta, std::cosh(pts[i].r), std::sinh(pts[i].r), std::cos(pts[i].theta),
                        std::sin(pts[i].theta), static_cast<uint32_t>(i)});
  }
  const double coshR = std::cosh(R);
  const int nt = omp_get_max_threads();
  std::vector<EdgeList> parts(nt);
#pragma omp parallel
  {
    EdgeList &mine = parts[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 1024)
    for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) {
      const double r = pts[i].r, th = pts[i].theta;
      const double chr = std::cosh(r), shr = std::sinh(r), ct = std::cos(th), st = std::sin(th);
      for (int b = 0; b < nb; ++b) {
        const auto &band = bands[b];
        if (band.empty()) continue;
        const double lo = bnd[b];
        double dtheta = M_PI;
        if (lo > 0 && shr > 0) {
          const double cb = (chr * std::cosh(lo) - coshR) / (shr * std::sinh(lo));
          if (cb >= 1.0) continue;
          if (cb > -1.0) dtheta = std::acos(cb);
        }
        auto scan = [&](double a0, double a1) {
          auto it = std::lower_bound(band.begin(), band.end(), a0, [](const BP &x, double v) { return x.theta < v; });
          for (; it != band.end() && it->theta <= a1; ++it) {
            if (it->id <= static_cast<uint32_t>(i)) continue;
            const double cosd = ct * it->cost + st * it->sint;
            if (chr * it->coshr - shr * it->sinhr * cosd <= coshR) mine.add(static_cast<uint32_t>(i), it->id);
          }
        };
        if (dtheta >= M_PI) {
          scan(-1.0, 7.0);
        } else {
          double a0 = th - dtheta, a1 = th + dtheta;
          if (a0 < 0) {
            scan(a0 + 2 * M_PI, 7.0);
            scan(0.0, a1);
          } else if (a1 > 2 * M_PI) {
            scan(a0, 7.0);
            scan(0.0, a1 - 2 * M_PI);
          } else {
            scan(a0, a1);
          }
        }
      }
    }
  }
  merge(parts, el);
}

void gen_ba(const kc_gen_params &p, EdgeList &el) {
  const uint64_t n = p.n;
  const uint64_t d = p.d >= 1 ? static_cast<uint64_t>(p.d) : 1;
  const Stream rs = kc::gen_stream(p.seed, 5);
  el.src.resize(n * d);
  el.dst.resize(n * d);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < static_cast<int64_t>(n * d); ++e) kc::ba_edge(rs, e, d, &el.src[e], &el.dst[e]);
}

}  // namespace

extern "C" void kc_gen_params_init(kc_gen_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->n = 1ull << 18;
  p->m = 1ull << 18;
  p->r = 0.001;
  p->p = 1.0;
  p->gamma = 2.2;
  p->d = 16.0;
  p->seed = 1;
  p->rmat_a = 0.57;
  p->rmat_b = 0.19;
  p->rmat_c = 0.19;
}

extern "C" int kc_generate(const char *name, const kc_gen_params *pp, kc_csr *out) {
  if (!name || !pp || !out) {
    kc::set_error("kc_generate: null argument");
    return KC_ERR_INVALID;
  }
  std::memset(out, 0, sizeof(*out));
  kc_gen_params p = *pp;
  const std::string g(name);
  EdgeList el;
  uint64_t n = 0;
  try {
    if (g == "gnm_undirected") {
      n = p.n;
      gen_gnm(p, el);
    } else if (g == "rmat") {
      uint32_t scale = 0;
      while ((1ull << scale) < p.n) ++scale;
      if ((1ull << scale) != p.n || scale > 31) {
        kc::set_error("rmat: n=%llu must be a power of two below 2^32", static_cast<unsigned long long>(p.n));
        return KC_ERR_INVALID;
      }
      n = p.n;
      gen_rmat(p, el, scale);
    } else if (g == "grid_2d") {
      if (!p.nx) p.nx = p.n;
      if (!p.ny) p.ny = p.m;
      n = p.nx * p.ny;
      gen_grid(p, el, 2);
    } else if (g == "grid_3d") {
      if (!p.nx) p.nx = p.n;
      if (!p.ny) p.ny = p.nx;
      if (!p.nz) p.nz = p.nx;
      n = p.nx * p.ny * p.nz;
      gen_grid(p, el, 3);
    } else if (g == "rgg_2d") {
      n = p.n;
      gen_rgg(p, el, 2);
    } else if (g == "rgg_3d") {
      n = p.n;
      gen_rgg(p, el, 3);
    } else if (g == "rhg") {
      n = p.n;
      if (!(p.gamma > 2.0)) {
        kc::set_error("rhg: gamma must exceed 2");
        return KC_ERR_INVALID;
      }
      gen_rhg(p, el);
    } else if (g == "ba") {
      n = p.n;
      gen_ba(p, el);
    } else if (g == "rdg_2d" || g == "rdg_3d") {
      kc::set_error("Generator not supported: %s (Delaunay graphs are not built by this library)", name);
      return KC_ERR_UNSUPPORTED;
    } else {
      kc::set_error("Generator not supported: %s", name);
      return KC_ERR_UNSUPPORTED;
    }
    if (n >= 0xFFFFFFFFull) {
      kc::set_error("%s: n=%llu exceeds 32-bit vertex ids", name, static_cast<unsigned long long>(n));
      return KC_ERR_INVALID;
    }
    kc::build_symmetric_csr(n, el.src.data(), el.dst.data(), el.src.size(), out);
  } catch (const std::bad_alloc &) {
    kc::set_error("%s: out of host memory", name);
    return KC_ERR_OOM;
  }
  if (!out->offsets) {
    kc::set_error("%s: out of host memory", name);
    return KC_ERR_OOM;
  }
  return KC_OK;
}
