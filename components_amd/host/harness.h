// Benchmark harness shared by the four command-line programs.  It mirrors the
// reference's app mains (app/exp_contraction.cpp:32-160, app/local_contraction.cpp,
// app/shortcuts.cpp, app/labels.cpp): parse the flags of ParseParameters
// (app/parse_parameters.h:30-100), print the CONFIG block, then for i < -i
// iterations set seed = seed + i + 1000, load or generate the graph (not
// timed), find the components (timed) and print one GLOBAL RESULT line with
// mean and standard deviation (Knuth's online update, lib/tools/benchmark.h:23-48).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "arg_parser.h"
#include "kacomp.h"

namespace kc {

struct Config {
  int seed = 1;
  std::string input_type = "null", input_file = "null", output_file = "out", debug_output_file = "tmp";
  bool print_verbose = false;
  unsigned iterations = 10, prop_iterations = 3, number_of_hitters = 0;
  double beta = 0.1;
  bool use_contraction = false, direct_contraction = false, single_level_contraction = false;
  uint64_t sequential_limit = 1;
  bool fixed_limit = false;
  uint64_t degree_limit = 100, degree_threshold = 16;
  float neighborhood_sampling_factor = 0.5f;
  bool replicate_high_degree = false, use_bfs = false, use_regular = false, overallocate = false;
  std::string gen = "null";
  int gen_k = 0;
  uint64_t gen_n = 0, gen_m = 0, gen_z = 0;
  float gen_r = 0.001f, gen_gamma = 2.2f, gen_d = 16.0f, gen_p = 1.0f;
  bool gen_periodic = false;
  // additions without a reference counterpart
  double rmat_a = 0.57, rmat_b = 0.19, rmat_c = 0.19;
  bool nofma = false, time_kernels = false;
  int tmax = 0, smax = 0, wmax = 0;
  uint64_t pes = 256;  // P of the relative sequential limit: one PE per compute unit
  bool check = false;  // compare with a host union-find after the last iteration
  bool hashed = false; // -stream hashed: SplitMix64 deviates instead of the reference draw
  int input_parts = 1;  // -parts: files "<in>_0" .. of a partitioned input (one rank per file in the reference)
  bool FromFile() const {
    return input_type == "file" || input_type == "edgefile" || input_type == "partition" || input_type == "binary";
  }
};

inline Config ParseParameters(int argc, char **argv) {
  ArgParser a(argc, argv);
  Config c;
  c.seed = a.Get<int>("seed", 1);
  c.input_type = a.Get<std::string>("input", "null");
  c.input_file = a.Get<std::string>("in", "null");
  c.input_parts = a.Get<int>("parts", 1);
  c.output_file = a.Get<std::string>("out", "out");
  c.debug_output_file = a.Get<std::string>("debug_out", "tmp");
  c.print_verbose = a.IsSet("verbose");
  c.iterations = a.Get<unsigned>("i", 10);
  c.prop_iterations = a.Get<unsigned>("pi", 3);
  c.number_of_hitters = a.Get<unsigned>("heavy-hitters", 0);
  c.beta = a.Get<double>("beta", 0.1);
  c.use_contraction = a.IsSet("contraction");
  c.direct_contraction = a.IsSet("direct");
  c.single_level_contraction = a.IsSet("single");
  c.sequential_limit = a.Get<unsigned>("seq", 1);
  c.fixed_limit = a.IsSet("fixed-seq");
  c.degree_limit = a.Get<unsigned>("deg", 100);
  c.degree_threshold = a.Get<unsigned>("hd-thres", 16);
  c.neighborhood_sampling_factor = a.Get<float>("hd-sf", 0.5f);
  c.replicate_high_degree = a.IsSet("replicate");
  c.use_bfs = a.IsSet("bfs");
  c.use_regular = a.IsSet("regular");
  c.overallocate = a.IsSet("overalloc");
  c.gen = a.Get<std::string>("gen", "null");
  c.gen_k = a.Get<int>("k", 0);
  const bool grid = c.gen == "grid_2d" || c.gen == "grid_3d";
  const bool exact_n = a.IsSet("exact_n") || grid;
  const bool exact_m = a.IsSet("exact_m") || grid;
  const uint64_t nv = a.Get<uint64_t>("n", 18), mv = a.Get<uint64_t>("m", 18);
  if ((!exact_n && nv > 62) || (!exact_m && mv > 62)) {  // -n / -m are log2 sizes (KaGen's convention)
    std::fprintf(stderr, "-n %llu / -m %llu: log2 sizes above 62 (use -exact_n / -exact_m for absolute sizes)\n",
                 static_cast<unsigned long long>(nv), static_cast<unsigned long long>(mv));
    std::exit(1);
  }
  c.gen_n = exact_n ? nv : (1ull << nv);
  c.gen_m = exact_m ? mv : (1ull << mv);
  c.gen_z = a.Get<uint64_t>("z", 0);
  c.gen_r = a.Get<float>("r", 0.001f);
  c.gen_gamma = a.Get<float>("gamma", 2.2f);
  c.gen_d = a.Get<float>("d", 16.0f);
  c.gen_p = a.Get<float>("p", 1.0f);
  c.gen_periodic = a.IsSet("periodic");
  c.rmat_a = a.Get<double>("rmat_a", 0.57);
  c.rmat_b = a.Get<double>("rmat_b", 0.19);
  c.rmat_c = a.Get<double>("rmat_c", 0.19);
  c.nofma = a.IsSet("nofma");
  c.time_kernels = a.IsSet("time_kernels");
  c.tmax = a.Get<int>("tmax", 0);
  c.smax = a.Get<int>("smax", 0);
  c.wmax = a.Get<int>("wmax", 0);
  c.pes = a.Get<uint64_t>("pes", 256);
  c.check = a.IsSet("check");
  c.hashed = a.Get<std::string>("stream", "reference") == "hashed";
  return c;
}

inline void PrintParameters(const Config &c) {
  std::printf("CONFIG\n");
  std::printf("type\t\t%s\ninfile\t\t%s\noutfile\t\t%s\nseed\t\t%d\niterations\t\t%u\n", c.input_type.c_str(),
              c.input_file.c_str(), c.output_file.c_str(), c.seed, c.iterations);
  std::printf("beta\t\t%g\nbfs\t\t%d\nlp iterations\t\t%u\ncontraction\t\t%d\ndirect\t\t%d\nsingle\t\t%d\n", c.beta,
              c.use_bfs, c.prop_iterations, c.use_contraction, c.direct_contraction, c.single_level_contraction);
  std::printf("sequential\t\t%llu\nfixed seq\t\t%d\nreplicate\t\t%d\ndeg thres\t\t%llu\nregular a2a\t\t%d\noveralloc\t\t%d\n",
              static_cast<unsigned long long>(c.sequential_limit), c.fixed_limit, c.replicate_high_degree,
              static_cast<unsigned long long>(c.degree_threshold), c.use_regular, c.overallocate);
  if (c.gen != "null") {
    std::printf("gen\t\t%s\nk\t\t%d\nn\t\t%llu\nm\t\t%llu\nr\t\t%g\nd\t\t%g\ngamma\t\t%g\nperiodic\t\t%d\n", c.gen.c_str(),
                c.gen_k, static_cast<unsigned long long>(c.gen_n), static_cast<unsigned long long>(c.gen_m), c.gen_r,
                c.gen_d, c.gen_gamma, c.gen_periodic);
  }
  std::fflush(stdout);
}

// Knuth's online mean / variance.
struct Statistics {
  uint64_t n = 0;
  double mean = 0, m2 = 0;
  void Push(double x) {
    ++n;
    const double d = x - mean;
    mean += d / n;
    m2 += d * (x - mean);
  }
  double Avg() const { return mean; }
  double Stddev() const { return n > 1 ? std::sqrt(m2 / (n - 1)) : 0.0; }
};

// IOUtility::LoadGraph (lib/io/io_utils.h:33-57).
inline int LoadGraph(const Config &c, kc_csr *g) {
  if (c.input_type == "file") return kc_read_metis(c.input_file.c_str(), g);
  if (c.input_type == "edgefile") return kc_read_edgefile(c.input_file.c_str(), g);
  if (c.input_type == "partition") return kc_read_partitioned_edgefile(c.input_file.c_str(), c.input_parts, g);
  if (c.input_type == "binary") return kc_read_binary(c.input_file.c_str(), g);
  if (c.gen != "null") {
    kc_gen_params p;
    kc_gen_params_init(&p);
    p.n = c.gen_n;
    p.m = c.gen_m;
    p.r = c.gen_r;
    p.p = c.gen_p;
    p.gamma = c.gen_gamma;
    p.d = c.gen_d;
    p.periodic = c.gen_periodic;
    p.seed = c.seed;
    p.rmat_a = c.rmat_a;
    p.rmat_b = c.rmat_b;
    p.rmat_c = c.rmat_c;
    if (c.gen == "grid_2d") {
      p.nx = c.gen_n;
      p.ny = c.gen_m;
    } else if (c.gen == "grid_3d") {
      p.nx = c.gen_n;
      p.ny = c.gen_m;
      p.nz = c.gen_z ? c.gen_z : c.gen_n;
    }
    return kc_generate(c.gen.c_str(), &p, g);
  }
  std::fprintf(stderr, "I/O type not supported\n");
  return KC_ERR_UNSUPPORTED;
}

// Textbook union-find with path halving: the canonical min-id labels used by -check.
inline std::vector<uint32_t> HostCanonicalLabels(const kc_csr &g) {
  std::vector<uint32_t> parent(g.n);
  for (uint64_t v = 0; v < g.n; ++v) parent[v] = static_cast<uint32_t>(v);
  auto find = [&](uint32_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  for (uint64_t v = 0; v < g.n; ++v)
    for (uint64_t e = g.offsets[v]; e < g.offsets[v + 1]; ++e) {
      uint32_t a = find(static_cast<uint32_t>(v)), b = find(g.targets[e]);
      if (a != b) parent[a < b ? b : a] = a < b ? a : b;
    }
  std::vector<uint32_t> lab(g.n);
  for (uint64_t v = 0; v < g.n; ++v) lab[v] = find(static_cast<uint32_t>(v));
  return lab;
}

// whole_graph_base_case: every edge goes to the base case (no contraction
// level), as the all_reduce program gathers the whole graph on one PE.
inline int RunApp(int argc, char **argv, int variant, const char *runner, bool whole_graph_base_case = false) {
  Config c = ParseParameters(argc, argv);
  PrintParameters(c);
  const int initial_seed = c.seed;
  Statistics t_stats, dev_stats, gteps_stats, level_stats, sweep_stats, radius_stats;
  uint64_t last_alg_bytes = 0;
  double last_dev_ms = 0;
  bool all_ok = true;
  kc_csr cached{};
  bool have_cached = false;
  for (unsigned i = 0; i < c.iterations; ++i) {
    c.seed = initial_seed + static_cast<int>(i) + 1000;
    kc_csr g{};
    if (c.FromFile() && have_cached) {
      g = cached;
    } else {
      int rc = LoadGraph(c, &g);
      if (rc != KC_OK) {
        std::fprintf(stderr, "graph load failed: %s\n", kc_last_error());
        return 1;
      }
      if (c.FromFile()) {  // a file graph is the same in every iteration
        cached = g;
        have_cached = true;
      }
    }
    if (i == 0) {
      uint64_t maxd = 0, nonisolated = 0;
      for (uint64_t v = 0; v < g.n; ++v) {
        const uint64_t d = g.offsets[v + 1] - g.offsets[v];
        maxd = d > maxd ? d : maxd;
        nonisolated += d > 0;
      }
      std::printf("GLOBAL INPUT s=%d p=1 n=%llu m=%llu c(min,max)=0,0 max_d=%llu non_isolated=%llu\n", c.seed,
                  static_cast<unsigned long long>(g.n), static_cast<unsigned long long>(g.m),
                  static_cast<unsigned long long>(maxd), static_cast<unsigned long long>(nonisolated));
    }
    kc_csr_host h{g.n, g.offsets, g.targets};
    kc_graph *dg = nullptr;
    if (kc_graph_create(&h, &dg) != KC_OK) {
      std::fprintf(stderr, "kc_graph_create failed: %s\n", kc_last_error());
      return 1;
    }
    kc_options o;
    kc_options_init(&o);
    o.variant = variant;
    o.seed = c.seed;
    o.beta = c.beta;
    o.use_bfs = c.use_bfs;
    o.deviate_fma = c.nofma ? 0 : 1;
    o.time_kernels = c.time_kernels;
    o.sweep_group_small = c.tmax;
    o.sweep_group_mid = c.smax;
    o.sweep_group_large = c.wmax;
    // ComputeSequentialLimit (exponential_contraction.h:1287-1292) with P = -pes
    // (one PE per compute unit); the relative limit is floored at 4096 edges,
    // below which the single-GPU base case is cheaper than further levels.
    // Both limits are in undirected edges: GatherNumberOfGlobalEdges sums the
    // adjacency entries of all PEs and halves the sum (static_graph.h:353-360,
    // dynamic_graph.h:335-342), and the level loop compares that count with
    // the limit (exponential_contraction.h:620-627).
    const uint64_t rel = (g.m / (c.pes ? c.pes : 1)) * c.sequential_limit;
    o.seq_limit = c.fixed_limit ? c.sequential_limit : (rel > 4096 ? rel : 4096);
    if (whole_graph_base_case) o.seq_limit = ~0ull;
    o.deviate_stream = c.hashed ? KC_STREAM_HASHED : KC_STREAM_REFERENCE;
    // -replicate / -hd-thres (exponential_contraction.h:235-237, parse_parameters.h:67-71): the
    // level-0 hubs split into proxies; the reference does it on the default branch only
    if (variant == KC_VARIANT_EXP && c.replicate_high_degree && !c.single_level_contraction && !c.use_contraction) {
      o.replicate = 1;
      o.hd_threshold = static_cast<double>(c.degree_threshold);
    }
    // -single takes precedence over -contraction (exponential_contraction.h:76,134): the
    // components inside -pes blocks of consecutive ids are contracted first, once or twice
    if (variant == KC_VARIANT_EXP && !whole_graph_base_case &&
        (c.single_level_contraction || c.use_contraction)) {
      o.pre_contract = c.single_level_contraction ? KC_PRE_SINGLE : KC_PRE_DOUBLE;
      o.pre_blocks = static_cast<uint32_t>(c.pes == 0 ? 1 : c.pes > 0xFFFFFFFFull ? 0xFFFFFFFFull : c.pes);
    }
    std::vector<uint32_t> labels(g.n);
    kc_stats st;
    if (kc_find_components(dg, &o, labels.data(), &st) != KC_OK) {
      std::fprintf(stderr, "kc_find_components failed: %s\n", kc_last_error());
      kc_graph_destroy(dg);
      return 1;
    }
    const double secs = st.total_ms / 1e3;
    t_stats.Push(secs);
    dev_stats.Push(st.time_ms / 1e3);
    gteps_stats.Push(st.time_ms > 0 ? static_cast<double>(g.m) / (st.time_ms / 1e3) / 1e9 : 0.0);
    level_stats.Push(st.levels);
    sweep_stats.Push(static_cast<double>(st.sweeps));
    radius_stats.Push(static_cast<double>(st.radius_sum));
    last_alg_bytes = st.alg_bytes;
    last_dev_ms = st.time_ms;
    if (c.print_verbose)
      std::printf("ITERATION i=%u seed=%d time=%g device_ms=%g levels=%u sweeps=%llu components=%u\n", i, c.seed, secs,
                  st.time_ms, st.levels, static_cast<unsigned long long>(st.sweeps), st.components);
    if (o.replicate && i == 0)
      std::printf("REPLICATE hubs=%u proxies=%llu\n", st.hubs, static_cast<unsigned long long>(st.replicas));
    if (variant == KC_VARIANT_LABELS || variant == KC_VARIANT_LOCAL) {
      // OutputComponents (lib/data_structures/static_graph.h:504-591), called by the
      // labels and local_contraction programs after every run; its vertex loop
      // skips the inactive degree-0 vertices, so size-1 components are not listed
      uint64_t np = 0;
      std::vector<uint64_t> pairs;
      if (kc_component_histogram(dg, nullptr, 0, &np) == KC_OK) {
        pairs.resize(2 * np + 2);
        kc_component_histogram(dg, pairs.data(), np, &np);
      }
      uint64_t total = 0;
      std::printf("COMPONENTS [ ");
      for (uint64_t k = 0; k < np; ++k) {
        if (pairs[2 * k] == 1) continue;
        std::printf("size=%llu (num=%llu) ", static_cast<unsigned long long>(pairs[2 * k]),
                    static_cast<unsigned long long>(pairs[2 * k + 1]));
        total += pairs[2 * k + 1];
      }
      std::printf("]\nNUM COMPONENTS %llu\n", static_cast<unsigned long long>(total));
    }
    if (c.check && i + 1 == c.iterations) {
      std::vector<uint32_t> ref = HostCanonicalLabels(g);
      const bool ok = ref == labels;
      all_ok = all_ok && ok;
      std::printf("CHECK labels_bit_exact=%s components=%u\n", ok ? "yes" : "no", st.components);
    }
    kc_graph_destroy(dg);
    if (!c.FromFile()) kc_csr_free(&g);
  }
  if (have_cached) kc_csr_free(&cached);
  const double hbm = last_dev_ms > 0 ? last_alg_bytes / (last_dev_ms / 1e3) / 1e9 : 0.0;
  std::printf(
      "GLOBAL RESULT runner=%s time=%g stddev=%g comm_time=0 stddev=0 send_volume=0 stddev=0 recv_volume=0 stddev=0 "
      "iterations=%u device_time=%g gteps=%g levels=%g sweeps=%g radius=%g alg_bytes=%llu hbm_gbs=%g roofline=%g\n",
      runner, t_stats.Avg(), t_stats.Stddev(), c.iterations, dev_stats.Avg(), gteps_stats.Avg(), level_stats.Avg(),
      sweep_stats.Avg(), radius_stats.Avg(), static_cast<unsigned long long>(last_alg_bytes), hbm, hbm / 8000.0);
  return all_ok ? 0 : 2;
}

}  // namespace kc
