/*
 * kacomp.h -- C ABI of libkacomp: connected components by exponential
 * contraction on one MI355X (gfx950).
 *
 * The whole graph lives on one GPU as CSR (u64 row offsets, u32 neighbour
 * ids, both directions of every undirected edge present).  Every entry point
 * returns KC_OK (0) or a negative KC_ERR_* code; the message of the last
 * failure on the calling thread is returned by kc_last_error().  No
 * exceptions cross this boundary and nothing calls exit().
 *
 * Reference interfaces replaced (sebalamm/components):
 *   kc_graph_create / kc_graph_destroy
 *       StaticGraph construction by IOUtility::LoadGraph
 *       (lib/io/io_utils.h:33-57, lib/data_structures/static_graph.h:80-134);
 *       the graph is copied to the device once.
 *   kc_find_components (variant KC_VARIANT_EXP)
 *       ExponentialContraction::FindComponents(G, labels)
 *       (lib/components/exponential_contraction.h:60-319)
 *   kc_find_components (variant KC_VARIANT_LOCAL)
 *       LocalContraction::FindComponents (lib/components/local_contraction.h:82-90)
 *   kc_find_components (variant KC_VARIANT_SHORTCUT)
 *       ShortcutPropagation::FindComponents (lib/components/shortcut_propagation.h:89-94)
 *   kc_find_components (variant KC_VARIANT_LABELS)
 *       Propagation::FindComponents (lib/components/propagation.h:48-53)
 *   kc_stats.comm_time_s / send_volume / recv_volume
 *       GetCommTime / GetSendVolume / GetReceiveVolume
 *       (lib/components/exponential_contraction.h:325-335); always 0 on one GPU.
 *   kc_read_metis / kc_write_metis
 *       GraphIO::ReadMETISFile (lib/io/graph_io.h:281-382, :1207-1248)
 *   kc_read_edgefile
 *       GraphIO::ReadSortedEdgeFile (lib/io/graph_io.h:385-669) with the
 *       line parser ParseEdgeFilestream (:1250-1300)
 *   kc_read_partitioned_edgefile
 *       GraphIO::ReadPartitionedSortedEdgeFile (lib/io/graph_io.h:672-907)
 *   kc_read_binary / kc_write_binary
 *       GraphIO::ReadSortedBinaryFile (lib/io/graph_io.h:910-1175); the
 *       writer produces that format (the reference has none)
 *   kc_generate
 *       IOUtility::GenerateSyntheticGraph (lib/io/io_utils.h:59-83); KaGen's
 *       generators are not available, so the generators are the project's own
 *       (counter-based SplitMix64 streams) with KaGen-style parameters.
 *
 * Labels written by kc_find_components are canonical: every vertex receives
 * the minimum vertex id of its connected component (isolated vertices receive
 * their own id).
 */
#ifndef KACOMP_H_
#define KACOMP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KC_OK 0
#define KC_ERR_INVALID (-1)     /* bad argument or malformed input */
#define KC_ERR_HIP (-2)         /* HIP runtime failure (no device, launch error) */
#define KC_ERR_OOM (-3)         /* device or host allocation failed */
#define KC_ERR_IO (-4)          /* file could not be opened or parsed */
#define KC_ERR_UNSUPPORTED (-5) /* option or generator not available */

typedef enum kc_variant {
  KC_VARIANT_EXP = 0,      /* exponential contraction (exp_contraction) */
  KC_VARIANT_LOCAL = 1,    /* 2-hop local contraction (local_contraction) */
  KC_VARIANT_SHORTCUT = 2, /* min-label propagation + pointer jumping (shortcuts) */
  KC_VARIANT_LABELS = 3    /* plain min-label propagation (labels) */
} kc_variant;

/* Host CSR handed over by the caller.  offsets has n+1 entries; targets has
 * offsets[n] entries, all < n.  Undirected edges appear in both rows. */
typedef struct kc_csr_host {
  uint64_t n;
  const uint64_t *offsets;
  const uint32_t *targets;
} kc_csr_host;

/* Host CSR owned by the library (results of kc_read_metis / kc_generate);
 * free with kc_csr_free. */
typedef struct kc_csr {
  uint64_t n;
  uint64_t m;        /* undirected edges = offsets[n] / 2 */
  uint64_t *offsets; /* n+1 */
  uint32_t *targets; /* offsets[n] */
} kc_csr;

/* Generator parameters of kc_generate / kc_graph_generate (generator names:
 * gnm_undirected, rgg_2d, rgg_3d, grid_2d, grid_3d, rmat, rhg, ba).  KaGen-style
 * (app/parse_parameters.h:83-99): n, m (exact counts), r, p, gamma, d,
 * periodic, plus rmat_a/b/c. */
typedef struct kc_gen_params {
  uint64_t n;
  uint64_t m;
  double r;
  double p;
  double gamma;
  double d;
  int32_t periodic;
  int32_t seed;
  double rmat_a, rmat_b, rmat_c;
  uint64_t nx, ny, nz; /* grid dimensions (grid_2d: nx=n, ny=m) */
} kc_gen_params;

/* Options of one kc_find_components call.  kc_options_init fills the
 * reference's defaults (app/parse_parameters.h:30-100). */
typedef struct kc_options {
  int32_t variant;      /* kc_variant */
  int32_t seed;         /* -seed (default 1); the benchmark harness passes seed+i+1000 */
  double beta;          /* -beta (default 0.1), rate of the exponential deviates */
  uint64_t seq_limit;   /* base case once remaining undirected edges <= seq_limit
                           (default 4096; 0 = contract until no edge is left) */
  int32_t use_bfs;      /* -bfs: time-stepped frontier schedule of the decomposition */
  int32_t deviate_fma;  /* 1: canonical draw summed with a fused multiply-add, as
                           libstdc++ compiled with -O3 -march=native on FMA hardware;
                           0: separate multiply and add */
  int32_t learned;      /* 1: deviates delayed by the scores of the MLP set with kc_graph_set_mlp */
  int32_t record_trace; /* keep per-level sizes/keys for kc_trace_* */
  int32_t time_kernels; /* bracket every launch with HIP events (per-kernel stats) */
  int32_t max_levels;   /* safety cap on contraction levels (default 4096) */
  int32_t sweep_group_small; /* tuning: degree bin thresholds (0 = built-in defaults) */
  int32_t sweep_group_mid;
  int32_t sweep_group_large;
  int32_t deviate_stream; /* KC_STREAM_REFERENCE (default) or KC_STREAM_HASHED */
  float learned_alpha;    /* learned: delay = round(learned_alpha * max(score, 0)) rounds (default 1) */
  int32_t input_order;    /* 1: level 0 keeps the input vertex order instead of the degree-descending
                             storage order (same results; for comparisons) */
  int32_t tail_sweeps;    /* -bfs: finish a level with pull sweeps once at most n >> k rows are pending:
                             0 = default k, k > 0 = that k, -1 = rounds to the end (same results) */
  int32_t pull_rounds;    /* -bfs: rounds 1..6 of a level each pushed by the rows at the round's
                             distance or pulled by the rows above it, whichever hold fewer edge slots:
                             0 = on levels of >= 2^16 vertices and average degree >= 8, 1 = on every
                             level, -1 = always pushed (same results) */
  int32_t dedup_huge;     /* dedup of the huge cluster-edge buckets: KC_DEDUP_AUTO (byte maps over the
                             cluster ids for the largest, within 8 GB; the others against bitmaps
                             while those take at most 2 GB, else in global hash tables), or one of
                             KC_DEDUP_TABLES, KC_DEDUP_BITMAPS, KC_DEDUP_BYTEMAPS (byte maps as far
                             as the cap allows, bitmaps for the rest), KC_DEDUP_MAPS_TABLES (byte
                             maps for the largest, tables for the rest) (same results) */
  int32_t time_class;     /* with time_kernels: 0 = every launch bracketed by HIP events, k > 0 = only
                             the launches of class k - 1 (the others are counted, not timed) */
  int32_t pre_contract;   /* local components contracted before the decomposition, the local phase
                             of ExponentialContraction::FindComponents (exponential_contraction.h:
                             60-319) with the vertices split over pre_blocks PEs by blocks of
                             consecutive ids: KC_PRE_NONE (default: every vertex its own PE),
                             KC_PRE_LOCAL, KC_PRE_SINGLE or KC_PRE_DOUBLE (below) */
  uint32_t pre_blocks;    /* PEs (blocks) of the pre-contraction, >= 1 */
  int32_t replicate;      /* -replicate (exponential contraction): every level-0 hub, a vertex of
                             degree >= hd_threshold * sqrt(vertices of degree > 0), is split into
                             floor(sqrt(degree)) proxies that take over its edges (each edge to the
                             proxy a seeded hash of the edge picks; an edge between two hubs joins
                             their proxies) and are joined to it; proxies draw their hub's deviate,
                             rank after every vertex in ties, and are dropped from the labels
                             (DistributeHighDegreeVertices, exponential_contraction.h:935-1253,
                             with every vertex its own PE; deviates :483-503) */
  double hd_threshold;    /* -hd-thres (default 16) */
  int32_t relabel_split;  /* relabel of the degree-ordered level 0: 1 = owner-split (its longest rows
                             cut by the owner part of their targets, each part looked up by its own
                             share of the workgroups), k > 1 = owner-split of the rows of more than k
                             slots, -1 = one pass over whole rows, 0 = default (owner-split)
                             (same results) */
} kc_options;

/* Pre-contraction (kc_options.pre_contract), each a level of the contraction
 * whose clusters are the components inside the blocks:
 *   LOCAL  - the default branch (:218-317, CAGBuilder::BuildLocalComponentGraph,
 *            cag_builder.h:160-230): vertices without a neighbour in another
 *            block go to their block component, the others stay themselves;
 *   SINGLE - -single (:76-133, BuildComponentAdjacencyGraph): every vertex goes
 *            to its block component;
 *   DOUBLE - -contraction (:134-217): SINGLE, then SINGLE again on the
 *            component graph with its vertices (in component order) split over
 *            the same number of blocks. */
#define KC_PRE_NONE 0
#define KC_PRE_LOCAL 1
#define KC_PRE_SINGLE 2
#define KC_PRE_DOUBLE 3

/* Dedup of the huge cluster-edge buckets (kc_options.dedup_huge). */
#define KC_DEDUP_AUTO 0
#define KC_DEDUP_TABLES 1
#define KC_DEDUP_BITMAPS 2
#define KC_DEDUP_BYTEMAPS 3
#define KC_DEDUP_MAPS_TABLES 4

/* Deviate streams.  REFERENCE: the draw of RunContractionLP bit for bit
 * (per-vertex minstd_rand0 seeds; its outputs are linear in the seed, so the
 * deviates of the surviving centres correlate and some levels contract
 * nothing).  HASHED: independent SplitMix64 draws of the same exponential law. */
#define KC_STREAM_REFERENCE 0
#define KC_STREAM_HASHED 1

#define KC_KERNEL_CLASSES 16
/* Kernel classes accounted in kc_stats.kernel_* (index = class). */
#define KC_K_DEVIATES 0  /* deviate sampling + key init */
#define KC_K_SWEEP 1     /* pull sweep (decomposition, base case) */
#define KC_K_BIN 2       /* degree binning */
#define KC_K_COUNT 3     /* cluster labels, radius, hub chunk folding */
#define KC_K_SCAN 4      /* prefix sums: selection and cluster segments */
#define KC_K_SCATTER 5   /* gather of cluster edges into their cluster's bucket */
#define KC_K_DEDUP 6     /* LDS-staged bucket dedup */
#define KC_K_CSR 7       /* next-level CSR compaction */
#define KC_K_UNDO 8      /* undo gather */
#define KC_K_CANON 9     /* canonical min-id labels */
#define KC_K_JUMP 10     /* pointer jumping */
#define KC_K_OTHER 11
#define KC_K_SORT 12     /* radix sort of (label, vertex) pairs */
#define KC_K_RELABEL 13  /* edge slots relabelled with their far end's cluster */
#define KC_K_ROUND 14    /* time-stepped push round (decomposition) */
#define KC_K_LEARNED 15  /* features and MLP scores of learned centre proposals */

typedef struct kc_stats {
  double time_ms;       /* device time of the whole call (HIP events), labels resident on device */
  double total_ms;      /* host wall time of the call including the label copy-out */
  double comm_time_s;   /* always 0: no inter-process communication on one GPU */
  uint64_t send_volume; /* always 0 */
  uint64_t recv_volume; /* always 0 */
  uint64_t n, m;        /* vertices, undirected edges of the input */
  uint32_t levels;      /* contraction levels (decompositions) executed */
  uint32_t components;  /* number of connected components (isolated vertices count) */
  uint64_t sweeps;      /* sum over levels of sweeps executed */
  uint64_t radius_sum;  /* sum over levels of the maximum cluster radius */
  uint64_t alg_bytes;   /* algorithmic bytes of all launched kernels */
  double kernel_ms[KC_KERNEL_CLASSES];        /* valid with time_kernels */
  uint64_t kernel_launches[KC_KERNEL_CLASSES];
  uint64_t kernel_bytes[KC_KERNEL_CLASSES];   /* algorithmic bytes per class */
  uint64_t replicas;    /* proxies added to level 0 by kc_options.replicate */
  uint32_t bin_small, bin_mid, bin_large; /* degree-bin thresholds the call used */
  uint32_t hubs;        /* hubs split by kc_options.replicate */
  double kernel_first_ms[KC_KERNEL_CLASSES];    /* with time_kernels: the class's first timed launch group
                                                   (the first level's, for the per-level classes) */
  uint64_t kernel_first_bytes[KC_KERNEL_CLASSES]; /* its algorithmic bytes */
} kc_stats;

typedef struct kc_graph kc_graph;

const char *kc_last_error(void);
void kc_options_init(kc_options *opt);

/* Device graph. */
int kc_graph_create(const kc_csr_host *csr, kc_graph **out);
int kc_graph_destroy(kc_graph *g);
int kc_graph_info(const kc_graph *g, uint64_t *n, uint64_t *m, uint64_t *non_isolated);
/* Generates gnm_undirected, rmat, grid_2d, grid_3d or ba directly in HBM with
 * the edge streams of kc_generate and builds the CSR on the device (same rows
 * as kc_generate; the order inside a row is not fixed).  Input preparation
 * only: IOUtility::LoadGraph's generator branch (lib/io/io_utils.h:46-51). */
int kc_graph_generate(const char *name, const kc_gen_params *p, kc_graph **out);
/* Reads a sorted binary graph file (the format of kc_read_binary) straight
 * into a device graph: the raw pairs are staged to HBM, converted there, and
 * the CSR is built on the device (same rows as kc_read_binary; the order
 * inside a row is not fixed).  GraphIO::ReadSortedBinaryFile
 * (lib/io/graph_io.h:910-1175). */
int kc_graph_read_binary(const char *path, kc_graph **out);
/* Reads a METIS graph file into a device graph: the header is read on the
 * host, the rest of the file is copied to HBM once and parsed there (lines,
 * comments, numbers), and the CSR is built on the device (same rows as
 * kc_read_metis; the order inside a row is not fixed).  A vertex line may hold
 * only digits and blanks: a sign, a letter or a 0 on one is an error here
 * (kc_read_metis stops the line there).  GraphIO::ReadMETISFile
 * (lib/io/graph_io.h:281-382) with ParseVertexFilestream (:1207-1248). */
int kc_graph_read_metis(const char *path, kc_graph **out);
/* Read a sorted edge file (the format of kc_read_edgefile) or its partitioned
 * form (parts files "<path>_<r>", as kc_read_partitioned_edgefile) straight into
 * a device graph. Each file goes to HBM in one copy. There its lines are
 * found, every edge line's "<type> u v" is parsed by one thread, and the CSR
 * is built, with the same rows as the host readers and no fixed order inside
 * a row. GraphIO::ReadSortedEdgeFile (lib/io/graph_io.h:385-669) and
 * ReadPartitionedSortedEdgeFile (:672-907) with ParseEdgeFilestream
 * (:1250-1300). */
int kc_graph_read_edgefile(const char *path, kc_graph **out);
int kc_graph_read_partitioned_edgefile(const char *path, int parts, kc_graph **out);
/* Copies the level-0 CSR of a graph to host memory (free with kc_csr_free). */
int kc_graph_download(const kc_graph *g, kc_csr *out);

/* Runs the selected variant.  labels_out (host, n entries) may be NULL: the
 * labels then stay on the device (kc_copy_labels fetches them). */
int kc_find_components(kc_graph *g, const kc_options *opt, uint32_t *labels_out, kc_stats *stats);

/* Degree-bin thresholds of the plain path: rows of at most `small` entries
 * take one lane, at most `mid` 16 lanes, at most `large` a wave; longer rows
 * are cut into chunks of `large` slots, one wave each. */
typedef struct kc_bin_tuning {
  int32_t small, mid, large;
  double ms;          /* mean device time of the chosen thresholds' calls */
  double default_ms;  /* the same with the built-in defaults (16, 256, 4096) */
  int32_t candidates; /* threshold triples timed */
} kc_bin_tuning;
/* Autotuner of the degree bins (no reference counterpart; the reference has no
 * bins): times every candidate triple with the options opt (their thresholds
 * ignored) on g and keeps the fastest in g (the defaults unless a triple is
 * more than 2% faster), where later kc_find_components
 * calls whose sweep_group_* fields are 0 use it.  Results do not depend on the
 * thresholds; only the time does.  chosen may be NULL. */
int kc_graph_autotune(kc_graph *g, const kc_options *opt, kc_bin_tuning *chosen);
/* Sets (or, with NULL, clears) the thresholds kept in g. */
int kc_graph_set_bins(kc_graph *g, const kc_bin_tuning *t);
int kc_copy_labels(kc_graph *g, uint32_t *labels_out);
/* Component-size histogram of the last kc_find_components call, computed on
 * the device (StaticGraph::OutputComponents, lib/data_structures/static_graph.h:504-591):
 * (size, number of components of that size) pairs in increasing size, every
 * vertex counted (a degree-0 vertex is a component of size 1).  Writes
 * min(capacity, *num_pairs) pairs (2 words each) to pairs; pairs may be NULL
 * to query *num_pairs. */
int kc_component_histogram(kc_graph *g, uint64_t *pairs, uint64_t capacity, uint64_t *num_pairs);

/* Per-level trace of the last kc_find_components call with record_trace=1.
 * level_sizes: n_l, m_l (undirected), sweeps_l, radius_l for level l. */
int kc_trace_levels(const kc_graph *g, uint32_t *levels);
int kc_trace_level_info(const kc_graph *g, uint32_t level, uint64_t info[4]);
/* keys_out: n_l u64 keys (deviate<<32 | centre index) at the fixed point;
 * gid_out: n_l original ids of the level's vertices. */
int kc_trace_level_keys(const kc_graph *g, uint32_t level, uint64_t *keys_out, uint32_t *gid_out);

/* Deviate of vertex gid at contraction iteration it (1-based) exactly as
 * RunContractionLP draws it (lib/components/exponential_contraction.h:467-471):
 * minstd_rand0 seeded with (unsigned)(gid + seed + it*(rng_offset+seed)),
 * exponential_distribution<double>(beta), truncated.  Computed on the device
 * for count consecutive gids starting at gid0 (test and trace helper);
 * stream KC_STREAM_HASHED gives the hashed draw instead. */
int kc_deviates(uint32_t gid0, uint32_t count, int32_t seed, uint32_t iteration, uint64_t rng_offset,
                double beta, int32_t fma, int32_t stream, uint32_t *out);

/* Allocation-free form of kc_find_components for the KC_VARIANT_LABELS and
 * KC_VARIANT_SHORTCUT variants.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  offsets (n+1 u64),
 * targets (offsets[n] u32, both directions of every edge) and labels (n u32,
 * written) are DEVICE pointers the caller owns; stream is a hipStream_t (NULL:
 * the null stream).  Every work buffer comes from a fixed device array of the
 * library (about 30 MB per GPU, held until the process ends), so graphs are
 * limited to KC_NOALLOC_MAX_VERTICES vertices and KC_NOALLOC_MAX_EDGE_SLOTS
 * edge slots (larger ones: KC_ERR_INVALID).  No device or pinned host memory
 * is allocated; the call reads a few words back to the host and returns
 * synchronised.  Labels equal kc_find_components' for the same variant (the
 * sweep count may differ: in-place sweeps race benignly on the keys).  Calls that overlap on one GPU share its
 * array and race: the caller serialises them.  stats (may be NULL) gets n, m,
 * sweeps, components and total_ms. */
#define KC_NOALLOC_MAX_VERTICES (1ull << 20)
#define KC_NOALLOC_MAX_EDGE_SLOTS (1ull << 24)
int kc_find_components_noalloc(uint64_t n, const uint64_t *offsets, const uint32_t *targets, const kc_options *opt,
                               uint32_t *labels, void *stream, kc_stats *stats);

/* ------------------------------------------------------------------------
 * Batched form: many independent graphs contracted in one persistent launch.
 * No reference counterpart as a batch; per graph it computes what
 * ExponentialContraction::FindComponents (lib/components/exponential_contraction.h:60-319)
 * computes on that graph alone, with labels local to the graph (the minimum
 * local vertex id of the component).
 *
 * Packed CSR segments: graph g owns vertices [vtx_off[g], vtx_off[g+1]) and
 * edge slots [edge_off[g], edge_off[g+1]); offsets are global (offsets[vtx_off[g]]
 * == edge_off[g]), targets are local ids (< n_g).  Every graph needs fewer than
 * 2^31 edge slots. */
typedef struct kc_batch_host {
  uint32_t graphs;
  const uint64_t *vtx_off;  /* graphs + 1 */
  const uint64_t *edge_off; /* graphs + 1 */
  const uint64_t *offsets;  /* vtx_off[graphs] + 1 */
  const uint32_t *targets;  /* edge_off[graphs] */
} kc_batch_host;

/* A packed batch owned by the library (kc_batch_generate); free with kc_batch_free. */
typedef struct kc_batch_csr {
  uint32_t graphs;
  uint64_t *vtx_off, *edge_off, *offsets;
  uint32_t *targets;
} kc_batch_csr;

/* Mixed batch: n_g = round(2^(log_n_min + (log_n_max - log_n_min) u)), average
 * degree uniform in [deg_min, deg_max], type uniform over the enabled ones
 * (types bit 0: gnm_undirected with m = n d / 2; bit 1: rgg_2d with
 * r = sqrt(d / (pi n)); bit 2: bond-percolated grid, 2D or 3D with equal
 * probability, side round(n^(1/dims)), keep probability min(1, d / (2 dims))). */
typedef struct kc_batch_gen_params {
  uint32_t graphs;
  double log_n_min, log_n_max;
  double deg_min, deg_max;
  int32_t types;
  int32_t seed;
} kc_batch_gen_params;

/* Launch shape of the persistent kernel: threads per workgroup (256, 512 or
 * 1024), workgroups per compute unit, and whether the launch sequence replays
 * a captured HIP graph.  Every row of a graph is processed by one lane, so the
 * batched form suits bounded-degree graphs; heavy-tailed graphs belong on the
 * plain path (kc_find_components), whose degree bins split long rows. */
typedef struct kc_batch_tuning {
  int32_t block;
  int32_t groups_per_cu;
  int32_t use_graph;
  uint32_t split_vertices; /* graphs with more vertices go through the device-wide path of
                              kc_find_components as one disjoint union, concurrently with the
                              persistent kernel (0 = every graph in the persistent kernel).
                              Their per-graph statistics are those of the union run. */
} kc_batch_tuning;

typedef struct kc_batch_stats {
  double time_ms;       /* device time of the launch sequence (HIP events) */
  double total_ms;      /* host wall time of the call */
  uint64_t n, m;        /* sums over the graphs (m undirected) */
  uint32_t graphs;
  uint32_t max_levels;  /* largest level count of a graph */
  uint64_t levels_sum;  /* sums over the graphs of levels, sweeps, max cluster radii */
  uint64_t sweeps_sum;
  uint64_t radius_sum;
  uint64_t components;  /* sum over the graphs */
  kc_batch_tuning tuning;
} kc_batch_stats;

typedef struct kc_batch kc_batch;

void kc_batch_gen_params_init(kc_batch_gen_params *p);
int kc_batch_generate(const kc_batch_gen_params *p, kc_batch_csr *out);
void kc_batch_free(kc_batch_csr *b);
/* Copies the batch to the device (input arrays are not retained). */
int kc_batch_create(const kc_batch_host *b, kc_batch **out);
int kc_batch_destroy(kc_batch *b);
/* Times every candidate launch shape on the batch and keeps the fastest. */
int kc_batch_autotune(kc_batch *b, const kc_options *opt, kc_batch_tuning *chosen);
int kc_batch_set_tuning(kc_batch *b, const kc_batch_tuning *t);
/* Exponential contraction of every graph (variant KC_VARIANT_EXP only; no
 * base case).  labels_out (host, vtx_off[graphs] entries) may be NULL. */
int kc_batch_run(kc_batch *b, const kc_options *opt, uint32_t *labels_out, kc_batch_stats *stats);
/* Per-graph levels, sweeps and summed max cluster radii of the last run (graphs entries each, any may be NULL). */
int kc_batch_graph_stats(const kc_batch *b, uint32_t *levels, uint32_t *sweeps, uint32_t *radius);
/* One-shot form: create, run with the default launch shape, destroy. */
int kc_batch_find_components(const kc_batch_host *b, const kc_options *opt, uint32_t *labels_out,
                             kc_batch_stats *stats);

/* ------------------------------------------------------------------------
 * Learned centre proposals.  No counterpart in the reference (its hook is the
 * commented-out TIEBREAK_DEGREE weighting in RunContractionLP,
 * lib/components/exponential_contraction.h:459-466).  A per-vertex MLP
 * 6 -> 64 -> 64 -> 1 (bf16 weights, f32 accumulation) scores every vertex of a
 * level from six features (log2(1+deg), log2(1+mean) and log2(1+max) of
 * sampled neighbour degrees, log2 of a 2-hop size estimate, a sampled
 * clustering rate, the round index); it is trained on traces of the plain
 * method to predict log2 of the size of the cluster the vertex ends in.  With
 * kc_options.learned the deviate of every vertex is delayed by
 * round(learned_alpha * max(score, 0)) rounds.  Labels stay exact: only the
 * choice of centres changes.
 *
 * Parameter order of kc_mlp_get_weights / set_weights / kc_mlp_gradients
 * (KC_MLP_PARAMS floats): W1[6][64] (input-major), b1[64], W2[64][64]
 * (hidden-1-major), b2[64], w3[64], b3. */
#define KC_MLP_FEATURES 6
#define KC_MLP_PARAMS 4673

typedef struct kc_mlp kc_mlp;

typedef struct kc_train_params {
  int32_t steps;  /* Adam steps (default 400) */
  int32_t batch;  /* samples per step (default 4096) */
  float lr;       /* default 3e-3 */
  float beta1, beta2, eps;
  int32_t seed;   /* mini-batch sampling stream */
} kc_train_params;

typedef struct kc_train_stats {
  double first_loss, last_loss; /* 0.5 * mean squared error of the first and last step */
  double time_ms;
  uint64_t samples;             /* dataset rows */
  int32_t steps;
} kc_train_stats;

int kc_mlp_create(int32_t seed, kc_mlp **out); /* He-uniform initial weights from seed */
int kc_mlp_destroy(kc_mlp *m);
int kc_mlp_get_weights(const kc_mlp *m, float *w);
int kc_mlp_set_weights(kc_mlp *m, const float *w);
/* Scores of n feature rows (n x 6, row-major). */
int kc_mlp_forward(kc_mlp *m, const float *features, uint64_t n, float *scores);
/* Gradient of 0.5 * mean((score - target)^2) over the n rows, without an update. */
int kc_mlp_gradients(kc_mlp *m, const float *features, const float *targets, uint64_t n, float *grads, float *loss);
/* Runs the plain method on g (opt with learned = 0) and appends the features
 * and targets of every contracted level (every stride-th vertex) to the
 * model's dataset. */
int kc_mlp_collect(kc_mlp *m, kc_graph *g, const kc_options *opt, uint64_t stride);
int kc_mlp_dataset_size(const kc_mlp *m, uint64_t *n);
/* Copies dataset rows [first, first + count): features (count x 6) and targets. */
int kc_mlp_get_dataset(const kc_mlp *m, uint64_t first, uint64_t count, float *features, float *targets);
int kc_mlp_clear_dataset(kc_mlp *m);
/* Appends n rows of caller features (n x 6) and targets to the dataset. */
int kc_mlp_add_dataset(kc_mlp *m, const float *features, const float *targets, uint64_t n);
void kc_train_params_init(kc_train_params *p);
/* Adam on the dataset with the project's MFMA forward and backward kernels. */
int kc_mlp_train(kc_mlp *m, const kc_train_params *p, kc_train_stats *stats);
/* Model used by kc_find_components on g when kc_options.learned is set (NULL detaches).
 * The features of g's level 0 (round index 0), hence its scores, depend on g
 * and the model only: the first learned call computes them and later calls
 * reuse them (each call still derives its delays with its own learned_alpha)
 * until the model is attached again or its weights change. */
int kc_graph_set_mlp(kc_graph *g, kc_mlp *m);
/* Features of the level-0 vertices of g (its vertices of degree > 0 in id
 * order), non_isolated x 6 floats. */
int kc_graph_features(kc_graph *g, float *features);

/* Host I/O and generators. */
int kc_read_metis(const char *path, kc_csr *out);
int kc_write_metis(const char *path, const kc_csr *g);
/* Sorted edge file: '%' comment lines, a header "<type> n m", then lines
 * "<type> u v" with 1-based ids, of which the first m are read. */
int kc_read_edgefile(const char *path, kc_csr *out);
/* The edge files "<path>_0" .. "<path>_<parts-1>" (same header in each,
 * every edge line of every part read), as one graph. */
int kc_read_partitioned_edgefile(const char *path, int parts, kc_csr *out);
/* Sorted binary file: n and m as u64, then m (u, v) u64 pairs, 1-based. */
int kc_read_binary(const char *path, kc_csr *out);
int kc_write_binary(const char *path, const kc_csr *g);
void kc_gen_params_init(kc_gen_params *p);
int kc_generate(const char *name, const kc_gen_params *p, kc_csr *out);
void kc_csr_free(kc_csr *g);

#ifdef __cplusplus
}
#endif

#endif /* KACOMP_H_ */
