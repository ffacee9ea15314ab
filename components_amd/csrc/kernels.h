// Launch wrappers of the device kernels (kernels.hip).  Every wrapper enqueues
// on the given stream and never synchronises.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace kc {

// A slice of a hub row (degree above the wave bin) handled by one wave.
struct HubChunk {
  uint32_t v;
  uint32_t k;  // index of the chunk within v's chunks (0: the first)
  uint64_t e0, e1;
};

struct BinCfg {
  uint32_t tmax;  // degree <= tmax: one lane per vertex
  uint32_t smax;  // degree <= smax: 16 lanes per vertex
  uint32_t wmax;  // degree <= wmax: one wave per vertex; above: wave chunks of wmax edges
};

constexpr uint32_t kMinWaveBin = 256;  // smallest one-wave bin bound

// Bin thresholds from the options' tuning fields (0 = built-in defaults 16,
// 256, 4096), clamped to tmax <= smax <= wmax and wmax >= kMinWaveBin.
inline BinCfg bin_config(int32_t small, int32_t mid, int32_t large) {
  BinCfg c;
  c.tmax = small > 0 ? static_cast<uint32_t>(small) : 16u;
  c.smax = mid > 0 ? static_cast<uint32_t>(mid) : 256u;
  c.wmax = large > 0 ? static_cast<uint32_t>(large) : 4096u;
  if (c.wmax < kMinWaveBin) c.wmax = kMinWaveBin;
  if (c.smax > c.wmax) c.smax = c.wmax;
  if (c.tmax > c.smax) c.tmax = c.smax;
  return c;
}
// Hub chunks a level of m2 edge slots can need: rows above wmax are cut into
// chunks of wmax slots, so at most m2 / (kMinWaveBin / 2) + 2 of them.
inline uint64_t hub_chunk_cap(uint64_t m2) { return m2 / (kMinWaveBin / 2) + 2; }

// Owner parts of a degree-ordered level 0: vertex t belongs to part
// owner_of(t), the 128-byte line of its 4-byte label taken modulo kParts, so
// each part's labels are one line in every kParts (see relabel_split).
constexpr uint32_t kParts = 8;
__host__ __device__ inline uint32_t owner_of(uint32_t t) { return (t >> 5) & (kParts - 1); }

// Degree bins of one level; lists and chunks live on the device.
struct Bins {
  uint32_t *listT = nullptr, *listS = nullptr, *listW = nullptr;
  HubChunk *hubs = nullptr;
  uint32_t nT = 0, nS = 0, nW = 0, nH = 0;
  // first block of each section in a binned launch (thread section starts at 0)
  uint32_t bS = 0, bW = 0, bH = 0, blocks = 0;
};

enum SweepMode : int {
  kSweepInPlace = 0,  // key[v] <- min(key[v], key[w] + hop), updates visible at once
  kSweepJacobi = 1,   // kout[v] <- min over closed neighbourhood of kin (kout pre-copied)
  kSweepCountOnly = 2, // no update: count neighbours with a different label
  kSweepBlock = 3      // in place, hop 0, over the neighbours in the row's block of consecutive indices
};

struct DeviateArgs {
  int32_t seed;
  uint32_t iteration;
  uint64_t rng_offset;
  int32_t fused;
  int32_t stream;  // 0: reference draw, 1: hashed draw
  const double *table;
  uint32_t k_max;
};

// Learned centre proposals: per-vertex delay added to the deviate (the vertex
// starts its wave that many rounds later).
struct LearnedArgs {
  const uint8_t *delay;  // null: plain deviates
};

// Thresholds W_k (k = 1..k_max) of the truncated exponential deviate with
// rate beta: the deviate of a draw u is the number of W_k >= 1 - u.
std::vector<double> build_deviate_table(double beta);

void fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s);
void fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t s);

void init_keys_exp(uint64_t n, const uint32_t *gid, const uint32_t *idx, DeviateArgs d, LearnedArgs la, uint64_t *key,
                   hipStream_t s);
void init_keys_index(uint64_t n, uint64_t *key, hipStream_t s);
void init_keys_uniform(uint64_t n, const uint32_t *gid, const uint32_t *idx, int32_t seed, uint32_t iteration,
                       uint64_t *key,
                       hipStream_t s);
void deviates_only(uint32_t gid0, uint32_t count, DeviateArgs d, uint32_t *out, hipStream_t s);

// Bins vertices of a CSR by degree (stable, no global atomics); counters[0..3]
// receive nT, nS, nW, nH.  deg_inter of degree-0 vertices is cleared.
// scratch: bin_scratch_words(n) words.
// With key != null only vertices at key distance >= tmin are binned.
// out[0] = edge slots of the rows at distance t, out[1] = of the rows above t.
void round_volumes(uint64_t n, const uint64_t *off, const uint64_t *key, uint32_t t, uint64_t *out, hipStream_t s);
void bin_vertices(uint64_t n, const uint64_t *off, BinCfg cfg, uint32_t *listT, uint32_t *listS,
                  uint32_t *listW, HubChunk *hubs, uint32_t *counters, uint32_t *deg_inter, uint64_t *scratch,
                  hipStream_t s, const uint64_t *key = nullptr, uint32_t tmin = 0);
void finish_bins(Bins &b);  // computes block offsets from the counts
uint64_t bin_scratch_words(uint64_t n);

// One binned sweep.  flag: set to 1 when a key decreased; prev_flag (may be
// null): when it reads 0 the launch returns at once (converged earlier).
void sweep(int mode, bool count, const uint64_t *off, const uint32_t *col, const uint64_t *kin,
           uint64_t *kout, uint64_t hop, const Bins &b, uint32_t *deg_inter, uint32_t *hub_inter,
           int *flag, const int *prev_flag, hipStream_t s, uint64_t block = 0, bool hubs_first = false);
// Packed keys (distance << 24 | label) for the pull sweeps of levels below
// 2^24 vertices: pack_keys writes them and sets *bad (zeroed first) when a
// distance exceeds kPackedMaxDist - 1 or a label 24 bits; sweep_packed is
// sweep(kSweepInPlace, no counts) over them; unpack_keys restores the keys.
constexpr uint32_t kPackedLabelBits = 24;
constexpr uint32_t kPackedHop = 1u << kPackedLabelBits;
constexpr uint32_t kPackedMaxDist = 255;
void pack_keys(uint64_t n, const uint64_t *key, uint32_t *pk, uint32_t *bad, hipStream_t s);
void unpack_keys(uint64_t n, const uint32_t *pk, uint64_t *key, hipStream_t s);
void sweep_packed(const uint64_t *off, const uint32_t *col, uint32_t *pk, const Bins &b, int *flag,
                  const int *prev_flag, hipStream_t s, bool hubs_first);
// Round t of the time-stepped schedule: vertices at key distance t push
// key + hop; *pending = 1 while a vertex is still above distance t;
// prev_pending (may be null) reading 0 makes the launch return at once.
// hub_pull_deg > 0: the hubs of at least that degree first pull round t's
// candidates, so that the pushes find them settled.
// hubs_first: the bins' sections dispatched from the hub chunks down to the
// one-lane rows (the same result).
void push_round(const uint64_t *off, const uint32_t *col, uint64_t *key, uint64_t hop, const Bins &b, uint32_t t,
                int *pending, const int *prev_pending, uint64_t hub_pull_deg, hipStream_t s, bool hubs_first = false);
// deg_inter of hubs from their chunk counts; hub_pos[c] = offset of chunk c in its hub's run
// (prefix: b.nH + 1 words; scratch: scan scratch).
void hub_fold(const Bins &b, const uint32_t *hub_inter, uint32_t *deg_inter, uint32_t *hub_pos, uint64_t *prefix,
              uint64_t *scratch, hipStream_t s);

// Pointer jumping on keys: key[v] <- min(key[v], key[label(v)]).
void jump(uint64_t n, uint64_t *key, int *flag, hipStream_t s);

// Interface vertices of blocks of bsize consecutive indices (a neighbour in
// another block) get their own index back as key: key[v] = v.
void block_interface(const uint64_t *off, const uint32_t *col, uint64_t n, uint64_t bsize, uint64_t *key,
                     hipStream_t s);

// lab[v] = low half of key[v]; max cluster radius (atomicMax into *radius);
// used (may be null, pre-zeroed): used[lab[v]] = 1, the used-label flags of
// cluster_ranks.
void labels_radius(uint64_t n, const uint64_t *key, const uint32_t *pos, uint32_t *lab, uint32_t *radius,
                   bool with_radius, hipStream_t s, uint32_t *used = nullptr);
// Dense cluster ids in label order from the used-label flags (labels_radius):
// rank[a] = number of used labels below a, written for the used ones, and
// *count (device) = number of clusters C; then cid[v] = rank[lab[v]] in [0, C).
void cluster_ranks(uint64_t n, const uint32_t *used, uint32_t *rank, uint64_t *count, uint64_t *scratch,
                   hipStream_t s);
void cluster_ids(uint64_t n, const uint32_t *lab, const uint32_t *rank, uint32_t *cid, hipStream_t s);
// Positions whose cluster ids the relabel of a degree-ordered level 0 keeps in
// LDS as 16-bit values.
constexpr uint64_t kHotClusters = 57344;
// cluster_ranks with the clusters of the first k positions (k <= 65536) numbered
// first, so that their ids are below 65536 (used: labels_radius's flags, changed
// in place; scratch: 8 words + scan scratch).  Same count and the same set of
// ids as cluster_ranks: only the order of the ids differs, and nothing after
// the relabel depends on it.
void cluster_ranks_hot(uint64_t n, uint64_t k, const uint32_t *lab, uint32_t *used, uint32_t *rank, uint64_t *count,
                       uint64_t *scratch, hipStream_t s);
// (label, vertex) of every vertex with a cluster edge; *count (device) = N.
void select_inter(uint64_t n, const uint64_t *key, const uint32_t *deg_inter, uint32_t *sk, uint32_t *sv,
                  uint64_t *count, uint64_t *scratch, hipStream_t s);
// Over the label-sorted list: pos_of[v] = write position of v's cluster edges,
// newidx[a] = next-level index of label a (caller pre-fills kNone),
// seg_of_cid[rank[a]] = the same index by dense cluster id, gid_next by
// next-level index, seg_off[0..n_next] bucket offsets; sizes[0] = n_next,
// sizes[1] = P (cluster edges).  cache: N words of scratch.
void segments(uint64_t N, const uint32_t *sk, const uint32_t *sv, const uint32_t *deg_inter, const uint32_t *gid,
              const uint32_t *rank, uint32_t *pos_of, uint32_t *newidx, uint32_t *seg_of_cid, uint32_t *gid_next,
              uint64_t *seg_off, uint64_t *sizes, uint32_t *cache, uint64_t *scratch, hipStream_t s);

// The labels lab[col[e]] of row v that differ from lab[v] (after a best-effort
// LDS filter of repeats), packed at relab[off[v] ...] (hub chunks: at the
// chunk start); deg_inter[v] (hub chunks: hub_inter) = their number.
// One persistent launch; lab[0 .. min(n_hot, 28672)) is served from LDS (for
// degree-ordered levels, whose first vertices take most endpoints); with
// hot16, lab[0 .. min(n_hot, kHotClusters)) as 16-bit values, which requires
// those positions' cluster ids to fit 16 bits (cluster_ranks_hot).
void relabel(const uint64_t *off, const uint32_t *col, const uint32_t *lab, const Bins &b, uint32_t *relab,
             uint32_t *deg_inter, uint32_t *hub_inter, uint64_t n_hot, hipStream_t s, bool hot16 = false);
// Owner-split relabel of a degree-ordered level 0 whose rows are grouped by
// owner (sort_rows): the one-lane and 16-lane bins of b as relabel does, and
// the pieces of the longer rows (write_pieces): the shared ones by every
// block with the first n_hot labels in LDS as 16-bit values (their ids must
// fit, cluster_ranks_hot), the others by the blocks of their
// owner part with that part's labels in LDS (part_at: kParts + 2 bounds into
// part_list); piece counts go to piece_inter.  Same cluster edges per row as relabel, split over the row's
// pieces (hub_fold with the pieces as chunks gives the rows' counts).
void relabel_split(const uint64_t *off, const uint32_t *col, const uint32_t *lab, uint64_t n, const Bins &b,
                   const HubChunk *pieces, const uint32_t *part_list, const uint32_t *part_at, uint32_t *relab,
                   uint32_t *deg_inter, uint32_t *piece_inter, uint64_t n_hot, hipStream_t s);
// out = b with every list (and the hub chunks, by their row) in increasing
// idx[position] order (stable; out's lists and chunks are the caller's
// buffers, its counts and block offsets b's).  k0, v0, k1, v1: max(nT, nS,
// nW, nH) words each; scratch: radix_sort_scratch_words of that many.
// Returns the median idx of the hub chunks' rows (kNone without hubs).
uint32_t order_bins_by_index(const Bins &b, const uint32_t *idx, uint64_t n, Bins &out, uint32_t *k0, uint32_t *v0,
                             uint32_t *k1, uint32_t *v1, uint64_t *scratch, hipStream_t s);
// Pieces of the rows of more than smax slots of an owner-grouped CSR: cnt[v]
// = the row's pieces (rows of at most `split` slots: the whole row, cut every
// wmax slots, shared by all blocks; longer rows: one per owner group they
// have, cut every wmax slots), start (n + 1 words) = their exclusive prefix,
// start[n] = the total.
void count_pieces(uint64_t n, const uint64_t *off, const uint32_t *col, uint32_t smax, uint32_t wmax, uint64_t split,
                  uint32_t *cnt, uint64_t *start, uint64_t *scratch, hipStream_t s);
// pieces[start[v] ...] = row v's pieces (HubChunk, k = index in the row),
// part[...] = their owner parts (kParts: shared).
void write_pieces(uint64_t n, const uint64_t *off, const uint32_t *col, uint32_t smax, uint32_t wmax, uint64_t split,
                  const uint64_t *start, HubChunk *pieces, uint8_t *part, hipStream_t s);
// Copies the deg_inter[v] packed relab slots of every row (hub chunks:
// hub_inter) to bucket[pos_of[v] + ...].
void scatter(const uint64_t *off, const uint32_t *relab, const Bins &b, const uint32_t *deg_inter,
             const uint32_t *hub_inter, const uint32_t *hub_pos, const uint32_t *pos_of, uint32_t *bucket,
             hipStream_t s);

// Dedup of every bucket a2 (elements [seg_off[a2], seg_off[a2+1]) of bucket):
// its uniq[a2] distinct values go to out[seg_off[a2] ...] in any order.  Small
// buckets in wave LDS tables (dedup_small, which skips the others); medium
// and huge ones are listed first by classify_buckets.
constexpr uint32_t kDedupWaveCap = 512;
constexpr uint32_t kDedupBlockCap = 8192;
constexpr uint64_t kHugeChunk = 16384;  // elements (or table slots) per huge-bucket chunk (tables)
constexpr uint32_t kBitsChunk = 16384;   // elements per huge-bucket chunk (bitmaps)
// A fixed-size slice of a huge bucket: elements or table slots [lo, hi).
struct HugeChunk {
  uint64_t base;      // bucket offset of the huge bucket
  uint64_t tab;       // table offset
  uint64_t lo, hi;
  uint32_t tab_mask;  // table size - 1
  uint32_t huge;      // index in the huge list
  uint32_t a2;        // bucket (next-level vertex)
  uint32_t pad;
};
// Medium (wave cap < size <= block cap) and huge (above) buckets among the
// first *n_dev (or n when n_dev is null; n bounds *n_dev), listed in
// increasing order: counts[0] = medium, counts[1] = huge.  scratch: scan scratch.
void classify_buckets(uint64_t n, const uint64_t *n_dev, const uint64_t *seg_off, uint32_t *medium, uint32_t *huge,
                      uint64_t *counts, uint64_t *scratch, hipStream_t s);
void dedup_small(uint64_t n_next, const uint64_t *seg_off, const uint32_t *bucket, uint32_t *out, uint32_t *uniq,
                 hipStream_t s);
void dedup_medium(uint32_t n_medium, const uint32_t *medium, const uint64_t *seg_off, const uint32_t *bucket,
                  uint32_t *out, uint32_t *uniq, hipStream_t s);
// out[2h] = bucket offset, out[2h+1] = size of huge bucket h.
void huge_sizes(uint32_t n_huge, const uint32_t *huge, const uint64_t *seg_off, uint64_t *out, hipStream_t s);
// Huge buckets over the label space [0, space): those of at least map_min
// elements (at most n_maps of them) against byte maps of huge_map_stride(space)
// bytes, all zero on entry and left all zero; the others against bitmaps of
// `words` u32 per bucket (with_bits false: the others are left alone).
// zero_uniq: uniq of every huge bucket is zeroed first (else the caller has
// zeroed it for the map buckets).  info: 3 n_huge + 6 words of scratch for
// the chunk plans (map_min = ~0: bitmaps only).
uint64_t huge_map_stride(uint64_t space);
void dedup_huge_mixed(uint32_t n_huge, const uint32_t *huge, const uint64_t *seg_off, uint64_t max_chunks,
                      uint64_t map_min, uint64_t n_maps, bool with_bits, bool zero_uniq, uint64_t *info,
                      uint32_t *bitmaps, uint64_t words, uint8_t *maps, uint64_t space, const uint32_t *bucket,
                      uint32_t *out, uint32_t *uniq, uint64_t *scratch, hipStream_t s);
// Huge buckets in global open-addressing tables (pre-filled with kNone by the
// caller, hcount zeroed), for label spaces whose bitmaps would not fit.
void dedup_huge(const HugeChunk *elem_chunks, uint32_t n_elem, const HugeChunk *slot_chunks, uint32_t n_slot,
                uint32_t n_huge, const uint32_t *huge, uint32_t *tables, const uint32_t *bucket, uint32_t *out,
                uint32_t *uniq, uint32_t *hcount, hipStream_t s);

// off_next = exclusive scan of uniq (n_next+1 entries); sizes[2] = total.
void csr_offsets(uint64_t n_next, const uint32_t *uniq, uint64_t *off_next, uint64_t *sizes,
                 uint64_t *scratch, hipStream_t s);
// col_next[off_next[a2] + i] = map[vals[seg_off[a2] + i]] for i < uniq[a2]
// (map null: the value itself): the unique runs of the dedup, concatenated.
// medium/huge: the listed buckets (classify_buckets); cstart: n_medium +
// n_huge + 1 words for their chunk plan; scratch: scan scratch.
void gather_uniques(uint64_t n_next, const uint64_t *off_next, const uint64_t *seg_off, const uint32_t *vals,
                    const uint32_t *map, const uint32_t *medium, uint64_t n_medium, const uint32_t *huge,
                    uint64_t n_huge, uint64_t *cstart, uint64_t *scratch, uint32_t *col_next, hipStream_t s);
uint64_t scan_scratch_words(uint64_t n);
void scan_exclusive_u32_to_u64(uint64_t n, const uint32_t *in, uint64_t *out, uint64_t *total, uint64_t *scratch,
                               hipStream_t s);

// Stable LSD radix sort of (key, value) pairs on the low `bits` key bits
// (sort.hip).  Returns 0 when keys/vals hold the result, 1 for tkeys/tvals.
uint64_t radix_sort_scratch_words(uint64_t n);
int radix_sort_pairs(uint64_t n, int bits, uint32_t *keys, uint32_t *vals, uint32_t *tkeys, uint32_t *tvals,
                     uint64_t *scratch, hipStream_t s);

void undo(uint64_t n, const uint32_t *lab, const uint32_t *newidx, const uint32_t *gid,
          const uint32_t *rep_up, uint32_t *rep_out, hipStream_t s);
// rep of a base-case level: rep[v] = gid[low(key[v])].
void rep_from_keys(uint64_t n, const uint64_t *key, const uint32_t *gid, uint32_t *rep, hipStream_t s);
// minid[rep[v]] = min over v of gid[v] (identity gid when null).
void canon_min(uint64_t n, const uint32_t *rep, const uint32_t *gid, uint32_t *minid, hipStream_t s);
// roots (device) = number of vertices that are their own label; partial needs
// ceil(n / 256) words.
// labels[gid[v]] = minid[rep[v]]; roots counts the vertices that are their own label.
void canon_label(uint64_t n, const uint32_t *rep, const uint32_t *gid, const uint32_t *minid, uint32_t *labels,
                 uint32_t *roots, uint32_t *partial, hipStream_t s);
// Level 0 without degree-0 vertices (the reference sets them inactive at
// load): newid (n), gid0 (kept), off_c (kept + 1), col_c (m2, mapped), *n_out.
void drop_isolated(uint64_t n, const uint64_t *off, const uint32_t *col, uint64_t m2, uint32_t *newid, uint32_t *gid0,
                   uint64_t *off_c, uint32_t *col_c, uint64_t *n_out, uint64_t *scratch, hipStream_t s);
void identity_u32(uint64_t n, uint32_t *p, hipStream_t s);
// Level 0 in degree-descending storage order (ties in index order): idx (position
// -> level index), pos (its inverse), gidp (position -> input id; gid null =
// identity), offp/colp (the CSR in storage order, neighbours as positions).
// k0/v0/k1/v1 are n-word sort buffers.
void level0_order(uint64_t n, const uint64_t *off, const uint32_t *col, const uint32_t *gid, uint32_t *idx,
                  uint32_t *pos, uint32_t *gidp, uint64_t *offp, uint32_t *colp, uint32_t *k0, uint32_t *v0,
                  uint32_t *k1, uint32_t *v1, uint64_t *sort_scratch, uint64_t *scan_scratch, hipStream_t s);
// Every row of a CSR (n rows, m2 slots) sorted in place by owner part, then
// ascending inside a part: the (owner key, row) pairs sorted by key, then
// stably by row.  k0, v0, k1, v1:
// m2 words each; scratch: radix_sort_scratch_words(m2) words.
void sort_rows(uint64_t n, const uint64_t *off, uint32_t *col, uint64_t m2, uint32_t *k0, uint32_t *v0, uint32_t *k1,
               uint32_t *v1, uint64_t *scratch, hipStream_t s);
// (size, number of components of that size) pairs of a canonical labelling in
// increasing size; count and hist are n and n + 1 words of scratch, pairs
// holds 2 * (distinct sizes) words, *num_pairs the number of pairs.
void component_histogram(uint64_t n, const uint32_t *labels, uint32_t *count, uint32_t *hist, uint64_t *pairs,
                         uint64_t *num_pairs, uint64_t *scratch, hipStream_t s);
void labels_from_keys(uint64_t n, const uint64_t *key, uint32_t *labels, uint32_t *roots, uint32_t *partial,
                      hipStream_t s);

// Hub replication (kc_options.replicate).  Hub h (vertex hubs[h], H of them in
// increasing order) has k[h] proxy slots at slot_base[h]; cnt[s] counts the
// hub edges a slot receives, rep_of_slot[s] is the proxy's level index (n0 +
// its rank among the slots with an edge) or kNone, nexist[h] / first_rep[h]
// the number and first index of h's proxies, slot_of_rep / root_of_rep the
// slot and hub of each proxy; hubidx (n0) = the hub index of a vertex or kNone.
struct ReplicaArgs {
  const uint64_t *off;
  const uint32_t *col;
  const uint32_t *gid;  // null: identity
  int32_t seed;
  const uint32_t *hubs, *k;
  const uint64_t *slot_base;
  uint32_t *hubidx, *cnt;
  const uint32_t *rep_of_slot, *nexist, *first_rep, *slot_of_rep, *root_of_rep;
};
// A slice of a hub's row: edges [e0, e1) of hub h.
struct HubEdgeChunk {
  uint32_t h, pad;
  uint64_t e0, e1;
};
// bounds[2i], bounds[2i+1] = off[rows[i]], off[rows[i] + 1].
void row_bounds(uint32_t count, const uint32_t *rows, const uint64_t *off, uint64_t *bounds, hipStream_t s);
// list[0 .. *count) = the vertices of degree >= limit (in no fixed order).
void find_hubs(uint64_t n, const uint64_t *off, uint64_t limit, uint32_t *list, uint32_t *count, hipStream_t s);
// hubidx from hubs, then cnt (pre-zeroed) from the hub edges.
void replica_count(const ReplicaArgs &r, uint64_t n0, uint32_t H, const HubEdgeChunk *chunks, uint32_t n_chunks,
                   hipStream_t s);
// Row lengths of the replicated level (n0 + R rows).
void replica_lengths(const ReplicaArgs &r, uint64_t n0, uint32_t R, uint32_t *len, hipStream_t s);
// Columns (row order of the proxies' hub edges not fixed) and ids of the
// replicated level; cursor (one per slot) pre-zeroed.
void replica_fill(const ReplicaArgs &r, uint64_t n0, uint32_t H, uint32_t R, const HubEdgeChunk *chunks,
                  uint32_t n_chunks, const uint64_t *off_r, uint32_t *cursor, uint32_t *col_r, uint32_t *gid_r,
                  hipStream_t s);

// Device generators (gen_kernels.hip): the edge streams of the host generators.
// 1-based (u, v) u64 pairs of a binary graph file -> 0-based u32 pairs
// (*bad = 1 when a pair lies outside [1, n]).
void binary_edges(const uint64_t *raw, uint64_t count, uint64_t n, uint32_t *src, uint32_t *dst, int *bad,
                  hipStream_t s);
void gen_gnm(uint64_t count, uint64_t n, int32_t seed, uint32_t *src, uint32_t *dst, hipStream_t s);
void gen_rmat(uint64_t count, uint32_t scale, double a, double b, double c, int32_t seed, uint32_t *src, uint32_t *dst,
              hipStream_t s);
void gen_ba(uint64_t count, uint64_t d, int32_t seed, uint32_t *src, uint32_t *dst, hipStream_t s);
// Lattices write n*dims slots: kept bonds, or self-loops for dropped ones.
void gen_grid(uint64_t nx, uint64_t ny, uint64_t nz, int dims, int periodic, double p, int32_t seed, uint32_t *src,
              uint32_t *dst, hipStream_t s);
// CSR construction from an edge list: degrees (self-loops skipped), scatter of
// both directions into row buckets.
void deg_count(uint64_t count, const uint32_t *src, const uint32_t *dst, uint32_t *deg, hipStream_t s);
void edge_scatter(uint64_t count, const uint32_t *src, const uint32_t *dst, const uint64_t *off, uint32_t *fill,
                  uint32_t *bucket, hipStream_t s);
void nonzero_rows(uint64_t n, const uint64_t *off, unsigned long long *counter, hipStream_t s);


// METIS text on the device.  byte_positions: the positions i < bytes of one
// byte class in increasing order into pos (null: count only); returns the
// count.  metis_vertex_lines: vline[j] = index of line j among the vertex
// lines (not starting with '%'), kNone for comments and lines past the n-th.
// metis_tokens: per number its vertex (kNone off the vertex lines) and value,
// *bad |= 1 for 0, > n or > 10 digits on a vertex line; metis_other: *bad |= 2
// for a byte other than digits and blanks on a vertex line; metis_edges: the
// (vertex, value - 1) pairs of the kept numbers in order, returns their count.
constexpr int kBytesNewline = 0, kBytesToken = 1, kBytesOther = 2;
uint64_t byte_positions(const uint8_t *b, uint64_t bytes, int what, uint64_t *pos, uint64_t *total, uint64_t *scratch,
                        hipStream_t s);
void metis_vertex_lines(const uint8_t *b, uint64_t bytes, const uint64_t *nlpos, uint64_t lines, uint64_t n,
                        uint32_t *vline, uint64_t *total, uint64_t *scratch, hipStream_t s);
void metis_tokens(const uint8_t *b, uint64_t bytes, const uint64_t *tokpos, uint64_t count, const uint64_t *nlpos,
                  uint64_t nl, const uint32_t *vline, uint64_t n, uint32_t *tv, uint32_t *tval, int *bad,
                  hipStream_t s);
void metis_other(const uint64_t *otherpos, uint64_t count, const uint64_t *nlpos, uint64_t nl, const uint32_t *vline,
                 int *bad, hipStream_t s);
uint64_t metis_edges(uint64_t count, const uint32_t *tv, const uint32_t *tval, uint32_t *src, uint32_t *dst,
                     uint64_t *total, uint64_t *scratch, hipStream_t s);

// Edge-list text on the device: edge_text_lines writes the indices of the
// lines that are neither empty nor comments into eline (null: count only) and
// returns their count; edge_text parses "<type> u v" from the first count of
// them into 0-based (src, dst), *bad |= 1 for a malformed line, |= 2 for an
// id outside [1, n].
uint64_t edge_text_lines(const uint8_t *b, uint64_t bytes, const uint64_t *nlpos, uint64_t nl, uint64_t lines,
                         uint64_t *eline, uint64_t *total, uint64_t *scratch, hipStream_t s);
void edge_text(const uint8_t *b, uint64_t bytes, const uint64_t *nlpos, uint64_t nl, const uint64_t *eline,
               uint64_t count, uint64_t n, uint32_t *src, uint32_t *dst, int *bad, hipStream_t s);

}  // namespace kc
